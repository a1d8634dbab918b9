"""Dataset output stage: the FDTD stage's complex float signal and the
AWGN stage (cli/fdtd_stage.c:20-24, engine/noise/awgn.c:16-44,
engine/core/rng.c:3-23) and the .iq writer (cli/main.c:70-85).

The oracle's restatement is pinned bit for bit against the reference
compiled from its own sources (its fdtd_stage and awgn transforms applied
to a signal_t, in the CLI's chain order); the device stage is checked
against the oracle on the same runs: bitwise without noise, within a few
float ulps of the noise with it (device logf, cosf, sinf); the writer
against the reference's own output file."""
import ctypes as C
import os

import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import NativeRun, c1_cpml_params, modgauss_params

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "example_000000.iq")


class Transform(C.Structure):
    _fields_ = [("apply", C.c_void_p), ("destroy", C.c_void_p), ("ctx", C.c_void_p)]


class Signal(C.Structure):
    _fields_ = [("samples", C.POINTER(C.c_float)), ("n_samples", C.c_size_t), ("sample_rate_hz", C.c_double),
                ("center_freq_hz", C.c_double), ("metadata", C.c_void_p), ("meta_copy", C.c_void_p),
                ("meta_free", C.c_void_p)]


APPLY = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)
DESTROY = C.CFUNCTYPE(None, C.c_void_p)


def _ref_chain(ref, stages, n, init=None):
    """Runs transforms of the reference on a fresh signal_t of n samples
    (optionally preset to init, [n, 2] float32); returns [n, 2] float32."""
    ref.signal_create.restype = C.POINTER(Signal)
    ref.signal_create.argtypes = [C.c_size_t, C.c_double, C.c_double]
    ref.signal_destroy.argtypes = [C.POINTER(Signal)]
    sig = ref.signal_create(n, 1e6, 915e6)
    assert sig
    if init is not None:
        flat = np.ascontiguousarray(init, dtype=np.float32).ravel()
        C.memmove(sig.contents.samples, flat.ctypes.data, flat.nbytes)
    for t in stages:
        assert APPLY(t.apply)(t.ctx, C.cast(sig, C.c_void_p)) == 0
        DESTROY(t.destroy)(t.ctx)
    m = sig.contents.n_samples
    out = np.ctypeslib.as_array(sig.contents.samples, shape=(2 * m,)).copy().reshape(m, 2)
    ref.signal_destroy(sig)
    return out


def _ref_awgn(ref, snr, seed, stream):
    ref.awgn_create.restype = Transform
    ref.awgn_create.argtypes = [C.c_float, C.c_uint64, C.c_uint64]
    t = ref.awgn_create(snr, seed, stream)
    assert t.apply
    return t


def _preset_cfg():
    # cli/presets.h:14-29 defaults (settings.c:6-33)
    return capi.SimConfig(50, 50, 50, 1e-3, 0.99, 200, 8, 25, 25, 25, capi.FIELD_EZ, 40.0, 8.0,
                          31, 25, 25, capi.FIELD_EZ)


def _oracle_awgn(oracle, iq, snr, seed, stream):
    a = np.ascontiguousarray(iq, dtype=np.float32).copy()
    oracle.oracle_awgn_iq(a.ctypes.data_as(C.POINTER(C.c_float)), a.shape[0], snr, seed, stream)
    return a


@pytest.mark.parametrize("stream", [0, 1, 5, 99])
def test_oracle_chain_matches_reference(ref, oracle, stream):
    """The CLI's default chain (FDTD stage, then AWGN at 15 dB with seed
    42 and the example index as stream): the reference's transforms against
    the oracle's echo and AWGN, bit for bit."""
    ref.fdtd_stage_create.restype = Transform
    ref.fdtd_stage_create.argtypes = [C.POINTER(capi.SimConfig)]
    cfg = _preset_cfg()
    want = _ref_chain(ref, [ref.fdtd_stage_create(C.byref(cfg)), _ref_awgn(ref, 15.0, 42, stream)], 1024)
    buf = (C.c_double * 200)()
    fs = C.c_double()
    assert oracle.fdtd_simulate_echo(C.byref(cfg), buf, C.byref(fs)) == 0
    iq = np.zeros((200, 2), dtype=np.float32)
    iq[:, 0] = np.array(buf[:]).astype(np.float32) + np.float32(0.0)
    got = _oracle_awgn(oracle, iq, 15.0, 42, stream)
    assert want.shape == (200, 2)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.any(got[:, 1] != 0)


@pytest.mark.parametrize("case", ["random", "zero", "low_snr"])
def test_oracle_awgn_matches_reference_on_signals(ref, oracle, case):
    """AWGN alone on complex signals (random, all zero: no noise power,
    and at -10 dB): bit for bit, signed zeros included."""
    rng = np.random.default_rng(7)
    n = 777
    x = rng.normal(0, 1, (n, 2)).astype(np.float32) if case != "zero" else np.zeros((n, 2), np.float32)
    snr = -10.0 if case == "low_snr" else 15.0
    want = _ref_chain(ref, [_ref_awgn(ref, snr, 42, 3)], n, init=x)
    got = _oracle_awgn(oracle, x, snr, 42, 3)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_write_iq_matches_reference_file(tmp_path):
    """The writer on the samples of the reference's own output file gives
    that file byte for byte."""
    import rfsyn_amd
    lib = rfsyn_amd.load_library()
    iq = np.fromfile(GOLDEN, dtype="<f4")
    path = str(tmp_path / "out.iq").encode()
    assert lib.fdtd_write_iq(path, iq.ctypes.data_as(C.POINTER(C.c_float)), iq.size // 2) == 0
    assert open(path, "rb").read() == open(GOLDEN, "rb").read()
    assert lib.fdtd_write_iq(b"/nonexistent-dir/x.iq", iq.ctypes.data_as(C.POINTER(C.c_float)), 1) == -1


def _dipole_run(lib, steps=150):
    run = NativeRun(lib, (24, 24, 24), 1e-3, 0.99, cpml=c1_cpml_params(6), precision=capi.FDTD_FP64)
    run.add_point_source(capi.FIELD_EZ, (12, 12, 12), "source_waveform_modulated_gaussian", modgauss_params(30e9, 20e9))
    pid = run.add_probe(capi.FIELD_EZ, (17, 12, 12), steps, None)
    run.run(steps)
    return run, pid


def _iq(lib, run, pid, n, noise, stream=0):
    out = np.zeros((n, 2), dtype=np.float32)
    assert lib.fdtd_sim_probe_iq(run.sim, pid, n, noise, 15.0, 42, stream,
                                 out.ctypes.data_as(C.POINTER(C.c_float))) == 0
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("stream", [0, 7])
def test_device_output_stage(gpu, oracle, stream):
    """Device I/Q of a dipole's probe series against the oracle: without
    noise bit for bit; with AWGN every sample within 4e-6 of the largest
    noise magnitude (logf, cosf, sinf of the device are within a few ulps
    of glibc's), and the noiseless part unchanged."""
    n = 150
    a, pa = _dipole_run(gpu)
    b, pb = _dipole_run(oracle)
    clean_g, clean_o = _iq(gpu, a, pa, n, 0), _iq(oracle, b, pb, n, 0)
    assert np.array_equal(clean_g.view(np.uint32), clean_o.view(np.uint32))
    assert np.max(np.abs(clean_o[:, 0])) > 0 and not np.any(clean_o[:, 1])
    noisy_g, noisy_o = _iq(gpu, a, pa, n, 1, stream), _iq(oracle, b, pb, n, 1, stream)
    noise = noisy_o - clean_o
    assert np.max(np.abs(noise)) > 0
    assert np.max(np.abs(noisy_g - noisy_o)) <= 4e-6 * np.max(np.abs(noise))
    # more samples than recorded: refused
    out = np.zeros((n + 1, 2), dtype=np.float32)
    assert gpu.fdtd_sim_probe_iq(a.sim, pa, n + 1, 1, 15.0, 42, 0, out.ctypes.data_as(C.POINTER(C.c_float))) == -1
    a.close()
    b.close()


STAGE_SO = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                        "libstage_native.so")


class EchoOptions(C.Structure):
    _fields_ = [("precision", C.c_int), ("add_noise", C.c_int), ("snr_db", C.c_float), ("seed", C.c_uint64),
                ("stream", C.c_uint64)]


@pytest.mark.gpu
@pytest.mark.parametrize("noise", [0, 1])
def test_native_stage_matches_reference_chain(gpu, ref, noise):
    """rfsyn_amd/cli/fdtd_stage_native.c, built with the reference's own
    signal code (oracle/Makefile), as the CLI would run it: the default
    preset's FDTD stage (and AWGN at 15 dB, seed 42, stream 3, done on the
    device) against the reference's fdtd_stage (+ awgn) transforms:
    bitwise without noise, within 4e-6 of the noise magnitude with it."""
    if not os.path.exists(STAGE_SO):
        pytest.skip("stage harness not built (oracle/Makefile target ref)")
    stage = C.CDLL(STAGE_SO)
    stage.fdtd_stage_native_create.restype = Transform
    stage.fdtd_stage_native_create.argtypes = [C.POINTER(capi.SimConfig), C.POINTER(EchoOptions)]
    ref.fdtd_stage_create.restype = Transform
    ref.fdtd_stage_create.argtypes = [C.POINTER(capi.SimConfig)]
    cfg = _preset_cfg()
    opt = EchoOptions(capi.FDTD_FP64, noise, 15.0, 42, 3)
    got = _ref_chain(stage, [stage.fdtd_stage_native_create(C.byref(cfg), C.byref(opt))], 1024)
    want = _ref_chain(ref, [ref.fdtd_stage_create(C.byref(cfg))] + ([_ref_awgn(ref, 15.0, 42, 3)] if noise else []),
                      1024)
    assert got.shape == want.shape == (200, 2)
    if not noise:
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    else:
        clean = _ref_chain(ref, [ref.fdtd_stage_create(C.byref(cfg))], 1024)
        nz = want - clean
        assert np.max(np.abs(nz)) > 0
        assert np.max(np.abs(got - want)) <= 4e-6 * np.max(np.abs(nz))
