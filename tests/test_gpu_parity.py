"""HIP library against the CPU oracle on the same seeded inputs.

FP64: bitwise on fields and time series (+0 and -0 compare equal); DFT
sums within 1e-12 relative of the largest accumulator (the device
evaluates cos/sin with its own libm, which may differ from glibc in the
last bit).  FP32: probe-DFT error <= 1e-4 of each probe's peak in band.
"""
import ctypes as C

import numpy as np
import pytest

from rfsyn_amd import capi
from lowlevel import wrapping_case, cpml_scene, tensor_scene
from scenes import same_values, run_c1, NativeRun, c1_cpml_params, modgauss_params, c1_freqs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("axes", range(8))
@pytest.mark.parametrize("mode", range(4))
def test_wrapping_cases(gpu, oracle, axes, mode):
    """tests/periodic_wrapping_unit.c: 8 periodic masks x 4 update modes."""
    a = wrapping_case(gpu, axes, mode, steps=3)
    b = wrapping_case(oracle, axes, mode, steps=3)
    for k in a:
        assert same_values(a[k], b[k]), (k, np.max(np.abs(a[k] - b[k])))


@pytest.mark.parametrize("periodic", [0, 3, 4, 7])
def test_cpml_materials_scene(gpu, oracle, periodic):
    a = cpml_scene(gpu, periodic=periodic)
    b = cpml_scene(oracle, periodic=periodic)
    for k in a:
        assert same_values(a[k], b[k]), (k, np.max(np.abs(a[k] - b[k])))


@pytest.mark.parametrize("kind", [2, 3])
def test_tensor_scene(gpu, oracle, kind):
    a = tensor_scene(gpu, kind=kind)
    b = tensor_scene(oracle, kind=kind)
    for k in a:
        assert same_values(a[k], b[k]), (k, np.max(np.abs(a[k] - b[k])))


def test_simulate_echo(gpu, oracle):
    cfg = capi.SimConfig(50, 50, 50, 1e-3, 0.99, 200, 8, 25, 25, 25, capi.FIELD_EZ, 40.0, 8.0,
                         31, 25, 25, capi.FIELD_EZ)
    out = []
    for lib in (gpu, oracle):
        buf = (C.c_double * 200)()
        fs = C.c_double()
        assert lib.fdtd_simulate_echo(C.byref(cfg), buf, C.byref(fs)) == 0
        out.append((np.array(buf[:]), fs.value))
    assert out[0][1] == out[1][1]
    assert same_values(out[0][0], out[1][0])
    # output_phase1_check/example_000000.iq: peak 7.8964634e-3 at sample 41
    td = out[0][0]
    assert int(np.argmax(np.abs(td))) == 41
    assert np.float32(td[41]) == np.float32(7.8964634e-3)


def _dft_close(a, b, rel=1e-12):
    scale = np.max(np.abs(b))
    return np.max(np.abs(a - b)) <= rel * scale


@pytest.mark.parametrize("block", [0, 7])
def test_c1_fp64_reduced(gpu, oracle, block):
    """64^3 dipole scene, 300 steps: series bitwise, DFT to 1e-12."""
    td_g, dft_g = run_c1(gpu, steps=300, block=block)
    td_o, dft_o = run_c1(oracle, steps=300)
    assert same_values(td_g, td_o)
    assert _dft_close(dft_g, dft_o)


def test_c1_fp32_vs_fp64(gpu):
    """FP32 fields with FP64 compensated DFT against the FP64 GPU run."""
    _, d64 = run_c1(gpu, steps=600)
    _, d32 = run_c1(gpu, steps=600, precision=capi.FDTD_FP32, compensated=1)
    peak = np.max(np.abs(d64), axis=1, keepdims=True)
    err = np.max(np.abs(d32 - d64) / peak)
    assert err <= 1e-4, err


def _dispersive_setup(lib, g, mat):
    d = capi.MaterialDispersive()
    d.eps_inf = 2.25
    d.sigma = 0.0
    d.n_lorentz = 2
    d.lorentz_deps[0], d.lorentz_deps[1] = 1.0, 0.5
    d.lorentz_w0[0], d.lorentz_w0[1] = 2 * np.pi * 60e9, 2 * np.pi * 90e9
    d.lorentz_delta[0], d.lorentz_delta[1] = 0.05 * 2 * np.pi * 60e9, 0.05 * 2 * np.pi * 90e9
    lid = C.c_uint8()
    assert lib.materials_register_dispersive(mat, C.byref(d), C.byref(lid)) == 0
    dr = capi.MaterialDispersive()
    dr.eps_inf = 1.0
    dr.n_drude = 1
    dr.drude_wp[0] = 1.37e16
    dr.drude_gamma[0] = 1e14
    did = C.c_uint8()
    assert lib.materials_register_dispersive(mat, C.byref(dr), C.byref(did)) == 0
    assert lib.materials_assign_box(mat, g, 0, 16, 0, 16, 20, 26, lid.value) == 0
    for i0 in range(0, 16, 8):
        assert lib.materials_assign_box(mat, g, i0, i0 + 4, 0, 16, 14, 17, did.value) == 0
    assert lib.materials_set_box(mat, g, 0, 16, 0, 16, 28, 34, 4.0, 0.0) == 0


def _run_periodic_stack(lib, precision, steps, compensated=0):
    run = NativeRun(lib, (16, 16, 44), 0.5e-3, 0.99, periodic=3, cpml=capi.CpmlParams(8, 3.0, 1.0, 1.0, 0.05, 1e-8),
                    precision=precision, compensated=compensated, material_setup=_dispersive_setup)
    run.add_plane_source(capi.FIELD_EX, 2, 10, "source_waveform_modulated_gaussian",
                         modgauss_params(40e9, 30e9))
    freqs = np.linspace(25e9, 55e9, 16)
    m_r = run.add_plane_monitor(2, 9, freqs)
    m_t = run.add_plane_monitor(2, 36, freqs)
    p = run.add_probe(capi.FIELD_EX, (3, 5, 24), steps, freqs)
    run.run(steps)
    out = {"td": run.td(p, steps), "dft": run.dft(p),
           "R": run.plane_dft(m_r, 0, (16, 16, 16)), "T": run.plane_dft(m_t, 0, (16, 16, 16)),
           "fields": run.fields(), "energy": run.energy()}
    run.close()
    return out


def test_dispersive_periodic_stack_fp64(gpu, oracle):
    """Periodic x/y, CPML on z, Ex plane source, Drude grating + 2-pole
    Lorentz layer + eps=4 substrate, plane monitors."""
    a = _run_periodic_stack(gpu, capi.FDTD_FP64, 160)
    b = _run_periodic_stack(oracle, capi.FDTD_FP64, 160)
    assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), k
    assert _dft_close(a["dft"], b["dft"])
    assert _dft_close(a["R"], b["R"])
    assert _dft_close(a["T"], b["T"])
    for x, y in zip(a["energy"], b["energy"]):
        assert abs(x - y) <= 1e-12 * abs(y)


def _c2_small(lib, precision, steps):
    from rfsyn_amd.driver import build_c2
    run, pids = build_c2(lib, precision, steps, n=48, npml=6, n_spheres=5, r_lo=3.0, r_hi=6.0, n_probes=4)
    run.run(steps)
    out = {"td": np.stack([run.td(p, steps) for p in pids]), "dft": np.stack([run.dft(p) for p in pids]),
           "fields": run.fields()}
    run.close()
    return out


def test_smoothed_dielectric_fp64(gpu, oracle):
    """Spheres + tilted slab with 3x3 subpixel tensors (48^3, 6-cell CPML)."""
    a = _c2_small(gpu, capi.FDTD_FP64, 150)
    b = _c2_small(oracle, capi.FDTD_FP64, 150)
    assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), k
    assert _dft_close(a["dft"], b["dft"])


def test_smoothed_dielectric_fp32_vs_fp64(gpu):
    a = _c2_small(gpu, capi.FDTD_FP64, 400)
    b = _c2_small(gpu, capi.FDTD_FP32, 400)
    peak = np.max(np.abs(a["dft"]), axis=1, keepdims=True)
    assert np.max(np.abs(b["dft"] - a["dft"]) / peak) <= 1e-4
