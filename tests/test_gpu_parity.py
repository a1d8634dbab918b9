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
from rfsyn_amd.driver import same_values, run_c1, NativeRun, c1_cpml_params, modgauss_params, c1_freqs, random_fill

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


def test_c1_fp64_full(gpu, oracle):
    """The whole 64^3 dipole workload (2000 steps, 8 probes x 64
    frequencies) in FP64: every probe series bit for bit against the
    oracle, every DFT accumulator within 1e-12 of the largest."""
    td_g, dft_g = run_c1(gpu, steps=2000)
    td_o, dft_o = run_c1(oracle, steps=2000)
    assert np.max(np.abs(td_o)) > 0
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


def test_shell_overlap_bitwise_fp32(gpu, monkeypatch):
    """The CPML shell's H update on a second stream beside the fused kernel
    (the default) and after it on the one stream give the same bits (c2's
    scene at 48^3: a shell on every face, z psi in the store, tensor
    edges)."""
    a = _c2_small(gpu, capi.FDTD_FP32, 120)
    monkeypatch.setenv("RFSYN_SHELL_OVERLAP", "0")
    b = _c2_small(gpu, capi.FDTD_FP32, 120)
    assert np.max(np.abs(a["td"])) > 0
    assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), k


def _wide_plane_run(lib, with_monitors, steps=40, n=(512, 512, 24), nf=128):
    """Periodic x/y, 4-cell CPML on z, Ex plane source; two full-plane
    monitors of nf frequencies (2 x 2 x nf x 512^2 doubles of accumulators)
    are allocated, zero-filled, just before the source table."""
    run = NativeRun(lib, n, 0.5e-3, 0.99, periodic=3, cpml=capi.CpmlParams(4, 3.0, 1.0, 1.0, 0.05, 1e-8),
                    precision=capi.FDTD_FP32, compensated=1)
    run.add_plane_source(capi.FIELD_EX, 2, 8, "source_waveform_modulated_gaussian", modgauss_params(40e9, 30e9))
    if with_monitors:
        freqs = np.linspace(25e9, 55e9, nf)
        run.add_plane_monitor(2, 6, freqs)
        run.add_plane_monitor(2, n[2] - 6, freqs)
    pid = run.add_probe(capi.FIELD_EX, (5, 7, 12), steps, None)
    run.run(steps)
    td = run.td(pid, steps)
    run.close()
    return td


def test_large_monitor_allocations_keep_source_table(gpu):
    """Zero-filling large monitor buffers must finish before later uploads
    (here the source waveform table) land in freshly allocated memory; the
    monitors read the fields and never change them."""
    a = _wide_plane_run(gpu, True)
    b = _wide_plane_run(gpu, False)
    assert np.max(np.abs(b)) > 0
    assert same_values(a, b)


def _region_materials(n):
    """A scalar dielectric block crossing the CPML boundary and a lossy one
    inside, so tags differ between the plain box and the shell."""
    def setup(lib, g, mat):
        assert lib.materials_set_box(mat, g, 0, n[0] // 2, 1, n[1] - 2, 2, n[2] // 3, 3.0, 0.0) == 0
        assert lib.materials_set_box(mat, g, n[0] // 3, n[0] - 1, n[1] // 3, n[1] - 1, n[2] // 2, n[2] - 3,
                                     2.0, 0.02) == 0
    return setup


REGION_CASES = {
    # n, periodic mask, CPML depth (None: no CPML)
    "all_shell": ((12, 11, 13), 0, 5),
    "x_periodic": ((20, 18, 70), 1, 4),
    "z_periodic": ((30, 26, 20), 4, 5),
    "no_cpml": ((16, 20, 70), 0, None),
    "wide_z": ((10, 9, 150), 3, 7),
}


def _region_run(lib, case, steps=40):
    n, per, npml = REGION_CASES[case]
    cp = capi.CpmlParams(npml, 3.0, 1.0, 2.5, 0.05, 1e-8) if npml else None
    run = NativeRun(lib, n, 1e-3, 0.99, periodic=per, cpml=cp, material_setup=_region_materials(n),
                    init_fields=lambda g: random_fill(g, 11))
    p = modgauss_params(30e9, 20e9)
    run.add_point_source(capi.FIELD_EZ, (n[0] // 2, n[1] // 2, n[2] // 2), "source_waveform_modulated_gaussian", p)
    run.add_point_source(capi.FIELD_EX, (1, 1, 1), "source_waveform_modulated_gaussian", p, amp=0.5)
    run.add_point_source(capi.FIELD_EY, (n[0] - 2, 2, n[2] - 2), "source_waveform_modulated_gaussian", p,
                         hard=True)
    run.add_plane_source(capi.FIELD_EY, 1, 2, "source_waveform_modulated_gaussian", p, amp=0.3)
    run.add_plane_source(capi.FIELD_EZ, 0, n[0] // 2, "source_waveform_modulated_gaussian", p, amp=0.2,
                         after_e=1)
    run.add_plane_source(capi.FIELD_EX, 2, n[2] - 3, "source_waveform_modulated_gaussian", p, amp=0.1,
                         after_e=1, hard=True)
    pid = run.add_probe(capi.FIELD_EZ, (n[0] // 2 + 1, n[1] // 2, n[2] // 2), steps, None)
    run.run(steps)
    out = {"td": run.td(pid, steps), "fields": run.fields()}
    run.close()
    return out


@pytest.mark.parametrize("case", sorted(REGION_CASES))
def test_plain_box_and_shell_fp64(gpu, oracle, case):
    """The fused step's split into the CPML-free box and the shell around
    it, with sources inside, across and outside the box, random initial
    fields and materials on both sides: bitwise against the oracle."""
    a = _region_run(gpu, case)
    b = _region_run(oracle, case)
    assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), (k, np.max(np.abs(a["fields"][k] - b["fields"][k])))


def _ade_region_setup(n):
    """A 2-pole Lorentz block and a Drude block that both cross the CPML
    slabs of a closed box, so the ADE state box is split between the fused
    box (ADE pass) and the shell (in-kernel ADE)."""
    def setup(lib, g, mat):
        d = capi.MaterialDispersive()
        d.eps_inf = 2.25
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
        assert lib.materials_assign_box(mat, g, 2, n[0] - 3, 1, n[1] - 2, n[2] // 3, n[2] // 3 + 6, lid.value) == 0
        assert lib.materials_assign_box(mat, g, 4, n[0] // 2, 6, n[1] - 1, n[2] // 2, n[2] - 2, did.value) == 0
        assert lib.materials_set_box(mat, g, 1, n[0] - 1, 2, 9, 3, 7, 3.0, 0.0) == 0
    return setup


def _ade_region_run(lib, src_in_ade, steps=50):
    n = (22, 20, 36)
    run = NativeRun(lib, n, 0.5e-3, 0.99, cpml=capi.CpmlParams(5, 3.0, 1.0, 2.0, 0.05, 1e-8),
                    material_setup=_ade_region_setup(n), init_fields=lambda g: random_fill(g, 5))
    p = modgauss_params(40e9, 30e9)
    run.add_point_source(capi.FIELD_EY, (11, 10, 4), "source_waveform_modulated_gaussian", p)
    if src_in_ade:   # inside the Lorentz block: the ADE update stays in the fused kernel
        run.add_point_source(capi.FIELD_EX, (10, 9, 14), "source_waveform_modulated_gaussian", p, amp=0.5)
    pid = run.add_probe(capi.FIELD_EX, (9, 9, 13), steps, None)
    run.run(steps)
    out = {"td": run.td(pid, steps), "fields": run.fields()}
    run.close()
    return out


@pytest.mark.parametrize("split", ["auto", "off"])
@pytest.mark.parametrize("src_in_ade", [False, True])
def test_ade_pass_box_and_shell_fp64(gpu, oracle, split, src_in_ade, monkeypatch):
    """Dispersive blocks across the CPML shell, random initial fields: the
    split ADE step (plain part in the fused kernel, ADE pass after it), the
    in-kernel ADE update (forced, or chosen because a source lies in the
    ADE box) and the shell's ADE update, bitwise against the oracle."""
    if split == "off":
        monkeypatch.setenv("RFSYN_ADE_SPLIT", "0")
    a = _ade_region_run(gpu, src_in_ade)
    b = _ade_region_run(oracle, src_in_ade)
    assert np.max(np.abs(b["td"])) > 0
    assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), (k, np.max(np.abs(a["fields"][k] - b["fields"][k])))


def test_ade_split_matches_inline_fp32(gpu, monkeypatch):
    """FP32 periodic dispersive stack: the split ADE step and the in-kernel
    update perform the same operations in the same order, so the fields
    agree bit for bit.  The split step's ADE pass also takes the plane
    captures (both monitor planes lie outside the dispersive layers); the
    in-kernel step captures in a launch of its own: the plane DFTs agree
    bit for bit too."""
    a = _run_periodic_stack(gpu, capi.FDTD_FP32, 120)
    monkeypatch.setenv("RFSYN_ADE_SPLIT", "0")
    b = _run_periodic_stack(gpu, capi.FDTD_FP32, 120)
    assert np.max(np.abs(a["td"])) > 0
    assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), k
    assert np.max(np.abs(a["T"])) > 0
    assert same_values(a["R"], b["R"]) and same_values(a["T"], b["T"])


def _monitor_runs(lib, precision):
    """Plane DFTs of two scenes: the periodic dispersive stack with z planes
    outside and inside the dispersive layers and a y plane across them,
    and a box without CPML, with lossy blocks and sources, with x, y and z
    planes (the last at the padded node nz)."""
    out = []
    freqs = np.linspace(25e9, 55e9, 8)
    run = NativeRun(lib, (16, 16, 44), 0.5e-3, 0.99, periodic=3, cpml=capi.CpmlParams(8, 3.0, 1.0, 1.0, 0.05, 1e-8),
                    precision=precision, material_setup=_dispersive_setup)
    run.add_plane_source(capi.FIELD_EX, 2, 10, "source_waveform_modulated_gaussian", modgauss_params(40e9, 30e9))
    mons = [(run.add_plane_monitor(2, k, freqs), (16, 16, 8)) for k in (9, 22, 36)]
    mons.append((run.add_plane_monitor(1, 7, freqs), (16, 44, 8)))
    run.run(150)
    out += [run.plane_dft(m, w, sh) for m, sh in mons for w in (0, 1)]
    run.close()
    n = (16, 20, 30)
    run = NativeRun(lib, n, 1e-3, 0.99, precision=precision, material_setup=_region_materials(n),
                    init_fields=lambda g: random_fill(g, 4))
    run.add_point_source(capi.FIELD_EZ, (8, 10, 15), "source_waveform_modulated_gaussian", modgauss_params(30e9, 20e9))
    run.add_plane_source(capi.FIELD_EY, 2, 4, "source_waveform_modulated_gaussian", modgauss_params(30e9, 20e9),
                         amp=0.3, after_e=1)
    mons = [(run.add_plane_monitor(0, 5, freqs), (20, 30, 8)), (run.add_plane_monitor(1, 12, freqs), (16, 30, 8)),
            (run.add_plane_monitor(2, 4, freqs), (16, 20, 8)), (run.add_plane_monitor(2, 30, freqs), (16, 20, 8))]
    run.run(90)
    out += [run.plane_dft(m, w, sh) for m, sh in mons for w in (0, 1)]
    run.close()
    return out


def test_plane_monitors_all_axes_fp64(gpu, oracle):
    """Eight plane monitors (x, y and z planes, one at the padded node)
    captured in one launch per step: DFTs against the oracle."""
    a = _monitor_runs(gpu, capi.FDTD_FP64)
    b = _monitor_runs(oracle, capi.FDTD_FP64)
    for q, (x, y) in enumerate(zip(a, b)):
        assert _dft_close(x, y), q
    for q in range(0, len(b), 2):   # each monitor sees a field (the stack's Ey is 0 by symmetry)
        assert max(np.max(np.abs(b[q])), np.max(np.abs(b[q + 1]))) > 0, q
