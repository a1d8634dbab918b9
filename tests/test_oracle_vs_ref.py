"""The CPU oracle against the rfsyn reference compiled from its own sources
(oracle/Makefile `ref`), bit for bit.  Skipped where the reference sources
are absent; there the reference's own output fixture (tests/golden/,
read by test_golden_echo.py and test_output_stage.py), the validators'
published figures and the closed forms of the other oracle tests pin it."""
import ctypes as C

import numpy as np
import pytest

from rfsyn_amd import capi
from lowlevel import wrapping_case, cpml_scene, tensor_scene
from rfsyn_amd.driver import same_values


def _default_preset():
    # cli/presets.h:14-29, cli/settings.c:23-32: 50^3, 200 steps, npml 8,
    # t0 = 40 dt, tau = 8 dt, probe at centre + (6,0,0).
    return capi.SimConfig(50, 50, 50, 1e-3, 0.99, 200, 8, 25, 25, 25, capi.FIELD_EZ, 40.0, 8.0,
                          31, 25, 25, capi.FIELD_EZ)


def _echo(lib, cfg):
    buf = (C.c_double * cfg.n_steps)()
    fs = C.c_double()
    assert lib.fdtd_simulate_echo(C.byref(cfg), buf, C.byref(fs)) == 0
    return np.array(buf[:]), fs.value


def test_simulate_echo_bitwise(oracle, ref):
    a, fa = _echo(oracle, _default_preset())
    b, fb = _echo(ref, _default_preset())
    assert fa == fb
    assert same_values(a, b)


@pytest.mark.parametrize("axes", range(8))
@pytest.mark.parametrize("mode", range(4))
def test_wrapping_cases_bitwise(oracle, ref, axes, mode):
    a = wrapping_case(oracle, axes, mode, steps=3)
    b = wrapping_case(ref, axes, mode, steps=3)
    for k in a:
        assert same_values(a[k], b[k]), k


@pytest.mark.parametrize("periodic", [0, 3, 4])
def test_cpml_materials_scene_bitwise(oracle, ref, periodic):
    a = cpml_scene(oracle, periodic=periodic)
    b = cpml_scene(ref, periodic=periodic)
    for k in a:
        assert same_values(a[k], b[k]), k


def test_tensor_scene_bitwise(oracle, ref):
    a = tensor_scene(oracle, kind=2)
    b = tensor_scene(ref, kind=2)
    for k in a:
        assert same_values(a[k], b[k]), k


def test_cpml_profiles_bitwise(oracle, ref):
    for lib_pair in [(oracle, ref)]:
        vals = []
        for lib in lib_pair:
            g = lib.fdtd_grid_create(30, 17, 24, 1e-3, 0.99)
            lib.fdtd_grid_set_periodic(g, 2)
            p = lib.cpml_create(g, capi.CpmlParams(6, 3.0, 1.0, 7.0, 0.05, 1e-8))
            arrs = {}
            for name, n in [("kappa_inv_m_x", 31), ("b_m_x", 31), ("a_m_x", 31), ("kappa_inv_h_x", 30),
                            ("b_h_x", 30), ("a_h_x", 30), ("a_m_y", 18), ("b_h_y", 17), ("a_h_z", 24),
                            ("kappa_inv_m_z", 25)]:
                arrs[name] = np.ctypeslib.as_array(getattr(p.contents, name), shape=(n,)).copy()
            vals.append(arrs)
            lib.cpml_destroy(p)
            lib.fdtd_grid_destroy(g)
        for k in vals[0]:
            assert np.array_equal(vals[0][k], vals[1][k]), k


def test_probe_dft_record_bitwise(oracle, ref):
    """validate_probe_dft.c:66-82 stepping (vacuum H/E with CPML, DFT at
    t=(n+1)dt) on a 24^3 grid."""
    out = []
    for lib in (oracle, ref):
        g = lib.fdtd_grid_create(24, 24, 24, 1e-3, 0.99)
        pml = lib.cpml_create(g, lib.cpml_default_params(5))
        p = capi.GaussDerivParams(40.0 * g.contents.dt, 8.0 * g.contents.dt)
        wf = capi.waveform_ptr(lib, "source_waveform_gaussian_derivative")
        src = lib.source_make_point(12, 12, 12, capi.FIELD_EZ, capi.SRC_SOFT, 1.0, wf,
                                    C.cast(C.pointer(p), C.c_void_p))
        freqs = (C.c_double * 3)(10e9, 15e9, 27e9)
        pr = capi.ProbeDFT()
        assert lib.probe_dft_create(C.byref(pr), 16, 12, 12, capi.FIELD_EZ, freqs, 3) == 0
        dt = g.contents.dt
        for n in range(120):
            lib.fdtd_update_h_cpml(g, pml)
            lib.source_apply(C.byref(src), g, n * dt)
            lib.fdtd_update_e_cpml(g, pml)
            lib.probe_dft_record(C.byref(pr), g, (n + 1) * dt, dt)
        out.append(np.array([pr.acc_re[q] for q in range(3)] + [pr.acc_im[q] for q in range(3)]))
        lib.probe_dft_destroy(C.byref(pr))
        lib.cpml_destroy(pml)
        lib.fdtd_grid_destroy(g)
    assert same_values(out[0], out[1])


def test_materials_queries_match(oracle, ref):
    """Coefficient lookups after a sequence of box/override/tensor edits."""
    res = []
    for lib in (oracle, ref):
        g = lib.fdtd_grid_create(12, 12, 12, 1e-3, 0.99)
        m = lib.materials_create(g)
        mid = C.c_uint8()
        assert lib.materials_register(m, 4.0, 0.0, C.byref(mid)) == 0
        assert lib.materials_assign_box(m, g, 5, 8, 5, 8, 5, 8, mid.value) == 0
        assert lib.materials_set_component(m, g, capi.FIELD_EX, 6, 6, 6, 9.0, 0.1) == 0
        assert lib.materials_set_component(m, g, capi.FIELD_EZ, 6, 6, 6, 16.0, 0.0) == 0
        assert lib.materials_set_component_tensor(m, g, capi.FIELD_EY, 4, 4, 4, 0.5, 0.1, 0.4) == 0
        assert lib.materials_set_box(m, g, 6, 7, 6, 7, 6, 7, 1.0, 0.0) == 0
        assert lib.materials_set_box(m, g, 0, 14, 0, 1, 0, 1, 4.0, 0.0) == -1
        assert lib.materials_set_component(m, g, capi.FIELD_HX, 6, 6, 6, 4.0, 0.0) == -1
        row = []
        for c in range(3):
            for (i, j, k) in [(6, 6, 6), (5, 5, 5), (4, 6, 6), (7, 7, 7), (8, 8, 8)]:
                co = capi.MaterialCoeff()
                assert lib.materials_get_coefficients(m, g, c, i, j, k, C.byref(co)) == 0
                row += [co.Ca, co.Cb]
        t = capi.MaterialTensor()
        row.append(lib.materials_get_tensor(m, g, capi.FIELD_EY, 4, 4, 4, C.byref(t)))
        row += [t.m11, t.m12, t.m22, t.scale]
        row.append(lib.materials_get_tensor(m, g, capi.FIELD_EX, 4, 4, 4, C.byref(t)))
        res.append(np.array(row, dtype=np.float64))
        lib.materials_destroy(m)
        lib.fdtd_grid_destroy(g)
    assert same_values(res[0], res[1])
