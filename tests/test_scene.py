"""Scene painter: the product library (painting on the device) and the
oracle (the host painter), given the same shapes, hold identical materials
(node coefficients, overrides and 3x3 tensors) and report the same counts;
and the oracle's smoothed tensors obey the closed forms of the mixing rule
(on the CPU)."""
import ctypes as C

import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import c2_shapes, c2_paint


def _materials_dump(lib, n, paint):
    g = lib.fdtd_grid_create(n[0], n[1], n[2], 1e-3, 0.99)
    if len(n) > 3:
        assert lib.fdtd_grid_set_periodic(g, n[3]) == 0
    m = lib.materials_create(g)
    stats = paint(lib, g, m)
    coeff = np.zeros((3,) + tuple(n[:3]) + (2,))
    ten = np.full((3,) + tuple(n[:3]) + (6,), np.nan)
    co = capi.MaterialCoeff()
    t6 = (C.c_double * 6)()
    for c in range(3):
        for i in range(n[0]):
            for j in range(n[1]):
                for k in range(n[2]):
                    assert lib.materials_get_coefficients(m, g, c, i, j, k, C.byref(co)) == 0
                    coeff[c, i, j, k] = (co.Ca, co.Cb)
                    if lib.materials_get_tensor3(m, g, c, i, j, k, t6) == 0:
                        ten[c, i, j, k] = t6[:]
    lib.materials_destroy(m)
    lib.fdtd_grid_destroy(g)
    return coeff, ten, stats


def _paint_shapes(shapes, ss, keep):
    def paint(lib, g, m):
        arr = (capi.Shape * len(shapes))(*shapes)
        opt = capi.SceneOptions(ss, keep)
        st = capi.SceneStats()
        assert lib.fdtd_scene_paint(m, g, arr, len(shapes), C.byref(opt), C.byref(st)) == 0
        return (st.nodes_painted, st.override_edges, st.tensor_edges, st.scalar_mixed_edges)
    return paint


def _shape(kind, eps, **kw):
    s = capi.Shape()
    s.kind, s.eps_r = kind, eps
    for k, v in kw.items():
        if k in ("radius", "half_thickness"):
            setattr(s, k, v)
        else:
            getattr(s, k)[:] = v
    return s


def _mixed_shapes():
    """Overlapping sphere, clipped tilted slab and boxes: later shapes win,
    a box face on integer coordinates, a slab clipped inside the grid, a
    sphere through a periodic boundary, two shapes of the same permittivity."""
    return [
        _shape(capi.SHAPE_BOX, 3.0, lo=(4.0, 3.0, 5.0), hi=(15.0, 12.0, 9.5)),
        _shape(capi.SHAPE_SLAB, 6.5, p=(12.0, 10.0, 14.0), normal=(0.3, -0.2, 1.0), half_thickness=2.25,
               lo=(2.5, 1.0, 0.0), hi=(20.0, 17.5, 27.0)),
        _shape(capi.SHAPE_SPHERE, 9.0, p=(7.3, 0.6, 20.2), radius=4.1),
        _shape(capi.SHAPE_BOX, 3.0, lo=(16.0, 12.0, 18.0), hi=(22.0, 19.0, 24.0)),
        _shape(capi.SHAPE_SPHERE, 2.0, p=(18.5, 15.5, 21.0), radius=3.0),
    ]


PAINT_CASES = {
    # the c2 scene's generator at 32^3 (spheres and a tilted slab, CPML band)
    "c2_32": ((32, 32, 32), lambda: c2_shapes(32, 6, 4, 2.5, 5.0), 4, 8),
    # mixed shapes, non-cubic grid, periodic x and y, 3^3 and 6^3 sub-samples
    "mixed_periodic_ss3": ((24, 20, 28, 3), _mixed_shapes, 3, 4),
    "mixed_open_ss6": ((24, 20, 28), _mixed_shapes, 6, 2),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(PAINT_CASES))
def test_painter_same_materials(gpu, oracle, case):
    n, shapes, ss, keep = PAINT_CASES[case]
    a = _materials_dump(gpu, n, _paint_shapes(shapes(), ss, keep))
    b = _materials_dump(oracle, n, _paint_shapes(shapes(), ss, keep))
    assert a[2] == b[2], (a[2], b[2])
    assert b[2][0] > 0 and b[2][2] + b[2][3] > 0
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(np.isnan(a[1]), np.isnan(b[1]))
    assert np.array_equal(np.nan_to_num(a[1]), np.nan_to_num(b[1]))


def test_smoothed_tensor_closed_form(oracle):
    """A single sphere: every tensor is lambda_n n n^T + lambda_t (I - n n^T)
    with eigenvalues between 1/eps_in and 1 and its eigenvector along the
    sphere normal at the edge."""
    n = 24
    g = oracle.fdtd_grid_create(n, n, n, 1e-3, 0.99)
    m = oracle.materials_create(g)
    s = capi.Shape()
    s.kind = capi.SHAPE_SPHERE
    s.eps_r = 9.0
    s.p[:] = (11.7, 12.2, 12.4)
    s.radius = 5.3
    arr = (capi.Shape * 1)(s)
    opt = capi.SceneOptions(6, 2)
    assert oracle.fdtd_scene_paint(m, g, arr, 1, C.byref(opt), None) == 0
    t6 = (C.c_double * 6)()
    seen = 0
    for c in range(3):
        for i in range(2, n - 2):
            for j in range(2, n - 2):
                for k in range(2, n - 2):
                    if oracle.materials_get_tensor3(m, g, c, i, j, k, t6) != 0:
                        continue
                    seen += 1
                    M = np.array([[t6[0], t6[1], t6[2]], [t6[1], t6[3], t6[4]], [t6[2], t6[4], t6[5]]])
                    x = np.array([i, j, k], dtype=float)
                    x[c] += 0.5
                    nv = (x - np.array(s.p[:])) / np.linalg.norm(x - np.array(s.p[:]))
                    w, v = np.linalg.eigh(M)
                    assert w[0] >= 1 / 9.0 - 1e-12 and w[-1] <= 1.0 + 1e-12
                    # the normal is an eigenvector (eigenvalue lambda_n)
                    ln = nv @ M @ nv
                    assert np.allclose(M @ nv, ln * nv, atol=1e-12)
                    # tangential eigenvalue lambda_t = 1/<eps>, lambda_n = <1/eps>, lambda_n >= lambda_t
                    lt = (np.trace(M) - ln) / 2
                    assert ln >= lt - 1e-12
                    f = (1 / lt - 1.0) / 8.0
                    assert abs(ln - ((1 - f) + f / 9.0)) < 1e-12
    assert seen > 100
    oracle.materials_destroy(m)
    oracle.fdtd_grid_destroy(g)
