"""Scene painter: the product library and the oracle, given the same
shapes, hold identical materials (node coefficients, overrides and 3x3
tensors); and the smoothed tensors obey the closed forms of the mixing
rule.  Host-side only (no GPU)."""
import ctypes as C

import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import c2_shapes, c2_paint


def _materials_dump(lib, n=32, npml=6):
    g = lib.fdtd_grid_create(n, n, n, 1e-3, 0.99)
    m = lib.materials_create(g)
    c2_paint(lib, g, m, c2_shapes(n, npml, 4, 2.5, 5.0), npml)
    coeff = np.zeros((3, n, n, n, 2))
    ten = np.full((3, n, n, n, 6), np.nan)
    co = capi.MaterialCoeff()
    t6 = (C.c_double * 6)()
    for c in range(3):
        for i in range(n):
            for j in range(n):
                for k in range(n):
                    assert lib.materials_get_coefficients(m, g, c, i, j, k, C.byref(co)) == 0
                    coeff[c, i, j, k] = (co.Ca, co.Cb)
                    if lib.materials_get_tensor3(m, g, c, i, j, k, t6) == 0:
                        ten[c, i, j, k] = t6[:]
    lib.materials_destroy(m)
    lib.fdtd_grid_destroy(g)
    return coeff, ten


def test_painter_same_materials(oracle):
    import rfsyn_amd
    prod = rfsyn_amd.load_library()
    a = _materials_dump(prod)
    b = _materials_dump(oracle)
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
