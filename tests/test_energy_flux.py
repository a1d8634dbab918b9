"""Energy audit and Poynting flux planes (audit_periodic_energy.c:43-71,
validation_common.h; extensions of the reference's host-side audits).

Pinned by mathematics, independently of any implementation:
- in a closed lossless grid the modified energy 0.5 eps E^n.E^n +
  0.5 mu0 H^{n-1/2}.H^{n+1/2} is invariant to rounding while the raw
  energy is not;
- between two flux planes the change of the slab's modified energy,
  computed here in numpy from downloaded fields, equals minus the
  difference of the planes' flux integrals (the discrete Poynting theorem
  of the Yee scheme, an exact identity for the pairing the planes use).
The oracle is checked against both; the device library against the
oracle (FP64) and against both identities.
"""
import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import NativeRun, random_fill

EPS0, MU0 = 8.8541878128e-12, 1.25663706212e-6   # include/constants.h


def _closed_run(lib, n=(14, 12, 10), eps_r=None, precision=capi.FDTD_FP64, seed=4):
    def mats(lb, g, mat):
        if eps_r:
            assert lb.materials_set_box(mat, g, 0, n[0] // 2, 0, n[1], 0, n[2], eps_r, 0.0) == 0
    return NativeRun(lib, n, 1e-3, 0.99, periodic=7, material_setup=mats if eps_r else None,
                     precision=precision, init_fields=lambda g: random_fill(g, seed))


def _audit_series(lib, steps, **kw):
    run = _closed_run(lib, **kw)
    rows = []
    for _ in range(steps):
        run.run(1)
        rows.append(run.energy_audit()[0])
    run.close()
    return np.array(rows)


def _check_invariant(rows, tol):
    modified = rows[:, 0] + rows[:, 4]
    raw = rows[:, 0] + rows[:, 1]
    drift = np.max(np.abs(modified - modified[0])) / modified[0]
    swing = np.max(np.abs(raw - raw[0])) / raw[0]
    assert drift <= tol, drift
    assert swing > 100 * max(drift, 1e-16), (swing, drift)
    # the synchronized energy lies between the two one-sided ones' mean
    # and is bounded by their max (convexity)
    assert np.all(rows[:, 3] <= 0.5 * (rows[:, 1] + rows[:, 2]) * (1 + 1e-12))


def _slab_energy(f0, f1, lo, hi, eps, dx, mu0, n):
    """Modified energy of the x slab between flux planes lo < hi: Ey, Ez
    at nodes (lo, hi], Ex at (lo..hi-1), Hy, Hz at array [lo, hi), Hx at
    (lo, hi]; f0 = state n, f1 = state n+1 (all y, z cells)."""
    shape = (n[0] + 1, n[1] + 1, n[2] + 1)
    f0 = {k: v.reshape(shape) for k, v in f0.items()}
    f1 = {k: v.reshape(shape) for k, v in f1.items()}
    ny, nz = n[1], n[2]
    e_t = slice(lo + 1, hi + 1)
    h_t = slice(lo, hi)
    u = 0.0
    u += 0.5 * eps * np.sum(f0["ey"][e_t, :ny, :nz] ** 2 + f0["ez"][e_t, :ny, :nz] ** 2)
    u += 0.5 * eps * np.sum(f0["ex"][h_t, :ny, :nz] ** 2)
    u += 0.5 * mu0 * np.sum(f0["hy"][h_t, :ny, :nz] * f1["hy"][h_t, :ny, :nz]
                            + f0["hz"][h_t, :ny, :nz] * f1["hz"][h_t, :ny, :nz])
    u += 0.5 * mu0 * np.sum(f0["hx"][e_t, :ny, :nz] * f1["hx"][e_t, :ny, :nz])
    return u * dx ** 3


def _balance(lib, precision=capi.FDTD_FP64, steps=(6, 40), lo=3, hi=11):
    """(slab energy change, flux difference) between steps[0] and steps[1]."""
    n3 = (16, 10, 9)
    run = _closed_run(lib, n=n3, precision=precision, seed=9)
    fl, fh = run.add_flux_plane(0, lo), run.add_flux_plane(0, hi)
    g = run.g.contents
    out = {}
    n = 0
    for target in steps:
        run.run(target - n)
        f0 = run.fields()
        w = (run.flux(fl), run.flux(fh))
        run.run(1)
        f1 = run.fields()
        n = target + 1
        out[target] = (_slab_energy(f0, f1, lo, hi, EPS0, g.dx, MU0, n3), w)
    run.close()
    (u0, w0), (u1, w1) = out[steps[0]], out[steps[1]]
    du = u1 - u0
    dflux = (w1[1] - w0[1]) - (w1[0] - w0[0])
    return du, dflux, abs(u0)


def test_oracle_modified_energy_is_invariant(oracle):
    _check_invariant(_audit_series(oracle, 60), 1e-13)
    _check_invariant(_audit_series(oracle, 60, eps_r=4.0), 1e-13)


def test_oracle_flux_balances_slab_energy(oracle):
    du, dflux, scale = _balance(oracle)
    assert abs(du) > 1e-6 * scale
    assert abs(du + dflux) <= 1e-12 * scale, (du, dflux)


@pytest.mark.gpu
def test_audit_matches_oracle_fp64(gpu, oracle):
    for eps_r in (None, 4.0):
        a = _audit_series(gpu, 25, eps_r=eps_r)
        b = _audit_series(oracle, 25, eps_r=eps_r)
        assert np.max(np.abs(a - b) / np.abs(b).max(axis=0)) <= 1e-12


@pytest.mark.gpu
def test_audit_invariant_on_device(gpu):
    _check_invariant(_audit_series(gpu, 60, eps_r=4.0), 1e-13)
    _check_invariant(_audit_series(gpu, 60, eps_r=4.0, precision=capi.FDTD_FP32), 2e-6)


@pytest.mark.gpu
def test_flux_matches_oracle_and_balances(gpu, oracle):
    du, dflux, scale = _balance(gpu)
    assert abs(du) > 1e-6 * scale
    assert abs(du + dflux) <= 1e-12 * scale, (du, dflux)
    du_o, dflux_o, _ = _balance(oracle)
    assert abs(dflux - dflux_o) <= 1e-12 * scale
