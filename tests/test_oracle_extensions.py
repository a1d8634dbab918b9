"""Pins for the oracle's extensions that have no reference implementation.

* Dispersive ADE: on a fully periodic grid filled with one dispersive
  medium and a uniform E field, the curl vanishes, H stays zero and each
  edge evolves as the trapezoidal rule applied to the linear ODE
  x' = A x, x = (E, J_1, P_1, ...):  eps0 eps_inf E' = -sigma E - sum J,
  J' = -w0^2 P - g J + eps0 A_p E, P' = J.  The trapezoidal (bilinear)
  propagator (I - dt/2 A)^-1 (I + dt/2 A) is built here from the ODE with
  numpy and compared with the oracle's field history.
* 3x3 tensors: the operator H -> (H update of (E update of H)) is
  -cb C^T Xi C; the pair rule must make it symmetric with zero row sums
  (uniform H is curl-free), the 3D form of the reference's operator test
  (tests/materials_tensor_unit.c:1057-1165).  A diagonal tensor 1/eps I is
  the scalar update (materials_tensor_unit.c:394-449).
"""
import ctypes as C

import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import field_arrays

EPS0 = 8.8541878128e-12


def _uniform_dispersive(oracle, disp, steps, e0=1.0, n=4):
    g = oracle.fdtd_grid_create(n, n, n, 1e-3, 0.99)
    assert oracle.fdtd_grid_set_periodic(g, 7) == 0
    m = oracle.materials_create(g)
    mid = C.c_uint8()
    assert oracle.materials_register_dispersive(m, C.byref(disp), C.byref(mid)) == 0
    assert oracle.materials_assign_box(m, g, 0, n, 0, n, 0, n, mid.value) == 0
    f = field_arrays(g)
    f["ex"][:] = e0
    hist = []
    for _ in range(steps):
        oracle.fdtd_update_h(g)
        oracle.fdtd_update_e_materials(g, m)
        hist.append(f["ex"][0])
        assert np.all(f["hx"] == 0.0) and np.all(f["hy"] == 0.0) and np.all(f["hz"] == 0.0)
    dt = g.contents.dt
    oracle.materials_destroy(m)
    oracle.fdtd_grid_destroy(g)
    return np.array(hist), dt


def _propagator(disp, dt):
    """(E, J_p, P_p ...) state matrix of the continuous ODE, bilinear map."""
    poles = []
    for q in range(disp.n_drude):
        poles.append((0.0, disp.drude_gamma[q], disp.drude_wp[q] ** 2))
    for q in range(disp.n_lorentz):
        w0 = disp.lorentz_w0[q]
        poles.append((w0 * w0, 2 * disp.lorentz_delta[q], disp.lorentz_deps[q] * w0 * w0))
    n = 1 + 2 * len(poles)
    A = np.zeros((n, n))
    ce = EPS0 * disp.eps_inf
    A[0, 0] = -disp.sigma / ce
    for p, (w0sq, gam, amp) in enumerate(poles):
        jj, pp = 1 + 2 * p, 2 + 2 * p
        A[0, jj] = -1.0 / ce
        A[jj, 0] = EPS0 * amp
        A[jj, jj] = -gam
        A[jj, pp] = -w0sq
        A[pp, jj] = 1.0
    I = np.eye(n)
    return np.linalg.solve(I - 0.5 * dt * A, I + 0.5 * dt * A)


def _disp(eps_inf=1.0, sigma=0.0, drude=(), lorentz=()):
    d = capi.MaterialDispersive()
    d.eps_inf = eps_inf
    d.sigma = sigma
    d.n_drude = len(drude)
    for q, (wp, gam) in enumerate(drude):
        d.drude_wp[q] = wp
        d.drude_gamma[q] = gam
    d.n_lorentz = len(lorentz)
    for q, (de, w0, dl) in enumerate(lorentz):
        d.lorentz_deps[q] = de
        d.lorentz_w0[q] = w0
        d.lorentz_delta[q] = dl
    return d


CASES = {
    "drude": _disp(1.0, 0.0, drude=[(2 * np.pi * 30e9, 2e10)]),
    "drude_stiff": _disp(1.0, 0.0, drude=[(1.37e16, 1e14)]),
    "lorentz2": _disp(2.25, 0.0, lorentz=[(1.0, 2 * np.pi * 60e9, 0.05 * 2 * np.pi * 60e9),
                                          (0.5, 2 * np.pi * 90e9, 0.05 * 2 * np.pi * 90e9)]),
    "mixed_lossy": _disp(3.0, 0.02, drude=[(2 * np.pi * 20e9, 5e9)],
                         lorentz=[(2.0, 2 * np.pi * 45e9, 1e9)]),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_ade_trapezoidal_propagator(oracle, name):
    disp = CASES[name]
    steps = 400
    hist, dt = _uniform_dispersive(oracle, disp, steps)
    M = _propagator(disp, dt)
    x = np.zeros(M.shape[0])
    x[0] = 1.0
    want = []
    for _ in range(steps):
        x = M @ x
        want.append(x[0])
    want = np.array(want)
    assert np.max(np.abs(hist - want)) <= 1e-9 * max(1.0, np.max(np.abs(want)))
    # unconditional stability: the propagator's spectral radius <= 1
    assert np.max(np.abs(np.linalg.eigvals(M))) <= 1.0 + 1e-12


def test_zero_pole_dispersive_is_plain_material(oracle):
    """eps_inf and sigma alone give the plain lossy update (to rounding)."""
    n = 6
    out = []
    for disp in (True, False):
        g = oracle.fdtd_grid_create(n, n, n, 1e-3, 0.9)
        oracle.fdtd_grid_set_periodic(g, 7)
        rng = np.random.default_rng(3)
        for a in field_arrays(g).values():
            a[:] = rng.uniform(-1, 1, a.shape)
        m = oracle.materials_create(g)
        if disp:
            mid = C.c_uint8()
            assert oracle.materials_register_dispersive(m, C.byref(_disp(4.0, 0.03)), C.byref(mid)) == 0
            assert oracle.materials_assign_box(m, g, 0, n, 0, n, 0, n, mid.value) == 0
        else:
            assert oracle.materials_set_box(m, g, 0, n, 0, n, 0, n, 4.0, 0.03) == 0
        for _ in range(30):
            oracle.fdtd_update_h(g)
            oracle.fdtd_update_e_materials(g, m)
        out.append({k: v.copy() for k, v in field_arrays(g).items()})
        oracle.materials_destroy(m)
        oracle.fdtd_grid_destroy(g)
    for k in out[0]:
        assert np.allclose(out[0][k], out[1][k], rtol=1e-12, atol=1e-12 * np.max(np.abs(out[1][k])))


def _operator_matrix(oracle, n, tensors):
    """Columns: H increment of one H update after one E update, for each
    H basis vector, on a fully periodic n^3 grid (E and H zero otherwise)."""
    g = oracle.fdtd_grid_create(n, n, n, 1e-3, 0.9)
    oracle.fdtd_grid_set_periodic(g, 7)
    m = oracle.materials_create(g)
    for (c, i, j, k, t6) in tensors:
        arr = (C.c_double * 6)(*t6)
        assert oracle.materials_set_component_tensor3(m, g, c, i, j, k, arr) == 0
    f = field_arrays(g)
    names = ("hx", "hy", "hz")
    dofs = [(q, i, j, k) for q in range(3) for i in range(n) for j in range(n) for k in range(n)]
    gg = g.contents

    def lin(i, j, k):
        return (i * (gg.ny + 1) + j) * (gg.nz + 1) + k

    A = np.zeros((len(dofs), len(dofs)))
    for col, (q, i, j, k) in enumerate(dofs):
        for a in f.values():
            a[:] = 0.0
        f[names[q]][lin(i, j, k)] = 1.0
        oracle.fdtd_update_e_materials(g, m)
        f[names[q]][lin(i, j, k)] = 0.0
        oracle.fdtd_update_h(g)
        A[:, col] = [f[names[qq]][lin(ii, jj, kk)] for (qq, ii, jj, kk) in dofs]
    oracle.materials_destroy(m)
    oracle.fdtd_grid_destroy(g)
    return A, dofs


def _random_tensors(n, seed, frac=0.6):
    rng = np.random.default_rng(seed)
    out = []
    for c in range(3):
        for i in range(n):
            for j in range(n):
                for k in range(n):
                    if rng.uniform() > frac:
                        continue
                    # random SPD inverse permittivity with eigenvalues in (0.1, 1)
                    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
                    M = Q @ np.diag(rng.uniform(0.1, 1.0, 3)) @ Q.T
                    out.append((c, i, j, k, [M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]]))
    return out


def test_tensor3_operator_symmetric(oracle):
    n = 4
    A, _ = _operator_matrix(oracle, n, _random_tensors(n, 11))
    scale = np.max(np.abs(A))
    assert scale > 0
    assert np.max(np.abs(A - A.T)) <= 1e-11 * scale
    # uniform H of each component is curl-free: zero row sums per component block
    m = n ** 3
    for q in range(3):
        assert np.max(np.abs(A[:, q * m:(q + 1) * m].sum(axis=1))) <= 1e-11 * scale
    # negative semi-definite (the H-H block of a lossless Yee operator)
    w = np.linalg.eigvalsh(0.5 * (A + A.T))
    assert w.max() <= 1e-9 * scale


def test_diagonal_tensor3_is_scalar(oracle):
    """A tensor (1/eps) I on every edge equals the scalar eps update."""
    n, eps = 5, 3.5
    res = []
    for use_tensor in (True, False):
        g = oracle.fdtd_grid_create(n, n, n, 1e-3, 0.9)
        oracle.fdtd_grid_set_periodic(g, 7)
        rng = np.random.default_rng(5)
        for a in field_arrays(g).values():
            a[:] = rng.uniform(-1, 1, a.shape)
        m = oracle.materials_create(g)
        if use_tensor:
            t6 = (C.c_double * 6)(1 / eps, 0, 0, 1 / eps, 0, 1 / eps)
            for c in range(3):
                for i in range(n):
                    for j in range(n):
                        for k in range(n):
                            assert oracle.materials_set_component_tensor3(m, g, c, i, j, k, t6) == 0
        else:
            assert oracle.materials_set_box(m, g, 0, n, 0, n, 0, n, eps, 0.0) == 0
        for _ in range(20):
            oracle.fdtd_update_h(g)
            oracle.fdtd_update_e_materials(g, m)
        res.append({k: v.copy() for k, v in field_arrays(g).items()})
        oracle.materials_destroy(m)
        oracle.fdtd_grid_destroy(g)
    for k in res[0]:
        assert np.allclose(res[0][k], res[1][k], rtol=1e-12, atol=1e-13 * np.max(np.abs(res[1][k])))


def test_probe_sampler_refuses_an_impossible_box():
    """The seeded probe sampler raises, rather than looping, when no node of
    its box lies far enough from the source (a 32^3 grid's [10, 22)^3)."""
    from rfsyn_amd.driver import c1_probes
    with pytest.raises(ValueError):
        c1_probes(32, (16, 16, 16), 2, lo=10, hi=22)
    assert len(c1_probes(48, (24, 24, 24), 2, lo=10, hi=38)) == 2


# ------------------------------------------------ dispersive slab R and T
C_LIGHT = 299792458.0


def _eps_model(disp, w):
    """eps(w) of a dispersive material as materials.h defines it (exp(-i w t))."""
    e = disp.eps_inf + 1j * disp.sigma / (w * EPS0)
    for q in range(disp.n_drude):
        e -= disp.drude_wp[q] ** 2 / (w * w + 1j * disp.drude_gamma[q] * w)
    for q in range(disp.n_lorentz):
        w0 = disp.lorentz_w0[q]
        e += disp.lorentz_deps[q] * w0 * w0 / (w0 * w0 - w * w - 2j * disp.lorentz_delta[q] * w)
    return e


def _modes(e1, e2, phase):
    """Forward and backward amplitudes at the first of two probes `phase`
    (radians of the wave number) apart: extended_test.h:202-215."""
    pos = np.exp(1j * phase)
    neg = np.conj(pos)
    return (e2 - e1 * neg) / (pos - neg), (e1 * pos - e2) / (pos - neg)


def _chain_left(eps_nodes, q, kdx, right, i1, i2, i3, i4):
    """Forward and backward amplitudes at node i1 of the 1D Yee chain
    E[i+1] - 2 E[i] + E[i-1] + q eps[i] E[i] = 0 (the grid's steady state
    at one frequency) whose modes at nodes i3, i4 are `right`: the chain
    is run from there to the left, node by node."""
    f, b = right
    E = {i3 + s: f * np.exp(1j * kdx * s) + b * np.exp(-1j * kdx * s) for s in (0, 1)}
    for i in range(i3, i1, -1):
        E[i - 1] = (2.0 - q * eps_nodes[i]) * E[i] - E[i + 1]
    return _modes(E[i1], E[i2], kdx * (i2 - i1))


def _slab_transfer(eps, kd):
    """r and t of a uniform slab (relative permittivity eps, thickness with
    vacuum wave-number product kd) between vacuum half spaces, from the
    characteristic matrix."""
    n = np.sqrt(eps)
    d = n * kd
    m11, m12 = np.cos(d), -1j * np.sin(d) / n
    m21, m22 = -1j * n * np.sin(d), np.cos(d)
    den = m11 + m12 + m21 + m22
    return (m11 + m12 - m21 - m22) / den, 2.0 / den


def _cw_slab(oracle, disp, f_hz, nx=700, npml=40, i_src=60, probes=(120, 136, 450, 466), slab=(250, 290),
             dx=1e-4):
    """validate_materials_high_contrast.c:66-203 on a dispersive slab in a
    line of Ez nodes (nx x 1 x 1, periodic in y and z, x CPML): a soft CW
    source with a sin^2 ramp of 8 periods, then two DFT windows of 8
    periods after the steady state, forward and backward waves separated
    from the probe pairs either side of the slab.  The period is made a
    whole number P of steps (f = 1/(P dt)) so that a window holds whole
    periods and the sum of the e^{-i w t} half of the cosine vanishes.
    Returns the used frequency, dt, and per window the left and right
    (forward, backward) pairs at the first probe of each pair."""
    g = oracle.fdtd_grid_create(nx, 1, 1, dx, 0.99)
    assert oracle.fdtd_grid_set_periodic(g, 6) == 0
    pml = oracle.cpml_create(g, oracle.cpml_default_params(npml))
    m = oracle.materials_create(g)
    mid = C.c_uint8()
    assert oracle.materials_register_dispersive(m, C.byref(disp), C.byref(mid)) == 0
    assert oracle.materials_assign_box(m, g, slab[0], slab[1], 0, 1, 0, 1, mid.value) == 0
    dt = g.contents.dt
    period = int(round(1.0 / (f_hz * dt)))
    f_hz = 1.0 / (period * dt)
    wdt = 2 * np.pi / period
    ramp, window = 8 * period, 8 * period
    settle = ramp + 6000
    ez = field_arrays(g)["ez"]
    sy = 2 * 2   # (ny + 1)(nz + 1): index of node (i, 0, 0) is i * sy
    idx = [p * sy for p in probes]
    sums = np.zeros((2, 4), dtype=complex)
    for n in range(settle + 2 * window):
        oracle.fdtd_update_h_cpml(g, pml)
        oracle.fdtd_update_e_cpml_materials(g, pml, m)
        phase = wdt * ((n + 1) % period)
        env = np.sin(0.5 * np.pi * (n + 1) / ramp) ** 2 if n < ramp else 1.0
        ez[i_src * sy] += env * np.cos(phase)
        if n >= settle:
            sums[(n - settle) // window] += ez[idx] * np.exp(1j * phase) / window
    oracle.materials_destroy(m)
    oracle.cpml_destroy(pml)
    oracle.fdtd_grid_destroy(g)
    kdx = 2 * np.arcsin(np.sin(wdt / 2) / (C_LIGHT * dt / dx))
    out = [(_modes(s[0], s[1], kdx * (probes[1] - probes[0])), _modes(s[2], s[3], kdx * (probes[3] - probes[2])))
           for s in sums]
    return f_hz, dt, kdx, out


def _lorentz_c3():
    d = capi.MaterialDispersive()
    d.eps_inf = 2.25
    d.n_lorentz = 2
    d.lorentz_deps[0], d.lorentz_deps[1] = 1.0, 0.5
    d.lorentz_w0[0], d.lorentz_w0[1] = 2 * np.pi * 60e9, 2 * np.pi * 90e9
    d.lorentz_delta[0], d.lorentz_delta[1] = 0.05 * 2 * np.pi * 60e9, 0.05 * 2 * np.pi * 90e9
    return d


def _drude_lossy():
    d = capi.MaterialDispersive()
    d.eps_inf = 1.0
    d.sigma = 2.0
    d.n_drude = 1
    d.drude_wp[0] = 2 * np.pi * 30e9
    d.drude_gamma[0] = 2 * np.pi * 4e9
    return d


@pytest.mark.parametrize("make,f_hz", [(_lorentz_c3, 40e9), (_lorentz_c3, 75e9), (_drude_lossy, 40e9)])
def test_dispersive_slab_rt_matches_transfer_matrix(oracle, make, f_hz):
    """A 40-cell dispersive slab (the 2-pole Lorentz layer of the periodic
    stack below, between and above its poles; a lossy Drude medium above
    its plasma frequency) in a line of cells, measured the way the
    reference's high-contrast validator measures an interface.  With the
    trapezoidal ADE the grid at angular frequency w is the plain Yee grid
    with permittivity eps(W), W = (2/dt) tan(w dt/2) (the bilinear map),
    at every Ez node of the slab: run from the measured waves right of the
    slab, that chain gives the measured waves left of it to the
    measurement's accuracy (the windows' drift), whatever the CPML
    reflects.  A wrong pole term, sign or time level moves them by far
    more.  The continuous slab of the analytic eps(w) gives |r| and |t|
    to a few percent (the Yee grid's dispersion over 40 cells)."""
    disp = make()
    slab, probes = (250, 290), (120, 136, 450, 466)
    f_hz, dt, kdx, out = _cw_slab(oracle, disp, f_hz, probes=probes, slab=slab)
    (fl, bl), (fr, br) = out[1]
    drift = max(abs(x - y) for x, y in zip(out[0][0] + out[0][1], out[1][0] + out[1][1])) / abs(fl)
    assert drift <= 1e-8, drift   # a steady state
    dx = 1e-4
    w = 2 * np.pi * f_hz
    W = (2 / dt) * np.tan(w * dt / 2)
    q = (2 * dx * np.sin(w * dt / 2) / (C_LIGHT * dt)) ** 2
    eps_nodes = np.ones(probes[3] + 2, dtype=complex)
    eps_nodes[slab[0]:slab[1]] = _eps_model(disp, W)
    def residual(eps):
        f_c, b_c = _chain_left(eps, q, kdx, (fr, br), *probes)
        return max(abs(f_c - fl), abs(b_c - bl)) / abs(fl)

    assert residual(eps_nodes) <= 1e-7
    # the plain eps(w) in place of eps(W) misses by 4e-4 (at 40 GHz, 131
    # steps a period) to 2e-2 (near the 60 GHz pole); a slab one node
    # longer on either side by 5e-2 or more
    wrong = eps_nodes.copy()
    wrong[slab[0]:slab[1]] = _eps_model(disp, w)
    assert residual(wrong) >= 1e-4
    for edge in (slab[0] - 1, slab[1]):
        wrong = eps_nodes.copy()
        wrong[edge] = wrong[slab[0]]
        assert residual(wrong) >= 1e-2
    # the continuous slab of the analytic permittivity
    r_a, t_a = _slab_transfer(_eps_model(disp, w), w / C_LIGHT * dx * (slab[1] - slab[0]))
    assert abs(abs(fr / fl) - abs(t_a)) <= 0.03 * abs(t_a), (abs(fr / fl), abs(t_a))
    assert abs(abs(bl / fl) - abs(r_a)) <= 0.03 * max(abs(r_a), 0.05), (abs(bl / fl), abs(r_a))
