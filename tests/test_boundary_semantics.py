"""Error semantics of the drop-in boundary on the product library itself,
against the reference compiled from its own sources (oracle/Makefile `ref`)
and the oracle.  CPU tests: every call here returns before any device work.

- fdtd_simulate_echo / _with_setup: the -1 returns of
  engine/fdtd/simulate.c:11-35 (NULL cfg or output, zero steps, grid
  failure, source or probe past n, CPML failure, failing setup callback)
  and the callback's arguments;
- the materials API (engine/fdtd/materials.c:50-120, 270-310, 388-799):
  a seeded script of edits, valid and invalid (out-of-range nodes, H
  components, eps_r < 1, sigma < 0, non-finite values, non-SPD tensors,
  tensors without their x/y margin, a materials object of another grid,
  the 256th material id), every return code and then every coefficient and
  tensor of every E edge compared bit for bit.
"""
import ctypes as C

import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import same_values


@pytest.fixture(scope="module")
def product():
    import rfsyn_amd
    return rfsyn_amd.load_library()


SETUP_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_void_p)


def _cfg(**kw):
    base = dict(nx=20, ny=20, nz=20, dx=1e-3, courant_safety=0.99, n_steps=30, npml=4, src_i=10, src_j=10,
                src_k=10, src_component=capi.FIELD_EZ, t0_over_dt=15.0, tau_over_dt=4.0, probe_i=13,
                probe_j=10, probe_k=10, probe_component=capi.FIELD_EZ)
    base.update(kw)
    return capi.SimConfig(*[base[f] for f, _ in capi.SimConfig._fields_])


# (name, config overrides or None for a NULL cfg, NULL output buffer)
ECHO_FAILURES = [
    ("null_cfg", None, False),
    ("null_out", {}, True),
    ("zero_steps", {"n_steps": 0}, False),
    ("grid_zero_nx", {"nx": 0}, False),
    ("grid_bad_dx", {"dx": 0.0}, False),
    ("src_i_past_n", {"src_i": 21}, False),
    ("src_k_past_n", {"src_k": 21}, False),
    ("probe_j_past_n", {"probe_j": 21}, False),
    ("probe_k_past_n", {"probe_k": 99}, False),
    ("npml_too_large", {"npml": 11}, False),
]


def _echo_rc(lib, over, null_out):
    buf = (C.c_double * 64)()
    fs = C.c_double(-7.0)
    cfg = None if over is None else C.byref(_cfg(**over))
    rc = lib.fdtd_simulate_echo(cfg, None if null_out else buf, C.byref(fs))
    return rc, fs.value


@pytest.mark.parametrize("name,over,null_out", ECHO_FAILURES, ids=[c[0] for c in ECHO_FAILURES])
def test_echo_failures_return_minus_one(product, oracle, ref, name, over, null_out):
    got = [_echo_rc(lib, over, null_out) for lib in (product, oracle, ref)]
    # simulate.c: -1, and the sample rate is left alone
    assert got[2] == (-1, -7.0), got[2]
    assert got[0] == got[2], (name, got)
    assert got[1] == got[2], (name, got)


def test_echo_boundary_indices_accepted_by_reference(ref):
    """Source and probe at index n are inside the padded arrays
    (simulate.c:17-21 rejects only > n); the reference runs them."""
    buf = (C.c_double * 30)()
    cfg = _cfg(src_i=20, probe_k=20)
    assert ref.fdtd_simulate_echo(C.byref(cfg), buf, None) == 0


@pytest.mark.parametrize("ret", [1, -1, 7])
def test_failing_setup_callback(product, oracle, ref, ret):
    """A setup callback returning nonzero makes the call return -1; it is
    called once with the created grid and materials (simulate.c:30-35)."""
    out = []
    for lib in (product, oracle, ref):
        seen = []

        def cb(g, m, ctx):
            grid = C.cast(g, C.POINTER(capi.Grid)).contents
            seen.append((grid.nx, grid.ny, grid.nz, grid.dx, grid.dt, bool(m), ctx))
            return ret

        fn = SETUP_FN(cb)
        buf = (C.c_double * 30)()
        rc = lib.fdtd_simulate_echo_with_setup(C.byref(_cfg()), buf, None, C.cast(fn, C.c_void_p),
                                               C.c_void_p(1234))
        out.append((rc, seen))
    assert out[2][0] == -1 and len(out[2][1]) == 1
    assert out[2][1][0][:3] == (20, 20, 20) and out[2][1][0][5] and out[2][1][0][6] == 1234
    assert out[0] == out[2] and out[1] == out[2]


def test_setup_not_called_after_bounds_failure(product, ref):
    """The bounds and CPML checks come before the callback."""
    for lib in (product, ref):
        calls = []
        fn = SETUP_FN(lambda g, m, ctx: calls.append(1) or 0)
        for over in ({"src_j": 21}, {"npml": 11}):
            buf = (C.c_double * 30)()
            assert lib.fdtd_simulate_echo_with_setup(C.byref(_cfg(**over)), buf, None, C.cast(fn, C.c_void_p),
                                                     None) == -1
        assert calls == []


# ---------------------------------------------------------------- materials

def _materials_script(lib, seed, periodic):
    """A seeded sequence of materials edits on a 9x8x7 grid, valid and
    invalid; returns every return code and the final coefficients."""
    rng = np.random.default_rng(seed)
    n = (9, 8, 7)
    g = lib.fdtd_grid_create(n[0], n[1], n[2], 1e-3, 0.99)
    assert g
    assert lib.fdtd_grid_set_periodic(g, periodic) == 0
    other = lib.fdtd_grid_create(n[0] + 1, n[1], n[2], 1e-3, 0.99)   # a grid of another size
    m = lib.materials_create(g)
    assert m
    rcs = []
    mid = C.c_uint8()
    ids = []
    for _ in range(6):
        eps = float(rng.choice([0.5, 1.0, 2.5, 13.0, 5.0, np.nan, np.inf]))
        sig = float(rng.choice([-0.1, 0.0, 0.03, 1e3, 0.0, np.nan]))
        rc = lib.materials_register(m, eps, sig, C.byref(mid))
        rcs.append(rc)
        if rc == 0:
            ids.append(mid.value)
            rcs.append(mid.value)
    rcs.append(lib.materials_register(None, 2.0, 0.0, C.byref(mid)))
    rcs.append(lib.materials_register(m, 2.0, 0.0, None))
    if not ids:
        ids.append(0)

    def box():
        # mostly ordered boxes inside the padded grid, some empty, reversed
        # or past the end
        out = []
        for d in n:
            a, b = sorted(int(x) for x in rng.integers(0, d + 2, size=2))
            if rng.random() < 0.1:
                a, b = b, a
            if rng.random() < 0.05:
                b = d + 2
            out += [a, b]
        return tuple(out)

    def node(c):
        # mostly inside the component's range, some past it
        hi = 3 if rng.random() < 0.1 else 1
        return [int(rng.integers(0, d + hi)) for d in n]

    for _ in range(200):
        op = int(rng.integers(0, 7))
        c = int(rng.integers(0, 6)) if rng.random() < 0.1 else int(rng.integers(0, 3))   # H invalid
        gg = other if rng.random() < 0.05 else g
        if op == 0:
            b = box()
            rcs.append(lib.materials_assign_box(m, gg, *b, int(rng.choice(ids + [250]))))
        elif op == 1:
            b = box()
            eps = float(rng.choice([0.9, 1.0, 3.0, 8.5, 2.0]))
            sig = float(rng.choice([-1.0, 0.0, 0.02, 0.0]))
            rcs.append(lib.materials_set_box(m, gg, *b, eps, sig))
        elif op == 2:
            i, j, k = node(c)
            eps = float(rng.choice([0.9, 1.0, 4.0, 11.0, 6.0, np.inf]))
            sig = float(rng.choice([-1.0, 0.0, 0.05, 0.0]))
            rcs.append(lib.materials_set_component(m, gg, c, i, j, k, eps, sig))
        elif op == 3:
            i, j, k = node(c)
            rcs.append(lib.materials_clear_component(m, gg, c, i, j, k))
        elif op == 4:
            i, j, k = node(c)
            # mostly SPD; some singular, indefinite or non-finite
            a = float(rng.choice([0.3, 0.8, 1.0, 0.5, -0.2, 0.0]))
            cc = float(rng.choice([0.25, 0.9, 0.6, 0.4, 0.0, np.nan]))
            b = float(rng.choice([0.0, 0.1, -0.2, 0.05, 0.6]))
            rcs.append(lib.materials_set_component_tensor(m, gg, c, i, j, k, a, b, cc))
        elif op == 5:
            i, j, k = node(c)
            rcs.append(lib.materials_clear_component_tensor(m, gg, c, i, j, k))
        else:
            i, j, k = node(c)
            co = capi.MaterialCoeff()
            rcs.append(lib.materials_get_coefficients(m, gg, c, i, j, k, C.byref(co)))
            t = capi.MaterialTensor()
            rcs.append(lib.materials_get_tensor(m, gg, c, i, j, k, C.byref(t)))
    rcs.append(lib.materials_get_coefficients(m, g, 0, 0, 0, 0, None))
    rcs.append(lib.materials_get_tensor(m, g, 0, 0, 0, 0, None))
    rcs.append(lib.materials_set_box(None, g, 0, 1, 0, 1, 0, 1, 2.0, 0.0))
    vals = []
    for c in range(3):
        for i in range(n[0] + 1):
            for j in range(n[1] + 1):
                for k in range(n[2] + 1):
                    co = capi.MaterialCoeff(-1.0, -1.0)
                    rc = lib.materials_get_coefficients(m, g, c, i, j, k, C.byref(co))
                    t = capi.MaterialTensor(-1.0, -1.0, -1.0, -1.0)
                    rt = lib.materials_get_tensor(m, g, c, i, j, k, C.byref(t))
                    vals += [rc, co.Ca, co.Cb, rt, t.m11, t.m12, t.m22, t.scale]
    lib.materials_destroy(m)
    lib.fdtd_grid_destroy(other)
    lib.fdtd_grid_destroy(g)
    return np.array(rcs, dtype=np.float64), np.array(vals, dtype=np.float64)


@pytest.mark.parametrize("seed", range(12))
def test_materials_script_bitwise(product, oracle, ref, seed):
    periodic = [0, 1, 2, 3, 4, 7][seed % 6]
    rp, vp = _materials_script(product, seed, periodic)
    rr, vr = _materials_script(ref, seed, periodic)
    ro, vo = _materials_script(oracle, seed, periodic)
    assert (rr == -1).any() and (rr == 0).any()
    assert np.array_equal(rp, rr), np.nonzero(rp != rr)
    assert np.array_equal(ro, rr), np.nonzero(ro != rr)
    assert same_values(vp, vr)
    assert same_values(vo, vr)


def test_material_table_full(product, ref):
    """256 ids: registering past the table fails with -1 (materials.c:9)."""
    res = []
    for lib in (product, ref):
        g = lib.fdtd_grid_create(4, 4, 4, 1e-3, 0.99)
        m = lib.materials_create(g)
        mid = C.c_uint8()
        rcs = [lib.materials_register(m, 1.0 + q * 0.01, 0.0, C.byref(mid)) for q in range(260)]
        res.append((rcs, mid.value))
        lib.materials_destroy(m)
        lib.fdtd_grid_destroy(g)
    assert res[0] == res[1]
    assert -1 in res[1][0]


def test_update_gpu_failure_recorded_not_aborting(product, monkeypatch):
    """fdtd_update_* return nothing, so a GPU failure inside them aborts by
    default; with RFSYN_GPU_ERRORS=record it is recorded instead and the
    call returns without touching the host fields.  Without a visible GPU
    (this suite's usual host) the upload fails and must be recorded; with
    one the update runs and records nothing."""
    monkeypatch.setenv("RFSYN_GPU_ERRORS", "record")
    g = product.fdtd_grid_create(6, 5, 4, 1e-3, 0.99)
    assert g
    gg = g.contents
    idx = (2 * (gg.ny + 1) + 2) * (gg.nz + 1) + 2
    gg.ez[idx] = 1.0
    before = product.fdtd_gpu_errors()
    product.fdtd_update_h(g)
    product.fdtd_update_e(g)
    after = product.fdtd_gpu_errors()
    if after > before:   # no device: both calls recorded their failure
        assert after == before + 2
        assert b"failed on the GPU" in product.fdtd_last_error()
        assert gg.ez[idx] == 1.0 and gg.hx[idx] == 0.0
    else:   # a device: the step ran (the curl of the lone Ez moved H)
        assert any(gg.hx[q] != 0.0 or gg.hy[q] != 0.0 for q in range(6 * 7 * 5))
    product.fdtd_grid_destroy(g)
