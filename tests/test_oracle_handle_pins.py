"""The oracle's device-handle API (fdtd_sim_* in oracle/oracle_sim.c) and
its 3x3 tensor path, pinned against the rfsyn reference compiled from its
own sources (oracle/_ref), bit for bit.

The handle API has no counterpart in the reference, which drives a run
through its low-level calls in a host loop; each test here runs that loop
on the reference library the way the reference's own callers do
(validate_probe_dft.c:66-82 for point-probe DFTs, validation_common.h:16-18
and 41-55 for the after-E plane injection) and compares it with the
oracle's handle on the same inputs.  The GPU library is checked against
the oracle elsewhere, so these pins carry through to it.  Skipped where
the reference sources are absent."""
import ctypes as C

import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import NativeRun, field_arrays, random_fill
from rfsyn_amd.driver import same_values


def _gauss(lib, g, t0=30.0, tau=7.0):
    p = capi.GaussDerivParams(t0 * g.contents.dt, tau * g.contents.dt)
    return p, capi.waveform_ptr(lib, "source_waveform_gaussian_derivative")


def _ref_loop(ref, n, steps, npml, periodic=0, src=None, sheet=None, probes=(), init_seed=None):
    """The reference's production step (simulate.c:48-54) on a vacuum grid
    with CPML, driven from the host: H update, point source at n dt, E
    update, then the after-E sheet injection at (n+1) dt on the plane
    x = sheet (validation_common.h:16-18) and every probe_dft_record at
    (n+1) dt (validate_probe_dft.c:77-80).  probes: (component, i, j, k,
    freqs); returns the fields, the DFT accumulators and the Ez series of
    the first probe."""
    g = ref.fdtd_grid_create(n[0], n[1], n[2], 1e-3, 0.99)
    assert g and ref.fdtd_grid_set_periodic(g, periodic) == 0
    if init_seed is not None:
        random_fill(g, init_seed)
    pml = ref.cpml_create(g, ref.cpml_default_params(npml))
    mat = ref.materials_create(g)
    assert pml and mat
    keep = []
    source = None
    if src is not None:
        p, wf = _gauss(ref, g)
        keep.append(p)
        source = ref.source_make_point(src[1], src[2], src[3], src[0], capi.SRC_SOFT, 1.0, wf,
                                       C.cast(C.pointer(p), C.c_void_p))
    sp = capi.GaussDerivParams(60 * g.contents.dt, 16 * g.contents.dt)
    dprobes = []
    for comp, i, j, k, freqs in probes:
        f = (C.c_double * len(freqs))(*freqs)
        keep.append(f)
        pr = capi.ProbeDFT()
        assert ref.probe_dft_create(C.byref(pr), i, j, k, comp, f, len(freqs)) == 0
        dprobes.append(pr)
    dt = g.contents.dt
    shape = (n[0] + 1, n[1] + 1, n[2] + 1)
    grid = {k: v.reshape(shape) for k, v in field_arrays(g).items()}
    series = []
    for s in range(steps):
        ref.fdtd_update_h_cpml(g, pml)
        if source is not None:
            ref.source_apply(C.byref(source), g, s * dt)
        ref.fdtd_update_e_cpml_materials(g, pml, mat)
        if sheet is not None:
            # pulse() of validation_common.h:11, -2u exp(-u^2): the
            # reference's Gaussian-derivative waveform, evaluated by it
            v = ref.source_waveform_gaussian_derivative((s + 1) * dt, C.cast(C.pointer(sp), C.c_void_p))
            grid["ez"][sheet, :n[1], :n[2]] += v
        for pr in dprobes:
            ref.probe_dft_record(C.byref(pr), g, (s + 1) * dt, dt)
        if probes:
            c0, i0, j0, k0, _ = probes[0]
            series.append(grid[("ex", "ey", "ez")[c0]][i0, j0, k0])
    acc = [np.array([pr.acc_re[q] + 1j * pr.acc_im[q] for q in range(pr.n_freqs)]) for pr in dprobes]
    fields = {k: v.copy() for k, v in field_arrays(g).items()}
    for pr in dprobes:
        ref.probe_dft_destroy(C.byref(pr))
    ref.materials_destroy(mat)
    ref.cpml_destroy(pml)
    ref.fdtd_grid_destroy(g)
    return fields, acc, np.array(series)


def test_handle_point_probe_dft_matches_reference_loop(oracle, ref):
    """Point sources at n dt between the updates, probe series and running
    DFTs at (n+1) dt after them, on a 24^3 CPML grid: the oracle's handle
    gives the reference loop's fields, series and accumulators bit for
    bit (plain sums, the reference's order)."""
    n, steps, npml = (24, 24, 24), 120, 5
    freqs = [10e9, 15e9, 27e9]
    probes = [(capi.FIELD_EZ, 16, 12, 12, freqs), (capi.FIELD_EX, 9, 14, 7, freqs), (capi.FIELD_EY, 12, 5, 18, freqs)]
    f_ref, acc_ref, ser_ref = _ref_loop(ref, n, steps, npml, src=(capi.FIELD_EZ, 12, 12, 12), probes=probes)
    run = NativeRun(oracle, n, 1e-3, 0.99, cpml=oracle.cpml_default_params(npml))
    p, _ = _gauss(oracle, run.g)
    run.add_point_source(capi.FIELD_EZ, (12, 12, 12), "source_waveform_gaussian_derivative", p)
    pids = [run.add_probe(c, (i, j, k), steps, fr) for c, i, j, k, fr in probes]
    run.run(steps)
    assert np.max(np.abs(ser_ref)) > 0
    assert same_values(run.td(pids[0], steps), ser_ref)
    for pid, a in zip(pids, acc_ref):
        assert np.max(np.abs(a)) > 0
        assert same_values(run.dft(pid), a)
    f = run.fields()
    run.close()
    for k in f_ref:
        assert same_values(f[k], f_ref[k]), k


def test_handle_after_e_plane_source_matches_sheet(oracle, ref):
    """The after-E plane source (fdtd_sim_add_plane_source with after_e)
    against the validators' injection: the pulse of validation_common.h
    added by the host to every Ez node of the plane x = 30 (j < ny, k <
    nz) after the E update at (n+1) dt, on the validators' periodic
    y-z plane-wave grid with x CPML; fields bit for bit."""
    n, steps, npml = (160, 4, 4), 300, 12
    f_ref, _, _ = _ref_loop(ref, n, steps, npml, periodic=6, sheet=30)
    run = NativeRun(oracle, n, 1e-3, 0.99, periodic=6, cpml=oracle.cpml_default_params(npml))
    p = capi.GaussDerivParams(60 * run.g.contents.dt, 16 * run.g.contents.dt)
    run.add_plane_source(capi.FIELD_EZ, 0, 30, "source_waveform_gaussian_derivative", p, amp=1.0, after_e=1)
    run.run(steps)
    f = run.fields()
    run.close()
    assert np.max(np.abs(f_ref["ez"])) > 0
    for k in f_ref:
        assert same_values(f[k], f_ref[k]), k


@pytest.mark.parametrize("axis,index", [(0, 5), (1, 7), (2, 9)])
def test_plane_monitor_matches_reference_point_dfts(oracle, ref, axis, index):
    """A full-plane DFT monitor of the oracle's handle, every cell of both
    tangential components in its [f][a][b] layout (a, b the other two axes
    in increasing order), against the reference's probe_dft_record at each
    of those nodes: the same accumulators bit for bit.  Random initial
    fields and a point source on a 12 x 14 x 16 CPML grid."""
    n, steps, npml = (12, 14, 16), 60, 3
    freqs = [18e9, 31e9, 44e9]
    others = [a for a in range(3) if a != axis]
    probes = []
    for w, comp in enumerate(others):   # the tangential E components are the other two axes
        for ia in range(n[others[0]]):
            for ib in range(n[others[1]]):
                x = [0, 0, 0]
                x[axis], x[others[0]], x[others[1]] = index, ia, ib
                probes.append((comp, x[0], x[1], x[2], freqs))
    _, acc_ref, _ = _ref_loop(ref, n, steps, npml, src=(capi.FIELD_EY, 6, 6, 8), probes=probes, init_seed=9)
    run = NativeRun(oracle, n, 1e-3, 0.99, cpml=oracle.cpml_default_params(npml), init_fields=lambda g: random_fill(g, 9))
    p, _ = _gauss(oracle, run.g)
    run.add_point_source(capi.FIELD_EY, (6, 6, 8), "source_waveform_gaussian_derivative", p)
    mid = run.add_plane_monitor(axis, index, freqs)
    run.run(steps)
    shape = (len(freqs), n[others[0]], n[others[1]])
    got = [run.plane_dft(mid, w, shape) for w in range(2)]
    run.close()
    q = 0
    for w in range(2):
        for ia in range(shape[1]):
            for ib in range(shape[2]):
                assert same_values(got[w][:, ia, ib], acc_ref[q]), (w, ia, ib)
                q += 1
    assert np.max(np.abs(got[0])) > 0 and np.max(np.abs(got[1])) > 0


def test_tensor3_reduces_to_reference_2x2(oracle, ref):
    """3x3 tensors whose xz and yz terms vanish, on a z-invariant grid (one
    periodic z cell: every z difference is zero), against the reference's
    2x2 tensors (materials_set_component_tensor, component-local (m11,
    m12, m22): (xx, xy, yy) for Ex, (yy, xy, xx) for Ey) on the same edges:
    fields bit for bit after 25 steps from random fields."""
    n = (20, 20, 1)
    runs = []
    for lib, three in ((ref, False), (oracle, True)):
        g = lib.fdtd_grid_create(n[0], n[1], n[2], 1e-3, 0.9)
        assert lib.fdtd_grid_set_periodic(g, 4) == 0
        random_fill(g, 13)
        mat = lib.materials_create(g)
        edges = 0
        for i in range(7, 13):
            for j in range(7, 13):
                for c in (0, 1):
                    if (i + 3 * j + c) % 4 == 0:
                        continue
                    xx = 0.30 + 0.04 * ((i + j) % 5)
                    yy = 0.45 + 0.03 * ((i * j + c) % 4)
                    xy = 0.06 * (((3 * i + j + c) % 5) - 2)
                    m11, m12, m22 = (xx, xy, yy) if c == 0 else (yy, xy, xx)
                    if three:
                        m = np.array([xx, xy, 0.0, yy, 0.0, 0.5], dtype=np.float64)
                        rc = lib.materials_set_component_tensor3(mat, g, c, i, j, 0, m.ctypes.data_as(C.POINTER(C.c_double)))
                    else:
                        rc = lib.materials_set_component_tensor(mat, g, c, i, j, 0, m11, m12, m22)
                    assert rc == 0
                    edges += 1
        for _ in range(25):
            lib.fdtd_update_h(g)
            lib.fdtd_update_e_materials(g, mat)
        runs.append({k: v.copy() for k, v in field_arrays(g).items()})
        lib.materials_destroy(mat)
        lib.fdtd_grid_destroy(g)
    assert edges > 40
    for k in runs[0]:
        assert same_values(runs[0][k], runs[1][k]), k
