"""Row mode of the FP32 fused step (RFSYN_LEAP_ROWS=1): a wavefront owns
one row of every z column, in one to four 64-lane chunks, and takes the z
neighbours of a chunk's edge lanes from the next or previous chunk.  The
arithmetic is that of the tiled step, so every field, probe series and
plane-monitor DFT must equal the tiled run's bit for bit (FP32; the tiled
FP32 step is itself checked bitwise against the two-kernel step)."""
import ctypes as C

import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import NativeRun, modgauss_params, random_fill, same_values

pytestmark = pytest.mark.gpu

FREQS = np.linspace(30e9, 50e9, 8)


def _stack_setup(nz):
    """Lorentz layer, Drude stripes and a plain dielectric, placed along z
    as in the periodic dispersive stack, at depths scaled to nz."""
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
        nx, ny = g.contents.nx, g.contents.ny
        z0 = nz // 2
        assert lib.materials_assign_box(mat, g, 0, nx, 0, ny, z0, z0 + 19, lid.value) == 0
        for i0 in range(0, nx, 8):
            assert lib.materials_assign_box(mat, g, i0, i0 + 4, 0, ny, z0 - 15, z0, did.value) == 0
        assert lib.materials_set_box(mat, g, 0, nx, 0, ny, z0 + 25, z0 + 45, 4.0, 0.0) == 0
    return setup


def _closed_setup(n):
    """Scalar blocks crossing the x/y CPML shell, one lossy, one reaching
    the upper z slab."""
    def setup(lib, g, mat):
        assert lib.materials_set_box(mat, g, 0, n[0] // 2, 1, n[1] - 2, 2, n[2] // 3, 3.0, 0.0) == 0
        assert lib.materials_set_box(mat, g, n[0] // 3, n[0] - 1, n[1] // 3, n[1] - 1, n[2] // 2, n[2] - 3,
                                     2.0, 0.02) == 0
    return setup


def _run(lib, case, rows, steps=40):
    p = modgauss_params(40e9, 30e9)
    kind, _, znum = case.partition("_nz")
    nz = int(znum) if znum else None
    if kind == "stack":
        n = (16, 12, nz or 256)
        run = NativeRun(lib, n, 0.5e-3, 0.99, periodic=3, cpml=capi.CpmlParams(16, 3.0, 1.0, 1.0, 0.05, 1e-8),
                        precision=capi.FDTD_FP32, material_setup=_stack_setup(n[2]),
                        init_fields=lambda g: random_fill(g, 5))
        run.add_plane_source(capi.FIELD_EX, 2, 24, "source_waveform_modulated_gaussian", p)
        mons = [(run.add_plane_monitor(2, k, FREQS), (n[0], n[1], len(FREQS))) for k in (20, n[2] - 22)]
        pid = run.add_probe(capi.FIELD_EX, (3, 5, n[2] // 2 + 8), steps, None)
    elif kind == "closed":
        n = (20, 18, nz or 256)
        run = NativeRun(lib, n, 1e-3, 0.99, cpml=capi.CpmlParams(6, 3.0, 1.0, 2.5, 0.05, 1e-8),
                        precision=capi.FDTD_FP32, material_setup=_closed_setup(n),
                        init_fields=lambda g: random_fill(g, 7))
        run.add_point_source(capi.FIELD_EZ, (10, 9, n[2] // 2), "source_waveform_modulated_gaussian", p)
        run.add_point_source(capi.FIELD_EY, (n[0] - 2, 2, n[2] - 2), "source_waveform_modulated_gaussian", p,
                             hard=True)
        run.add_plane_source(capi.FIELD_EZ, 0, n[0] // 2, "source_waveform_modulated_gaussian", p, amp=0.2,
                             after_e=1)
        mons = [(run.add_plane_monitor(1, 7, FREQS), (n[0], n[2] + 1, len(FREQS)))]
        pid = run.add_probe(capi.FIELD_EZ, (11, 9, n[2] // 2 + 1), steps, None)
    else:   # vacuum, x periodic, CPML on y and z
        n = (14, 22, nz or 240)
        run = NativeRun(lib, n, 1e-3, 0.99, periodic=1, cpml=capi.CpmlParams(10, 3.0, 1.0, 1.0, 0.05, 1e-8),
                        precision=capi.FDTD_FP32, init_fields=lambda g: random_fill(g, 9))
        run.add_point_source(capi.FIELD_EX, (7, 11, n[2] // 4), "source_waveform_modulated_gaussian", p)
        mons = []
        pid = run.add_probe(capi.FIELD_EX, (7, 12, n[2] // 4 + 2), steps, None)
    run.run(steps)
    out = {"rows": run.lib.fdtd_sim_leap_rows(run.sim), "td": run.td(pid, steps), "fields": run.fields(),
           "mons": [run.plane_dft(m, w, shape) for m, shape in mons for w in range(2)]}
    run.close()
    return out


@pytest.mark.parametrize("chunk", ["0", "5"])
@pytest.mark.parametrize("case", ["stack", "stack_nz230", "closed", "vacuum",
                                  # one to three 64-column chunks, with and
                                  # without column nz past the last chunk
                                  "vacuum_nz48", "vacuum_nz64", "closed_nz100", "stack_nz120",
                                  "closed_nz192"])
def test_row_mode_bitwise_tiled(gpu, case, chunk, monkeypatch):
    monkeypatch.setenv("RFSYN_LEAP_CHUNK", chunk)
    monkeypatch.setenv("RFSYN_LEAP_ROWS", "4")
    a = _run(gpu, case, 4)
    monkeypatch.setenv("RFSYN_LEAP_ROWS", "1")
    b = _run(gpu, case, 1)
    assert a["rows"] == 4 and b["rows"] == 1
    assert np.max(np.abs(a["td"])) > 0
    assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), k
    for x, y in zip(a["mons"], b["mons"]):
        assert np.max(np.abs(x)) > 0
        assert same_values(x.real, y.real) and same_values(x.imag, y.imag)


def test_row_mode_refused_where_unsupported(gpu, monkeypatch):
    """z periodic (no z CPML in the box): row mode is refused with an error
    naming the tiling, and the default tiling still runs."""
    monkeypatch.setenv("RFSYN_LEAP_ROWS", "1")
    run = NativeRun(gpu, (12, 12, 256), 1e-3, 0.99, periodic=7, precision=capi.FDTD_FP32)
    assert gpu.fdtd_sim_run(run.sim, 2) == -1
    assert b"tiling" in gpu.fdtd_last_error()
    run.close()
    monkeypatch.setenv("RFSYN_LEAP_ROWS", "0")
    run = NativeRun(gpu, (12, 12, 256), 1e-3, 0.99, periodic=7, precision=capi.FDTD_FP32)
    run.run(2)
    assert gpu.fdtd_sim_leap_rows(run.sim) == 4
    run.close()
