"""Tiling autotuner: the chosen tiling changes which wavefront updates a
node, never the arithmetic, and the state it touches while timing is
restored, so a tuned run equals an untuned one bit for bit."""
import ctypes as C

import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import NativeRun, modgauss_params, random_fill, same_values
from test_gpu_parity import _ade_region_run, _ade_region_setup, _dispersive_setup, _region_run

pytestmark = pytest.mark.gpu


def _stack(lib, precision, tune):
    run = NativeRun(lib, (16, 20, 44), 0.5e-3, 0.99, periodic=3, cpml=capi.CpmlParams(8, 3.0, 1.0, 1.0, 0.05, 1e-8),
                    precision=precision, material_setup=_dispersive_setup, init_fields=lambda g: random_fill(g, 2))
    run.add_plane_source(capi.FIELD_EX, 2, 10, "source_waveform_modulated_gaussian", modgauss_params(40e9, 30e9))
    pid = run.add_probe(capi.FIELD_EX, (3, 5, 24), 90, None)
    chosen = (C.c_int * 4)(-1, -1, -1, -1)
    if tune:
        assert run.lib.fdtd_sim_autotune(run.sim, 2, chosen) == 0, run.lib.fdtd_last_error()
    run.run(90)
    out = {"td": run.td(pid, 90), "fields": run.fields()}
    run.close()
    return out, list(chosen)


def _closed(lib, precision, tune):
    n = (22, 20, 36)
    run = NativeRun(lib, n, 0.5e-3, 0.99, cpml=capi.CpmlParams(5, 3.0, 1.0, 2.0, 0.05, 1e-8), precision=precision,
                    material_setup=_ade_region_setup(n), init_fields=lambda g: random_fill(g, 6))
    run.add_point_source(capi.FIELD_EY, (11, 10, 4), "source_waveform_modulated_gaussian", modgauss_params(40e9, 30e9))
    if tune:
        assert run.lib.fdtd_sim_autotune(run.sim, 1, None) == 0
        two = C.c_int(-1)
        assert run.lib.fdtd_sim_select_path(run.sim, 2, C.byref(two)) == 0
        assert two.value in (0, 1)
    run.run(60)
    out = {"fields": run.fields()}
    run.close()
    return out, None


@pytest.mark.parametrize("precision", [capi.FDTD_FP32, capi.FDTD_FP64])
@pytest.mark.parametrize("scene", [_stack, _closed])
def test_tuned_run_is_bitwise_untuned(gpu, precision, scene):
    a, _ = scene(gpu, precision, False)
    b, chosen = scene(gpu, precision, True)
    if chosen is not None:
        assert chosen[0] in ((1, 4, 8) if precision == capi.FDTD_FP32 else (2, 4))
    if "td" in a:
        assert np.max(np.abs(a["td"])) > 0
        assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), k


def test_autotune_after_first_step_is_refused(gpu):
    run = NativeRun(gpu, (12, 12, 12), 1e-3, 0.99, periodic=7)
    run.run(1)
    assert gpu.fdtd_sim_autotune(run.sim, 1, None) == -1
    run.close()


@pytest.mark.parametrize("coop", [1, 3, 4, 20, 35])
@pytest.mark.parametrize("case", ["wide_z", "x_periodic", "z_periodic", "all_shell"])
def test_cooperative_blocks_fp64(gpu, oracle, case, coop, monkeypatch):
    """Cooperative blocks (coop & 15 neighbouring row tiles of one z tile,
    planes in lockstep or within coop >> 4 planes), with short chunks so
    that the last chunk is shorter and some wavefronts of a block have no
    tile: bitwise against the oracle."""
    monkeypatch.setenv("RFSYN_LEAP_COOP", str(coop))
    monkeypatch.setenv("RFSYN_LEAP_CHUNK", "3")
    a = _region_run(gpu, case)
    b = _region_run(oracle, case)
    assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), k


@pytest.mark.parametrize("coop", [4, 18])
def test_cooperative_blocks_ade_fp64(gpu, oracle, coop, monkeypatch):
    """Dispersive blocks across the CPML shell with cooperative blocks."""
    monkeypatch.setenv("RFSYN_LEAP_COOP", str(coop))
    a = _ade_region_run(gpu, False)
    b = _ade_region_run(oracle, False)
    assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), k


@pytest.mark.parametrize("coop", [2, 4, 20])
def test_cooperative_blocks_fp32_bitwise_independent(gpu, coop, monkeypatch):
    """FP32 periodic dispersive stack: cooperative blocks give the bits of
    independent wavefronts."""
    a, _ = _stack(gpu, capi.FDTD_FP32, False)
    monkeypatch.setenv("RFSYN_LEAP_COOP", str(coop))
    b, _ = _stack(gpu, capi.FDTD_FP32, False)
    assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), k


@pytest.mark.parametrize("precision", [capi.FDTD_FP32, capi.FDTD_FP64])
@pytest.mark.parametrize("scene", [_stack, _closed])
def test_two_kernel_step_bitwise_fused(gpu, precision, scene, monkeypatch):
    """The autotuner may keep the two-kernel step instead of the fused one:
    both give the same bits, in FP32 as in FP64."""
    monkeypatch.setenv("RFSYN_FUSED", "1")
    a, _ = scene(gpu, precision, False)
    monkeypatch.setenv("RFSYN_FUSED", "0")
    b, _ = scene(gpu, precision, False)
    if "td" in a:
        assert same_values(a["td"], b["td"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), k
