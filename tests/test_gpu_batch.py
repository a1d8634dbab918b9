"""Batch of independent scenes (fdtd_sim_run_many): every handle ends in
the state its own fdtd_sim_run would give it, bit for bit."""
import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import build_c4_scene, c4_scene, run_many, same_values

pytestmark = pytest.mark.gpu


def _specs(count=4):
    return [c4_scene(q, seed=7, npml=6, n_lo=24, n_hi=44) for q in range(count)]


def _outputs(run, pids, steps):
    return [run.td(p, steps) for p in pids], np.stack([run.dft(p) for p in pids]), run.fields()


@pytest.mark.parametrize("precision", [capi.FDTD_FP32, capi.FDTD_FP64])
def test_run_many_equals_separate_runs(gpu, precision):
    steps = 150   # not a multiple of the DFT block: graph blocks and single steps
    specs = _specs()
    alone = []
    for sp in specs:
        run, pids = build_c4_scene(gpu, sp, precision, steps, npml=6)
        run.run(steps)
        alone.append(_outputs(run, pids, steps))
        run.close()
    runs = [build_c4_scene(gpu, sp, precision, steps, npml=6) for sp in specs]
    run_many(gpu, [r for r, _ in runs], 100)
    run_many(gpu, [r for r, _ in runs], steps - 100)
    for (run, pids), (td, dft, f) in zip(runs, alone):
        td2, dft2, f2 = _outputs(run, pids, steps)
        assert all(same_values(a, b) for a, b in zip(td, td2))
        assert same_values(dft, dft2)
        for k in f:
            assert same_values(f[k], f2[k]), k
        run.close()


def test_run_many_rejects_repeated_handle(gpu):
    run, _ = build_c4_scene(gpu, _specs(1)[0], capi.FDTD_FP32, 10, npml=6)
    import ctypes as C
    arr = (C.c_void_p * 2)(run.sim, run.sim)
    assert gpu.fdtd_sim_run_many(arr, 2, 5) == -1
    run.close()
