"""Divergence detection on the device (fdtd_sim_fields_finite and the
per-scene flags of fdtd_batch_fields_finite): the device scan agrees with
numpy's isfinite over the downloaded state, for a NaN or an infinity
planted in any one field component, in FP32 and FP64, alone and in a
batch.  The reference's validators gate on the same condition on the host
(isfinite checks, e.g. engine/fdtd/validation_common.h:53)."""
import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import Batch, NativeRun, field_arrays, modgauss_params, random_fill

pytestmark = pytest.mark.gpu


def _scene(lib, precision, plant=None, seed=3):
    """A 20 x 18 x 24 CPML scene from random fields; plant = (field, flat
    index, value) puts a non-finite value into the initial state."""
    def init(g):
        random_fill(g, seed)
        if plant:
            name, idx, val = plant
            field_arrays(g)[name][idx] = val
    run = NativeRun(lib, (20, 18, 24), 1e-3, 0.99, cpml=capi.CpmlParams(4, 3.0, 1.0, 1.0, 0.05, 1e-8),
                    precision=precision, init_fields=init)
    run.add_point_source(capi.FIELD_EZ, (10, 9, 12), "source_waveform_modulated_gaussian",
                         modgauss_params(30e9, 20e9))
    return run


def _host_finite(run):
    return int(all(np.all(np.isfinite(v)) for v in run.fields().values()))


@pytest.mark.parametrize("precision", [capi.FDTD_FP32, capi.FDTD_FP64])
@pytest.mark.parametrize("plant", [None, ("ex", 1234, np.nan), ("hy", 7, np.inf), ("hz", -1, -np.inf)])
def test_fields_finite_matches_host(gpu, precision, plant):
    """hz[-1] is a padding node no update touches."""
    run = _scene(gpu, precision, plant)
    run.run(5)
    got = gpu.fdtd_sim_fields_finite(run.sim)
    assert got == _host_finite(run)
    assert got == (1 if plant is None else 0)
    run.close()


def test_batch_flags_per_scene(gpu):
    runs = [_scene(gpu, capi.FDTD_FP32, plant, seed=s)
            for s, plant in enumerate([None, ("hx", 99, np.nan), None, ("ey", 555, np.inf)])]
    b = Batch(gpu, runs)
    b.run(8)
    flags = b.finite()
    b.close()
    assert flags == [1, 0, 1, 0]
    assert flags == [_host_finite(r) for r in runs]
    for r in runs:
        r.close()
