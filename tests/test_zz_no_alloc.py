"""The output stage's entry point that allocates no device memory
(fdtd_sim_probe_iq_noalloc) against the allocating one
(fdtd_sim_probe_iq), through the C ABI: the same bits, with and without
noise, on two calls in a row (the second shorter, so the shared array
holds the first call's longer result), and the size limit refused."""
import ctypes as C

import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import NativeRun, modgauss_params

pytestmark = pytest.mark.gpu

MAX = 65536   # FDTD_IQ_NOALLOC_MAX


def _iq(fn, run, n, noise, stream):
    out = (C.c_float * (2 * n))()
    assert fn(run.sim, 0, n, noise, C.c_float(12.0), 42, stream, out) == 0, run.lib.fdtd_last_error()
    return np.frombuffer(out, dtype=np.float32).copy()


def test_noalloc_iq_matches_allocating_stage(gpu):
    steps = 700
    run = NativeRun(gpu, (20, 20, 20), 1e-3, 0.99, cpml=capi.CpmlParams(5, 3.0, 1.0, 1.0, 0.05, 1e-8))
    run.add_point_source(capi.FIELD_EZ, (10, 10, 10), "source_waveform_modulated_gaussian",
                         modgauss_params(30e9, 20e9))
    run.add_probe(capi.FIELD_EZ, (13, 10, 10), steps, None)
    run.run(steps)
    for noise in (0, 1):
        for n, stream in ((steps, 3), (150, 4)):
            a = _iq(gpu.fdtd_sim_probe_iq, run, n, noise, stream)
            b = _iq(gpu.fdtd_sim_probe_iq_noalloc, run, n, noise, stream)
            assert np.max(np.abs(a)) > 0
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (noise, n)
    run.close()


def test_noalloc_iq_refuses_more_than_its_array(gpu):
    run = NativeRun(gpu, (8, 8, 8), 1e-3, 0.99, periodic=7)
    run.add_probe(capi.FIELD_EZ, (4, 4, 4), MAX + 1, None)
    run.run(MAX + 1)
    out = (C.c_float * (2 * (MAX + 1)))()
    assert gpu.fdtd_sim_probe_iq_noalloc(run.sim, 0, MAX + 1, 0, C.c_float(10.0), 1, 0, out) == -1
    assert b"at most" in gpu.fdtd_last_error()
    assert gpu.fdtd_sim_probe_iq_noalloc(run.sim, 0, MAX, 0, C.c_float(10.0), 1, 0, out) == 0
    run.close()
