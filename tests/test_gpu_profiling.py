"""Per-kernel timing (fdtd_sim_set_profiling): the profiled steps launch
their kernels between HIP event records on the simulation stream.  A
profiled run gives the same bits as an unprofiled one (which replays
captured step graphs) and times every step it ran, on the fused path
(periodic dispersive stack: leapfrog, ADE pass) and on the two-kernel
path, for runs of whole and partial DFT blocks."""
import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import same_values, NativeRun, c1_cpml_params, modgauss_params, c1_freqs

pytestmark = pytest.mark.gpu


def _stack(lib, profile, runs):
    from test_gpu_parity import _dispersive_setup
    run = NativeRun(lib, (16, 16, 44), 0.5e-3, 0.99, periodic=3, cpml=capi.CpmlParams(8, 3.0, 1.0, 1.0, 0.05, 1e-8),
                    precision=capi.FDTD_FP32, block=8, material_setup=_dispersive_setup)
    run.add_plane_source(capi.FIELD_EX, 2, 10, "source_waveform_modulated_gaussian",
                         modgauss_params(40e9, 30e9))
    freqs = np.linspace(25e9, 55e9, 8)
    m = run.add_plane_monitor(2, 36, freqs)
    p = run.add_probe(capi.FIELD_EX, (3, 5, 24), sum(runs), freqs)
    times = []
    if profile:
        run.set_profiling(True)
    for n in runs:
        run.run(n)
        if profile:
            times.append(run.kernel_times3())
    out = {"td": run.td(p, sum(runs)), "T": run.plane_dft(m, 0, (16, 16, 8)), "fields": run.fields(),
           "fused": run.lib.fdtd_sim_fused(run.sim)}
    run.close()
    return out, times


@pytest.mark.parametrize("runs", [(8, 8, 8, 8, 8), (5, 19, 40), (3,)])
def test_profiled_run_same_bits_and_every_step_timed(gpu, runs):
    a, times = _stack(gpu, True, runs)
    b, _ = _stack(gpu, False, runs)
    assert a["fused"] == 1
    assert max(np.max(np.abs(v)) for v in b["fields"].values()) > 0
    assert same_values(a["td"], b["td"]) and same_values(a["T"], b["T"])
    for k in a["fields"]:
        assert same_values(a["fields"][k], b["fields"][k]), k
    for (ms, n), steps in zip(times, runs):
        assert n == steps
        assert ms[0] > 0 and ms[1] > 0      # leapfrog and ADE pass timed in every step
        assert ms[0] / n < 1e3               # per-step milliseconds, not garbage


def test_profiled_two_kernel_path(gpu, monkeypatch):
    """A vacuum box on the two-kernel step (the fused kernel switched off):
    every step timed, the same bits as unprofiled."""
    monkeypatch.setenv("RFSYN_FUSED", "0")

    def one(profile):
        run = NativeRun(gpu, (24, 24, 24), 1e-3, 0.99, cpml=c1_cpml_params(6), precision=capi.FDTD_FP64, block=8)
        run.add_point_source(capi.FIELD_EZ, (12, 12, 12), "source_waveform_modulated_gaussian",
                             modgauss_params(30e9, 20e9))
        p = run.add_probe(capi.FIELD_EZ, (15, 12, 12), 40, c1_freqs())
        if profile:
            run.set_profiling(True)
        run.run(40)
        t = run.kernel_times3() if profile else None
        out = (run.td(p, 40), run.fields(), run.lib.fdtd_sim_fused(run.sim))
        run.close()
        return out, t
    a, t = one(True)
    b, _ = one(False)
    assert a[2] == 0 and np.max(np.abs(b[0])) > 0
    assert same_values(a[0], b[0])
    for k in a[1]:
        assert same_values(a[1][k], b[1][k]), k
    assert t[1] == 40 and t[0][0] > 0
