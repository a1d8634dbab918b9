"""Point sources on the two-kernel step.  The step applies them to E^n
inside the E update when every source sits on a node and component the
update writes and no plane source runs between them; otherwise (a source on
a boundary node the update skips, a plane source before the E update) they
keep their own kernel.  Either way the step is source_apply(n dt) followed
by the E update (the reference's order): FP64 fields and probe series
bitwise the oracle's.  On the fused step, sources on the CPML slab nodes
are applied by the shell sweeps, in the wavefronts that hold one."""
import numpy as np
import pytest

from rfsyn_amd import capi
from rfsyn_amd.driver import NativeRun, c1_cpml_params, c1_freqs, modgauss_params, same_values

pytestmark = pytest.mark.gpu

ORACLE_LIB = [None]   # the oracle library (it has no fused step to report)

WAVE = ("source_waveform_modulated_gaussian", modgauss_params(30e9, 20e9))

# (component, node, hard) per source, in insertion order
CASES = {
    # interior nodes: applied inside the E update
    "interior": [(capi.FIELD_EZ, (10, 10, 10), False), (capi.FIELD_EX, (7, 12, 9), False)],
    # a soft and then a hard source on one node and component: insertion
    # order decides the value
    "soft_then_hard": [(capi.FIELD_EY, (9, 8, 11), False), (capi.FIELD_EY, (9, 8, 11), True)],
    # Ey on the x = 0 plane, which the E update does not write: the sources
    # keep their own kernel
    "boundary": [(capi.FIELD_EY, (0, 10, 10), False), (capi.FIELD_EZ, (10, 10, 10), False)],
}


def _run(lib, sources, plane=False, steps=40):
    run = NativeRun(lib, (20, 20, 20), 1e-3, 0.99, cpml=c1_cpml_params(5), precision=capi.FDTD_FP64, block=8)
    for comp, node, hard in sources:
        run.add_point_source(comp, node, WAVE[0], WAVE[1], hard=hard)
    if plane:   # a plane source before the E update, crossing the first source's node
        run.add_plane_source(sources[0][0], 2, sources[0][1][2], WAVE[0], WAVE[1], amp=0.5)
    p = run.add_probe(capi.FIELD_EZ, (12, 11, 10), steps, c1_freqs(8))
    run.run(steps)
    out = (run.td(p, steps), run.fields())
    fused = run.lib.fdtd_sim_fused(run.sim) if lib is not ORACLE_LIB[0] else None
    run.close()
    return out + (fused,)


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("plane", [False, True])
def test_two_kernel_point_sources_fp64_vs_oracle(gpu, oracle, case, plane, monkeypatch):
    monkeypatch.setenv("RFSYN_FUSED", "0")
    ORACLE_LIB[0] = oracle
    td, fields, fused = _run(gpu, CASES[case], plane)
    td_ref, fields_ref, _ = _run(oracle, CASES[case], plane)
    assert fused == 0
    assert np.max(np.abs(td_ref)) > 0
    assert same_values(td, td_ref)
    for k in fields_ref:
        assert same_values(fields[k], fields_ref[k]), k


# ------------------------------------------------------- fused step, shell

SHELL_CASES = {
    # soft Ez in the low-x slab, hard Ex in the high-y slab (both updated
    # by the shell sweeps, which apply sources only in wavefronts holding one)
    "slabs": ([(capi.FIELD_EZ, (2, 10, 10), False), (capi.FIELD_EX, (10, 17, 9), True)], False),
    # and a plane source across the box and the slabs
    "slabs_and_plane": ([(capi.FIELD_EZ, (2, 10, 10), False), (capi.FIELD_EX, (10, 17, 9), True)], True),
}


@pytest.mark.parametrize("case", sorted(SHELL_CASES))
def test_fused_step_shell_sources_fp64_vs_oracle(gpu, oracle, case, monkeypatch):
    """Sources on CPML slab nodes, which the shell kernels update around the
    fused box: FP64 bitwise the oracle."""
    monkeypatch.setenv("RFSYN_FUSED", "1")
    sources, plane = SHELL_CASES[case]
    ORACLE_LIB[0] = oracle
    td, fields, fused = _run(gpu, sources, plane)
    td_ref, fields_ref, _ = _run(oracle, sources, plane)
    assert fused == 1
    assert np.max(np.abs(td_ref)) > 0
    assert same_values(td, td_ref)
    for k in fields_ref:
        assert same_values(fields[k], fields_ref[k]), k
