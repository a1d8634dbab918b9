"""Test-side access to the seeded scene builders of rfsyn_amd.driver."""
from rfsyn_amd.driver import *  # noqa: F401,F403
from rfsyn_amd.driver import (Pcg32, NativeRun, c1_cpml_params, c1_freqs, c1_probes,  # noqa: F401
                              field_arrays, modgauss_params, pattern_fill, random_fill, run_c1,
                              same_values)
