"""The C-ABI boundary: the built library loads without a GPU and exports
every function that include/*.h declares (no compute calls here)."""
import ctypes
import glob
import os
import re

import pytest

import rfsyn_amd

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")


def declared_functions():
    """(header, name) of every function prototype in include/*.h."""
    out = []
    for h in sorted(glob.glob(os.path.join(INCLUDE, "*.h"))):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"^(?!\s*typedef)[A-Za-z_][\w \t\*]*?\b(\w+)\s*\(([^;{]*?)\)\s*;", text, flags=re.M):
            out.append((os.path.basename(h), m.group(1)))
    return out


def test_headers_declare_the_engine_api():
    names = {n for _, n in declared_functions()}
    # the reference's engine/fdtd entry points the drop-in boundary replaces
    for n in ("fdtd_grid_create", "fdtd_update_h", "fdtd_update_e", "fdtd_update_h_cpml",
              "fdtd_update_e_cpml_materials", "cpml_create", "materials_create", "source_apply",
              "probe_dft_record", "fdtd_simulate_echo", "fdtd_sim_create", "fdtd_sim_run"):
        assert n in names, n


def test_library_exports_every_declared_function():
    if not os.path.exists(rfsyn_amd.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(rfsyn_amd.LIB_PATH)
    decl = declared_functions()
    assert len(decl) > 50
    missing = [(h, n) for h, n in decl if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_library():
    """capi.bind resolves every prototype it knows (raises otherwise)."""
    if not os.path.exists(rfsyn_amd.LIB_PATH):
        pytest.skip("library not built")
    lib = rfsyn_amd.capi.bind(ctypes.CDLL(rfsyn_amd.LIB_PATH))
    assert lib.fdtd_sim_kernel_times3 is not None


def test_batch_argument_checks():
    """The packed-batch entry points refuse bad arguments before touching
    the device: no handles, NULL handles, NULL batches."""
    if not os.path.exists(rfsyn_amd.LIB_PATH):
        pytest.skip("library not built")
    lib = rfsyn_amd.capi.bind(ctypes.CDLL(rfsyn_amd.LIB_PATH))
    assert not lib.fdtd_batch_create(None, 0)
    arr = (ctypes.c_void_p * 2)(None, None)
    assert not lib.fdtd_batch_create(arr, 2)
    assert b"NULL handle" in lib.fdtd_last_error()
    assert lib.fdtd_batch_run(None, 10) == -1
    assert lib.fdtd_batch_size(None) == 0
    assert lib.fdtd_batch_launches(None, None, None, None) == -1
    lib.fdtd_batch_destroy(None)
    assert lib.fdtd_sim_dft_block(None) == -1
