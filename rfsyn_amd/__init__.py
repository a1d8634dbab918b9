"""rfsyn_amd: MI355X-native FDTD time-stepping core for rfsyn.

The product is the C library ``librfsyn_fdtd.so`` (C host code plus HIP
kernels for gfx950) exporting the engine/fdtd C API declared in
``include/*.h``.  This package only locates and binds it; there is no
Python or CPU fallback for any computation.
"""
import ctypes
import os

from . import capi

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "librfsyn_fdtd.so")
INCLUDE_DIR = os.path.join(REPO, "include")

_lib = None


def build(jobs=8):
    """Compile librfsyn_fdtd.so for gfx950 (always recompiles)."""
    import subprocess
    subprocess.run(["make", "-s", "-B", f"-j{jobs}", "-C", HERE], check=True)


def load_library():
    """Load and bind the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with rfsyn_amd.build() "
                "(there is no fallback implementation)")
        _lib = capi.bind(ctypes.CDLL(LIB_PATH))
    return _lib
