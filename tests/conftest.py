"""Shared fixtures: the product library (HIP), the CPU oracle and, where
the rfsyn sources are present, the reference compiled from them."""
import ctypes
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from rfsyn_amd import capi  # noqa: E402

ORACLE_SO = os.path.join(REPO, "oracle", "liboracle_fdtd.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "libref_fdtd.so")
PRODUCT_SO = os.path.join(REPO, "rfsyn_amd", "librfsyn_fdtd.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def _load(path):
    return capi.bind(ctypes.CDLL(path))


@pytest.fixture(scope="session")
def oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)
    return _load(ORACLE_SO)


@pytest.fixture(scope="session")
def ref():
    """The rfsyn reference compiled from its own sources (oracle/Makefile
    target `ref`); skipped where those sources are absent."""
    if not os.path.exists(REF_SO):
        if not os.path.isdir("/root/reference/engine/fdtd"):
            pytest.skip("reference sources not present")
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"], check=True)
    return _load(REF_SO)


@pytest.fixture(scope="session")
def gpu():
    import rfsyn_amd
    return rfsyn_amd.load_library()
