"""Shared pytest set-up: import paths, the ``gpu`` marker, fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "peakpatch.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")
REFERENCE = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden_dir():
    return GOLDEN


@pytest.fixture(scope="session")
def homeltab_path():
    return os.path.join(GOLDEN, "HomelTab_fortran.dat")


@pytest.fixture(scope="module")
def ctx():
    from peakpatch_hip import Context
    c = Context(0)
    yield c
    c.close()


def ipp1(i, j, k, n):
    """1-based column-major flat index of 0-based (i, j, k) (PeakFind.jl:66)."""
    return (np.asarray(i, np.int64) + 1) + np.asarray(j, np.int64) * n + np.asarray(k, np.int64) * n * n
