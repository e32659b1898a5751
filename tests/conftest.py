"""Shared test setup: import paths, the `gpu` marker, graph builders."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gpm-artifact_amd"))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def gpm():
    import gpm_amd
    gpm_amd.lib()  # raises if libgpm.so is missing: no silent fallback
    return gpm_amd


@pytest.fixture(scope="session")
def workloads(gpm):
    import workloads as W
    return W


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    oracle.lib()
    return oracle
