"""Shared pytest set-up: the `gpu` marker and import paths.

GPU tests need libgcvol.so built in-tree (python3 __graft_entry__.py build)
and a visible HIP device; they fail loudly if the library is missing.
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gpu-out-of-core-volume-data_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libgcvol.so")


@pytest.fixture(scope="session")
def oracle():
    import oracle_lib
    oracle_lib.lib()
    return oracle_lib


@pytest.fixture(scope="session")
def gcv():
    import gcvol
    gcvol.lib()
    return gcvol


@pytest.fixture(scope="session")
def gpu(gcv):
    if gcv.device_count() < 1:
        pytest.fail("no HIP device visible to a gpu-marked test")
    return gcv
