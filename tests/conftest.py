"""Shared test configuration: import paths, the `gpu` marker and fixtures."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "engineer-thesis_amd")
TESTS = os.path.join(ROOT, "tests")
for p in (PKG, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(TESTS, "golden")
LAYOUT_DIR = os.path.join(GOLDEN, "layouts")
# (file, master) pairs: the reference's own fixtures (tests/ and generatedtests/)
FIXTURES = [("test", 10), ("test1", 0), ("test2", 5), ("test100.txt", 0),
            ("test1000.txt", 0), ("test10000.txt", 0)]


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


def layout_path(name):
    return os.path.join(LAYOUT_DIR, name)


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle_ffi
    oracle_ffi.build_oracle(ref=os.path.isdir(oracle_ffi.REFERENCE_SRC))
    return oracle_ffi


def require_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    return torch
