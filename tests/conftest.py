"""Shared fixtures: the `gpu` marker, reference fixture loading, oracle build."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

FIXTURES = os.path.join(ROOT, "tests", "golden", "reference_fixtures")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


def load_fixture(name, fname):
    """Reference fixture format (source/tests/tests/test_matrix.c:52-62,113-125): the order n,
    then the values in %e text, row-major."""
    with open(os.path.join(FIXTURES, name, fname)) as f:
        vals = f.read().split()
    n = int(vals[0])
    return n, np.array(vals[1:], dtype=np.float64)


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle


@pytest.fixture(scope="session")
def hip():
    """The product library on a real GPU (fails loudly if libhipla.so is missing)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import python_amd
    python_amd.lib()
    assert python_amd.lib().hipla_init() == 0, python_amd.last_error_detail()
    return python_amd
