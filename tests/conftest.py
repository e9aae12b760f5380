import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def golden():
    import json
    with open(os.path.join(ROOT, "tests", "golden", "reference_tests.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def ctx():
    """One device context for all GPU tests."""
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    from mymoonmd_amd import parmoon as pm
    return pm.Context(0)


def iterative_7x7(g):
    """[[A, B^T], [B, D]] of iterative_solver_test (test/iterative_solver_test.c++:9-70) as CSR."""
    M = np.zeros((7, 7))
    M[np.arange(4), np.arange(4)] = g["A"]["diag"]
    B = g["B"]
    for r in range(3):
        for k in range(B["row_ptr"][r], B["row_ptr"][r + 1]):
            M[4 + r, B["col"][k]] = B["val"][k]
            M[B["col"][k], 4 + r] = B["val"][k]
    D = g["D"]
    for r in range(3):
        for k in range(D["row_ptr"][r], D["row_ptr"][r + 1]):
            M[4 + r, 4 + D["col"][k]] = D["val"][k]
    rp, col, val = [0], [], []
    for r in range(7):
        nz = np.nonzero(M[r])[0]
        col += list(nz)
        val += list(M[r, nz])
        rp.append(len(col))
    return M, np.array(rp, np.int32), np.array(col, np.int32), np.array(val)
