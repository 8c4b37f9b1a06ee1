"""Shared test setup. GPU tests are marked ``gpu``; everything else runs on a CPU-only host."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "ref_vectors.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def tol_ref(log2n: int) -> float:
    """Bound on rel-L2 against the reference's own CPU transform: 3e-7 * log2 n (at least 3e-7)."""
    return 3e-7 * max(log2n, 1)


def tol_fp64(log2n: int) -> float:
    """The reference tests' bound against an FP64 transform (tests/test_onefft.c:99-101)."""
    return 4e-7 + 1e-7 * log2n
