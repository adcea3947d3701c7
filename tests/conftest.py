import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "skewlinearalgebra.jl_amd")
for p in (ROOT, PKG_DIR):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def golden():
    with open(os.path.join(ROOT, "tests", "golden", "fixtures.json")) as f:
        return json.load(f)


def skew_gauss(n, seed, dtype=np.float64):
    """A = B - B^T with B iid N(0,1) (complex: (B - B^H)/2 with N(0,1) real and imaginary parts)."""
    rng = np.random.default_rng(seed)
    if np.dtype(dtype).kind == "c":
        B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
        return np.asfortranarray((B - B.conj().T) / 2)
    B = rng.standard_normal((n, n))
    return np.asfortranarray(B - B.T)


def sparse_to_dense(d):
    n = d["n"]
    A = np.zeros((n, n))
    for i, j, v in zip(d["I"], d["J"], d["V"]):
        A[i - 1, j - 1] += v
    return A
