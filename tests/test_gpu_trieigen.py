"""GPU application of the skew-tridiagonal QR's eigenvector rotations.

Reference behaviour: skewtrieigen_divided! / skewtrieigen_merged! (src/skeweigen.jl:322-457,
:195-320), whose Givens rotations accumulate into the halves Qodd and Qeven of the eigenvector
matrix.  skh_skewtrieigen_d_gpu runs the same host QR but applies the logged rotations on the GPU;
the arithmetic is the host's (separately rounded products and sums), so eigenvalues and
eigenvectors are bit-identical to the oracle and to the host-only entry point.
"""
import time

import numpy as np
import pytest

import oracle
from conftest import skew_gauss

pytestmark = pytest.mark.gpu
skl = pytest.importorskip("skewlinearalgebra")
from skewlinearalgebra import _lib  # noqa: E402


def say(msg):
    print("[case] " + msg, flush=True)


def trieigen(eta, variant, gpu):
    ev = np.array(eta, dtype=np.float64)
    N = ev.size + 1
    vals = np.zeros(N)
    vr = np.zeros((N, N), order="F")
    vi = np.zeros((N, N), order="F")
    fn = _lib.lib().skh_skewtrieigen_d_gpu if gpu else _lib.lib().skh_skewtrieigen_d
    rc = fn(N, ev.ctypes.data, vals.ctypes.data, vr.ctypes.data, vi.ctypes.data, variant)
    assert rc == 0
    return vals, vr + 1j * vi


@pytest.mark.parametrize("variant", [0, 1])
@pytest.mark.parametrize("n", [2, 3, 7, 64, 257, 512, 1001])
def test_trieigen_gpu_bitwise_vs_oracle(n, variant):
    A = skew_gauss(n, 1300 + n)
    _, _, eta = oracle.skewblockedhess(A)
    say("trieigen gpu n=%d variant=%d" % (n, variant))
    vals, V = trieigen(eta, variant, True)
    ov, oV = oracle.skewtrieigen(eta, "divided" if variant == 0 else "merged")
    np.testing.assert_array_equal(vals, ov)
    np.testing.assert_array_equal(V, oV)


def test_trieigen_gpu_equals_host_n1600():
    """A larger size against the host-only entry point (also timed, printed)."""
    n = 1600
    rng = np.random.default_rng(16)
    eta = rng.standard_normal(n - 1)
    say("trieigen gpu vs host n=%d" % n)
    t0 = time.perf_counter()
    v1, V1 = trieigen(eta, 0, True)
    t1 = time.perf_counter()
    v2, V2 = trieigen(eta, 0, False)
    t2 = time.perf_counter()
    print("[time] n=%d gpu rotations %.3f s, host rotations %.3f s" % (n, t1 - t0, t2 - t1), flush=True)
    np.testing.assert_array_equal(v1, v2)
    np.testing.assert_array_equal(V1, V2)


def test_eigen_through_gpu_rotations():
    """eigen(A) at n >= 256 takes the GPU rotations: A X = X diag(lambda), X unitary."""
    n = 600
    A = skew_gauss(n, 61)
    say("eigen n=%d" % n)
    E = skl.eigen(A)
    X, lam = E.vectors, E.values
    nA = np.linalg.norm(A)
    assert np.linalg.norm(A @ X - X * lam) <= 1e-12 * n * nA
    assert np.linalg.norm(X.conj().T @ X - np.eye(n)) <= 1e-12 * n
    T = skl.SkewHermTridiagonal(oracle.skewblockedhess(A)[2])
    Et = skl.eigen(T)
    ov, oV = oracle.skewtrieigen(T.ev, "merged")
    np.testing.assert_array_equal(Et.values.imag, ov)
    np.testing.assert_array_equal(Et.vectors, oV)
