"""GPU parity of the skew-Cholesky factorization against the oracle.

Reference behaviour: _skewchol!(A::SkewHermitian{<:Real}, ::RowMaximum) and (..., ::NoPivot) in
src/cholesky.jl:58-107 and :28-56, behind skewchol / skewchol! (:109-151); the reference's test is
R^T J R = A[p, p] (test/runtests.jl:436-455).  The GPU kernel and the oracle perform the same
floating-point operations in the same order (products rounded separately, IEEE division and square
root), so R, p and the step count agree bit for bit; the stop threshold 1e-15 ||A||_F is summed in a
different order, which could only matter for a maximum within rounding of the threshold.
"""
import ctypes

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
skl = pytest.importorskip("skewlinearalgebra")
from skewlinearalgebra import _lib  # noqa: E402


def say(msg):
    print("[case] " + msg, flush=True)


def gskew(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n))
    return np.asfortranarray(M - M.T)


def residual(A, C):
    n = A.shape[0]
    J = np.asarray(C.J)
    return np.abs(C.R.T @ J @ C.R - A[np.ix_(C.p, C.p)]).max() / max(1.0, np.abs(A).max())


@pytest.mark.parametrize("n", [1, 2, 3, 10, 11, 64, 65, 200, 256, 257, 700, 1024])
@pytest.mark.parametrize("pivot", [True, False])
def test_skewchol_bitwise_vs_oracle(n, pivot):
    A = gskew(n, 100 + n)
    say("skewchol n=%d pivot=%s" % (n, pivot))
    C = skl.skewchol(A, skl.RowMaximum() if pivot else skl.NoPivot())
    R, p, steps = oracle.skewchol(A, pivot)
    assert np.array_equal(C.p, p) and C.steps == steps
    assert np.array_equal(C.R, R, equal_nan=True)
    if pivot:
        assert residual(A, C) <= 1e-13 * max(1, n)


@pytest.mark.parametrize("n", [1, 2, 10, 11])
@pytest.mark.parametrize("kind", ["int", "float"])
def test_skewchol_reference_cases(n, kind):
    """test/runtests.jl:436-455: skewhermitian(rand(-10:10, n, n) * 2) and skewhermitian(randn),
    R^T J R = A[p, p] for the default and RowMaximum, = A for NoPivot."""
    rng = np.random.default_rng(7 + n)
    M = rng.integers(-10, 11, size=(n, n)).astype(float) * 2 if kind == "int" else rng.standard_normal((n, n))
    A = skl.skewhermitian(M)
    say("skewchol reference case n=%d %s" % (n, kind))
    for C in (skl.skewchol(A), skl.skewchol(A, skl.RowMaximum())):
        assert residual(A.data, C) <= 1e-13
    C = skl.skewchol(A, skl.NoPivot())
    assert np.array_equal(C.p, np.arange(n)) and residual(A.data, C) <= 1e-12
    B = np.array(A.data)
    assert residual(B, skl.skewchol(B)) <= 1e-13


def test_skewchol_tie_rule_and_closed_forms():
    TIE4 = np.array([[0, 3, -5, 0], [-3, 0, 0, 5], [5, 0, 0, 5], [0, -5, -5, 0]], dtype=float)
    say("skewchol ties")
    C = skl.skewchol(TIE4)
    assert list(C.p[:2]) == [2, 0] and C.R[0, 0] == np.sqrt(5.0)
    R, p, _ = oracle.skewchol(TIE4)
    assert np.array_equal(C.R, R) and np.array_equal(C.p, p)
    for a in (2.0, -3.0):
        A = np.array([[0.0, a], [-a, 0.0]])
        r = np.sqrt(abs(a))
        C = skl.skewchol(A)
        assert np.array_equal(C.R, np.diag([r, r])) and list(C.p) == ([0, 1] if a > 0 else [1, 0])
        C = skl.skewchol(A, skl.NoPivot())
        assert np.array_equal(C.R, np.diag([r, np.sign(a) * r]))


def test_skewchol_rank_deficient():
    rng = np.random.default_rng(8)
    X = rng.standard_normal((10, 4))
    A = X @ oracle.jmatrix(4) @ X.T
    A = (A - A.T) / 2
    say("skewchol rank 4 of 10")
    C = skl.skewchol(A)
    R, p, steps = oracle.skewchol(A)
    assert C.steps == steps == 2
    assert np.array_equal(C.R, R) and np.array_equal(C.p, p)
    assert residual(A, C) <= 1e-13


def test_skewchol_batch_independent():
    """Sizes 0..120 plus one 300 (the large-LDS variant): every matrix gets the bits of its own
    call and of the oracle."""
    rng = np.random.default_rng(12)
    sizes = [int(s) for s in rng.integers(0, 121, size=150)] + [0, 1, 300]
    mats = [gskew(s, 500 + i) for i, s in enumerate(sizes)]
    say("skewchol batch of %d" % len(mats))
    out = skl.skewchol_batched(mats)
    for i in (0, 5, 77, len(mats) - 1):
        one = skl.skewchol(mats[i])
        assert np.array_equal(one.R, out[i].R) and np.array_equal(one.p, out[i].p)
    for A, C in zip(mats, out):
        R, p, steps = oracle.skewchol(A)
        assert np.array_equal(C.R, R) and np.array_equal(C.p, p) and C.steps == steps
    nop = skl.skewchol_batched(mats[:40], skl.NoPivot())
    for A, C in zip(mats[:40], nop):
        R, p, steps = oracle.skewchol(A, False)
        assert np.array_equal(C.R, R, equal_nan=True)


def test_skewchol_mixed_large_and_small():
    """Members of 512 and more take the cooperative whole-GPU kernel, the rest the per-matrix
    kernel, in one call; all bitwise equal to the oracle (with and without pivoting)."""
    sizes = [600, 40, 513, 7, 512, 130]
    mats = [gskew(s, 900 + i) for i, s in enumerate(sizes)]
    for pivot in (skl.RowMaximum(), skl.NoPivot()):
        say("skewchol mixed %s pivot=%s" % (sizes, type(pivot).__name__))
        out = skl.skewchol_batched(mats, pivot)
        for A, C in zip(mats, out):
            R, p, steps = oracle.skewchol(A, isinstance(pivot, skl.RowMaximum))
            assert np.array_equal(C.R, R, equal_nan=True) and np.array_equal(C.p, p) and C.steps == steps


def test_skewchol_n4096_property():
    """The largest supported size: R^T J R = A[p, p] and R upper triangular with r > 0 blocks."""
    n = 4096
    A = gskew(n, 4096)
    say("skewchol n=4096")
    C = skl.skewchol(A)
    assert C.steps == n // 2 and sorted(C.p) == list(range(n))
    assert np.array_equal(C.R, np.triu(C.R))
    d = np.diag(C.R)
    assert np.all(d > 0) and np.array_equal(d[0::2], d[1::2])
    assert residual(A, C) <= 1e-11


def test_skewchol_device_entry():
    torch = pytest.importorskip("torch")
    ns = np.array([64, 0, 33, 128], dtype=np.int64)
    mats = [gskew(int(s), 60 + i) for i, s in enumerate(ns)]
    offs = np.zeros(len(ns), dtype=np.int64)
    offs[1:] = np.cumsum(ns[:-1] ** 2)
    packed = np.concatenate([M.ravel(order="F") for M in mats])
    dA = torch.from_numpy(packed.copy()).cuda()
    dR = torch.zeros_like(dA)
    dp = torch.zeros(int(ns.sum()), dtype=torch.int64, device="cuda")
    ds = torch.zeros(len(ns), dtype=torch.int64, device="cuda")
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    say("skewchol device entry")
    rc = _lib.lib().skh_skewchol_d_batched_device(len(ns), ctypes.c_void_p(ns.ctypes.data),
                                                  ctypes.c_void_p(offs.ctypes.data), ctypes.c_void_p(dA.data_ptr()),
                                                  ctypes.c_void_p(dR.data_ptr()), ctypes.c_void_p(dp.data_ptr()),
                                                  ctypes.c_void_p(ds.data_ptr()), 1, st, None)
    _lib.check("skh_skewchol_d_batched_device", rc)
    torch.cuda.synchronize()
    host = skl.skewchol_batched(mats)
    R = dR.cpu().numpy()
    p = dp.cpu().numpy()
    pacc = 0
    for k, C in enumerate(host):
        n = int(ns[k])
        assert np.array_equal(R[offs[k]:offs[k] + n * n].reshape((n, n), order="F"), C.R)
        assert np.array_equal(p[pacc:pacc + n], C.p)
        pacc += n
    assert list(ds.cpu().numpy()) == [C.steps for C in host]


def test_skewchol_inplace_and_errors():
    A = gskew(12, 3)
    B = A.copy()
    say("skewchol! and errors")
    C = skl.skewchol_(B)
    assert C.R is B and np.array_equal(B, skl.skewchol(A).R)
    with pytest.raises(skl.ArgumentError):
        skl.skewchol(np.ones((3, 3)))
    with pytest.raises(ValueError):
        skl.skewchol(np.zeros((4097, 4097)))
