"""GPU parity of the matrix functions of a skew matrix (exp, log, cis, cos, sin, sincos, sinh, cosh, tan,
tanh, cot, coth; reference src/exp.jl:3-296) against the oracle's restatement and scipy's dense functions.

The reference's test (test/runtests.jl:208-281) compares each function with Julia's dense one on n in
{1, 2, 10, 11} (SkewHermitian) and {2, 10, 11} (SkewHermTridiagonal), integer-valued, Float64 and complex
inputs; these tests do the same against scipy.linalg (expm, logm, cosm, sinm, tanm, tanhm, sinhm, coshm)
and add larger sizes, where the eigenvector rotations run on the GPU, against the oracle.  The GPU
products and LU solves sum in another order than numpy's, so the bar is a tolerance: max |GPU - ref| <=
1e-11 max(1, max |ref|) (cot, coth: 1e-9, whose divisor's conditioning grows with n).  Inputs are scaled
to spectral radius about 1: the reference's formula for cos / sin / cis forms e^{-lambda} and e^{lambda}
separately, which loses accuracy for large |lambda| in the reference as well.
"""
import numpy as np
import pytest
import scipy.linalg as sl

import oracle

pytestmark = pytest.mark.gpu
skl = pytest.importorskip("skewlinearalgebra")

FUNS = ["exp", "log", "cis", "cos", "sin", "sincos", "sinh", "cosh", "tan", "tanh", "cot", "coth"]


def say(msg):
    print("[case] " + msg, flush=True)


def rel(X, R):
    X = np.asarray(X.data if isinstance(X, skl.SkewHermitian) else X)
    return np.abs(X - R).max() / max(1.0, np.abs(R).max())


def tol(f):
    return 1e-9 if f in ("cot", "coth") else 1e-11


def real_skew(n, seed, integer=False):
    rng = np.random.default_rng(seed)
    if integer:
        M = rng.integers(-10, 11, size=(n, n)).astype(float) * 2
        return (M - M.T) / 2 / (10 * np.sqrt(n))
    M = rng.standard_normal((n, n))
    return np.asfortranarray((M - M.T) / (2 * np.sqrt(n)))


def complex_skew(n, seed):
    rng = np.random.default_rng(seed)
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return np.asfortranarray((M - M.conj().T) / (2 * np.sqrt(n)))


def dense_ref(A, f):
    inv = np.linalg.inv
    return {"exp": lambda: sl.expm(A), "log": lambda: sl.logm(A), "cis": lambda: sl.expm(1j * A),
            "cos": lambda: sl.cosm(A), "sin": lambda: sl.sinm(A), "sinh": lambda: sl.sinhm(A),
            "cosh": lambda: sl.coshm(A), "tan": lambda: sl.tanm(A), "tanh": lambda: sl.tanhm(A),
            "cot": lambda: sl.cosm(A) @ inv(sl.sinm(A)), "coth": lambda: sl.coshm(A) @ inv(sl.sinhm(A))}[f]()


def check(A, f, got, want_oracle):
    if f == "sincos":
        S, C = got
        So, Co = want_oracle
        assert rel(S, So) <= 1e-11 and rel(C, Co) <= 1e-11
        assert rel(S, sl.sinm(A)) <= 1e-11 and rel(C, sl.cosm(A)) <= 1e-11
        return
    assert rel(got, want_oracle) <= tol(f), (f, A.shape, rel(got, want_oracle))
    if A.shape[0] <= 64:
        assert rel(got, dense_ref(A, f)) <= tol(f), (f, A.shape, rel(got, dense_ref(A, f)))


def applicable(f, n, cplx):
    if f in ("log", "cot", "coth") and not cplx and n % 2 == 1:
        return False   # singular: log is a DomainError, cot / coth divide by a singular sin / sinh
    return True


@pytest.mark.parametrize("n", [1, 2, 10, 11])
@pytest.mark.parametrize("kind", ["int", "real", "complex"])
def test_skewfun_reference_cases(n, kind):
    """The reference's exp.jl test sizes and input kinds, every function."""
    A = complex_skew(n, 10 + n) if kind == "complex" else real_skew(n, 20 + n, kind == "int")
    S = skl.SkewHermitian(A)
    for f in FUNS:
        if not applicable(f, n, kind == "complex") or (n == 1 and f in ("cot", "coth")):
            continue
        say("skewfun %s n=%d %s" % (f, n, kind))
        check(A, f, getattr(skl, f)(S), oracle.skewfun(A, f))


@pytest.mark.parametrize("n", [2, 10, 11])
def test_skewfun_tridiagonal_cases(n):
    rng = np.random.default_rng(30 + n)
    ev = rng.standard_normal(n - 1) * 0.6
    T = skl.SkewHermTridiagonal(ev)
    A = np.diag(ev, -1) - np.diag(ev, 1)
    for f in FUNS:
        if not applicable(f, n, False):
            continue
        say("skewfun tridiagonal %s n=%d" % (f, n))
        check(A, f, getattr(skl, f)(T), oracle.skewfun_tri(ev, f))


@pytest.mark.parametrize("n", [300, 601])
def test_skewfun_large_vs_oracle(n):
    """Sizes whose eigenvector rotations run on the GPU: every function against the oracle."""
    A = real_skew(n, 7 + n)
    S = skl.SkewHermitian(A)
    for f in FUNS:
        if not applicable(f, n, False):
            continue
        say("skewfun %s n=%d" % (f, n))
        check(A, f, getattr(skl, f)(S), oracle.skewfun(A, f))


def test_skewfun_complex_large_and_tridiagonal_rotations():
    n = 260
    A = complex_skew(n, 5)
    for f in ("exp", "log", "cos", "sin", "tanh"):
        say("skewfun complex %s n=%d" % (f, n))
        check(A, f, getattr(skl, f)(A), oracle.skewfun(A, f))
    ev = np.random.default_rng(6).standard_normal(299) * 0.3
    for f in ("exp", "cis", "coth"):
        say("skewfun tridiagonal %s n=300" % f)
        check(np.diag(ev, -1) - np.diag(ev, 1), f, getattr(skl, f)(skl.SkewHermTridiagonal(ev)),
              oracle.skewfun_tri(ev, f))


def test_skewfun_identities():
    """exp(A) is orthogonal with det 1 and exp(log(A)) = A (test/runtests.jl:225-227); sin^2 + cos^2 = I;
    cis(A) = cos(A) + i sin(A) for real A; sinh and cosh add up to exp."""
    n = 512
    A = real_skew(n, 99)
    say("skewfun identities n=%d" % n)
    E = skl.exp(A)
    assert np.abs(E.T @ E - np.eye(n)).max() <= 1e-12
    # log(A) is a plain matrix, as in the reference, so its exp is the general one (scipy here)
    assert np.abs(sl.expm(skl.log(A)) - A).max() <= 1e-10
    S, C = skl.sincos(A)
    assert np.abs(S.data @ S.data + C @ C - np.eye(n)).max() <= 1e-12
    assert np.abs(skl.cis(A) - (C + 1j * S.data)).max() <= 1e-12
    assert np.abs(skl.sinh(A).data + skl.cosh(A) - E).max() <= 1e-12
    # the wrappers: sin is exactly skew, cos exactly symmetric
    assert np.array_equal(S.data, -S.data.T) and np.array_equal(C, C.T)


def test_skewfun_errors():
    say("skewfun errors")
    with pytest.raises(skl.DomainError):
        skl.log(skl.SkewHermTridiagonal(np.array([1.0, 0.0, 0.0])))   # eigenvalues +-i, 0, 0 (exact zeros)
    with pytest.raises(skl.SingularException):
        skl.cot(np.zeros((1, 1)))                                       # sin(0) = 0: an exactly zero pivot
    with pytest.raises(skl.SingularException):
        skl.coth(skl.SkewHermitian(np.zeros((2, 2))))                   # exp(0) - I = 0
    assert np.array_equal(skl.exp(np.zeros((1, 1))), [[1.0]])
    assert np.array_equal(skl.cis(np.zeros((1, 1))), [[1.0 + 0j]])


@pytest.mark.parametrize("n", [2, 10, 11])
def test_skewfun_complex_tridiagonal_cases(n):
    """The reference's complex tridiagonal inputs (test/runtests.jl:246-251: SkewHermTridiagonal(rand(T, n - 1),
    rand(real(T), n)) for T = ComplexF32), every function, against the oracle on Matrix(A) and scipy."""
    rng = np.random.default_rng(80 + n)
    ev = (rng.random(n - 1) + 1j * rng.random(n - 1)) * 0.5
    dv = rng.random(n) * 0.5
    T = skl.SkewHermTridiagonal(ev, dv)
    A = T.to_dense()
    for f in FUNS:
        say("skewfun complex tridiagonal %s n=%d" % (f, n))
        check(A, f, getattr(skl, f)(T), oracle.skewfun(A, f))
