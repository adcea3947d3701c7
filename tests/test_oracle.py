"""Pins the CPU oracle (oracle/skew_oracle.c) before it is trusted as the GPU checker.

Each check is independent of the oracle's own formulas: the reference's hard-coded fixtures,
LAPACK gehrd (the oracle the reference's hessenberg test uses), exact rational Pfaffians,
dense eigensolvers, closed forms and invariants.  A dropped term, a wrong sign or index or a
transposed operand in the restatement fails at least one of them (see test_planted_mistakes).
"""
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from conftest import skew_gauss, sparse_to_dense

EPS = np.finfo(float).eps


def tri(eta):
    return np.diag(eta, -1) - np.diag(eta, 1)


def check_factorization(A, F, tau, eta, tol=1e-13):
    Q = oracle.form_q(F, tau)
    n = A.shape[0]
    if np.iscomplexobj(A):
        T = np.diag(eta.astype(complex), -1) - np.diag(eta.astype(complex), 1) + np.diag(1j * np.imag(np.diag(F)))
        R = Q @ T @ Q.conj().T - A
        O = Q.conj().T @ Q - np.eye(n)
    else:
        R = Q @ tri(eta) @ Q.T - A
        O = Q.T @ Q - np.eye(n)
    nA = max(np.linalg.norm(A), 1e-300)
    assert np.linalg.norm(R) / nA <= tol
    assert np.linalg.norm(O) <= tol * max(n, 1)
    return Q


def test_setnb():
    # src/hessenberg.jl:133-142
    assert [oracle.setnb(n) for n in (1, 2, 5, 12, 13, 100, 101, 8192)] == [1, 1, 1, 8, 10, 10, 60, 60]


def test_readme_fixture(golden):
    g = golden["readme"]
    A = np.array(g["A"], float)
    F, tau, eta = oracle.skewblockedhess(A)
    np.testing.assert_allclose(eta, g["eta"], rtol=1e-14)
    Q = oracle.form_q(F, tau)
    np.testing.assert_allclose(Q, np.array(g["Q"]), atol=1e-14)
    check_factorization(A, F, tau, eta)
    lam = np.sort(oracle.skewtrieigvals(eta))
    p = g["eig_pos"]
    np.testing.assert_allclose(lam, np.sort([p[0], p[1], -p[1], -p[0]]), rtol=1e-14)
    assert oracle.logabspfaffian(A) == (g["logabspfaffian"][0], g["logabspfaffian"][1])
    assert abs(oracle.pfaffian(A) - g["pfaffian_exact"]) < 1e-13 * 9


def test_gehrd_fixtures(golden):
    for c in golden["gehrd"]:
        A = np.array(c["A"])
        F, tau, eta = oracle.skewblockedhess(A)
        nA = np.linalg.norm(A)
        n = A.shape[0]
        np.testing.assert_allclose(eta, c["eta"], atol=1e-13 * nA, rtol=0)
        Q = check_factorization(A, F, tau, eta)
        if n > 2:  # gehrd's Q: the reference's last reflector is the identity (tau=0)
            np.testing.assert_allclose(Q, np.array(c["Q"]), atol=1e-12)


@pytest.mark.parametrize("n", [1, 2, 3, 10, 11, 12, 13, 59, 60, 61, 62, 63, 100, 101, 128, 161, 200])
def test_factorization_random(n):
    A = skew_gauss(n, 1000 + n)
    F, tau, eta = oracle.skewblockedhess(A)
    assert eta.shape == (max(n - 1, 0),)
    if n >= 2:
        check_factorization(A, F, tau, eta)
        H = sl.hessenberg(A)
        np.testing.assert_allclose(eta, np.diag(H, -1), atol=1e-12 * np.linalg.norm(A))
        # the real final 1x1 step is no reflection (hessenberg.jl:27-29): tau = 0, factor keeps eta
        assert tau[-1] == 0.0 and F[n - 1, n - 2] == eta[-1]
        # every other reflector is a true reflection (tau * v'v = 2)
        for i in range(n - 2):
            if tau[i] != 0:
                v = np.concatenate([[1.0], F[i + 2:, i]])
                assert abs(tau[i] * (v @ v) - 2.0) < 1e-12


@pytest.mark.parametrize("nb", [1, 2, 3, 7, 16, 33])
def test_blocking_invariance(nb):
    """Any panel width gives the same factorization (to rounding): pins the W/V bookkeeping."""
    A = skew_gauss(90, 7)
    _, tau0, eta0 = oracle.skewblockedhess(A)
    F, tau, eta = oracle.skewblockedhess(A, nb=nb)
    np.testing.assert_allclose(eta, eta0, atol=1e-12 * np.linalg.norm(A))
    np.testing.assert_allclose(tau, tau0, atol=1e-12)
    check_factorization(A, F, tau, eta)


def test_zero_column_conventions():
    """householder! special cases (hessenberg.jl:26-44, Appendix A of the design notes):
    all-zero x -> tau=0, beta=0; xnorm=0 with alpha != 0 -> tau=2 (LAPACK would give 0)."""
    # column of ones (test/runtests.jl:165-174)
    A = np.zeros((4, 4))
    A[1:, 0] = 1
    A[0, 1:] = -1
    F, tau, eta = oracle.skewblockedhess(A)
    check_factorization(A, F, tau, eta)
    H = sl.hessenberg(A)
    np.testing.assert_allclose(tri(eta), H, atol=1e-14)
    # issue #121: column 0 has x = [-1, 0, 0] (alpha != 0, xnorm = 0) -> tau = 2, beta = +1;
    # column 1 has x = [0, 0] (all zero) -> tau = 0, beta = 0; the last 1x1 step keeps A[4,3]
    J = np.array([[0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1], [0, 0, -1, 0]], float)
    F, tau, eta = oracle.skewblockedhess(J)
    np.testing.assert_array_equal(tau, [2.0, 0.0, 0.0])
    np.testing.assert_array_equal(eta, [1.0, 0.0, -1.0])
    check_factorization(J, F, tau, eta)
    assert oracle.pfaffian(J) == 1.0
    # alpha != 0, xnorm = 0 in a reflector of length 2 -> tau = 2, beta = -alpha
    A = np.zeros((3, 3))
    A[1, 0] = 3.0
    A[0, 1] = -3.0
    F, tau, eta = oracle.skewblockedhess(A)
    assert tau[0] == 2.0 and eta[0] == -3.0
    check_factorization(A, F, tau, eta)
    # alpha = +0 exactly with xnorm > 0: reference takes beta = +xnorm (LAPACK: -xnorm)
    A = np.zeros((3, 3))
    A[2, 0] = 2.0
    A[0, 2] = -2.0
    F, tau, eta = oracle.skewblockedhess(A)
    assert eta[0] == 2.0 and tau[0] == 1.0
    check_factorization(A, F, tau, eta)


def test_small_sizes():
    F, tau, eta = oracle.skewblockedhess(np.zeros((1, 1)))
    assert tau.size == 0 and eta.size == 0
    A = np.array([[0.0, -2.5], [2.5, 0.0]])
    F, tau, eta = oracle.skewblockedhess(A)
    assert eta[0] == 2.5 and tau[0] == 0.0
    assert oracle.logabspfaffian(A) == (np.log(2.5), -1.0)  # Pf = a_12 = -2.5
    assert oracle.logabspfaffian(np.zeros((3, 3)))[0] == -np.inf
    assert oracle.logabspfaffian(np.zeros((3, 3)))[1] == 0.0


def test_exact_pfaffians(golden):
    cases = golden["integer_pf"] + golden["analytic4"] + [golden["issue49"], golden["issue121"]]
    for c in cases:
        A = np.array(c["A"], float)
        pf = Fraction(c["pfaffian"])
        n = A.shape[0]
        got = oracle.pfaffian(A)
        lp, sg = oracle.logabspfaffian(A)
        if n % 2 == 1:
            assert got == 0.0 and lp == -np.inf and sg == 0.0
            continue
        if pf == 0:
            continue
        scale = max(1.0, np.linalg.norm(A)) ** (n // 2)
        assert abs(got - float(pf)) <= 1e-12 * scale
        assert sg == float(np.sign(float(pf)))
        assert abs(lp - np.log(abs(float(pf)))) <= 1e-10 * max(1.0, abs(lp)) + 1e-12 * scale / abs(float(pf))
        if c.get("det") is not None:
            assert int(pf * pf) == c["det"]


def test_constructed_logpf(golden):
    for c in golden["constructed"]:
        A = np.array(c["A"])
        lp, sg = oracle.logabspfaffian(A)
        assert sg == c["sign"]
        assert abs(lp - c["logpf"]) <= 1e-12 * max(1.0, abs(c["logpf"])) + 1e-13 * A.shape[0]


@pytest.mark.parametrize("n", [2, 3, 4, 9, 10, 11, 50, 51, 128, 257])
def test_eigvals_vs_dense(n):
    A = skew_gauss(n, 77 + n)
    _, _, eta = oracle.skewblockedhess(A)
    lam = np.sort(oracle.skewtrieigvals(eta))
    ref = np.sort(np.linalg.eigvals(A).imag)
    assert np.max(np.abs(lam - ref)) <= 10 * n * EPS * np.linalg.norm(A, 2)
    # closed form: lambda(T) = +-sigma(C), C lower bidiagonal built from eta
    if n % 2 == 0:
        k = n // 2
        C = np.zeros((k, k))
        for j in range(k):
            C[j, j] = -eta[2 * j]
            if j > 0:
                C[j, j - 1] = eta[2 * j - 1]
        sv = np.linalg.svd(C, compute_uv=False)
        np.testing.assert_allclose(np.sort(np.abs(lam)), np.sort(np.concatenate([sv, sv])),
                                   atol=10 * n * EPS * np.linalg.norm(A, 2))


def test_eigvals_pairs_order():
    """Order of the unsorted output: deflated 2x2 blocks write (+i*e, -i*e) (skeweigen.jl:64-67)."""
    lam = oracle.skewtrieigvals([3.0, 0.0, 5.0])
    np.testing.assert_array_equal(lam, [3.0, -3.0, 5.0, -5.0])


def test_issue116_118(golden):
    for ev in golden["issue116_ev"] + golden["issue118_ev"]:
        ev = np.array(ev, float)
        T = tri(ev)
        a = np.sort(oracle.skewtrieigvals(ev))
        b = np.sort((np.linalg.eigvals(1j * T) / 1j).imag)
        np.testing.assert_allclose(a, b, atol=1e-9)
        # eigen(::SkewHermTridiagonal) runs the merged variant (src/tridiag.jl:506)
        vals, V = oracle.skewtrieigen(ev, "merged")
        R = V @ np.diag(1j * vals) @ V.conj().T - T
        assert np.linalg.norm(R) <= 1e-9 * max(1.0, np.linalg.norm(T))
    # issue #118 also runs eigen(SkewHermitian(Matrix(A))) -> hessenberg + divided variant
    for ev in golden["issue118_ev"]:
        T = tri(np.array(ev))
        F, tau, eta = oracle.skewblockedhess(T)
        Q = check_factorization(T, F, tau, eta)
        vals, V = oracle.skewtrieigen(eta, "divided")
        X = Q @ V
        R = X @ np.diag(1j * vals) @ X.conj().T - T
        assert np.linalg.norm(R) <= 1e-9 * np.linalg.norm(T)


def test_divided_initial_mem_quirk():
    """skewtrieigen_divided! starts with mem = 1 (src/skeweigen.jl:343) and deflates when
    |ev[n-1] - mem| < tol (:355-358), so an exact ev[n-1] == 1 after the first sweep deflates a
    block too early.  The restatement reproduces this; the merged variant does not have it."""
    vals_d, _ = oracle.skewtrieigen([0.0, 1.0, 1.0], "divided")
    vals_m, _ = oracle.skewtrieigen([0.0, 1.0, 1.0], "merged")
    np.testing.assert_allclose(np.sort(vals_d), [-1, 0, 0, 1], atol=1e-15)
    np.testing.assert_allclose(np.sort(vals_m), [-np.sqrt(2), 0, 0, np.sqrt(2)], atol=1e-15)


def test_issue130_eigen(golden):
    for key in ("m10", "m66"):
        A = sparse_to_dense(golden["issue130"][key])
        np.testing.assert_array_equal(A, -A.T)
        F, tau, eta = oracle.skewblockedhess(A)
        Q = check_factorization(A, F, tau, eta)
        vals, V = oracle.skewtrieigen(eta, "divided")
        X = Q @ V
        R = X @ np.diag(1j * vals) @ X.conj().T - A
        assert np.linalg.norm(R) <= 1e-10 * np.linalg.norm(A)
        np.testing.assert_allclose(X.conj().T @ X, np.eye(A.shape[0]), atol=1e-10)


@pytest.mark.parametrize("n", [1, 2, 5, 8, 31, 64])
def test_eigen_random(n):
    A = skew_gauss(n, 5 + n)
    if n == 1:
        return
    F, tau, eta = oracle.skewblockedhess(A)
    Q = oracle.form_q(F, tau)
    for variant in ("divided", "merged"):
        vals, V = oracle.skewtrieigen(eta.copy(), variant)
        X = Q @ V
        R = X @ np.diag(1j * vals) @ X.conj().T - A
        assert np.linalg.norm(R) <= 1e-12 * n * np.linalg.norm(A)


def test_complex_factorization():
    for n in [2, 3, 4, 7, 12, 64, 130]:
        A = skew_gauss(n, 40 + n, np.complex128)
        F, tau, eta = oracle.skewblockedhess(A)
        check_factorization(A, F, tau, eta)
        # LAPACK zgehrd agrees on the first n-2 subdiagonal entries (it leaves the last complex)
        H = sl.hessenberg(A)
        if n > 2:
            np.testing.assert_allclose(eta[:-1], np.diag(H, -1)[:-1].real, atol=1e-12 * np.linalg.norm(A))
            assert np.max(np.abs(np.diag(H, -1)[:-1].imag)) < 1e-12 * np.linalg.norm(A)
        np.testing.assert_allclose(abs(eta[-1]), abs(np.diag(H, -1)[-1]), atol=1e-12 * np.linalg.norm(A))
        # eigenvalues of T (imaginary diagonal + real subdiagonal) match the dense ones
        Tm = np.diag(eta.astype(complex), -1) - np.diag(eta.astype(complex), 1) + np.diag(1j * np.imag(np.diag(F)))
        np.testing.assert_allclose(np.sort(np.linalg.eigvalsh(1j * Tm)), np.sort(np.linalg.eigvalsh(1j * A)),
                                   atol=1e-12 * n * np.linalg.norm(A))


def test_without_blas_matches():
    """The plain-C path and the OpenBLAS path agree (guards the BLAS call-shape restatement)."""
    A = skew_gauss(150, 3)
    L = oracle.lib()
    _, tau1, eta1 = oracle.skewblockedhess(A)
    L.orc_blas_unload()
    try:
        F, tau2, eta2 = oracle.skewblockedhess(A)
        l2 = oracle.skewtrieigvals(eta2)
    finally:
        L.orc_blas_load(oracle.openblas_path().encode())
    np.testing.assert_allclose(eta2, eta1, atol=1e-12 * np.linalg.norm(A))
    np.testing.assert_allclose(np.sort(l2), np.sort(oracle.skewtrieigvals(eta1)), atol=1e-11)


def test_planted_mistakes():
    """A plausible slip in the restatement is caught by the invariants used above.

    The slips are planted in a numpy copy of the unblocked step (the oracle itself is not
    modified): each one breaks Q T Q^T = A by far more than the tolerance."""
    A = skew_gauss(12, 11)

    def unblocked(A, slip=None):
        A = A.copy()
        n = A.shape[0]
        taus, etas = np.zeros(n - 1), np.zeros(n - 1)
        for i in range(n - 2):
            x = A[i + 1:, i].copy()
            al, xn = x[0], np.linalg.norm(x[1:])
            b = (-1 if al > 0 else 1) * np.hypot(al, xn)
            if slip == "sign":
                b = -b
            t = 1 - al / b
            v = np.concatenate([[1.0], x[1:] / (al - b)])
            w = t * (A[i + 1:, i + 1:] @ v)
            if slip == "transpose":
                w = t * (A[i + 1:, i + 1:].T @ v)
            upd = np.outer(v, w) - np.outer(w, v)
            if slip == "drop":
                upd = np.outer(v, w)
            A[i + 1:, i + 1:] += upd
            A[i + 1:, i] = v
            taus[i], etas[i] = t, b
        etas[n - 2] = A[n - 1, n - 2]
        return A, taus, etas

    def full_check(F, tau, eta):
        check_factorization(A, F, tau, eta)
        np.testing.assert_allclose(eta, np.diag(sl.hessenberg(A), -1), atol=1e-12 * np.linalg.norm(A))

    full_check(*unblocked(A))
    for slip in ("sign", "transpose", "drop"):
        with pytest.raises(AssertionError):
            full_check(*unblocked(A, slip))


# ---- complex skew-symmetric Pfaffian (Parlett-Reid, src/pfaffian.jl:95-137, 193-233) ----

def _complex_case(c):
    return np.array(c["re"]) + 1j * np.array(c["im"])


def _exact(c):
    return complex(float(Fraction(c["pf"][0])), float(Fraction(c["pf"][1])))


def test_complex_pfaffian_exact(golden):
    """Oracle against exact Gaussian-rational Pfaffians (the reference's own test inputs, plus
    cases that pivot and a singular one)."""
    for c in golden["complex_pf"]:
        A = _complex_case(c)
        ex = _exact(c)
        pf = oracle.pfaffian(A)
        assert abs(pf - ex) <= 1e-12 * max(1.0, abs(ex)), (A.shape, pf, ex)
        lp, ph = oracle.logabspfaffian(A)
        if ex == 0:
            assert lp == -np.inf and ph == 0
        else:
            assert abs(lp - np.log(abs(ex))) <= 1e-12 * max(1.0, abs(np.log(abs(ex))))
            assert abs(ph - ex / abs(ex)) <= 1e-12


@pytest.mark.parametrize("n", [2, 8, 16, 33, 64])
def test_complex_pfaffian_det_and_congruence(n):
    """Pf(A)^2 = det(A), and Pf(B A B^T) = det(B) Pf(A) for any complex B."""
    rng = np.random.default_rng(500 + n)
    R = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = R - R.T
    pf = oracle.pfaffian(A)
    if n % 2:
        assert pf == 0
        return
    d = np.linalg.det(A)
    assert abs(pf * pf - d) <= 1e-12 * abs(d) * n
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    pfc = oracle.pfaffian(B @ A @ B.T)
    assert abs(pfc - np.linalg.det(B) * pf) <= 1e-11 * abs(pfc) * n


def _parlett_reid(A, flip_on_swap=True, transposed_update=False):
    """Plain-loop Parlett-Reid used only to show the complex fixtures catch plausible slips."""
    A = np.array(A, dtype=complex)
    n = A.shape[0]
    pf = 1.0 + 0j
    for k in range(0, n - 1, 2):
        kp = k + 1 + int(np.argmax(np.abs(A[k + 1:, k])))
        if kp != k + 1:
            A[[k + 1, kp], k:] = A[[kp, k + 1], k:]
            A[k:, [k + 1, kp]] = A[k:, [kp, k + 1]]
            if flip_on_swap:
                pf = -pf
        if A[k + 1, k] == 0:
            return 0j
        tau = A[k, k + 2:] / A[k, k + 1]
        pf *= A[k, k + 1]
        col = A[k + 2:, k + 1]
        upd = np.outer(tau, col) - np.outer(col, tau)
        A[k + 2:, k + 2:] += upd.T if transposed_update else upd
    return pf


def test_complex_pfaffian_planted_slips(golden):
    """A missing sign flip on a pivot swap, or the update applied transposed, fails the exact
    fixtures, while the plain restatement passes them."""
    def passes(fn):
        return all(abs(fn(_complex_case(c)) - _exact(c)) <= 1e-10 * max(1.0, abs(_exact(c)))
                   for c in golden["complex_pf"] if len(c["re"]) % 2 == 0)
    assert passes(_parlett_reid)
    assert not passes(lambda A: _parlett_reid(A, flip_on_swap=False))
    assert not passes(lambda A: _parlett_reid(A, transposed_update=True))


# ---------------------------------------------------------------------------------------------
# skew-Cholesky (src/cholesky.jl:28-107)
# ---------------------------------------------------------------------------------------------
def _skewchol_py(A, pivot=True, slip=None):
    """Pure-Python restatement of _skewchol! with the oracle's rounding order (small n only).
    slip plants one mistake: 'row_sign' drops the minus of the second new row, 'row_major_ties'
    resolves equal maxima in row-major order, 'rows_only' swaps rows but not columns."""
    B = np.array(A, dtype=np.float64)
    n = B.shape[0]
    p = list(range(n))
    if n <= 1:
        return np.triu(B), np.array(p), 0
    tol = 1e-15 * np.sqrt(sum(float(B[r, c]) ** 2 for c in range(n) for r in range(n))) if pivot else 0.0

    def swap(x, y):
        B[[x, y], :] = B[[y, x], :]
        if slip != "rows_only":
            B[:, [x, y]] = B[:, [y, x]]

    steps = 0
    for j in range(n // 2):
        i0, i1 = 2 * j, 2 * j + 1
        s = 1.0
        if pivot:
            order = [(r, c) for c in range(i0, n) for r in range(i0, n)]
            if slip == "row_major_ties":
                order = [(r, c) for r in range(i0, n) for c in range(i0, n)]
            ii, jj = order[0]
            for r, c in order:
                if B[r, c] > B[ii, jj]:
                    ii, jj = r, c
            if abs(B[ii, jj]) < tol:
                break
            kk = ii if jj == i0 else jj
            if ii != i0:
                p[ii], p[i0] = p[i0], p[ii]
                swap(ii, i0)
            if kk != i1:
                p[kk], p[i1] = p[i1], p[kk]
                swap(kk, i1)
        v = B[i0, i1]
        if pivot:
            r = np.sqrt(v)
        else:
            s = np.sign(v)
            r = np.sqrt(abs(v))
        q = -1.0 / r
        b0 = B[i0, i1 + 1:].copy()
        b1 = B[i1, i1 + 1:].copy()
        n0 = (b1 if pivot else s * b1) * q
        n1 = (b0 if slip == "row_sign" else -b0) * q
        B[i0, i0] = r
        B[i1, i1] = r if pivot else s * r
        B[i0, i1] = 0.0
        B[i0, i1 + 1:] = n0
        B[i1, i1 + 1:] = n1
        B[i1 + 1:, i1 + 1:] -= np.outer(n0, n1) - np.outer(n1, n0)
        steps += 1
    return np.triu(B), np.array(p), steps


def _chol_residual(A, R, p):
    n = A.shape[0]
    return np.abs(R.T @ oracle.jmatrix(n) @ R - A[np.ix_(p, p)]).max() / max(1.0, np.abs(A).max())


TIE4 = np.array([[0, 3, -5, 0], [-3, 0, 0, 5], [5, 0, 0, 5], [0, -5, -5, 0]], dtype=float)


@pytest.mark.parametrize("n", [1, 2, 10, 11])
@pytest.mark.parametrize("kind", ["int", "float"])
def test_skewchol_reference_cases(n, kind):
    """The reference's own test (test/runtests.jl:436-455): R^T J R = A[p, p] with RowMaximum,
    R^T J R = A with NoPivot, for skewhermitian(rand(-10:10) * 2) and skewhermitian(randn)."""
    rng = np.random.default_rng(40 + n)
    M = rng.integers(-10, 11, size=(n, n)).astype(float) * 2 if kind == "int" else rng.standard_normal((n, n))
    A = (M - M.T) / 2
    R, p, steps = oracle.skewchol(A)
    assert np.array_equal(R, np.triu(R)) and sorted(p) == list(range(n))
    assert _chol_residual(A, R, p) <= 1e-13
    R, p, steps = oracle.skewchol(A, pivot=False)
    assert np.array_equal(p, np.arange(n))
    assert _chol_residual(A, R, p) <= 1e-12


@pytest.mark.parametrize("a", [2.0, -3.0])
def test_skewchol_closed_forms(a):
    """n = 2: RowMaximum gives R = sqrt|a| I, with p swapped when a < 0; NoPivot gives
    R = diag(r, sign(a) r)."""
    A = np.array([[0.0, a], [-a, 0.0]])
    r = np.sqrt(abs(a))
    R, p, steps = oracle.skewchol(A)
    assert steps == 1 and np.array_equal(R, np.diag([r, r])) and list(p) == ([0, 1] if a > 0 else [1, 0])
    R, p, steps = oracle.skewchol(A, pivot=False)
    assert np.array_equal(R, np.diag([r, np.sign(a) * r])) and list(p) == [0, 1]


def test_skewchol_tie_rule():
    """The maximum 5 sits at (2,0), (1,3) and (2,3) of the full square; findmax takes the first
    in column-major order, (2,0): index 2 moves to position 0, then its partner 0 (now at 2) to
    position 1, so p starts [2, 0]."""
    R, p, steps = oracle.skewchol(TIE4)
    assert list(p[:2]) == [2, 0] and R[0, 0] == np.sqrt(5.0)
    assert _chol_residual(TIE4, R, p) <= 1e-15


def test_skewchol_rank_deficient():
    """Rank 4 in n = 10: the factorization stops after 2 steps (max below 1e-15 ||A||_F) and
    R^T J R still reproduces A[p, p]."""
    rng = np.random.default_rng(8)
    X = rng.standard_normal((10, 4))
    A = X @ oracle.jmatrix(4) @ X.T
    A = (A - A.T) / 2
    R, p, steps = oracle.skewchol(A)
    assert steps == 2
    assert _chol_residual(A, R, p) <= 1e-13


@pytest.mark.parametrize("pivot", [True, False])
def test_skewchol_restatement_matches(pivot):
    """The pure-Python restatement reproduces the C oracle bit for bit (same rounding order)."""
    rng = np.random.default_rng(3)
    for n in (2, 7, 12):
        M = rng.standard_normal((n, n))
        A = M - M.T
        R, p, s = oracle.skewchol(A, pivot)
        R2, p2, s2 = _skewchol_py(A, pivot)
        assert np.array_equal(R, R2) and np.array_equal(p, p2) and s == s2


def test_skewchol_planted_slips():
    """Each planted mistake fails a check the oracle passes."""
    rng = np.random.default_rng(9)
    M = rng.standard_normal((12, 12))
    A = M - M.T
    R, p, _ = _skewchol_py(A, slip="row_sign")
    assert not _chol_residual(A, R, p) <= 1e-3
    with np.errstate(invalid="ignore"):
        R, p, _ = _skewchol_py(A, slip="rows_only")
    assert not _chol_residual(A, R, p) <= 1e-3
    R, p, _ = _skewchol_py(TIE4, slip="row_major_ties")
    assert list(p[:2]) != [2, 0]


# ----------------------------------------------------------------------------------------
# matrix functions (src/exp.jl): the oracle's restatement against scipy's dense matrix functions,
# as the reference's own exp.jl test compares against Julia's dense exp / cos / sin / ... (test/runtests.jl:208-281)
# ----------------------------------------------------------------------------------------
def _dense_funs(A):
    inv = np.linalg.inv
    n = A.shape[0]
    ref = {"exp": sl.expm(A), "cis": sl.expm(1j * A), "cos": sl.cosm(A), "sin": sl.sinm(A),
           "sinh": sl.sinhm(A), "cosh": sl.coshm(A), "tan": sl.tanm(A), "tanh": sl.tanhm(A)}
    if np.iscomplexobj(A) or n % 2 == 0:
        ref["log"] = sl.logm(A)
        ref["cot"] = sl.cosm(A) @ inv(sl.sinm(A))
        ref["coth"] = sl.coshm(A) @ inv(sl.sinhm(A))
    return ref


def _rel(X, R):
    return np.abs(np.asarray(X) - R).max() / max(1.0, np.abs(R).max())


@pytest.mark.parametrize("n", [1, 2, 5, 10, 11, 32])
@pytest.mark.parametrize("kind", ["real", "int", "complex", "tri"])
def test_skewfun_vs_dense_functions(n, kind):
    rng = np.random.default_rng(40 + n)
    if kind == "tri":
        if n == 1:
            return
        ev = rng.standard_normal(n - 1) * 0.6
        A = tri(ev)
        got = lambda f: oracle.skewfun_tri(ev, f)   # noqa: E731
    else:
        if kind == "int":
            M = rng.integers(-10, 11, size=(n, n)).astype(float) * 2
            A = (M - M.T) / 2 / (10 * np.sqrt(n))   # a power-of-ten scale keeps cosh, tan well conditioned
        elif kind == "real":
            M = rng.standard_normal((n, n))
            A = (M - M.T) / (2 * np.sqrt(n))
        else:
            M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
            A = (M - M.conj().T) / (2 * np.sqrt(n))
        got = lambda f: oracle.skewfun(A, f)   # noqa: E731
    for f, R in _dense_funs(A).items():
        tol = 1e-12 if f not in ("cot", "coth") else 1e-10
        assert _rel(got(f), R) <= tol, (kind, n, f, _rel(got(f), R))
    S, C = got("sincos")
    assert _rel(S, sl.sinm(A)) <= 1e-12 and _rel(C, sl.cosm(A)) <= 1e-12


def test_skewfun_wrappers_and_domain():
    """skewhermitian! keeps the lower triangle's rounding, Hermitian reads the upper one; log of an
    odd-sized or singular real skew matrix is a DomainError (src/exp.jl:52, :68)."""
    X = np.array([[1.0, 2.0], [5.0, 3.0]])
    assert np.array_equal(oracle._skewhermitian_(X), [[0.0, -1.5], [1.5, 0.0]])
    assert np.array_equal(oracle._hermitian_view(X), [[1.0, 2.0], [2.0, 3.0]])
    Z = np.array([[1 + 2j, 3 + 1j], [5 - 4j, 2j]])
    assert np.array_equal(oracle._hermitian_view(Z), [[1, 3 + 1j], [3 - 1j, 0]])
    assert np.array_equal(oracle._skewhermitian_(Z), [[2j, -1 - 1.5j], [1 - 1.5j, 2j]])
    with pytest.raises(ValueError, match="DomainError"):
        oracle.skewfun(tri(np.ones(2)), "log")
    with pytest.raises(ValueError, match="DomainError"):
        oracle.skewfun_tri(np.array([1.0, 0.0, 0.0]), "log")


def test_skewfun_planted_slips():
    """A sign slip in the eigenvalue convention or swapped real / imaginary eigenvector parts fails the
    dense comparison that pins the restatement."""
    rng = np.random.default_rng(3)
    M = rng.standard_normal((8, 8))
    A = (M - M.T) / 4
    vals, Qr, Qim = oracle.skeweigen(A)
    assert _rel(oracle._fun_from_eigs_real(vals, Qr, Qim, "exp", 8), sl.expm(A)) <= 1e-13
    assert _rel(oracle._fun_from_eigs_real(-vals, Qr, Qim, "exp", 8), sl.expm(A)) > 1e-3
    assert _rel(oracle._fun_from_eigs_real(vals, Qim, Qr, "exp", 8), sl.expm(A)) > 1e-3
    assert _rel(oracle._fun_from_eigs_real(vals, Qr, Qim, "cis", 8), sl.expm(-1j * A)) > 1e-3
