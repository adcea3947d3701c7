"""GPU parity of the tridiagonalization path against the CPU oracle (calls go through the C ABI).

Bars (written in each test):
  * eta vs the oracle's eta within 1e-11 * ||A||_F (same reflectors; blocking reorders sums)
  * ||Q T Q^T - A||_F / ||A||_F <= 1e-13 and ||Q^T Q - I||_F <= 1e-13 * n
  * eigenvalues (sorted) within 10 * n * eps * ||A||_2 of the oracle's
  * log|Pf| relative error <= 1e-10 and the sign exactly equal
"""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle
from conftest import skew_gauss, sparse_to_dense

pytestmark = pytest.mark.gpu

skl = pytest.importorskip("skewlinearalgebra")
from skewlinearalgebra import _lib  # noqa: E402

EPS = np.finfo(float).eps


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tri(eta):
    return np.diag(eta, -1) - np.diag(eta, 1)


def check_fact(A, F, tau, eta, tol=1e-13):
    Q = oracle.form_q(F, tau)
    n = A.shape[0]
    nA = max(np.linalg.norm(A), 1e-300)
    res = np.linalg.norm(Q @ tri(eta) @ Q.T - A) / nA
    orth = np.linalg.norm(Q.T @ Q - np.eye(n))
    assert res <= tol, res
    assert orth <= tol * max(n, 1), orth
    return Q


def test_mfma_lane_layout():
    out = np.zeros(768)
    _lib.check("probe", _lib.lib().skh_diag_mfma_probe(out.ctypes.data))
    lanes = np.arange(64)
    d1 = out[:256].reshape(64, 4)
    d2 = out[256:512].reshape(64, 4)
    d3 = out[512:].reshape(64, 4)
    # accumulator item q of lane l holds D[4q + l/16][l%16] (the layout rank2k_tile relies on)
    rows = 4 * np.arange(4)[None, :] + (lanes // 16)[:, None]
    cols = np.repeat((lanes % 16)[:, None], 4, axis=1)
    print("probe rows (first lane of each group):", d1[::16].tolist(), "cols:", d2[:17:4].tolist(), flush=True)
    np.testing.assert_array_equal(d1, rows)
    np.testing.assert_array_equal(d2, cols)
    Am = np.array([[np.sin(i * 4 + k) for k in range(4)] for i in range(16)])
    Bm = np.array([[np.cos(k * 16 + j) for j in range(16)] for k in range(4)])
    D = Am @ Bm
    np.testing.assert_allclose(d3, D[rows, cols], rtol=1e-14, atol=1e-14)


def gpu_sytrd(A, nb=0, engine=skl.ENGINE_LARGE):
    F = np.array(A, dtype=np.float64, order="F", copy=True)
    n = F.shape[0]
    tau = np.zeros(max(n - 1, 0))
    eta = np.zeros(max(n - 1, 0))
    nz = ctypes.c_int64(0)
    o = _lib.opts(nb, engine)
    rc = _lib.lib().skh_sytrd_d(n, F.ctypes.data, n, tau.ctypes.data, eta.ctypes.data, ctypes.byref(nz),
                                ctypes.byref(o))
    _lib.check("skh_sytrd_d", rc)
    return F, tau, eta, nz.value


@pytest.mark.parametrize("nb", [8, 32, 64])
@pytest.mark.parametrize("n", [2, 3, 4, 5, 10, 63, 64, 65, 66, 127, 130, 200, 257])
def test_sytrd_large_vs_oracle(n, nb):
    say("sytrd large n=%d nb=%d" % (n, nb))
    A = skew_gauss(n, 10 * n + nb)
    F, tau, eta, nz = gpu_sytrd(A, nb)
    Fo, tauo, etao = oracle.skewblockedhess(A)
    nA = np.linalg.norm(A)
    np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-11 * nA)
    np.testing.assert_allclose(tau, tauo, rtol=0, atol=1e-11)
    assert nz == int(np.count_nonzero(tau))
    check_fact(A, F, tau, eta)
    # unit stored at the head of each reflector, last entry keeps eta (hessenberg.jl:27-29)
    for i in range(n - 2):
        assert F[i + 1, i] == 1.0
    assert F[n - 1, n - 2] == eta[-1] and tau[-1] == 0.0
    # the upper triangle and diagonal are left as given
    iu = np.triu_indices(n)
    np.testing.assert_array_equal(F[iu], A[iu])


@pytest.mark.parametrize("n", [2, 4, 6, 10, 62, 64, 66, 128, 130, 200, 256, 512, 1024])
def test_sytrd_compact_vs_oracle(n):
    """The compact engine (one workgroup per matrix, unblocked fused update + matvec)."""
    say("sytrd compact n=%d" % n)
    A = skew_gauss(n, 20 * n + 1)
    F, tau, eta, nz = gpu_sytrd(A, 0, skl.ENGINE_TEAM1)
    Fo, tauo, etao = oracle.skewblockedhess(A)
    nA = np.linalg.norm(A)
    np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-11 * nA)
    np.testing.assert_allclose(tau, tauo, rtol=0, atol=1e-11)
    assert nz == int(np.count_nonzero(tau))
    check_fact(A, F, tau, eta)
    for i in range(n - 2):
        assert F[i + 1, i] == 1.0
    assert F[n - 1, n - 2] == eta[-1] and tau[-1] == 0.0
    iu = np.triu_indices(n)
    np.testing.assert_array_equal(F[iu], A[iu])


def test_compact_edge_conventions(golden):
    say("compact edge conventions")
    cases = [np.array(golden["readme"]["A"], float)]
    A = np.zeros((4, 4))
    A[1:, 0] = 1
    A[0, 1:] = -1
    cases.append(A)
    A = np.zeros((6, 6))
    A[1, 0], A[0, 1] = 3.0, -3.0        # alpha != 0, xnorm = 0 -> tau = 2
    A[5, 2], A[2, 5] = 2.0, -2.0        # alpha = +0 with xnorm > 0 -> beta = +xnorm
    cases.append(A)
    cases.append(np.zeros((8, 8)))      # all-zero columns -> tau = 0, v = 0
    for A in cases:
        F, tau, eta, nz = gpu_sytrd(A, 0, skl.ENGINE_TEAM1)
        Fo, tauo, etao = oracle.skewblockedhess(A)
        np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-12 * max(1.0, np.linalg.norm(A)))
        np.testing.assert_array_equal(tau != 0, tauo != 0)
        np.testing.assert_allclose(tau, tauo, rtol=0, atol=1e-12)
        check_fact(A, F, tau, eta)
    A = sparse_to_dense(golden["issue130"]["m66"])
    F, tau, eta, nz = gpu_sytrd(A, 0, skl.ENGINE_TEAM1)
    check_fact(A, F, tau, eta)
    g = golden["readme"]
    F, tau, eta, nz = gpu_sytrd(np.array(g["A"], float), 0, skl.ENGINE_TEAM1)
    np.testing.assert_allclose(eta, g["eta"], rtol=1e-14)
    np.testing.assert_allclose(oracle.form_q(F, tau), np.array(g["Q"]), atol=1e-14)


@pytest.mark.parametrize("n", [512, 1000, 2049])
def test_sytrd_medium(n):
    say("sytrd medium n=%d" % n)
    A = skew_gauss(n, n)
    F, tau, eta, nz = gpu_sytrd(A)
    _, _, etao = oracle.skewblockedhess(A)
    nA = np.linalg.norm(A)
    np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-11 * nA)
    check_fact(A, F, tau, eta)


def degenerate_skew(n, seed):
    """Skew matrices whose reflectors hit the householder! special cases inside multi-column panels:
    an already tridiagonal one (every tail zero, tau = 2 where alpha != 0), a block-diagonal one
    (zero coupling at the block edges: tau = 0) and a tridiagonal one with a dense trailing block
    (zero tails up to the block, then full reflectors in the same panels).  Zeroed rows and columns
    are avoided: they make the matrix singular, and past a numerically zero eta the reflectors are
    fixed by rounding noise (the issue-130 case below)."""
    rng = np.random.default_rng(seed)
    e = rng.standard_normal(n - 1)
    e[rng.integers(0, n - 1, size=max(1, n // 50))] = 0.0
    T = np.diag(e, -1) - np.diag(e, 1)
    B = np.zeros((n, n))
    r = 0
    for size in rng.integers(1, 140, size=n):
        if r >= n:
            break
        m = int(min(size, n - r))
        G = rng.standard_normal((m, m))
        B[r:r + m, r:r + m] = G - G.T
        r += m
    e2 = rng.standard_normal(n - 1)
    Z = np.diag(e2, -1) - np.diag(e2, 1)
    h = n // 2 + 7
    Z[h:, h:] += skew_gauss(n - h, seed + 1)
    return [T, B, Z]


@pytest.mark.parametrize("n,nb", [(700, 32), (1000, 64), (513, 128)])
def test_sytrd_degenerate_columns_large(n, nb):
    """Zero reflector tails and zero columns in the middle of the large engine's panels (the matvec
    runs on the unscaled column and its consumer applies scale and corr = unit - scale * alpha, which
    must give the reference's tau = 2 and tau = 0 reflectors): the oracle's eta, the same zero
    pattern of tau, and the factorization."""
    for k, A in enumerate(degenerate_skew(n, 40 + n)):
        say("sytrd degenerate n=%d nb=%d case %d" % (n, nb, k))
        F, tau, eta, nz = gpu_sytrd(A, nb)
        Fo, tauo, etao = oracle.skewblockedhess(A)
        nA = max(1.0, np.linalg.norm(A))
        np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-11 * nA)
        np.testing.assert_array_equal(tau != 0, tauo != 0)
        np.testing.assert_allclose(tau, tauo, rtol=0, atol=1e-11)
        assert nz == int(np.count_nonzero(tau))
        check_fact(A, F, tau, eta)


def test_edge_conventions(golden):
    say("edge conventions")
    cases = [np.array(golden["issue121"]["A"], float), np.array(golden["readme"]["A"], float)]
    A = np.zeros((4, 4))
    A[1:, 0] = 1
    A[0, 1:] = -1
    cases.append(A)
    A = np.zeros((6, 6))
    A[1, 0], A[0, 1] = 3.0, -3.0        # alpha != 0, xnorm = 0 -> tau = 2
    A[5, 2], A[2, 5] = 2.0, -2.0        # alpha = +0 with xnorm > 0 -> beta = +xnorm
    cases.append(A)
    cases.append(np.zeros((7, 7)))      # all-zero columns -> tau = 0, v = 0
    for A in cases:
        for nb in (1, 2, 3, 8):
            F, tau, eta, nz = gpu_sytrd(A, nb)
            Fo, tauo, etao = oracle.skewblockedhess(A)
            np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-12 * max(1.0, np.linalg.norm(A)))
            np.testing.assert_array_equal(tau != 0, tauo != 0)
            np.testing.assert_allclose(tau, tauo, rtol=0, atol=1e-12)
            check_fact(A, F, tau, eta)
    # issue #130 matrices are numerically singular (an eta of ~1e-16 appears): past that pivot the
    # reflectors are fixed by rounding noise, so only the factorization itself and the spectrum
    # are comparable with the oracle
    for key in ("m10", "m66"):
        A = sparse_to_dense(golden["issue130"][key])
        for nb in (1, 8, 32):
            F, tau, eta, nz = gpu_sytrd(A, nb)
            check_fact(A, F, tau, eta)
            _, _, etao = oracle.skewblockedhess(A)
            np.testing.assert_allclose(np.sort(np.abs(oracle.skewtrieigvals(eta))),
                                       np.sort(np.abs(oracle.skewtrieigvals(etao))), atol=1e-12 * np.linalg.norm(A))
    g = golden["readme"]
    F, tau, eta, nz = gpu_sytrd(np.array(g["A"], float))
    np.testing.assert_allclose(eta, g["eta"], rtol=1e-14)
    np.testing.assert_allclose(oracle.form_q(F, tau), np.array(g["Q"]), atol=1e-14)


@pytest.mark.parametrize("n", [2, 4, 10, 64, 130, 512])
def test_logabspfaffian_vs_oracle(n):
    say("logpf n=%d" % n)
    A = skew_gauss(n, 500 + n)
    lp, sg = skl.logabspfaffian(A)
    lpo, sgo = oracle.logabspfaffian(A)
    assert sg == sgo
    assert abs(lp - lpo) <= 1e-10 * max(1.0, abs(lpo))


def test_pfaffian_exact(golden):
    say("exact pfaffians")
    from fractions import Fraction
    cases = golden["integer_pf"] + golden["analytic4"] + [golden["issue49"], golden["issue121"]]
    for c in cases:
        A = np.array(c["A"], float)
        pf = Fraction(c["pfaffian"])
        n = A.shape[0]
        lp, sg = skl.logabspfaffian(A)
        got = skl.pfaffian(A)
        if n % 2 == 1:
            assert (lp, sg, got) == (-np.inf, 0.0, 0.0)
            continue
        if pf == 0:
            continue
        assert sg == float(np.sign(float(pf)))
        assert abs(lp - np.log(abs(float(pf)))) <= 1e-10 * max(1.0, abs(lp))
        assert abs(got - float(pf)) <= 1e-10 * abs(float(pf))
    for c in golden["constructed"]:
        lp, sg = skl.logabspfaffian(np.array(c["A"]))
        assert sg == c["sign"]
        assert abs(lp - c["logpf"]) <= 1e-10 * max(1.0, abs(c["logpf"]))
    lp, sg = skl.logabspfaffian(np.array(golden["readme"]["A"], float))   # docs/src/pfaffian.md:55-57
    assert sg == -1.0 and abs(lp - 2.1972245773362196) <= 4 * EPS * 2.2


@pytest.mark.parametrize("n", [4, 11, 128, 512, 513])
def test_eigvals_vs_oracle(n):
    say("eigvals n=%d" % n)
    A = skew_gauss(n, 60 + n)
    lam = np.sort(skl.eigvals(A).imag)
    _, _, etao = oracle.skewblockedhess(A)
    lamo = np.sort(oracle.skewtrieigvals(etao))
    assert np.max(np.abs(lam - lamo)) <= 10 * n * EPS * np.linalg.norm(A, 2)


def test_batched_vs_single():
    say("batched logpf")
    rng = np.random.default_rng(4)
    sizes = [0, 1, 2, 3, 4, 6, 31, 32, 64, 66, 96, 128, 130, 200, 256, 257, 320, 512] + list(rng.integers(1, 40, 20) * 2)
    mats = []
    for k, n in enumerate(sizes):
        B = rng.standard_normal((n, n)) / np.sqrt(max(n, 1))
        mats.append(np.asfortranarray(B - B.T))
    lp, sg = skl.logabspfaffian_batched(mats)
    for k, M in enumerate(mats):
        n = M.shape[0]
        if n == 0:
            assert (lp[k], sg[k]) == (0.0, 1.0)
            continue
        lpo, sgo = oracle.logabspfaffian(M)
        assert sg[k] == sgo, (k, n)
        if n % 2 == 1:
            assert lp[k] == -np.inf
        else:
            assert abs(lp[k] - lpo) <= 1e-10 * max(1.0, abs(lpo)), (k, n, lp[k], lpo)


def test_batched_large_members():
    """Matrices above the compact engine's size limit go through the large engine in the same call."""
    say("batched with n > 1024 members")
    rng = np.random.default_rng(5)
    mats = []
    for n in (1026, 64, 1030, 7, 250):
        B = rng.standard_normal((n, n)) / np.sqrt(n)
        mats.append(B - B.T)
    lp, sg = skl.logabspfaffian_batched(mats)
    for k, M in enumerate(mats):
        lpo, sgo = oracle.logabspfaffian(M)
        assert sg[k] == sgo
        if M.shape[0] % 2:
            assert lp[k] == -np.inf
        else:
            assert abs(lp[k] - lpo) <= 1e-10 * max(1.0, abs(lpo))


@pytest.mark.parametrize("n", [128, 256])
def test_batched_equal_sizes_deterministic(n):
    """Same matrices in different batch positions and workgroups give bit-identical results."""
    say("batched equal sizes n=%d" % n)
    rng = np.random.default_rng(n)
    mats = []
    for _ in range(6):
        B = rng.standard_normal((n, n)) / np.sqrt(n)
        mats.append(B - B.T)
    lp1, sg1 = skl.logabspfaffian_batched(mats + mats[::-1])
    np.testing.assert_array_equal(lp1[:6], lp1[6:][::-1])
    np.testing.assert_array_equal(sg1[:6], sg1[6:][::-1])
    for k, M in enumerate(mats):
        lpo, sgo = oracle.logabspfaffian(M)
        assert sg1[k] == sgo and abs(lp1[k] - lpo) <= 1e-10 * max(1.0, abs(lpo))


@pytest.mark.parametrize("n", [2, 3, 5, 8, 65, 130, 300])
def test_eigen_vs_oracle(n):
    say("eigen n=%d" % n)
    A = skew_gauss(n, 700 + n)
    E = skl.eigen(A)
    X = E.vectors
    lam = E.values
    nA = np.linalg.norm(A)
    assert np.linalg.norm(X @ np.diag(lam) @ X.conj().T - A) <= 1e-12 * n * nA
    assert np.linalg.norm(X.conj().T @ X - np.eye(n)) <= 1e-12 * n
    Fo, tauo, etao = oracle.skewblockedhess(A)
    vo, Vo = oracle.skewtrieigen(etao, "divided")
    np.testing.assert_allclose(np.sort(lam.imag), np.sort(vo), atol=10 * n * EPS * np.linalg.norm(A, 2))


def test_eigen_issue130_118(golden):
    say("eigen fixtures")
    mats = [sparse_to_dense(golden["issue130"][k]) for k in ("m10", "m66")]
    for ev in golden["issue118_ev"]:
        mats.append(tri(np.array(ev)))
    for A in mats:
        E = skl.eigen(A)
        X = E.vectors
        assert np.linalg.norm(X @ np.diag(E.values) @ X.conj().T - A) <= 1e-9 * np.linalg.norm(A)


def test_orgtr_device_vs_numpy():
    say("orgtr")
    import torch
    n = 333
    A = skew_gauss(n, 3)
    F, tau, eta, _ = gpu_sytrd(A)
    dF = torch.tensor(F.T.copy(), device="cuda")   # row-major (n x n) tensor of F^T = column-major F
    dtau = torch.tensor(tau, device="cuda")
    dQ = torch.empty((n, n), dtype=torch.float64, device="cuda")
    rc = _lib.lib().skh_orgtr_d_device(n, dF.data_ptr(), n, dtau.data_ptr(), dQ.data_ptr(), n,
                                       torch.cuda.current_stream().cuda_stream)
    _lib.check("skh_orgtr_d_device", rc)
    torch.cuda.synchronize()
    Q = dQ.cpu().numpy().T
    np.testing.assert_allclose(Q, oracle.form_q(F, tau), atol=1e-13 * n)


def _constructed(n, rng):
    Z = rng.standard_normal((n, n))
    Q, R = np.linalg.qr(Z)
    Q = Q * np.sign(np.diag(R))
    s = np.exp(rng.uniform(np.log(0.1), np.log(10.0), size=n // 2))
    Bm = np.zeros((n, n))
    for k in range(n // 2):
        Bm[2 * k, 2 * k + 1] = s[k]
        Bm[2 * k + 1, 2 * k] = -s[k]
    A = Q @ Bm @ Q.T
    A = 0.5 * (A - A.T)
    return A, float(np.sum(np.log(s))), float(np.sign(np.linalg.det(Q)))


def test_grad_batched_vs_inverse_and_fd():
    """Gradient bar: vs -1/2 inv(A) (LU, the reference's inv(A::SkewHermitian) route) within
    1e-10 relative (Frobenius), and vs central differences h = 1e-6 on skew pairs within 1e-6."""
    say("grad batched")
    rng = np.random.default_rng(11)
    mats, ref = [], []
    for n in (2, 4, 6, 16, 64, 256, 100):
        A, lpf, sgn = _constructed(n, rng)
        mats.append(A)
        ref.append((lpf, sgn))
    lp, sg, grads = skl.logabspfaffian_grad_batched(mats)
    for k, A in enumerate(mats):
        n = A.shape[0]
        assert sg[k] == ref[k][1]
        assert abs(lp[k] - ref[k][0]) <= 1e-10 * max(1.0, abs(ref[k][0]))
        G0 = -0.5 * np.linalg.inv(A)
        assert np.linalg.norm(grads[k] - G0) <= 1e-10 * np.linalg.norm(G0), (n, np.linalg.norm(grads[k] - G0))
        # central differences on a few skew pairs: d/dh log|Pf(A + h(E_ij - E_ji))| = 2 * grad_ij
        for (i, j) in [(1, 0), (n - 1, 0), (n - 1, n - 2)]:
            if i <= j:
                continue
            h = 1e-6
            Ap = A.copy(); Ap[i, j] += h; Ap[j, i] -= h
            Am = A.copy(); Am[i, j] -= h; Am[j, i] += h
            fd = (oracle.logabspfaffian(Ap)[0] - oracle.logabspfaffian(Am)[0]) / (2 * h)
            assert abs(fd - 2 * grads[k][i, j]) <= 1e-6 * max(1.0, abs(fd)), (n, i, j, fd, 2 * grads[k][i, j])


def test_grad_mixed_and_odd():
    say("grad mixed sizes")
    rng = np.random.default_rng(12)
    mats = []
    for n in (3, 8, 8, 10, 0, 1, 32):
        B = rng.standard_normal((n, n))
        mats.append(B - B.T)
    lp, sg, grads = skl.logabspfaffian_grad_batched(mats)
    for k, A in enumerate(mats):
        n = A.shape[0]
        if n % 2 == 1:
            assert lp[k] == -np.inf and sg[k] == 0.0 and not np.any(grads[k])
            continue
        if n == 0:
            continue
        G0 = -0.5 * np.linalg.inv(A)
        assert np.linalg.norm(grads[k] - G0) <= 1e-9 * np.linalg.norm(G0)


def test_autograd_function():
    say("autograd")
    import torch
    rng = np.random.default_rng(13)
    As = np.stack([_constructed(32, rng)[0] for _ in range(5)])
    A = torch.tensor(As, device="cuda", requires_grad=True)
    lp, sg = skl.LogAbsPfaffian.apply(A)
    w = torch.arange(1, 6, dtype=torch.float64, device="cuda")
    (lp * w).sum().backward()
    G = A.grad.cpu().numpy()
    for k in range(5):
        np.testing.assert_allclose(G[k], (k + 1) * -0.5 * np.linalg.inv(As[k]), rtol=0,
                                   atol=1e-10 * np.abs(np.linalg.inv(As[k])).max() * (k + 1))
        assert abs(lp[k].item() - oracle.logabspfaffian(As[k])[0]) <= 1e-10 * max(1.0, abs(lp[k].item()))


@pytest.mark.parametrize("nb", [32, 64])
def test_sytrd_n4096_vs_oracle(nb):
    """Large enough that every workgroup of the panel grid exchanges data each column."""
    say("sytrd n=4096 nb=%d" % nb)
    n = 4096
    A = skew_gauss(n, 4096 + nb)
    F, tau, eta, nz = gpu_sytrd(A, nb)
    oracle.set_threads(16)
    _, tauo, etao = oracle.skewblockedhess(A)
    nA = np.linalg.norm(A)
    np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-11 * nA)
    np.testing.assert_allclose(tau, tauo, rtol=0, atol=1e-10)
    assert abs(np.sqrt(2 * np.sum(eta ** 2)) - nA) <= 1e-13 * nA


def test_plan_tune_and_graph_replay():
    """A batched plan gives the same bits as the one-shot device call, after tuning and when its
    launch is captured in a HIP graph and replayed (eager work and synchronisation in between)."""
    say("plan + graph replay")
    import torch
    rng = np.random.default_rng(21)
    sizes = [0, 3, 2, 64, 96, 130, 256, 510, 512, 31] + list(rng.integers(1, 40, 30) * 2)
    mats = []
    for n in sizes:
        B = rng.standard_normal((n, n)) / np.sqrt(max(n, 1))
        mats.append(np.asfortranarray(B - B.T))
    ns = np.array(sizes, dtype=np.int64)
    offs = np.zeros(len(sizes), dtype=np.int64)
    tot = 0
    for k, n in enumerate(sizes):
        offs[k] = tot
        tot += (n * n + 1) & ~1
    host = np.zeros(max(tot, 1))
    for k, M in enumerate(mats):
        host[offs[k]:offs[k] + M.size] = M.ravel(order="F")
    dA = torch.tensor(host, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    lp0 = torch.empty(len(sizes), dtype=torch.float64, device="cuda")
    sg0 = torch.empty_like(lp0)
    o = _lib.opts()
    _lib.check("device", _lib.lib().skh_logpf_batched_d_device(len(sizes), ns.ctypes.data, offs.ctypes.data,
                                                               dA.data_ptr(), lp0.data_ptr(), sg0.data_ptr(), st,
                                                               ctypes.byref(o)))
    torch.cuda.synchronize()
    for k, M in enumerate(mats):
        n = M.shape[0]
        lpo, sgo = (0.0, 1.0) if n == 0 else oracle.logabspfaffian(M)
        assert sg0[k].item() == sgo
        if n % 2 == 0 and n > 0:
            assert abs(lp0[k].item() - lpo) <= 1e-10 * max(1.0, abs(lpo))
    plan = skl.LogPfPlan(ns, offs)
    lp = torch.empty_like(lp0)
    sg = torch.empty_like(lp0)
    plan.tune(dA.data_ptr(), lp.data_ptr(), sg.data_ptr(), st)
    plan.status(st)
    assert plan.info()["tuned"]
    assert torch.equal(lp, lp0) and torch.equal(sg, sg0)
    lp.zero_()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run(dA.data_ptr(), lp.data_ptr(), sg.data_ptr(), side.cuda_stream)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run(dA.data_ptr(), lp.data_ptr(), sg.data_ptr(), torch.cuda.current_stream().cuda_stream)
    for r in range(3):
        lp.fill_(-1.0)
        sg.fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        plan.status(st)
        assert torch.equal(lp, lp0) and torch.equal(sg, sg0), r
    plan.close()


def test_eigvals_ranges(golden):
    """eigvals(A, vl, vh) and eigvals(A, irange) (src/eigen.jl:8-16,57-73) after the GPU
    tridiagonalization: the README values (test/runtests.jl:24-25) and a random n = 130 matrix
    against the same selection on the oracle's eta."""
    say("eigvals ranges")
    g = golden["readme"]
    A = np.array(g["A"], float)
    l1, l2 = g["eig_pos"]
    np.testing.assert_allclose(skl.eigvals(A, vl=0, vh=15).imag, [l1, l2], rtol=1e-12)
    np.testing.assert_allclose(skl.eigvals(A, irange=(1, 3)).imag, [l1, l2, -l2], rtol=1e-12)
    assert skl.eigvals(np.zeros((1, 1)), irange=(1, 1)).size == 1
    assert skl.eigvals(np.zeros((1, 1)), vl=1.0, vh=2.0).size == 0
    n = 130
    A = skew_gauss(n, 130)
    _, _, etao = oracle.skewblockedhess(A)
    H = np.diag(etao, -1) + np.diag(etao, 1)
    lam = np.linalg.eigvalsh(H)
    tol = 10 * n * EPS * np.linalg.norm(A, 2)
    np.testing.assert_allclose(skl.eigvals(A, irange=(10, 60)).imag, -lam[9:60], atol=tol)
    sel = lam[(lam > -2.0) & (lam <= 1.0)]
    got = skl.eigvals(A, vl=-1.0, vh=2.0).imag
    assert got.size == sel.size
    np.testing.assert_allclose(got, -sel, atol=tol)


def test_complex_eigvals_value_range():
    say("complex eigvals (vl, vh)")
    rng = np.random.default_rng(5)
    n = 64
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = (B - B.conj().T) / 2
    lam = np.linalg.eigvalsh(1j * A)
    got = skl.eigvals(A, vl=-2.0, vh=3.0).imag
    ref = -lam[(lam > -3.0) & (lam <= 2.0)]
    assert got.size == ref.size
    np.testing.assert_allclose(got, ref, atol=10 * n * EPS * np.linalg.norm(A, 2))


@pytest.mark.parametrize("n,cap", [(1000, 4), (2049, 16)])
def test_sytrd_small_grid_several_blocks(n, cap):
    """A capped panel grid makes workgroups own several 64-row blocks (the load-based row
    pass beside the LDS-resident one); the factorization must not change beyond rounding."""
    say("sytrd small grid n=%d cap=%d" % (n, cap))
    A = skew_gauss(n, 77)
    os.environ["SKH_PANEL_GRID_CAP"] = str(cap)
    try:
        F, tau, eta, nz = gpu_sytrd(A, 64)
    finally:
        del os.environ["SKH_PANEL_GRID_CAP"]
    _, _, etao = oracle.skewblockedhess(A)
    np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-11 * np.linalg.norm(A))
    check_fact(A, F, tau, eta)


@pytest.mark.parametrize("n,nb", [(1000, 32), (2049, 64), (700, 128)])
def test_sytrd_cooperative_launch_same_bits(n, nb):
    """The panel kernels' cooperative launch (SKH_PANEL_COOPERATIVE=1) and the default plain launch
    run the same grid, so the factorization must agree bit for bit."""
    say("sytrd cooperative launch n=%d nb=%d" % (n, nb))
    A = skew_gauss(n, 91)
    F0, tau0, eta0, _ = gpu_sytrd(A, nb)
    os.environ["SKH_PANEL_COOPERATIVE"] = "1"
    try:
        F1, tau1, eta1, _ = gpu_sytrd(A, nb)
    finally:
        del os.environ["SKH_PANEL_COOPERATIVE"]
    assert np.array_equal(F0, F1) and np.array_equal(tau0, tau1) and np.array_equal(eta0, eta1)
    _, _, etao = oracle.skewblockedhess(A)
    np.testing.assert_allclose(eta1, etao, rtol=0, atol=1e-11 * np.linalg.norm(A))


def tri_fourth_moment(e):
    """tr(T^4) = ||T^2||_F^2 for the skew tridiagonal T with subdiagonal e: T^2 is symmetric
    pentadiagonal with (T^2)_ii = -(e_{i-1}^2 + e_i^2) and (T^2)_{i,i+2} = e_i e_{i+1}."""
    e2 = np.concatenate(([0.0], e * e, [0.0]))
    d = e2[:-1] + e2[1:]
    off = e[:-1] * e[1:]
    return float(np.sum(d * d) + 2.0 * np.sum(off * off))


@pytest.mark.parametrize("nb", [32, 64, 128])
def test_sytrd_n8192_invariants(nb):
    """Full-size factorization, the whole panel grid (512 workgroups) in the first panels.
    A = Q T Q^T fixes ||T||_F = ||A||_F, tr(T^4) = ||A^2||_F^2 and log|det A| = 2 log|Pf T|;
    the right-hand sides come from A itself on the GPU (a product and an LU), so a reflector
    built from stale or misplaced data fails one of them."""
    torch = pytest.importorskip("torch")
    say("sytrd n=8192 nb=%d invariants" % nb)
    n = 8192
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(8192 + nb)
    B = torch.randn((n, n), generator=g, dtype=torch.float64, device=dev)
    A = B - B.T
    del B
    work = A.T.contiguous()   # column-major storage of A
    tau = torch.empty(n, dtype=torch.float64, device=dev)
    eta = torch.empty(n, dtype=torch.float64, device=dev)
    o = _lib.opts(nb, skl.ENGINE_LARGE)
    st = torch.cuda.current_stream().cuda_stream
    rc = _lib.lib().skh_sytrd_d_device(n, work.data_ptr(), n, tau.data_ptr(), eta.data_ptr(), st, ctypes.byref(o))
    _lib.check("skh_sytrd_d_device", rc)
    _lib.check("skh_device_status", _lib.lib().skh_device_status(st))   # no bounded wait abandoned
    torch.cuda.synchronize()
    del work
    e = eta[: n - 1].cpu().numpy()
    nA = float(torch.linalg.norm(A))
    assert abs(np.sqrt(2.0 * np.sum(e * e)) - nA) <= 1e-13 * nA
    A2 = A @ A
    m4 = float(torch.sum(A2 * A2))
    del A2
    assert abs(tri_fourth_moment(e) - m4) <= 1e-12 * m4, (tri_fourth_moment(e), m4)
    _, logdet = torch.linalg.slogdet(A)
    lp = float(np.sum(np.log(np.abs(e[0::2]))))
    assert abs(2.0 * lp - float(logdet)) <= 1e-10 * abs(float(logdet)), (2.0 * lp, float(logdet))


@pytest.mark.parametrize("nb", [32, 64])
def test_sytrd_repeatable_and_clock_neutral(nb):
    """The large engine hands partial sums between workgroups through polled slots and
    parity buffers: every result must be independent of timing.  Three factorizations of the same
    matrix (one with the per-kernel events and the phase clock on) give bit-identical factors,
    tau and eta."""
    torch = pytest.importorskip("torch")
    say("sytrd repeatability n=3000 nb=%d" % nb)
    n = 3000
    A = skew_gauss(n, 3000 + nb)
    dev = torch.device("cuda:0")
    A0 = torch.tensor(A.T.copy(), device=dev)   # row-major tensor of A^T = column-major A
    st = torch.cuda.current_stream().cuda_stream
    outs = []
    for flags in (skl.ENGINE_LARGE, skl.ENGINE_LARGE, skl.ENGINE_LARGE | _lib.FLAG_TIMING | _lib.FLAG_PHASES):
        work = A0.clone()
        tau = torch.empty(n, dtype=torch.float64, device=dev)
        eta = torch.empty(n, dtype=torch.float64, device=dev)
        o = _lib.opts(nb, flags)
        rc = _lib.lib().skh_sytrd_d_device(n, work.data_ptr(), n, tau.data_ptr(), eta.data_ptr(), st, ctypes.byref(o))
        _lib.check("skh_sytrd_d_device", rc)
        torch.cuda.synchronize()
        outs.append((work.cpu().numpy(), tau[: n - 1].cpu().numpy(), eta[: n - 1].cpu().numpy()))
    for F, tau, eta in outs[1:]:
        assert np.array_equal(np.tril(F, -1), np.tril(outs[0][0], -1))
        assert np.array_equal(tau, outs[0][1])
        assert np.array_equal(eta, outs[0][2])
    # and they are the orthogonal similarity they should be (Frobenius norm kept)
    assert abs(np.sqrt(2.0 * np.sum(outs[0][2] ** 2)) - np.linalg.norm(A)) <= 1e-13 * np.linalg.norm(A)


@pytest.mark.parametrize("nb", [8, 16])
@pytest.mark.parametrize("n", [2, 4, 6, 8, 10, 16, 18, 30, 62, 64, 66, 96, 128, 130, 200, 256, 258, 320, 512])
def test_sytrd_blocked_vs_oracle(n, nb):
    """The blocked one-workgroup engine (panels of nb reflectors, MFMA rank-2k update per panel):
    the same reflectors as the oracle within the blocking tolerance, the factor layout, and the
    orthogonal similarity."""
    if nb == 16 and n > 256:
        pytest.skip("panel width 16 is built for n <= 256")
    say("sytrd blocked n=%d nb=%d" % (n, nb))
    A = skew_gauss(n, 30 * n + nb)
    F, tau, eta, nz = gpu_sytrd(A, nb, skl.ENGINE_BLOCKED)
    Fo, tauo, etao = oracle.skewblockedhess(A)
    nA = np.linalg.norm(A)
    np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-11 * nA)
    np.testing.assert_allclose(tau, tauo, rtol=0, atol=1e-11)
    assert nz == int(np.count_nonzero(tau))
    check_fact(A, F, tau, eta)
    for i in range(n - 2):
        assert F[i + 1, i] == 1.0
    assert F[n - 1, n - 2] == eta[-1] and tau[-1] == 0.0
    iu = np.triu_indices(n)
    np.testing.assert_array_equal(F[iu], A[iu])


def test_blocked_batch_vs_compact():
    """A mixed batch (odd, empty and every even size class up to 512) through the blocked engine
    against the one-workgroup unblocked engine and the oracle."""
    say("blocked batch")
    rng = np.random.default_rng(41)
    sizes = [0, 1, 2, 3, 4, 32, 64, 96, 130, 256, 300, 384, 510, 512, 7] + list(rng.integers(1, 40, 24) * 2)
    mats = []
    for n in sizes:
        B = rng.standard_normal((n, n)) / np.sqrt(max(n, 1))
        mats.append(np.asfortranarray(B - B.T))
    lp_b, sg_b = skl.logabspfaffian_batched(mats, engine=skl.ENGINE_BLOCKED)
    lp_c, sg_c = skl.logabspfaffian_batched(mats, engine=skl.ENGINE_TEAM1)
    for k, M in enumerate(mats):
        n = M.shape[0]
        assert sg_b[k] == sg_c[k], (k, n)
        if n % 2 == 1:
            assert lp_b[k] == -np.inf
            continue
        if n == 0:
            assert lp_b[k] == 0.0
            continue
        lpo, sgo = oracle.logabspfaffian(M)
        assert sg_b[k] == sgo
        assert abs(lp_b[k] - lpo) <= 1e-10 * max(1.0, abs(lpo)), (n, lp_b[k], lpo)


def test_plan_tune_table(tmp_path):
    """The tuner over engines and panel widths, with its JSON table: a second plan of the same layout
    takes the stored configuration without timing and gives the same bits; a configuration set by
    hand (the unblocked engine) agrees with the oracle too."""
    say("plan tuning table")
    import torch
    rng = np.random.default_rng(23)
    sizes = list(rng.integers(1, 17, 300) * 32) + [2, 3, 0]
    mats = []
    for n in sizes:
        B = rng.standard_normal((n, n)) / np.sqrt(max(n, 1))
        mats.append(np.asfortranarray(B - B.T))
    ns = np.array(sizes, dtype=np.int64)
    offs = np.zeros(len(sizes), dtype=np.int64)
    tot = 0
    for k, n in enumerate(sizes):
        offs[k] = tot
        tot += (n * n + 1) & ~1
    host = np.zeros(max(tot, 1))
    for k, M in enumerate(mats):
        host[offs[k]:offs[k] + M.size] = M.ravel(order="F")
    dA = torch.tensor(host, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    path = str(tmp_path / "tune.json")
    p1 = skl.LogPfPlan(ns, offs)
    lp1 = torch.empty(len(sizes), dtype=torch.float64, device="cuda")
    sg1 = torch.empty_like(lp1)
    ms1 = p1.tune(dA.data_ptr(), lp1.data_ptr(), sg1.data_ptr(), st, table=path)
    p1.run(dA.data_ptr(), lp1.data_ptr(), sg1.data_ptr(), st)
    p1.status(st)
    p2 = skl.LogPfPlan(ns, offs)
    lp2 = torch.empty_like(lp1)
    sg2 = torch.empty_like(lp1)
    ms2 = p2.tune(dA.data_ptr(), lp2.data_ptr(), sg2.data_ptr(), st, table=path)
    assert ms2 == ms1 and p2.config() == p1.config(), (p1.config(), p2.config())
    p2.run(dA.data_ptr(), lp2.data_ptr(), sg2.data_ptr(), st)
    p2.status(st)
    assert torch.equal(lp1, lp2) and torch.equal(sg1, sg2)
    say("tuned configuration: %s, %.2f ms" % (p1.config(), ms1))
    p2.set_config({"engine": "unblocked", "kernel_nmax": 512, "panel_width": 0, "workgroups": 256})
    p2.run(dA.data_ptr(), lp2.data_ptr(), sg2.data_ptr(), st)
    p2.status(st)
    hl, hs = lp2.cpu().numpy(), sg2.cpu().numpy()
    for k in range(0, len(sizes), 7):
        n = sizes[k]
        lpo, sgo = (0.0, 1.0) if n == 0 else oracle.logabspfaffian(mats[k])
        assert hs[k] == sgo
        if n % 2 == 0 and n > 0:
            assert abs(hl[k] - lpo) <= 1e-10 * max(1.0, abs(lpo))
    p1.close()
    p2.close()
