"""GPU parity of the complex (ComplexF64) skew-Hermitian path against the oracle.

Reference behaviour: the complex branches of skewblockedhess! (src/hessenberg.jl:92-131, :96-104,
:123-126, :161, tail :70-90), hessenberg! (:5-13) and the complex eigvals!/eigen! that call LAPACK
on Hermitian(iA) (src/eigen.jl:18-41, :101-106).  Tolerances: eta, tau, dvim, factors within
1e-11 * ||A||; ||Q T Q^H - A|| / ||A|| <= 1e-13; eigenvalues within 10 n eps ||A||.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
skl = pytest.importorskip("skewlinearalgebra")
from skewlinearalgebra import _lib  # noqa: E402

EPS = np.finfo(float).eps


def say(msg):
    print("[case] " + msg, flush=True)


def zskew(n, seed):
    """A = (B - B^H) / 2, Re and Im of B iid N(0, 1)."""
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    return np.asfortranarray((B - B.conj().T) / 2)


def ztri(eta, dvim):
    n = dvim.size
    T = np.diag(1j * dvim).astype(complex)
    k = np.arange(n - 1)
    T[k + 1, k] = eta
    T[k, k + 1] = -eta
    return T


def gpu_zsytrd(A, engine=None):
    F = np.array(A, dtype=np.complex128, order="F", copy=True)
    n = F.shape[0]
    tau = np.zeros(max(n - 1, 0), dtype=np.complex128)
    eta = np.zeros(max(n - 1, 0))
    dvim = np.zeros(n)
    o = _lib.opts() if engine is None else _lib.opts(0, engine)
    rc = _lib.lib().skh_sytrd_z(n, F.ctypes.data, max(n, 1), tau.ctypes.data, eta.ctypes.data, dvim.ctypes.data,
                                ctypes.byref(o))
    _lib.check("skh_sytrd_z", rc)
    return F, tau, eta, dvim


def check_zfact(A, F, tau, eta, dvim, tol=1e-13):
    n = A.shape[0]
    Q = oracle.form_q(F, tau)
    nA = max(np.linalg.norm(A), 1e-300)
    res = np.linalg.norm(Q @ ztri(eta, dvim) @ Q.conj().T - A) / nA
    orth = np.linalg.norm(Q.conj().T @ Q - np.eye(n))
    assert res <= tol, res
    assert orth <= tol * max(n, 1), orth


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 10, 63, 64, 65, 66, 130, 257, 512])
def test_zsytrd_vs_oracle(n):
    say("zsytrd n=%d" % n)
    A = zskew(n, 100 + n)
    F, tau, eta, dvim = gpu_zsytrd(A)
    Fo, tauo, etao = oracle.skewblockedhess(A)
    nA = max(1.0, np.linalg.norm(A))
    np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-11 * nA)
    np.testing.assert_allclose(tau, tauo, rtol=0, atol=1e-11)
    np.testing.assert_allclose(dvim, np.imag(np.diag(Fo)), rtol=0, atol=1e-11 * nA)
    np.testing.assert_allclose(np.tril(F, -1), np.tril(Fo, -1), rtol=0, atol=1e-11)
    iu = np.triu_indices(n, 1)
    np.testing.assert_array_equal(F[iu], A[iu])   # upper triangle left as given
    if n >= 2:
        check_zfact(A, F, tau, eta, dvim)
    # the Python mirror: hessenberg(A) -> SkewHermTridiagonal(complex(eta), dvim)
    H = skl.hessenberg(A)
    np.testing.assert_allclose(H.H.to_dense(), ztri(eta, dvim), rtol=0, atol=1e-12 * nA)


@pytest.mark.parametrize("engine", ["blocked", "unblocked"])
@pytest.mark.parametrize("n", [200, 513, 700, 1111])
def test_zsytrd_engines_vs_oracle(n, engine):
    """Both complex engines against the oracle: the blocked panels (complex skewlatrd! with the
    complex rank-2k update, the unblocked engine finishing the last columns) and the unblocked one."""
    say("zsytrd %s n=%d" % (engine, n))
    A = zskew(n, 3 * n + 1)
    eng = _lib.ENGINE_LARGE if engine == "blocked" else _lib.ENGINE_TEAM1
    F, tau, eta, dvim = gpu_zsytrd(A, eng)
    Fo, tauo, etao = oracle.skewblockedhess(A)
    nA = max(1.0, np.linalg.norm(A))
    np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-11 * nA)
    np.testing.assert_allclose(tau, tauo, rtol=0, atol=1e-11)
    np.testing.assert_allclose(dvim, np.imag(np.diag(Fo)), rtol=0, atol=1e-11 * nA)
    np.testing.assert_allclose(np.tril(F, -1), np.tril(Fo, -1), rtol=0, atol=1e-10)
    iu = np.triu_indices(n, 1)
    np.testing.assert_array_equal(F[iu], A[iu])
    check_zfact(A, F, tau, eta, dvim)


def test_zsytrd_cooperative_launch_same_bits():
    """The blocked complex panels under the cooperative launch (SKH_PANEL_COOPERATIVE=1) and the
    default plain launch: same grid, so the same bits."""
    say("zsytrd cooperative launch n=700")
    A = zskew(700, 71)
    r0 = gpu_zsytrd(A, _lib.ENGINE_LARGE)
    os.environ["SKH_PANEL_COOPERATIVE"] = "1"
    try:
        r1 = gpu_zsytrd(A, _lib.ENGINE_LARGE)
    finally:
        del os.environ["SKH_PANEL_COOPERATIVE"]
    for a, b in zip(r0, r1):
        assert np.array_equal(a, b)
    check_zfact(A, *r1)


def test_finalize_and_reinit_same_bits():
    """skh_finalize releases every cached workspace (the blocked complex panels' included) and the
    stream; the next call initialises the library again and must give the same bits."""
    say("finalize and re-init")
    A = zskew(600, 17)
    r0 = gpu_zsytrd(A, _lib.ENGINE_LARGE)
    Ar = np.asfortranarray(np.real(A) - np.real(A).T)
    Fr0 = np.array(Ar, order="F", copy=True)
    n = Ar.shape[0]
    tau0 = np.zeros(n - 1)
    eta0 = np.zeros(n - 1)
    o = _lib.opts(0, _lib.ENGINE_LARGE)
    _lib.check("skh_sytrd_d", _lib.lib().skh_sytrd_d(n, Fr0.ctypes.data, n, tau0.ctypes.data, eta0.ctypes.data, None,
                                                      ctypes.byref(o)))
    assert _lib.lib().skh_finalize() == 0
    r1 = gpu_zsytrd(A, _lib.ENGINE_LARGE)
    Fr1 = np.array(Ar, order="F", copy=True)
    tau1 = np.zeros(n - 1)
    eta1 = np.zeros(n - 1)
    _lib.check("skh_sytrd_d", _lib.lib().skh_sytrd_d(n, Fr1.ctypes.data, n, tau1.ctypes.data, eta1.ctypes.data, None,
                                                      ctypes.byref(o)))
    for a, b in zip(r0, r1):
        assert np.array_equal(a, b)
    assert np.array_equal(Fr0, Fr1) and np.array_equal(tau0, tau1) and np.array_equal(eta0, eta1)


def test_zsytrd_structured():
    say("zsytrd structured")
    cases = [np.zeros((6, 6), complex)]
    D = np.diag(1j * np.arange(1.0, 6.0))          # diagonal only: eta = 0, tau = 0
    cases.append(D.astype(complex))
    a = 0.3 - 1.7j                                  # n = 2: one length-1 complex reflector
    cases.append(np.array([[0.5j, -np.conj(a)], [a, -0.25j]]))
    A = np.zeros((5, 5), complex)                   # alpha != 0 with xnorm = 0 in column 0
    A[1, 0], A[0, 1] = 2 + 1j, -(2 - 1j)
    A[4, 2], A[2, 4] = 1j, 1j
    cases.append(A)
    for A in cases:
        F, tau, eta, dvim = gpu_zsytrd(A)
        Fo, tauo, etao = oracle.skewblockedhess(A)
        np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-13 * max(1.0, np.linalg.norm(A)))
        np.testing.assert_allclose(tau, tauo, rtol=0, atol=1e-13)
        np.testing.assert_allclose(dvim, np.imag(np.diag(Fo)), rtol=0, atol=1e-13 * max(1.0, np.linalg.norm(A)))
        check_zfact(A, F, tau, eta, dvim)


def zdegenerate(n, seed):
    """Skew-Hermitian matrices whose reflectors hit the householder! special cases inside the blocked
    panels: a tridiagonal one with an imaginary diagonal (every tail zero: x = alpha e_1, where the
    unscaled matvec's m takes q = conj(alpha) A[j+1, j+1]), the same plus a dense trailing block, and
    a block-diagonal one (zero columns at the block edges: tau = 0)."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(n - 1) + 1j * rng.standard_normal(n - 1)
    T = np.diag(1j * rng.standard_normal(n)) + np.diag(a, -1) - np.diag(np.conj(a), 1)
    Z = T.copy()
    h = n // 2 + 5
    Z[h:, h:] += zskew(n - h, seed + 1)
    B = np.zeros((n, n), complex)
    r = 0
    for size in rng.integers(1, 140, size=n):
        if r >= n:
            break
        m = int(min(size, n - r))
        B[r:r + m, r:r + m] = zskew(m, seed + 2 + r)
        r += m
    return [T, Z, B]


@pytest.mark.parametrize("n", [700, 1111])
def test_zsytrd_blocked_degenerate(n):
    for k, A in enumerate(zdegenerate(n, 60 + n)):
        say("zsytrd blocked degenerate n=%d case %d" % (n, k))
        F, tau, eta, dvim = gpu_zsytrd(A, _lib.ENGINE_LARGE)
        Fo, tauo, etao = oracle.skewblockedhess(A)
        nA = max(1.0, np.linalg.norm(A))
        np.testing.assert_allclose(eta, etao, rtol=0, atol=1e-11 * nA)
        np.testing.assert_array_equal(tau != 0, tauo != 0)
        np.testing.assert_allclose(tau, tauo, rtol=0, atol=1e-11)
        np.testing.assert_allclose(dvim, np.imag(np.diag(Fo)), rtol=0, atol=1e-11 * nA)
        check_zfact(A, F, tau, eta, dvim)


def test_zorgtr_device_vs_oracle():
    say("zorgtr")
    import torch
    n = 333
    A = zskew(n, 9)
    F, tau, eta, dvim = gpu_zsytrd(A)
    dF = torch.tensor(F.T.copy(), device="cuda")   # row-major storage of F^T = column-major F
    dtau = torch.tensor(tau, device="cuda")
    dQ = torch.empty((n, n), dtype=torch.complex128, device="cuda")
    rc = _lib.lib().skh_orgtr_z_device(n, dF.data_ptr(), n, dtau.data_ptr(), dQ.data_ptr(), n,
                                       torch.cuda.current_stream().cuda_stream)
    _lib.check("skh_orgtr_z_device", rc)
    torch.cuda.synchronize()
    Q = dQ.cpu().numpy().T
    np.testing.assert_allclose(Q, oracle.form_q(F, tau), rtol=0, atol=1e-13 * n)


def test_zhessenberg_n4096_device():
    """ComplexF64 n = 4096, A = (B - B^H)/2: device tridiagonalization and Q, residual and
    orthogonality on the GPU; eta against the oracle."""
    say("zsytrd n=4096 device")
    import torch
    n = 4096
    A = zskew(n, 4096)
    dA = torch.tensor(A.T.copy(), device="cuda")       # column-major A as a row-major tensor of A^T
    dtau = torch.empty(n - 1, dtype=torch.complex128, device="cuda")
    deta = torch.empty(n - 1, dtype=torch.float64, device="cuda")
    ddv = torch.empty(n, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    o = _lib.opts()
    rc = _lib.lib().skh_sytrd_z_device(n, dA.data_ptr(), n, dtau.data_ptr(), deta.data_ptr(), ddv.data_ptr(), st,
                                       ctypes.byref(o))
    _lib.check("skh_sytrd_z_device", rc)
    dQ = torch.empty((n, n), dtype=torch.complex128, device="cuda")
    rc = _lib.lib().skh_orgtr_z_device(n, dA.data_ptr(), n, dtau.data_ptr(), dQ.data_ptr(), n, st)
    _lib.check("skh_orgtr_z_device", rc)
    torch.cuda.synchronize()
    Qc = dQ.T.contiguous()                                # Q as a row-major tensor
    T = torch.diag(1j * ddv.to(torch.complex128))
    k = torch.arange(n - 1, device="cuda")
    T[k + 1, k] = deta.to(torch.complex128)
    T[k, k + 1] = -deta.to(torch.complex128)
    At = torch.tensor(A, device="cuda")
    res = float(torch.linalg.norm(Qc @ T @ Qc.conj().T - At) / torch.linalg.norm(At))
    orth = float(torch.linalg.norm(Qc.conj().T @ Qc - torch.eye(n, dtype=torch.complex128, device="cuda")))
    say("n=4096 residual %.3e orthogonality %.3e" % (res, orth))
    assert res <= 1e-13
    assert orth <= 1e-12
    Fo, tauo, etao = oracle.skewblockedhess(A)
    eta = deta.cpu().numpy()
    assert np.max(np.abs(eta - etao)) <= 1e-11 * np.linalg.norm(A)


@pytest.mark.parametrize("n", [1, 2, 3, 8, 65, 300])
def test_zeigen_vs_lapack(n):
    say("zeigen n=%d" % n)
    A = zskew(n, 300 + n)
    E = skl.eigen(A)
    lam = np.linalg.eigvalsh(1j * A)                     # the reference's route (src/eigen.jl:101-106)
    nA = np.linalg.norm(A, 2) if n > 0 else 1.0
    np.testing.assert_allclose(E.values.imag, -lam, rtol=0, atol=10 * n * EPS * max(nA, 1.0))
    assert np.all(E.values.real == 0)
    X = E.vectors
    assert np.linalg.norm(A @ X - X @ np.diag(E.values)) <= 1e-12 * n * max(nA, 1.0)
    assert np.linalg.norm(X.conj().T @ X - np.eye(n)) <= 1e-12 * n


@pytest.mark.parametrize("n", [2, 7, 1024])
def test_zeigvals_vs_lapack(n):
    say("zeigvals n=%d" % n)
    A = zskew(n, 1300 + n)
    vals = skl.eigvals(A)
    lam = np.linalg.eigvalsh(1j * A)
    nA = np.linalg.norm(A, 2)
    np.testing.assert_allclose(vals.imag, -lam, rtol=0, atol=10 * n * EPS * nA)
    # sortby (src/eigen.jl:23-26): sort the eigenvalues of iA by the key, reverse, negate
    vs = skl.eigvals(A, sortby=abs)
    ref = -np.array(sorted(lam, key=abs))[::-1]
    np.testing.assert_allclose(vs.imag, ref, rtol=0, atol=10 * n * EPS * nA)
