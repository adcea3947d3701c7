"""GPU checks of the path at the full sizes of its workloads, through the C ABI.

  * n = 8192 real, default panel width (32): ||Q T Q^T - A||_F / ||A||_F <= 1e-13 with Q from
    skh_orgtr_d_device, ||Q^T Q - I||_F <= 1e-13 n, and the host implicit-QR eigenvalues of the GPU
    tridiagonal against a dense Hermitian solve of iA on the device within 10 n eps ||A||_2;
  * n = 512 real (the plain CPU/GPU comparison): hessenberg + eigvals + logabspfaffian against the oracle;
  * ComplexF64 n = 4096 skew-Hermitian: eigen (src/eigen.jl:101-106) by ||A X - X Lambda|| and
    ||X^H X - I||, eigenvalues purely imaginary;
  * 4096 uneven matrices n in {32, ..., 512}: one plan replayed from a HIP graph, every 16th matrix
    against the oracle's logabspfaffian (src/pfaffian.jl:167-179): relative 1e-10, sign exact.
"""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from conftest import skew_gauss

pytestmark = pytest.mark.gpu

skl = pytest.importorskip("skewlinearalgebra")
from skewlinearalgebra import _lib  # noqa: E402

EPS = np.finfo(float).eps


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gauss_skew_device(torch, n, seed):
    """A = B - B^T, B iid N(0, 1) generated on the device; returned as a row-major tensor of A
    (A^T = -A, so A.T.contiguous() is the column-major storage)."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    B = torch.randn((n, n), generator=g, dtype=torch.float64, device=dev)
    A = B - B.T
    del B
    return A


def test_n8192_default_width_residual_orthogonality_eigvals():
    torch = pytest.importorskip("torch")
    n = 8192
    say("n=8192 default panel width: factor, Q, residual, eigenvalues")
    A = _gauss_skew_device(torch, n, 314159)
    dev = A.device
    work = A.T.contiguous()
    tau = torch.empty(n, dtype=torch.float64, device=dev)
    eta = torch.empty(n, dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    L = _lib.lib()
    o = _lib.opts(0, skl.ENGINE_LARGE)
    _lib.check("skh_sytrd_d_device", L.skh_sytrd_d_device(n, work.data_ptr(), n, tau.data_ptr(), eta.data_ptr(), st,
                                                         ctypes.byref(o)))
    _lib.check("skh_device_status", L.skh_device_status(st))   # the factorization abandoned no bounded wait
    dQ = torch.empty((n, n), dtype=torch.float64, device=dev)
    _lib.check("skh_orgtr_d_device", L.skh_orgtr_d_device(n, work.data_ptr(), n, tau.data_ptr(), dQ.data_ptr(), n, st))
    torch.cuda.synchronize()
    del work
    Q = dQ.T   # column-major storage read back as Q
    e = eta[: n - 1]
    # Q T with T[j+1, j] = e_j, T[j, j+1] = -e_j: column j = e_j Q[:, j+1] - e_{j-1} Q[:, j-1]
    QT = torch.zeros((n, n), dtype=torch.float64, device=dev)
    QT[:, : n - 1] += Q[:, 1:] * e[None, :]
    QT[:, 1:] -= Q[:, : n - 1] * e[None, :]
    R = QT @ Q.T
    del QT
    R -= A
    nA = float(torch.linalg.norm(A))
    res = float(torch.linalg.norm(R)) / nA
    del R
    I = torch.eye(n, dtype=torch.float64, device=dev)
    orth = float(torch.linalg.norm(Q.T @ Q - I))
    del I, dQ, Q
    say("residual %.3e orthogonality %.3e" % (res, orth))
    assert res <= 1e-13, res
    assert orth <= 1e-13 * n, orth
    # host implicit QR on the GPU tridiagonal against the dense Hermitian solve of iA
    ev = e.cpu().numpy().copy()
    vals = np.zeros(n)
    _lib.check("skh_skewtrieigvals_d", L.skh_skewtrieigvals_d(n, ev.ctypes.data, vals.ctypes.data))
    H = torch.complex(torch.zeros_like(A), A)   # iA, Hermitian
    del A
    mu = torch.linalg.eigvalsh(H)   # eigenvalues of iA = -imag(lambda)
    del H
    ref = np.sort(-mu.cpu().numpy())
    norm2 = float(np.max(np.abs(ref)))
    err = float(np.max(np.abs(np.sort(vals) - ref)))
    say("eigenvalues: max |diff| %.3e against 10 n eps ||A||_2 = %.3e" % (err, 10 * n * EPS * norm2))
    assert err <= 10 * n * EPS * norm2, err


def test_n512_plain_method_vs_oracle():
    """The n = 512 comparison workload: hessenberg + eigvals + logabspfaffian/sign (A = B - B^T, seed
    314159), each against the oracle's restatement of the reference CPU path."""
    n = 512
    say("n=512 hessenberg + eigvals + logabspfaffian")
    rng = np.random.default_rng(314159)
    B = rng.standard_normal((n, n))
    A = np.asfortranarray(B - B.T)
    H = skl.hessenberg(A)
    Fo, tauo, etao = oracle.skewblockedhess(A)
    nA = np.linalg.norm(A)
    np.testing.assert_allclose(H.H.ev, etao, rtol=0, atol=1e-11 * nA)
    Q = oracle.form_q(H.factors, H.tau)
    T = np.diag(H.H.ev, -1) - np.diag(H.H.ev, 1)
    assert np.linalg.norm(Q @ T @ Q.T - A) / nA <= 1e-13
    lam = np.sort(skl.eigvals(A).imag)
    lamo = np.sort(oracle.skewtrieigvals(etao))
    assert np.max(np.abs(lam - lamo)) <= 10 * n * EPS * np.linalg.norm(A, 2)
    lp, sg = skl.logabspfaffian(A)
    lpo, sgo = oracle.logabspfaffian(A)
    assert sg == sgo
    assert abs(lp - lpo) <= 1e-10 * abs(lpo)


def test_complex_n4096_eigen():
    torch = pytest.importorskip("torch")
    n = 4096
    say("complex n=4096 eigen")
    A = skew_gauss(n, 4096, np.complex128)
    E = skl.eigen(A)
    vals = np.asarray(E.values)
    assert np.all(vals.real == 0.0)   # purely imaginary (src/eigen.jl:101-106)
    dev = torch.device("cuda:0")
    dA = torch.tensor(A, device=dev)
    X = torch.tensor(np.ascontiguousarray(E.vectors), device=dev)
    lam = torch.tensor(vals, device=dev)
    nA = float(torch.linalg.norm(dA))
    res = float(torch.linalg.norm(dA @ X - X * lam[None, :])) / nA
    orth = float(torch.linalg.norm(X.conj().T @ X - torch.eye(n, dtype=torch.complex128, device=dev)))
    say("complex eigen: residual %.3e orthogonality %.3e" % (res, orth))
    assert res <= 1e-13, res
    assert orth <= 1e-13 * n, orth
    # the eigenvalues against a dense Hermitian solve of iA
    mu = torch.linalg.eigvalsh(1j * dA).cpu().numpy()
    err = float(np.max(np.abs(np.sort(-vals.imag) - np.sort(mu))))
    norm2 = float(np.max(np.abs(mu)))
    assert err <= 10 * n * EPS * norm2, err


def _cfg4_batch(torch, count=4096, seed=314159):
    """4096 real skew matrices, n uniform over {32, 64, ..., 512}, A = B - B^T with B iid N(0, 1/n),
    packed column-major at even offsets of one device buffer (generated on the device)."""
    rng = np.random.default_rng(seed)
    ns = (rng.integers(1, 17, size=count) * 32).astype(np.int64)
    offs = np.zeros(count, dtype=np.int64)
    offs[1:] = np.cumsum(ns[:-1] ** 2)
    total = int(offs[-1] + ns[-1] ** 2)
    dev = torch.device("cuda:0")
    buf = torch.empty(total, dtype=torch.float64, device=dev)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        Bm = torch.randn((len(idx), int(n), int(n)), generator=g, dtype=torch.float64, device=dev) / (float(n) ** 0.5)
        Am = Bm - Bm.transpose(1, 2)
        for q, k in enumerate(idx):
            buf[offs[k]:offs[k] + n * n] = Am[q].T.reshape(-1)
    return ns, offs, buf


def test_batch4096_graph_replay_vs_oracle():
    torch = pytest.importorskip("torch")
    say("4096-matrix batch: plan, graph replays, oracle on every 16th matrix")
    ns, offs, buf = _cfg4_batch(torch)
    count = len(ns)
    dev = buf.device
    st = torch.cuda.current_stream().cuda_stream
    plan = skl.LogPfPlan(ns, offs)
    lp0 = torch.empty(count, dtype=torch.float64, device=dev)
    sg0 = torch.empty_like(lp0)
    plan.run(buf.data_ptr(), lp0.data_ptr(), sg0.data_ptr(), st)
    plan.status(st)
    lp = torch.empty_like(lp0)
    sg = torch.empty_like(lp0)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        plan.run(buf.data_ptr(), lp.data_ptr(), sg.data_ptr(), torch.cuda.current_stream().cuda_stream)
    for r in range(3):
        lp.fill_(-1.0)
        sg.fill_(7.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        plan.status(st)
        assert torch.equal(lp, lp0) and torch.equal(sg, sg0), r
    plan.close()
    sample = np.arange(0, count, 16)
    say("oracle on %d matrices" % len(sample))
    glp = lp0.cpu().numpy()
    gsg = sg0.cpu().numpy()
    worst = 0.0
    for k in sample:
        n = int(ns[k])
        M = buf[int(offs[k]):int(offs[k]) + n * n].cpu().numpy().reshape((n, n), order="F")
        lpo, sgo = oracle.logabspfaffian(M)
        assert gsg[k] == sgo, (k, n)
        rel = abs(glp[k] - lpo) / max(1.0, abs(lpo))
        worst = max(worst, rel)
        assert rel <= 1e-10, (k, n, glp[k], lpo)
    say("worst relative deviation %.3e" % worst)
