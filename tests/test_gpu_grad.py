"""The log|Pf| gradient -1/2 A^{-1} through gradient plans (skh_logpf_grad_batched_plan_*):
the hand-written backward pass replayed from a HIP graph against the eager device entry (bit for
bit) and against -1/2 inv(A) by LU (the reference's CPU route, inv(A::SkewHermitian),
src/skewhermitian.jl:233) within 1e-10 relative (Frobenius); log|Pf| against the exact value of
constructed matrices A = Q blockdiag([0 s; -s 0]) Q^T (log|Pf| = sum log s, sign = det Q)."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

skl = pytest.importorskip("skewlinearalgebra")
from skewlinearalgebra import _lib  # noqa: E402


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _constructed(n, rng):
    Q, R = np.linalg.qr(rng.standard_normal((n, n)))
    Q = Q * np.sign(np.diag(R))
    s = np.exp(rng.uniform(np.log(0.1), np.log(10.0), size=n // 2))
    B = np.zeros((n, n))
    k = np.arange(n // 2)
    B[2 * k, 2 * k + 1] = s
    B[2 * k + 1, 2 * k] = -s
    A = Q @ B @ Q.T
    return 0.5 * (A - A.T), float(np.sum(np.log(s))), float(np.sign(np.linalg.det(Q)))


def _pack(mats):
    ns = np.array([M.shape[0] for M in mats], dtype=np.int64)
    offs = np.zeros(len(mats), dtype=np.int64)
    tot = 0
    for k, n in enumerate(ns):
        offs[k] = tot
        tot += (int(n) * int(n) + 1) & ~1
    host = np.zeros(max(tot, 1))
    for k, M in enumerate(mats):
        host[offs[k]:offs[k] + M.size] = M.ravel(order="F")
    return ns, offs, host


def test_grad_plan_graph_replay_mixed_sizes():
    torch = pytest.importorskip("torch")
    say("gradient plan, mixed sizes, graph replays")
    rng = np.random.default_rng(77)
    sizes = [2, 0, 3, 64, 256, 130, 1, 512, 32, 96, 1024, 256, 7]
    mats, ref = [], []
    for n in sizes:
        if n % 2 == 0 and n > 0:
            A, lp, sg = _constructed(n, rng)
        else:
            B = rng.standard_normal((n, n))
            A, lp, sg = B - B.T, (0.0 if n == 0 else -np.inf), (1.0 if n == 0 else 0.0)
        mats.append(A)
        ref.append((lp, sg))
    ns, offs, host = _pack(mats)
    dev = torch.device("cuda:0")
    dA = torch.tensor(host, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    count = len(sizes)
    lp0 = torch.empty(count, dtype=torch.float64, device=dev)
    sg0 = torch.empty_like(lp0)
    g0 = torch.full_like(dA, 3.0)
    o = _lib.opts()
    _lib.check("device", _lib.lib().skh_logpf_grad_batched_d_device(count, ns.ctypes.data, offs.ctypes.data,
                                                                    dA.data_ptr(), lp0.data_ptr(), sg0.data_ptr(),
                                                                    g0.data_ptr(), st, ctypes.byref(o)))
    torch.cuda.synchronize()
    G0 = g0.cpu().numpy()
    for k, A in enumerate(mats):
        n = A.shape[0]
        assert sg0[k].item() == ref[k][1], (k, n)
        if n % 2 == 0 and n > 0:
            assert abs(lp0[k].item() - ref[k][0]) <= 1e-10 * max(1.0, abs(ref[k][0])), (k, n)
            G = G0[offs[k]:offs[k] + n * n].reshape((n, n), order="F")
            Gi = -0.5 * np.linalg.inv(A)
            assert np.linalg.norm(G - Gi) <= 1e-10 * np.linalg.norm(Gi), (n, np.linalg.norm(G - Gi))
        elif n % 2 == 1:
            assert lp0[k].item() == -np.inf
            assert not np.any(G0[offs[k]:offs[k] + n * n])
    plan = skl.LogPfGradPlan(ns, offs)
    lp = torch.empty_like(lp0)
    sg = torch.empty_like(lp0)
    g = torch.empty_like(dA)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run(dA.data_ptr(), lp.data_ptr(), sg.data_ptr(), g.data_ptr(), side.cuda_stream)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        plan.run(dA.data_ptr(), lp.data_ptr(), sg.data_ptr(), g.data_ptr(), torch.cuda.current_stream().cuda_stream)
    for r in range(3):
        lp.fill_(-1.0)
        sg.fill_(7.0)
        g.fill_(5.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        plan.status(st)
        assert torch.equal(lp, lp0) and torch.equal(sg, sg0), r
        # every member's block (the padding between members is not written)
        for k in range(count):
            n = int(ns[k])
            a, b = int(offs[k]), int(offs[k]) + n * n
            assert torch.equal(g[a:b], g0[a:b]), (r, k, n)
    plan.close()


def test_grad_batch1024_n256():
    """The gradient workload at its full size: 1024 constructed matrices n = 256 through one plan;
    every matrix's log|Pf| and sign against the exact values, every 64th gradient against the LU
    inverse."""
    torch = pytest.importorskip("torch")
    say("gradient plan, 1024 x 256")
    rng = np.random.default_rng(1024)
    count, n = 1024, 256
    mats, ref = [], []
    for _ in range(count):
        A, lp, sg = _constructed(n, rng)
        mats.append(A)
        ref.append((lp, sg))
    ns, offs, host = _pack(mats)
    dev = torch.device("cuda:0")
    dA = torch.tensor(host, device=dev)
    lp = torch.empty(count, dtype=torch.float64, device=dev)
    sg = torch.empty_like(lp)
    g = torch.empty_like(dA)
    st = torch.cuda.current_stream().cuda_stream
    plan = skl.LogPfGradPlan(ns, offs)
    plan.run(dA.data_ptr(), lp.data_ptr(), sg.data_ptr(), g.data_ptr(), st)
    plan.status(st)
    hl, hs = lp.cpu().numpy(), sg.cpu().numpy()
    for k in range(count):
        assert hs[k] == ref[k][1], k
        assert abs(hl[k] - ref[k][0]) <= 1e-10 * max(1.0, abs(ref[k][0])), k
    G = g.cpu().numpy()
    for k in range(0, count, 64):
        Gk = G[offs[k]:offs[k] + n * n].reshape((n, n), order="F")
        Gi = -0.5 * np.linalg.inv(mats[k])
        assert np.linalg.norm(Gk - Gi) <= 1e-10 * np.linalg.norm(Gi), k
    plan.close()
