"""Benchmark: n=8192 real Float64 skew tridiagonalization on one MI355X (TFLOP/s).

One step = one blocked Householder tridiagonalization (the reference's skewblockedhess!,
src/hessenberg.jl:144-174) of a fresh copy of A = B - B^T (B iid N(0,1), seed 314159) held in
HBM, through the C ABI (skh_sytrd_d_device) on torch's current stream.
TFLOP/s = (4/3) n^3 / time per step (the standard sytrd operation count).

Also reported:
  roofline      the panel kernel (k_panel, HBM-bound skew matvec): algorithmic bytes per
                launch / mean launch time from HIP events recorded around each launch
  mfma_update   the rank-2k trailing update kernel (FP64 MFMA) against the FP64 matrix peak
  cpu_baseline  the oracle (C restatement of the reference algorithm, OpenBLAS, all host
                threads): median of 3 full factorizations at n=8192 after a warm-up
  plain512      n = 512 hessenberg + eigvals + logabspfaffian through the host API, against the
                oracle's same calls
  nb_sweep      one factorization at each panel width nb in {32, 64, 128}: ms, TFLOP/s, panel
                HBM GB/s, rank-2k TFLOP/s
  host_eigvals  the host implicit-QR eigenvalues of the GPU tridiagonal (sum of lambda^2 against
                ||A||_F^2 as the check)
  batched       4096 uneven-size matrices (n in {32,...,512}), log|Pf| + sign through one tuned
                plan replayed from a HIP graph, matrices/s, against the per-matrix GPU loop (the
                library's single-matrix engine choice) and the oracle's per-matrix loop
  grad          1024 matrices n = 256, log|Pf| + gradient, matrices/s, against the per-matrix GPU
                loop and the oracle's logabspfaffian + LU inverse per matrix
  complex       ComplexF64 n = 4096 skew-Hermitian tridiagonalization, TFLOP/s, with Q formation and
                the residual / orthogonality of A = Q T Q^H
  complex_pfaffian  4096 complex skew-symmetric matrices (n in {32,...,256}), log|Pf| and phase by
                pivoted Parlett-Reid in one batched call, matrices/s, against the oracle on a sample,
                plus one n = 4096 matrix alone
  skewchol      4096 real skew matrices (n in {32,...,256}), skew-Cholesky with complete pivoting in
                one batched call, matrices/s, plus one n = 4096 matrix alone, against the oracle
  matfun        (--matfun) exp(A) and sincos(A) of a real skew matrix at n = 1024 and 4096, whole
                host-API calls (ms), against the oracle's exp at n = 1024

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 8192] [--nb 0] [--no-cpu] [--cpu-n 8192]
                       [--cpu-runs 3] [--no-plain] [--no-batched] [--no-complex] [--matfun]
                       [--only complex_pfaffian|skewchol|matfun]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "skewlinearalgebra.jl_amd"))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TF = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


CSRC = os.path.join(HERE, "skewlinearalgebra.jl_amd", "csrc")
# the sources a kernel's code comes from (its traffic profile is valid for exactly these bytes)
KERNEL_SOURCES = {
    "k_panel": ("skh_device.hpp", "skh_sync.hpp", "skh_fill.hpp", "skh_kernels.hip"),
    "k_compact": ("skh_device.hpp", "skh_sync.hpp", "skh_compact.hpp", "skh_compact.hip"),
    "k_cblock": ("skh_device.hpp", "skh_sync.hpp", "skh_cblock.hpp", "skh_compact.hip"),
}


def kernel_source_sha(kernel):
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES[kernel]:
        with open(os.path.join(CSRC, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(profile, kernel):
    """HBM bytes per launch of `kernel` from a committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (scaled by the copy-kernel factors), or (None, reason) when the profile is missing or was taken
    on other kernel sources than the ones built here (its kernel_source_sha256 line)."""
    path = os.path.join(HERE, "profiles", profile)
    value, sha = None, None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    value = float(line.split(":")[1])
                elif line.startswith("kernel_source_sha256:"):
                    sha = line.split(":")[1].strip()
    except OSError:
        return None, "no profile %s" % profile
    if value is None:
        return None, "no traffic line in %s" % profile
    if sha != kernel_source_sha(kernel):
        return None, "%s was taken on other %s sources (sha %s, built %s)" % (profile, kernel, sha,
                                                                             kernel_source_sha(kernel))
    return value, "%s (kernel sources sha %s)" % (profile, sha)


def cpu_baseline(n_cpu, runs):
    """The oracle (C restatement of the reference's CPU path: setnb = 60, the full-square gemv per
    column, the Omega gemm and the serial skew combine, OpenBLAS on all host threads) on the same
    generator: one warm-up factorization at n = 2048, then the median of `runs` at n_cpu."""
    import numpy as np
    import oracle
    threads = len(os.sched_getaffinity(0))
    oracle.lib()
    oracle.set_threads(threads)
    rng = np.random.default_rng(314159)
    W = rng.standard_normal((2048, 2048))
    oracle.skewblockedhess(np.asfortranarray(W - W.T))
    del W
    rng = np.random.default_rng(314159)
    B = rng.standard_normal((n_cpu, n_cpu))
    A = np.asfortranarray(B - B.T)
    del B
    times = []
    for r in range(runs):
        t0 = time.perf_counter()
        oracle.skewblockedhess(A)
        times.append(time.perf_counter() - t0)
        log("cpu baseline run %d: %.2f s" % (r, times[-1]))
    dt = float(np.median(times))
    flops = 4.0 / 3.0 * n_cpu ** 3
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": "full factorizations at n=%d (same generator), median of %d after a warm-up: %s s, "
                      "OpenBLAS threads=%d" % (n_cpu, runs, ", ".join("%.2f" % t for t in times), threads)}


def plain_bench(torch, skl, reps=5, n=512, seed=314159):
    """The plain n = 512 comparison: hessenberg + eigvals + logabspfaffian/sign of A = B - B^T
    (B iid N(0,1), seed 314159) through the host API (host arrays in and out, as the reference's
    calls), each call timed (median of `reps`), against the oracle's same three calls (C restatement
    of the reference's CPU path, OpenBLAS on all host threads).  Each call factors A itself, as in the
    reference (src/hessenberg.jl:15, src/eigen.jl:45, src/pfaffian.jl:236)."""
    import oracle
    threads = len(os.sched_getaffinity(0))
    oracle.set_threads(threads)
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((n, n))
    A = np.asfortranarray(B - B.T)

    def med(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3

    g = {"hessenberg": med(lambda: skl.hessenberg(A)), "eigvals": med(lambda: skl.eigvals(A)),
         "logabspfaffian": med(lambda: skl.logabspfaffian(A))}
    c = {"hessenberg": med(lambda: oracle.skewblockedhess(A)),
         "eigvals": med(lambda: oracle.skewtrieigvals(oracle.skewblockedhess(A)[2])),
         "logabspfaffian": med(lambda: oracle.logabspfaffian(A))}
    lam = np.sort(skl.eigvals(A).imag)
    lamo = np.sort(oracle.skewtrieigvals(oracle.skewblockedhess(A)[2]))
    lp, sg = skl.logabspfaffian(A)
    lpo, sgo = oracle.logabspfaffian(A)
    gt, ct = sum(g.values()), sum(c.values())
    return {"metric": "n=512 hessenberg + eigvals + logabspfaffian, whole host-API calls", "unit": "ms",
            "value": gt, "gpu_ms": g,
            "eigvals_max_abs_diff_vs_oracle": float(np.max(np.abs(lam - lamo))),
            "eigvals_tolerance_10_n_eps_norm2": 10 * n * np.finfo(float).eps * float(np.linalg.norm(A, 2)),
            "logpf_rel_diff_vs_oracle": abs(lp - lpo) / abs(lpo), "sign_equal": bool(sg == sgo),
            "cpu_baseline": {"value": ct, "unit": "ms (the three calls)", "cores": threads, "kind": "port",
                             "per_call_ms": c, "sample": "oracle, median of %d per call" % reps,
                             "cpu_over_gpu": ct / gt}}


def cfg4_batch(torch, dev, count=4096, seed=314159):
    """4096 real skew matrices, n uniform over {32, 64, ..., 512}, A = B - B^T, B iid N(0, 1/n),
    packed column-major at even offsets in one HBM buffer."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ns = (rng.integers(1, 17, size=count) * 32).astype(np.int64)
    offs = np.zeros(count, dtype=np.int64)
    offs[1:] = np.cumsum(ns[:-1] ** 2)
    total = int(offs[-1] + ns[-1] ** 2)
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


def cfg5_batch(torch, dev, count=1024, n=256, seed=314159):
    """1024 real skew matrices n = 256, A = Q blockdiag([0 s_i; -s_i 0]) Q^T, Q Haar (QR of a
    Gaussian with the sign fix, Mezzadri 2007), s_i log-uniform on [0.1, 10]; exact
    log|Pf| = sum log s_i.  Built on the host (numpy, seeded) and copied to HBM once."""
    rng = np.random.default_rng(seed)
    A = np.empty((count, n, n))
    exact = np.empty(count)
    Bm = np.zeros((n, n))
    k = np.arange(n // 2)
    for b in range(count):
        Q, R = np.linalg.qr(rng.standard_normal((n, n)))
        Q = Q * np.sign(np.diag(R))
        s = np.exp(rng.uniform(np.log(0.1), np.log(10.0), size=n // 2))
        Bm[2 * k, 2 * k + 1] = s
        Bm[2 * k + 1, 2 * k] = -s
        M = Q @ Bm @ Q.T
        A[b] = 0.5 * (M - M.T)
        exact[b] = np.sum(np.log(s))
    return torch.tensor(A, device=dev), torch.tensor(exact, device=dev)


def timed(torch, fn, reps):
    st = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


BATCH_BYTES_RULE = {
    True: "per matrix, the strictly lower triangle of each column's trailing block read once (the matvec) "
          "and, per panel, that of the block after it read and written once (the rank-2k update)",
    False: "per matrix, the strictly lower triangle of each column's trailing block read and written once",
}


def batch_bytes(ns, nb):
    """Algorithmic bytes of one batched launch: the blocked engine (panel width nb) reads the strictly
    lower triangle of each column's trailing block [j+1, n)^2 once and reads and writes that of the
    block [c0+nb, n)^2 after each panel; the unblocked engine (nb = 0) reads and writes each
    column's trailing triangle once."""
    total = 0.0
    for n in ns:
        n = int(n)
        if n < 2 or n % 2:
            continue
        cols = 0.0
        for j in range(n - 2):
            m = n - j - 1
            cols += m * (m - 1) / 2.0
        if nb == 0:
            total += 16.0 * cols
            continue
        upd = 0.0
        for c0 in range(0, n - 1, nb):
            e = min(c0 + nb, n - 1)
            if e >= n - 1:
                continue
            m = n - e
            upd += m * (m - 1) / 2.0
        total += 8.0 * cols + 16.0 * upd
    return total


def batched_bench(torch, skl, _lib, steps):
    """cfg4: the whole batch through one tuned plan replayed from a HIP graph, against the
    per-matrix GPU loop (same kernels, one launch per matrix) on a stratified sample."""
    import numpy as np
    dev = torch.device("cuda:0")
    ns, offs, buf = cfg4_batch(torch, dev)
    count = len(ns)
    lp = torch.empty(count, dtype=torch.float64, device=dev)
    sg = torch.empty(count, dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    plan = skl.LogPfPlan(ns, offs)
    tuned_ms = plan.tune(buf.data_ptr(), lp.data_ptr(), sg.data_ptr(), st)
    plan.status(st)
    cfg = plan.config()
    graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run(buf.data_ptr(), lp.data_ptr(), sg.data_ptr(), side.cuda_stream)   # warm-up off-graph
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    with torch.cuda.graph(graph):
        plan.run(buf.data_ptr(), lp.data_ptr(), sg.data_ptr(), torch.cuda.current_stream().cuda_stream)
    graph.replay()
    torch.cuda.synchronize()
    plan.status(st)
    ref_lp = lp.clone()
    ms = timed(torch, graph.replay, max(3, steps))
    plan.status(st)
    same = bool(torch.equal(lp, ref_lp))
    # per-matrix GPU loop on every 16th matrix (all 16 sizes represented): one skh_logpf_d_device
    # call per matrix, the library's single-matrix engine choice (one workgroup to n = 320, the
    # large engine above), as a user looping logabspfaffian over device matrices would
    o = _lib.opts()
    L = _lib.lib()
    sample = np.arange(0, count, 16)

    def loop():
        for k in sample:
            n = int(ns[k])
            rc = L.skh_logpf_d_device(n, buf.data_ptr() + 8 * int(offs[k]), n, lp.data_ptr() + 8 * int(k),
                                      sg.data_ptr() + 8 * int(k), st, ctypes.byref(o))
            _lib.check("skh_logpf_d_device", rc)

    loop()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loop()
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    plain_rate = len(sample) / loop_s
    flops = float(np.sum(4.0 / 3.0 * ns.astype(np.float64) ** 3))
    lp_h = lp.cpu().numpy()
    sg_h = sg.cpu().numpy()
    ref_h = ref_lp.cpu().numpy()
    loop_dev = float(np.max(np.abs(lp_h[sample] - ref_h[sample]) / np.maximum(1.0, np.abs(ref_h[sample]))))
    # CPU leg: the oracle's logabspfaffian per matrix (the reference user's loop), same sample
    import oracle
    threads = len(os.sched_getaffinity(0))
    oracle.set_threads(threads)
    host = [buf[int(offs[k]):int(offs[k]) + int(ns[k]) ** 2].cpu().numpy().reshape((int(ns[k]),) * 2, order="F")
            for k in sample]
    t0 = time.perf_counter()
    outs = [oracle.logabspfaffian(M) for M in host]
    cpu_s = time.perf_counter() - t0
    worst = max(abs(lp_h[k] - lo) / max(1.0, abs(lo)) for k, (lo, so) in zip(sample, outs))
    signs = all(sg_h[k] == so for k, (lo, so) in zip(sample, outs))
    tflops = flops / (ms * 1e-3) / 1e12
    kname = "k_cblock" if cfg["engine"] == "blocked" else "k_compact"
    traffic, tnote = measured_traffic("%s_batch4096_traffic.md" % kname, kname)
    abytes = batch_bytes(ns, cfg["panel_width"] if cfg["engine"] == "blocked" else 0)
    gbs = abytes / (ms * 1e-3) / 1e9
    return {"metric": "batched log|Pf|+sign, 4096 matrices n in {32,...,512}", "value": count / (ms * 1e-3),
            "unit": "matrices/s", "ms_per_batch": ms, "gflops": flops / (ms * 1e-3) / 1e9,
            "launch": "one tuned plan replayed from a HIP graph (%s engine, kernel NMAX %d, panel width %d, %d "
                      "workgroups); tuner best %.2f ms" % (cfg["engine"], cfg["kernel_nmax"], cfg["panel_width"],
                                                         cfg["workgroups"], tuned_ms),
            "graph_replays_identical": same,
            "roofline": {"bound": "hbm", "kernel": kname, "achieved": gbs,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                         "bytes_per_launch": abytes, "traffic": traffic, "traffic_source": tnote,
                         "fp64_tflops": tflops, "fp64_frac": tflops / FP64_MFMA_PEAK_TF,
                         "work": "%.4e bytes per launch: %s; sum (4/3) n^3 = %.4e flop (0.8 flop/byte)"
                                 % (abytes, BATCH_BYTES_RULE[cfg["engine"] == "blocked"], flops),
                         "note": "the graph replay (a fill kernel and the one-workgroup-per-matrix launch) timed "
                                 "with HIP events"},
            "plain": {"value": plain_rate, "unit": "matrices/s",
                      "sample": "per-matrix GPU loop, skh_logpf_d_device per matrix (library engine choice), every "
                                "16th matrix (%d)" % len(sample), "max_rel_dev_vs_batch": loop_dev},
            "speedup_vs_plain": (count / (ms * 1e-3)) / plain_rate,
            "oracle_check": {"matrices": len(sample), "max_rel_dev_logpf": worst, "signs_equal": bool(signs)},
            "cpu_baseline": {"value": len(sample) / cpu_s, "unit": "matrices/s (same size mix)", "cores": threads,
                             "kind": "port", "sample": "oracle logabspfaffian per matrix on every 16th matrix "
                                                        "(%d matrices, %.2f s), OpenBLAS threads=%d"
                                                        % (len(sample), cpu_s, threads)}}


def grad_bench(torch, skl, _lib, steps):
    """Forward log|Pf| + hand-written gradient for 1024 x n=256: one gradient plan replayed from a HIP
    graph (the timed figure; the eager device call is timed beside it), against the per-matrix GPU loop
    (the device entry, one matrix per call) on a sample and the oracle's per-matrix route on the CPU."""
    import numpy as np
    dev = torch.device("cuda:0")
    A, exact = cfg5_batch(torch, dev)
    count, n = A.shape[0], A.shape[1]
    At = A.transpose(1, 2).contiguous()   # column-major storage of each matrix
    ns = np.full(count, n, dtype=np.int64)
    offs = np.arange(count, dtype=np.int64) * n * n
    lp = torch.empty(count, dtype=torch.float64, device=dev)
    sg = torch.empty(count, dtype=torch.float64, device=dev)
    gr = torch.empty_like(At)
    st = torch.cuda.current_stream().cuda_stream
    o = _lib.opts()
    L = _lib.lib()

    def run():
        rc = L.skh_logpf_grad_batched_d_device(count, ns.ctypes.data, offs.ctypes.data, At.data_ptr(), lp.data_ptr(),
                                               sg.data_ptr(), gr.data_ptr(), st, ctypes.byref(o))
        _lib.check("skh_logpf_grad_batched_d_device", rc)

    run()
    torch.cuda.synchronize()
    eager_ms = timed(torch, run, max(3, steps))
    ref_g = gr.clone()
    # the same work through a gradient plan captured in a HIP graph (the timed figure)
    plan = skl.LogPfGradPlan(ns, offs)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        plan.run(At.data_ptr(), lp.data_ptr(), sg.data_ptr(), gr.data_ptr(), side.cuda_stream)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        plan.run(At.data_ptr(), lp.data_ptr(), sg.data_ptr(), gr.data_ptr(), torch.cuda.current_stream().cuda_stream)
    graph.replay()
    torch.cuda.synchronize()
    plan.status(st)
    ms = timed(torch, graph.replay, max(3, steps))
    plan.status(st)
    graph_same = bool(torch.equal(gr, ref_g))
    lp_err = float(torch.max(torch.abs(lp - exact) / torch.clamp(torch.abs(exact), min=1.0)))
    # gradient check on 4 matrices: -1/2 A^{-1} (LU on the GPU, float64)
    G = gr[:4].transpose(1, 2)
    G0 = -0.5 * torch.linalg.inv(A[:4])
    g_err = float(torch.max(torch.linalg.norm(G - G0, dim=(1, 2)) / torch.linalg.norm(G0, dim=(1, 2))))
    sample = 32
    one_n = np.array([n], dtype=np.int64)
    one_off = np.zeros(1, dtype=np.int64)

    def loop():
        for k in range(sample):
            rc = L.skh_logpf_grad_batched_d_device(1, one_n.ctypes.data, one_off.ctypes.data,
                                                   At.data_ptr() + 8 * k * n * n, lp.data_ptr() + 8 * k,
                                                   sg.data_ptr() + 8 * k, gr.data_ptr() + 8 * k * n * n, st,
                                                   ctypes.byref(o))
            _lib.check("skh_logpf_grad_batched_d_device", rc)

    loop()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loop()
    torch.cuda.synchronize()
    plain_rate = sample / (time.perf_counter() - t0)
    # CPU leg, the reference user's per-matrix route: logabspfaffian (src/pfaffian.jl:167-179) and the
    # gradient -1/2 inv(A) by LU (inv(A::SkewHermitian), src/skewhermitian.jl:233)
    import oracle
    threads = len(os.sched_getaffinity(0))
    oracle.set_threads(threads)
    ncpu = 64
    host = A[:ncpu].cpu().numpy()
    t0 = time.perf_counter()
    for k in range(ncpu):
        M = np.asfortranarray(host[k])
        oracle.logabspfaffian(M)
        G = -0.5 * np.linalg.inv(M)
        G = 0.5 * (G - G.T)
    cpu_s = time.perf_counter() - t0
    # work per matrix: forward (4/3) n^3, backward (7/3) n^3 (Q by blocked reflectors (4/3) n^3 and
    # K = Q_E C^-1 Q_O^T, n^3)
    flops = count * (11.0 / 3.0) * n ** 3
    tflops = flops / (ms * 1e-3) / 1e12
    return {"metric": "batched log|Pf| + gradient, 1024 matrices n=256", "value": count / (ms * 1e-3),
            "unit": "matrices/s", "ms_per_batch": ms, "logpf_max_rel_err_vs_exact": lp_err,
            "grad_max_rel_err_vs_inverse": g_err,
            "launch": "one gradient plan (forward launch + hand-written backward kernels) replayed from a HIP graph; "
                      "the eager device call takes %.2f ms; graph replay bitwise equal to it: %s" % (eager_ms, graph_same),
            "roofline": {"bound": "mfma", "kernel": "whole call (factorization + backward)", "achieved": tflops,
                         "peak": FP64_MFMA_PEAK_TF, "unit": "TFLOP/s", "frac": tflops / FP64_MFMA_PEAK_TF,
                         "work": "(11/3) n^3 flop per matrix: forward (4/3) n^3, backward (7/3) n^3",
                         "traffic": None},
            "plain": {"value": plain_rate, "unit": "matrices/s",
                      "sample": "per-matrix GPU loop (one device call per matrix), first %d matrices" % sample},
            "speedup_vs_plain": (count / (ms * 1e-3)) / plain_rate,
            "cpu_baseline": {"value": ncpu / cpu_s, "unit": "matrices/s", "cores": threads, "kind": "port",
                             "sample": "oracle logabspfaffian + numpy LU inverse (-1/2 inv(A)) per matrix, first %d "
                                       "matrices (%.2f s), OpenBLAS threads=%d" % (ncpu, cpu_s, threads)}}


def complex_bench(torch, _lib, steps, n=4096, seed=314159):
    """ComplexF64 n = 4096 skew-Hermitian A = (B - B^H)/2, Re and Im of B iid N(0,1): the device
    tridiagonalization (the reference's skewblockedhess! on SkewHermitian{ComplexF64}) timed over
    fresh HBM copies, then Q from the factors and the residual / orthogonality of A = Q T Q^H."""
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    B = torch.complex(torch.randn((n, n), generator=g, dtype=torch.float64, device=dev),
                      torch.randn((n, n), generator=g, dtype=torch.float64, device=dev))
    A = (B - B.conj().T) / 2
    del B
    A0 = A.T.contiguous()   # column-major storage of A
    work = torch.empty_like(A0)
    dtau = torch.empty(n - 1, dtype=torch.complex128, device=dev)
    deta = torch.empty(n - 1, dtype=torch.float64, device=dev)
    ddv = torch.empty(n, dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    L = _lib.lib()
    o = _lib.opts()

    def step():
        work.copy_(A0)
        _lib.check("skh_sytrd_z_device", L.skh_sytrd_z_device(n, work.data_ptr(), n, dtau.data_ptr(), deta.data_ptr(),
                                                             ddv.data_ptr(), st, ctypes.byref(o)))

    step()
    torch.cuda.synchronize()
    ms = timed(torch, step, max(3, steps))
    _lib.check("skh_device_status", L.skh_device_status(st))   # the last timed call abandoned no bounded wait
    dQ = torch.empty((n, n), dtype=torch.complex128, device=dev)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    _lib.check("skh_orgtr_z_device", L.skh_orgtr_z_device(n, work.data_ptr(), n, dtau.data_ptr(), dQ.data_ptr(), n, st))
    e1.record()
    torch.cuda.synchronize()
    q_ms = e0.elapsed_time(e1)
    Q = dQ.T.contiguous()
    T = torch.diag(1j * ddv.to(torch.complex128))
    k = torch.arange(n - 1, device=dev)
    T[k + 1, k] = deta.to(torch.complex128)
    T[k, k + 1] = -deta.to(torch.complex128)
    nA = float(torch.linalg.norm(A))
    res = float(torch.linalg.norm(Q @ T @ Q.conj().T - A)) / nA
    orth = float(torch.linalg.norm(Q.conj().T @ Q - torch.eye(n, dtype=torch.complex128, device=dev)))
    flops = 16.0 / 3.0 * n ** 3
    return {"metric": "complex skew-Hermitian n=4096 tridiagonalization (ComplexF64)",
            "value": flops / (ms * 1e-3) / 1e12, "unit": "TFLOP/s (16/3 n^3 real flops)", "ms": ms,
            "q_formation_ms": q_ms, "residual_QTQh_minus_A_rel": res, "orthogonality_QhQ_minus_I": orth}


def zpfaff_bench(torch, _lib, steps, count=4096, seed=2718):
    """Complex skew-symmetric Pfaffians (Parlett-Reid with pivoting, the reference's
    _logabspfaffian! on Matrix{ComplexF64}): `count` matrices A = R - R^T, n cycling through
    32, 64, ..., 256, Re and Im of R iid N(0, 1), resident in HBM; one batched device call per step.
    The oracle (C, one core) on every 65th matrix (each size 8 times, the batch's mix) gives the
    CPU rate and the agreement."""
    import numpy as np
    import oracle
    dev = torch.device("cuda:0")
    ns = np.array([32 * (1 + k % 8) for k in range(count)], dtype=np.int64)
    offs = np.zeros(count, dtype=np.int64)
    offs[1:] = np.cumsum(ns[:-1] ** 2)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    tot = int(offs[-1] + ns[-1] ** 2)
    buf = torch.empty(2 * tot, dtype=torch.float64, device=dev)
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        n = int(n)
        Rm = torch.complex(torch.randn((len(idx), n, n), generator=g, dtype=torch.float64, device=dev),
                           torch.randn((len(idx), n, n), generator=g, dtype=torch.float64, device=dev))
        Am = torch.view_as_real((Rm - Rm.transpose(1, 2)).transpose(1, 2).contiguous())   # column-major each
        for q, k in enumerate(idx):
            buf[2 * int(offs[k]):2 * int(offs[k] + n * n)] = Am[q].reshape(-1)
    lp = torch.empty(count, dtype=torch.float64, device=dev)
    ph = torch.empty(2 * count, dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    L = _lib.lib()

    def step():
        _lib.check("skh_logabspfaffian_z_batched_device", L.skh_logabspfaffian_z_batched_device(
            count, ns.ctypes.data, offs.ctypes.data, buf.data_ptr(), lp.data_ptr(), ph.data_ptr(), st, None))

    step()
    torch.cuda.synchronize()
    ms = timed(torch, step, max(3, steps))
    flops = float(np.sum(4.0 / 3.0 * ns.astype(np.float64) ** 3))
    # CPU leg: the oracle on every 65th matrix (65 = 1 mod 8: every size equally often)
    host = buf.cpu().numpy()
    glp = lp.cpu().numpy()
    sample = np.arange(0, count, 65)
    worst = 0.0
    t0 = time.perf_counter()
    outs = []
    for k in sample:
        n = int(ns[k])
        M = host[2 * int(offs[k]):2 * int(offs[k] + n * n)].view(np.complex128).reshape((n, n), order="F")
        outs.append(oracle.logabspfaffian_z(M)[0])
    cpu_s = time.perf_counter() - t0
    for k, lr in zip(sample, outs):
        worst = max(worst, abs(glp[k] - lr) / abs(lr))
    cpu_rate = len(sample) / cpu_s
    # one large matrix alone (the whole-GPU kernel)
    big_n = 4096
    Rb = torch.complex(torch.randn((big_n, big_n), generator=g, dtype=torch.float64, device=dev),
                       torch.randn((big_n, big_n), generator=g, dtype=torch.float64, device=dev))
    Ab = torch.view_as_real((Rb - Rb.T).T.contiguous()).reshape(-1)
    del Rb
    one_n = np.array([big_n], dtype=np.int64)
    one_off = np.zeros(1, dtype=np.int64)

    def big():
        _lib.check("skh_logabspfaffian_z_batched_device", L.skh_logabspfaffian_z_batched_device(
            1, one_n.ctypes.data, one_off.ctypes.data, Ab.data_ptr(), lp.data_ptr(), ph.data_ptr(), st, None))

    big()
    torch.cuda.synchronize()
    big_ms = timed(torch, big, 1)
    return {"metric": "batched complex skew-symmetric log|Pf| (Parlett-Reid), %d matrices n in {32,...,256}" % count,
            "value": count / (ms * 1e-3), "unit": "matrices/s", "ms_per_batch": ms,
            "gflops": flops / (ms * 1e-3) / 1e9, "flop_count": "4/3 n^3 real flops per matrix (lower triangle)",
            "max_rel_dev_logpf_vs_oracle": worst, "single_n%d_ms" % big_n: big_ms,
            "cpu_baseline": {"value": cpu_rate, "unit": "matrices/s (same size mix)", "cores": 1, "kind": "port",
                             "sample": "oracle on every 65th matrix (%d matrices, %.2f s)" % (len(sample), cpu_s)}}


def skewchol_bench(torch, _lib, steps, count=4096, seed=1618, big_n=4096):
    """Skew-Cholesky with complete pivoting (the reference's skewchol(A), RowMaximum): `count`
    real matrices A = M - M^T, n cycling through 32, 64, ..., 256, M iid N(0, 1), resident in HBM,
    one batched device call per step; then one n = 4096 matrix alone (one workgroup).  The oracle
    (C, one core) on every 65th matrix gives the CPU rate and the bitwise agreement."""
    import numpy as np
    import oracle
    dev = torch.device("cuda:0")
    ns = np.array([32 * (1 + k % 8) for k in range(count)], dtype=np.int64)
    offs = np.zeros(count, dtype=np.int64)
    offs[1:] = np.cumsum(ns[:-1] ** 2)
    tot = int(offs[-1] + ns[-1] ** 2)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    buf = torch.empty(tot, dtype=torch.float64, device=dev)
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        Mm = torch.randn((len(idx), int(n), int(n)), generator=g, dtype=torch.float64, device=dev)
        Am = Mm - Mm.transpose(1, 2)
        for q, k in enumerate(idx):
            buf[offs[k]:offs[k] + n * n] = Am[q].T.reshape(-1)
    R = torch.empty_like(buf)
    p = torch.empty(int(ns.sum()), dtype=torch.int64, device=dev)
    stp = torch.empty(count, dtype=torch.int64, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    L = _lib.lib()

    def step():
        _lib.check("skh_skewchol_d_batched_device", L.skh_skewchol_d_batched_device(
            count, ns.ctypes.data, offs.ctypes.data, buf.data_ptr(), R.data_ptr(), p.data_ptr(), stp.data_ptr(), 1,
            st, None))

    step()
    torch.cuda.synchronize()
    ms = timed(torch, step, max(3, steps))
    # n^3/3 flops: sum over steps of 2 products + 2 subtractions per entry of the trailing triangle
    flops = float(np.sum(ns.astype(np.float64) ** 3 / 3.0))
    host = buf.cpu().numpy()
    hR = R.cpu().numpy()
    sample = np.arange(0, count, 65)
    same = True
    t0 = time.perf_counter()
    outs = []
    for k in sample:
        n = int(ns[k])
        outs.append(oracle.skewchol(host[int(offs[k]):int(offs[k] + n * n)].reshape((n, n), order="F")))
    cpu_s = time.perf_counter() - t0
    for k, (Ro, po, so) in zip(sample, outs):
        n = int(ns[k])
        same = same and bool(np.array_equal(hR[int(offs[k]):int(offs[k] + n * n)].reshape((n, n), order="F"), Ro))
    # one large matrix
    Mb = torch.randn((big_n, big_n), generator=g, dtype=torch.float64, device=dev)
    Ab = (Mb - Mb.T).T.contiguous()
    del Mb
    Rb = torch.empty_like(Ab)
    pb = torch.empty(big_n, dtype=torch.int64, device=dev)
    one_n = np.array([big_n], dtype=np.int64)
    one_off = np.zeros(1, dtype=np.int64)

    def big():
        _lib.check("skh_skewchol_d_batched_device", L.skh_skewchol_d_batched_device(
            1, one_n.ctypes.data, one_off.ctypes.data, Ab.data_ptr(), Rb.data_ptr(), pb.data_ptr(), None, 1, st,
            None))

    big()
    torch.cuda.synchronize()
    big_ms = timed(torch, big, 1)
    return {"metric": "batched skew-Cholesky with complete pivoting, %d matrices n in {32,...,256}" % count,
            "value": count / (ms * 1e-3), "unit": "matrices/s", "ms_per_batch": ms,
            "gflops": flops / (ms * 1e-3) / 1e9, "flop_count": "n^3/3 flops per matrix",
            "bitwise_equal_to_oracle_on_sample": same,
            "single_n%d_ms" % big_n: big_ms,
            "cpu_baseline": {"value": len(sample) / cpu_s, "unit": "matrices/s (same size mix)", "cores": 1,
                             "kind": "port",
                             "sample": "oracle on every 65th matrix (%d matrices, %.2f s)" % (len(sample), cpu_s)}}


def matfun_bench(skl, sizes=(1024, 4096), cpu_n=1024, seed=577):
    """Matrix functions of a real skew matrix (the reference's exp(A) and sincos(A), src/exp.jl:3-48 and
    :208-256): whole host-API calls, host arrays in and out, A = (M - M^T) / (2 sqrt(n)), M iid N(0, 1).
    The GPU runs the tridiagonalization, Q, the eigenvector rotations, the products and the wrappers;
    the host runs the QR sweeps.  The oracle (numpy + C restatement) times exp at n = cpu_n."""
    import numpy as np
    import oracle
    out = {"metric": "exp(A) / sincos(A) of a real skew matrix, whole call (host arrays)", "unit": "ms"}
    for n in sizes:
        rng = np.random.default_rng(seed + n)
        M = rng.standard_normal((n, n))
        A = np.asfortranarray((M - M.T) / (2 * np.sqrt(n)))
        del M
        skl.exp(A[:64, :64] - A[:64, :64].T)   # library and BLAS start-up outside the timing
        t0 = time.perf_counter()
        E = skl.exp(A)
        t_exp = time.perf_counter() - t0
        t0 = time.perf_counter()
        S, C = skl.sincos(A)
        t_sc = time.perf_counter() - t0
        cols = E[:, :64]
        orth = float(np.abs(cols.T @ cols - np.eye(64)).max())
        pyth = float(np.abs(S.data[:64] @ S.data[:, :64] + C[:64] @ C[:, :64] - np.eye(64)).max())
        out["n%d" % n] = {"exp_ms": t_exp * 1e3, "sincos_ms": t_sc * 1e3, "exp_orthogonality_64cols": orth,
                          "sin2_plus_cos2_minus_I_64x64": pyth}
        if n == cpu_n:
            t0 = time.perf_counter()
            Eo = oracle.skewfun(A, "exp")
            cpu_s = time.perf_counter() - t0
            out["exp_max_abs_diff_vs_oracle_n%d" % n] = float(np.abs(E - Eo).max())
            out["cpu_baseline"] = {"value": cpu_s * 1e3, "unit": "ms per exp(A) at n=%d" % n, "cores": 16,
                                   "kind": "port", "sample": "one exp(A) by the oracle at n=%d (OpenBLAS "
                                   "threads for its products)" % n, "gpu_over_cpu": cpu_s / t_exp}
        log("matfun n=%d: exp %.1f ms, sincos %.1f ms" % (n, t_exp * 1e3, t_sc * 1e3))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nb", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=8192)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the n=512 plain-method line")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched log|Pf| and gradient lines")
    ap.add_argument("--no-complex", action="store_true", help="skip the complex n=4096 and complex Pfaffian lines")
    ap.add_argument("--matfun", action="store_true", help="also run the matrix-function (exp, sincos) line")
    ap.add_argument("--only", choices=["complex_pfaffian", "skewchol", "matfun", "batched", "grad"],
                    help="run just this secondary line and print its JSON (for profiling it alone)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the single-matrix tridiagonalization is a one-GPU "
              "kernel here (no multi-GPU path is built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import skewlinearalgebra as skl
    from skewlinearalgebra import _lib

    if args.only == "matfun":
        print(json.dumps(matfun_bench(skl)), flush=True)
        return
    if args.only in ("batched", "grad"):
        fn = batched_bench if args.only == "batched" else grad_bench
        print(json.dumps(fn(torch, skl, _lib, args.steps)), flush=True)
        return
    if args.only:
        fn = zpfaff_bench if args.only == "complex_pfaffian" else skewchol_bench
        print(json.dumps(fn(torch, _lib, args.steps)), flush=True)
        return

    n = args.n
    dev = torch.device("cuda:0")
    ld = n + (n & 1)
    g = torch.Generator(device=dev)
    g.manual_seed(314159)
    B = torch.randn((n, n), generator=g, dtype=torch.float64, device=dev)
    A0 = torch.zeros((n, ld), dtype=torch.float64, device=dev)   # column-major: row index fastest
    A0[:, :n] = (B - B.T).T
    del B
    work = torch.empty_like(A0)
    tau = torch.empty(n, dtype=torch.float64, device=dev)
    eta = torch.empty(n, dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    L = _lib.lib()
    o = _lib.opts(args.nb, _lib.ENGINE_LARGE | _lib.FLAG_TIMING)

    def step():
        work.copy_(A0)
        rc = L.skh_sytrd_d_device(n, work.data_ptr(), ld, tau.data_ptr(), eta.data_ptr(), st, ctypes.byref(o))
        _lib.check("skh_sytrd_d_device", rc)

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log("warmup %d done" % w)
    tm = [0.0] * 10
    out6 = (ctypes.c_double * 10)()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record()
    for k in range(args.steps):
        step()
        _lib.check("skh_diag_timings", L.skh_diag_timings(out6))
        _lib.check("skh_device_status", L.skh_device_status(st))   # no bounded wait abandoned
        tm = [a + b for a, b in zip(tm, out6)]
        log("step %d done" % k)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    flops = 4.0 / 3.0 * n ** 3
    tflops = flops / (ms * 1e-3) / 1e12

    panel_ms, panel_cnt, upd_ms, upd_cnt, panel_bytes, upd_flops = tm[:6]
    # one more factorization, outside the timed steps, with the panel kernel's phase clock on
    # (workgroup 0 stamps its phases, which slows that workgroup and so the whole launch)
    o_ph = _lib.opts(args.nb, _lib.ENGINE_LARGE | _lib.FLAG_TIMING | _lib.FLAG_PHASES)
    work.copy_(A0)
    _lib.check("skh_sytrd_d_device",
               L.skh_sytrd_d_device(n, work.data_ptr(), ld, tau.data_ptr(), eta.data_ptr(), st, ctypes.byref(o_ph)))
    clk = (ctypes.c_double * 13)()
    _lib.check("skh_diag_phase_clock", L.skh_diag_phase_clock(clk, 13))
    _lib.check("skh_diag_timings", L.skh_diag_timings(out6))
    phases = {"note": "separate instrumented factorization (phase clock on); ms of workgroup 0",
              "factorization_ms_with_clock": out6[0] + out6[2]}
    phases.update(dict(zip(("first_row_pass", "barrier1_arrive", "phase_b", "row_pass", "row_loads_and_wait",
                            "row_correct_w", "row_update_x", "row_handover", "b_loads_scalars", "b_t", "b_matvec",
                            "b_first_tile", "b_wait"), clk[:13])))
    panel_traffic, panel_tnote = None, "profiles exist for n = 8192 at nb 32 (default) and 64 only"
    if n == 8192 and args.nb in (0, 32, 64):
        panel_traffic, panel_tnote = measured_traffic(
            "k_panel_n8192_nb64_traffic.md" if args.nb == 64 else "k_panel_n8192_traffic.md", "k_panel")
    panel_avg_s = panel_ms / panel_cnt * 1e-3
    panel_bytes_per = panel_bytes / panel_cnt
    achieved = panel_bytes_per / panel_avg_s / 1e9
    upd_tf = (upd_flops / upd_cnt) / (upd_ms / upd_cnt * 1e-3) / 1e12 if upd_cnt else 0.0

    # sanity: the factorization of the last step reproduces the Frobenius norm (orthogonal similarity)
    eta_h = eta[: n - 1].double()
    fro_t = float(torch.sqrt(2.0 * torch.sum(eta_h * eta_h)))
    fro_a = float(torch.linalg.norm(A0[:, :n]))
    rel = abs(fro_t - fro_a) / fro_a

    res = {
        "metric": "n=8192 skew tridiag FP64 TFLOP/s",
        "value": tflops,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: A = B - B^T, B iid N(0,1), seed 314159, resident in HBM",
        "config": {"workload": "single real Float64 skew-symmetric matrix, blocked tridiagonalization",
                   "n": n, "nb": args.nb or "auto", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": "k_panel (skew matvec panel)", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": panel_traffic, "traffic_source": panel_tnote,
                     "bytes_per_launch": panel_bytes_per, "launch_ms": panel_avg_s * 1e3,
                     "launches_per_step": panel_cnt / args.steps},
        "mfma_update": {"kernel": "k_rank2k", "achieved": upd_tf, "peak": FP64_MFMA_PEAK_TF, "unit": "TFLOP/s",
                        "frac": upd_tf / FP64_MFMA_PEAK_TF, "ms_per_step": upd_ms / args.steps},
        "time_split_ms_per_step": {"panel": panel_ms / args.steps, "rank2k": upd_ms / args.steps,
                                   "panel_phases": phases},
        "check": {"frobenius_rel_err": rel},
        "wall_s": wall,
    }
    # panel width sweep over nb in {32, 64, 128} (one factorization each, outside the timed steps),
    # then the host implicit-QR eigenvalues of the tridiagonal (the reference's skewtrieigvals!)
    sweep = {}
    for nbv in (32, 64, 128):
        o_nb = _lib.opts(nbv, _lib.ENGINE_LARGE | _lib.FLAG_TIMING)
        work.copy_(A0)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        rc = L.skh_sytrd_d_device(n, work.data_ptr(), ld, tau.data_ptr(), eta.data_ptr(), st, ctypes.byref(o_nb))
        e1.record()
        torch.cuda.synchronize()
        _lib.check("skh_sytrd_d_device", rc)
        _lib.check("skh_device_status", L.skh_device_status(st))
        _lib.check("skh_diag_timings", L.skh_diag_timings(out6))
        t_ms = e0.elapsed_time(e1)
        sweep[str(nbv)] = {"ms": t_ms, "tflops": flops / (t_ms * 1e-3) / 1e12,
                           "panel_gbs": out6[4] / (out6[0] * 1e-3) / 1e9,
                           "rank2k_tflops": out6[5] / (out6[2] * 1e-3) / 1e12}
        log("nb=%d: %.1f ms" % (nbv, t_ms))
    res["nb_sweep"] = sweep
    step()
    torch.cuda.synchronize()
    ev_h = eta[: n - 1].cpu().numpy().copy()
    vals = np.zeros(n)
    t0 = time.perf_counter()
    rc = L.skh_skewtrieigvals_d(n, ev_h.ctypes.data, vals.ctypes.data)
    t_eig = time.perf_counter() - t0
    _lib.check("skh_skewtrieigvals_d", rc)
    lam2 = float(np.sum(vals * vals))
    res["host_eigvals"] = {"ms": t_eig * 1e3, "n": n,
                           "note": "host implicit QR on the GPU-computed tridiagonal (one core)",
                           "sum_lambda2_vs_frobenius2_rel": abs(lam2 - fro_a * fro_a) / (fro_a * fro_a)}
    log("host eigvals: %.1f ms" % (t_eig * 1e3))
    if not args.no_batched:
        log("batched log|Pf| (4096 matrices) ...")
        res["batched"] = batched_bench(torch, skl, _lib, args.steps)
        log(json.dumps(res["batched"]))
        log("batched gradient (1024 x 256) ...")
        res["grad"] = grad_bench(torch, skl, _lib, args.steps)
        log(json.dumps(res["grad"]))
    if not args.no_complex:
        log("complex n=4096 ...")
        res["complex"] = complex_bench(torch, _lib, args.steps)
        log(json.dumps(res["complex"]))
        log("complex skew-symmetric Pfaffians (4096 matrices) ...")
        res["complex_pfaffian"] = zpfaff_bench(torch, _lib, args.steps)
        log(json.dumps(res["complex_pfaffian"]))
    if not args.no_batched:
        log("skew-Cholesky (4096 matrices, one n=4096) ...")
        res["skewchol"] = skewchol_bench(torch, _lib, args.steps)
        log(json.dumps(res["skewchol"]))
    if not args.no_plain:
        log("plain method n=512 (hessenberg, eigvals, logabspfaffian) ...")
        res["plain512"] = plain_bench(torch, skl)
        log(json.dumps(res["plain512"]))
    if args.matfun:
        log("matrix functions (exp, sincos; n = 1024, 4096) ...")
        res["matfun"] = matfun_bench(skl)
        log(json.dumps(res["matfun"]))
    if not args.no_cpu:
        log("cpu baseline (n=%d, %d runs) ..." % (args.cpu_n, args.cpu_runs))
        res["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_runs)
        res["cpu_baseline"]["gpu_over_cpu"] = tflops / res["cpu_baseline"]["value"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
