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
                threads) timed on a bounded sample: one full factorization at n=4096
  batched       (with --batched) 4096 uneven-size matrices, log|Pf| + sign, matrices/s

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 8192] [--nb 0] [--no-cpu] [--batched]
"""
import argparse
import ctypes
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "skewlinearalgebra.jl_amd"))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0
FP64_MFMA_PEAK_TF = 78.6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n_cpu):
    import numpy as np
    import oracle
    threads = len(os.sched_getaffinity(0))
    oracle.lib()
    oracle.set_threads(threads)
    rng = np.random.default_rng(314159)
    B = rng.standard_normal((n_cpu, n_cpu))
    A = np.asfortranarray(B - B.T)
    t0 = time.perf_counter()
    oracle.skewblockedhess(A)
    dt = time.perf_counter() - t0
    flops = 4.0 / 3.0 * n_cpu ** 3
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": "one full factorization at n=%d (same generator), %.2f s, OpenBLAS threads=%d"
                      % (n_cpu, dt, threads)}


def batched_bench(torch, skl, _lib, steps):
    """4096 real skew matrices, n uniform over {32,64,...,512}, B iid N(0,1/n), in HBM."""
    import numpy as np
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(314159)
    rng = np.random.default_rng(314159)
    count = 4096
    ns = (rng.integers(1, 17, size=count) * 32).astype(np.int64)
    offs = np.zeros(count, dtype=np.int64)
    offs[1:] = np.cumsum(ns[:-1] ** 2)
    total = int(offs[-1] + ns[-1] ** 2)
    buf = torch.empty(total, dtype=torch.float64, device=dev)
    for k in range(count):
        n = int(ns[k])
        Bm = torch.randn((n, n), generator=g, dtype=torch.float64, device=dev) / (n ** 0.5)
        buf[offs[k]:offs[k] + n * n] = (Bm - Bm.T).T.reshape(-1)
    lp = torch.empty(count, dtype=torch.float64, device=dev)
    sg = torch.empty(count, dtype=torch.float64, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    o = _lib.opts()
    L = _lib.lib()

    def run():
        rc = L.skh_logpf_batched_d_device(count, ns.ctypes.data, offs.ctypes.data, buf.data_ptr(), lp.data_ptr(),
                                          sg.data_ptr(), st, ctypes.byref(o))
        _lib.check("skh_logpf_batched_d_device", rc)

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    return {"metric": "batched log|Pf|+sign, 4096 matrices n in {32..512}", "value": count / dt,
            "unit": "matrices/s", "ms_per_batch": dt * 1e3, "flop_per_batch": float(np.sum(4.0 / 3.0 * ns ** 3.0))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--nb", type=int, default=0)
    ap.add_argument("--cpu-n", type=int, default=4096)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the single-matrix tridiagonalization is a one-GPU "
              "kernel here (no multi-GPU path is built); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import skewlinearalgebra as skl
    from skewlinearalgebra import _lib

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
        tm = [a + b for a, b in zip(tm, out6)]
        log("step %d done" % k)
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    flops = 4.0 / 3.0 * n ** 3
    tflops = flops / (ms * 1e-3) / 1e12

    panel_ms, panel_cnt, upd_ms, upd_cnt, panel_bytes, upd_flops = tm[:6]
    phases = {k: v / args.steps for k, v in zip(("row_pass", "barrier1", "matvec", "barrier2"), tm[6:])}
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
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "bytes_per_launch": panel_bytes_per, "launch_ms": panel_avg_s * 1e3,
                     "launches_per_step": panel_cnt / args.steps},
        "mfma_update": {"kernel": "k_rank2k", "achieved": upd_tf, "peak": FP64_MFMA_PEAK_TF, "unit": "TFLOP/s",
                        "frac": upd_tf / FP64_MFMA_PEAK_TF, "ms_per_step": upd_ms / args.steps},
        "time_split_ms_per_step": {"panel": panel_ms / args.steps, "rank2k": upd_ms / args.steps,
                                   "panel_phases_wg0": phases},
        "check": {"frobenius_rel_err": rel},
        "wall_s": wall,
    }
    if args.batched:
        res["batched"] = batched_bench(torch, skl, _lib, max(1, args.steps))
    if not args.no_cpu:
        log("cpu baseline (n=%d) ..." % args.cpu_n)
        res["cpu_baseline"] = cpu_baseline(args.cpu_n)
        res["cpu_baseline"]["gpu_over_cpu"] = tflops / res["cpu_baseline"]["value"]
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
