#!/usr/bin/env python3
"""Benchmark of the dense matrix-multiply hot path on one MI355X.

Workload: Matrix.mult at 8192 x 8192 x 8192 in FP64 (entries 2r-1 from
java.util.Random seeds 1 and 2), operands resident in HBM.  A step is one
product C = A*B.  The metric is FP64-accurate GFLOP/s, crediting 2*M*N*K flops
per product whatever arithmetic computes it.

Forms measured on the same inputs: plain FP64 MFMA, and the split-precision
forms (BF16 digits, 6 slices; FP16 digits, 5 slices), each timed over split,
slice products, recombination and the FP64 residual check, with its error
measured against the compensated oracle on sampled entries.  `value` is the
fastest form whose sampled error is <= 1e-15 of (|A||B|)_ij.  The batched
workload (10,000 products with M, N, K log-uniform on [16, 512], N(0,1)
entries) is reported beside it.

Prints one JSON line.  `roofline` is for the dominant kernel of the form that
gives `value`, timed with HIP events on the stream it is launched on;
`cpu_baseline` times the oracle (a C restatement of the kernel Flag4j's
dispatcher selects at this shape, RealDenseMatMult.concurrentBlockedReordered)
on a bounded sample of the same product on this host's cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6     # MI355X dense FP64 matrix peak
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-split", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--profile-copy", action="store_true",
                    help="after the timed region, copy 512 MiB on the device once (calibrates HBM byte counters)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(A, B, n, threads):
    """Time the oracle's restatement of concurrentBlockedReordered (one task per
    64-row block on a `threads`-thread pool, ThreadManager.java:202-228) on the
    first R rows of the product; rows are independent, so GFLOP/s scales
    linearly to the full product."""
    import numpy as np
    from oracle import oracle as o
    rows = min(n, threads * 64 * 8)
    C = np.empty((rows, n))
    t0 = time.perf_counter()
    o.real_mt_blocked_reordered(A[:rows], B, C=C, rows=(0, rows), threads=threads, bs=64)
    dt = time.perf_counter() - t0
    return {"value": round(2.0 * rows * n * n / dt / 1e9, 3), "unit": "GFLOP/s", "cores": threads,
            "kind": "port",
            "sample": f"rows 0..{rows - 1} of the {n}^3 product ({2.0 * rows * n * n:.3g} flop, {dt:.1f} s), "
                      f"C restatement of RealDenseMatMult.concurrentBlockedReordered, {threads} threads"}


def exact_entries(np, A, B, I, J):
    """(AB)_ij correctly rounded, and (|A||B|)_ij, for sampled (i, j): every
    product a*b is split exactly into p + e (Dekker's TwoProduct with Veltkamp
    splitting) and the 2K terms are summed with math.fsum (exact, then
    rounded once)."""
    import math

    def split(x):
        c = 134217729.0 * x
        hi = c - (c - x)
        return hi, x - hi

    out, ab = np.empty(len(I)), np.empty(len(I))
    for s, (i, j) in enumerate(zip(I, J)):
        a, b = A[i, :], B[:, j]
        p = a * b
        ah, al = split(a)
        bh, bl = split(b)
        e = ((ah * bh - p) + ah * bl + al * bh) + al * bl
        out[s] = math.fsum(np.concatenate([p, e]))
        ab[s] = float(np.sum(np.abs(a) * np.abs(b)))
    return out, ab


def bench_batched(np, torch, f4j, rng, dev, count=10000, reps=10):
    """10,000 FP64 products, M, N, K log-uniform on [16, 512] (java.util.Random
    seed 4), N(0,1) entries drawn on the device; one grouped plan (tuned), its
    hipGraph replayed `reps` times."""
    from flag4j_amd import grouped
    dims = rng.log_uniform_ints(4, 3 * count, 16, 512).reshape(count, 3)
    M, N, K = dims[:, 0], dims[:, 1], dims[:, 2]
    oa = np.cumsum(np.r_[0, M * K])
    ob = np.cumsum(np.r_[0, K * N])
    oc = np.cumsum(np.r_[0, M * N])
    g = torch.Generator(device=dev)
    g.manual_seed(1)
    dA = torch.randn(int(oa[-1]), dtype=torch.float64, device=dev, generator=g)
    dB = torch.randn(int(ob[-1]), dtype=torch.float64, device=dev, generator=g)
    dC = torch.empty(int(oc[-1]), dtype=torch.float64, device=dev)
    plan = grouped.GroupedPlan(list(zip(M.tolist(), N.tolist(), K.tolist())),
                               [dA.data_ptr() + 8 * int(x) for x in oa[:-1]],
                               [dB.data_ptr() + 8 * int(x) for x in ob[:-1]],
                               [dC.data_ptr() + 8 * int(x) for x in oc[:-1]], K.tolist(), N.tolist(), N.tolist())
    plan.tune(2)
    st = torch.cuda.Stream(device=dev)
    plan.run(st.cuda_stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(reps):
        plan.run(st.cuda_stream)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    flops = float(np.sum(2.0 * M * N * K))
    roof = float(np.sum(np.maximum(2.0 * M * N * K / (PEAK_FP64_TFLOPS * 1e12), 8.0 * (M * K + K * N + M * N)
                                   / (PEAK_HBM_GBS * 1e9))))
    info = plan.info()
    plan.close()
    return {"workload": "10000 FP64 products, M,N,K log-uniform on [16,512], N(0,1)", "ms": round(ms, 4),
            "tflops": round(flops / ms / 1e9, 2), "per_problem_roofline_ms": round(roof * 1e3, 4),
            "roofline_frac": round(roof * 1e3 / ms, 4), "tile_cfg_per_bucket": info["tile_cfg"],
            "tiles_per_bucket": info["tiles"]}


def load_traffic(name):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import flag4j_amd as f4j
    from flag4j_amd import rng

    n = args.size
    log(f"[bench] generating {n}x{n} inputs")
    A = rng.uniform(1, n, n)
    B = rng.uniform(2, n, n)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    dA = torch.from_numpy(A).to(dev)
    dB = torch.from_numpy(B).to(dev)
    dC = torch.empty((n, n), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    sh = stream.cuda_stream

    def step():
        f4j.dgemm_dev("N", "N", n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, dC.data_ptr(), n,
                      precision="fp64", stream=sh)

    log("[bench] warmup")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()

    # Timed region: K back-to-back products, events on the launch stream.
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(stream)
        step()
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [e0.elapsed_time(e1) for e0, e1 in evs]
    ms_per_step = wall * 1e3 / args.steps
    flops = 2.0 * n * n * n
    value = flops / (ms_per_step * 1e-3) / 1e9
    kavg = float(np.mean(kern_ms))
    achieved = flops / (kavg * 1e-3) / 1e12
    log(f"[bench] kernel avg {kavg:.3f} ms, {achieved:.2f} TFLOP/s")
    import ctypes
    kbuf = ctypes.create_string_buffer(64)
    f4j._native.check(f4j._native.load().f4j_dispatch_name(0, 0, 0, n, n, n, 0, kbuf, 64))
    fp64_kernel = kbuf.value.decode()

    # Accuracy of the measured product on sampled entries, against exact sums.
    r = np.random.default_rng(0)
    I, J = r.integers(0, n, 512), r.integers(0, n, 512)
    exact, ab = exact_entries(np, A, B, I, J)

    def sampled_err():
        Ch = dC.cpu().numpy()
        return float(np.max(np.abs(Ch[I, J] - exact) / ab))

    max_err = sampled_err()
    forms = {"fp64": {"ms": round(ms_per_step, 3), "gflops": round(value, 1), "max_err": max_err}}
    best = ("fp64", value)
    if not args.no_split:
        for prec, S in (("split_bf16", 6), ("split_fp16", 5)):
            def split_step():
                f4j.dgemm_dev("N", "N", n, n, n, dA.data_ptr(), n, dB.data_ptr(), n, dC.data_ptr(), n,
                              precision=prec, slices=S, stream=sh)
            split_step()
            torch.cuda.synchronize()
            reps = max(2, args.steps // 4)
            t0 = time.perf_counter()
            for _ in range(reps):
                split_step()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / reps
            rep = f4j._native.last_split_report()
            err = sampled_err()
            gf = flops / (ms * 1e-3) / 1e9
            forms[prec] = {"slices": S, "products": rep.products, "ms": round(ms, 3), "gflops": round(gf, 1),
                           "max_err": err, "residual_check_err": rep.max_rel_err,
                           "delivered": "split" if rep.precision != 0 else "fp64 fallback"}
            log(f"[bench] {prec} S={S}: {ms:.2f} ms, {gf:.0f} GFLOP/s, err {err:.2e}")
            if err <= 1e-15 and rep.precision != 0 and gf > best[1]:
                best = (prec, gf)

    if args.profile_copy:
        src = torch.empty(64 * 1024 * 1024, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    batched = None
    if not args.no_batched:
        batched = bench_batched(np, torch, f4j, rng, dev)
        log(f"[bench] batched: {batched}")

    # End-to-end through the host-array ABI (PCIe included), one call.
    Ce = np.empty((n, n))
    t1 = time.perf_counter()
    f4j._native.check(f4j._native.load().f4j_dgemm(0, 0, n, n, n, A.ctypes.data, n, B.ctypes.data, n,
                                                   Ce.ctypes.data, n, 0, 0))
    e2e_s = time.perf_counter() - t1

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or min(len(os.sched_getaffinity(0)), 16)
        log(f"[bench] CPU baseline on {threads} threads")
        cpu = cpu_baseline(A, B, n, threads)

    out = {
        "metric": "FP64-accurate GFLOP/s, Matrix.mult 8192^3",
        "value": round(best[1], 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: java.util.Random seeds 1 (A) and 2 (B), entries 2r-1 in [-1,1)",
        "config": {"workload": f"Matrix.mult FP64 {n}x{n}x{n} (C = A*B, row-major, device-resident)",
                   "M": n, "N": n, "K": n, "form": best[0], "fp64_kernel": fp64_kernel},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                     "traffic": load_traffic(f"dgemm_f64_{n}_pmc.txt")},
        "cpu_baseline": cpu,
        "forms": forms,
        "accuracy": {"max_err_over_absAabsB_sampled": max_err, "samples": 512,
                     "reference": "exact sums (TwoProduct + fsum) on sampled entries"},
        "batched": batched,
        "end_to_end_gflops_host_arrays": round(flops / e2e_s / 1e9, 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
