#!/usr/bin/env python3
"""Benchmark of the dense matrix-multiply hot path on one MI355X.

Workload: Matrix.mult at 8192 x 8192 x 8192 in FP64 (entries 2r-1 from
java.util.Random seeds 1 and 2), operands resident in HBM.  A step is one
product C = A*B.  The metric is FP64-accurate GFLOP/s, crediting 2*M*N*K flops
per product.

Prints one JSON line.  `roofline` is for the dominant kernel, timed with HIP
events on the stream it is launched on; `cpu_baseline` times the oracle (a C
restatement of the kernel Flag4j's dispatcher selects at this shape,
RealDenseMatMult.concurrentBlockedReordered) on a bounded sample of the same
product on this host's cores.
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

    # Accuracy of the measured product on sampled entries (compensated oracle).
    Ch = dC.cpu().numpy()
    r = np.random.default_rng(0)
    I, J = r.integers(0, n, 512), r.integers(0, n, 512)
    from oracle import oracle as o
    hi, lo, ab = o.real_dot2(A, B, I, J)
    max_err = float(np.max(np.abs((Ch[I, J] - hi) - lo) / ab))

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
        "value": round(value, 2),
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
                   "M": n, "N": n, "K": n, "form": "plain FP64 MFMA", "kernel": "MFMA_F64_128x128x16"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
                     "traffic": load_traffic(f"dgemm_f64_{n}_pmc.txt")},
        "cpu_baseline": cpu,
        "accuracy": {"max_err_over_absAabsB_sampled": max_err, "samples": 512},
        "end_to_end_gflops_host_arrays": round(flops / e2e_s / 1e9, 2),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
