#!/usr/bin/env python3
"""Benchmark of the aclfft FP32 Stockham C2C path on one MI355X.

Workload (the headline metric's configuration): 1D C2C forward FFT, complex64, N = 1024,
batch = 65536, contiguous, re/im ~ N(0, 1/2) from seed 0 (synthetic), input resident in HBM.
A step is one aclfftExecC2C over the whole batch (one kernel launch).

Prints ONE JSON line:
  value      = C2C GFLOP/s = 5 N log2 N batch / t
  roofline   = algorithmic HBM bytes (16 B per point) / average kernel time vs 8 TB/s
  cpu_baseline = the FP64 oracle (oracle/, a restatement of the reference's Stockham) timed on the
                 host on a bounded sample of the same workload
  cpu_golden_path = torch.fft.fft on host cores (the reference's golden path) on the full workload
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ops-fft_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N = 1024
BATCH = 65536


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def read_traffic(name):
    """HBM bytes per launch from a committed rocprofv3 --pmc summary (profiles/<name>), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch:"):
                return float(line.split(":", 1)[1].strip())
    return None


def cpu_baseline(x_sample_np, budget_s=10.0):
    """FP64 oracle on a bounded sample: transforms per second, scaled to GFLOP/s."""
    from oracle import oracle
    import numpy as np
    n = x_sample_np.shape[-1]
    done = 0
    t0 = time.perf_counter()
    while True:
        oracle.c2c(x_sample_np)
        done += x_sample_np.shape[0]
        el = time.perf_counter() - t0
        if el >= budget_s:
            break
    gflops = 5.0 * n * math.log2(n) * done / el / 1e9
    return dict(value=round(gflops, 4), unit="GFLOP/s", cores=1, kind="port",
                sample="%d transforms of N=%d (FP64 oracle, single thread), %.1f s" % (done, n, el))


def cpu_golden(x_cpu, torch):
    threads = len(os.sched_getaffinity(0))
    torch.set_num_threads(threads)
    ts = []
    for _ in range(3):
        t0 = time.perf_counter()
        torch.fft.fft(x_cpu, dim=1)
        ts.append(time.perf_counter() - t0)
    t = sorted(ts)[len(ts) // 2]
    n, b = x_cpu.shape[1], x_cpu.shape[0]
    return dict(value=round(5.0 * n * math.log2(n) * b / t / 1e9, 3), unit="GFLOP/s", cores=threads,
                kind="torch.fft (reference golden path)", sample="full workload, median of 3, %.3f s" % t)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the FFT path has no multi-GPU mode in this build "
              "(batch sharding across GPUs needs no communication; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import aclfft
    from oracle import oracle

    dev = torch.device("cuda:0")
    g = torch.Generator().manual_seed(0)
    x_cpu = torch.randn(BATCH, N, dtype=torch.complex64, generator=g)
    x = x_cpu.to(dev)
    y = torch.empty_like(x)
    stream = torch.cuda.current_stream(dev)
    plan = aclfft.Plan.plan1d(N, aclfft.C2C, BATCH)
    plan.set_stream(stream)
    log("bench: plan ready, warmup %d" % args.warmup)
    for _ in range(args.warmup):
        plan.exec_c2c(x, y, aclfft.FORWARD)
    torch.cuda.synchronize()

    # parity on a subset of the output against the FP64 oracle
    sub = 256
    ref = oracle.c2c(x_cpu[:sub].numpy().astype(np.complex128))
    err = oracle.rel_l2(y[:sub].cpu().numpy(), ref)
    log("bench: rel_l2 vs FP64 oracle on %d transforms = %.3e" % (sub, err))

    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        plan.exec_c2c(x, y, aclfft.FORWARD)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    log("bench: %.4f ms per step" % ms)

    flops = 5.0 * N * math.log2(N) * BATCH
    gflops = flops / (ms * 1e-3) / 1e9
    algo_bytes = 16.0 * N * BATCH
    achieved = algo_bytes / (ms * 1e-3) / 1e9
    traffic = read_traffic("traffic_c2c_n1024_b65536.txt")

    out = {
        "metric": "C2C GFLOP/s (5NlogN) + %HBM peak, MFMA-split vs FP32 Stockham at rel-L2<=1e-6",
        "value": round(gflops, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (torch.randn complex64, seed 0), resident in HBM",
        "config": {"workload": "1D C2C forward FFT, complex64, N=1024, batch=65536, contiguous",
                   "n": N, "batch": BATCH, "path": "FP32 Stockham (aclfftExecC2C, device pointers)"},
        "rel_l2_vs_fp64": err,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
    }
    if not args.no_cpu:
        sample = x_cpu[:512].numpy().astype(np.complex128)
        out["cpu_baseline"] = cpu_baseline(sample)
        out["cpu_golden_path"] = cpu_golden(x_cpu, torch)
    plan.destroy()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
