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
import ctypes
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


def warm(torch, fn, warmup, min_seconds=0.1):
    """`warmup` untimed calls, continued until at least `min_seconds` of device time has passed:
    the clocks settle over tens of milliseconds of sustained load, so a few calls of a
    sub-millisecond workload are not enough."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < min_seconds:
        for _ in range(10):
            fn()
        torch.cuda.synchronize()


def time_steps(torch, fn, steps, warmup, stream):
    warm(torch, fn, warmup)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def extra_copy(torch, aclfft, x, steps, stream, headline_ms):
    """Device-to-device copies of the same 1 GiB (512 MiB read + 512 MiB written): torch's copy and
    the library's streaming copy (16-byte non-temporal loads and stores, the access pattern no
    transform that reads and writes every byte once can beat): the achievable HBM bandwidth
    against which the headline kernel's fraction of the published peak is read."""
    from aclfft._lib import stream_copy
    y = torch.empty_like(x)
    b = 2.0 * x.numel() * x.element_size()
    out = {}
    ms = time_steps(torch, lambda: y.copy_(x), steps, 3, stream)
    out["torch_copy"] = {"ms": round(ms, 5), "GB/s": round(b / ms / 1e6, 1),
                         "frac_of_peak": round(b / ms / 1e6 / HBM_PEAK_GBS, 4)}
    h = ctypes.c_void_p(stream.cuda_stream)
    nb = x.numel() * x.element_size()
    ms = time_steps(torch, lambda: stream_copy(x.data_ptr(), y.data_ptr(), nb, h), steps, 3, stream)
    out["stream_copy"] = {"ms": round(ms, 5), "GB/s": round(b / ms / 1e6, 1),
                          "frac_of_peak": round(b / ms / 1e6 / HBM_PEAK_GBS, 4)}
    out["headline_vs_stream_copy"] = round(ms / headline_ms, 4)
    if not torch.equal(x, y):
        out["error"] = "stream copy mismatch"
    return out


def sixty_db_signal(np, batch, seed):
    """The precision workload's 60 dB case: per transform, 8 complex tones at random bins
    (amplitudes 1 x4 and 1e-3 x4, random phases) plus complex Gaussian noise of deviation 1e-6."""
    rng = np.random.default_rng(seed)
    t = np.arange(N)
    x = np.empty((batch, N), np.complex128)
    amp = np.array([1, 1, 1, 1, 1e-3, 1e-3, 1e-3, 1e-3])
    for b in range(batch):
        f = rng.choice(N, 8, replace=False)
        ph = rng.uniform(0, 2 * np.pi, 8)
        x[b] = (amp * np.exp(1j * (2 * np.pi * np.outer(t, f) / N + ph))).sum(1)
    x += 1e-6 * (rng.standard_normal((batch, N)) + 1j * rng.standard_normal((batch, N)))
    return x.astype(np.complex64)


def extra_mfma(torch, aclfft, oracle, np, x, x_cpu, y, steps, stream, e32):
    out = {}
    bytes_ = 16.0 * N * BATCH
    # the 60 dB case: 64 transforms, every path against the FP64 oracle
    s60 = sixty_db_signal(np, 64, 3)
    ref60 = oracle.c2c(s60.astype(np.complex128))
    s60d = torch.from_numpy(s60).cuda()

    def err60(mode):
        q = aclfft.Plan.plan1d(N, aclfft.C2C, 64)
        if mode:
            q.set_precision(mode)
        o = torch.empty_like(s60d)
        q.exec_c2c(s60d, o, aclfft.FORWARD)
        torch.cuda.synchronize()
        q.destroy()
        return oracle.rel_l2(o.cpu().numpy(), ref60)

    out["sixty_db_fp32_rel_l2_vs_fp64"] = err60(0)
    for mode, name in ((aclfft.PRECISION_BF16, "bf16"), (aclfft.PRECISION_FP16X2, "fp16x2"),
                       (aclfft.PRECISION_BF16X3, "bf16x3")):
        p = aclfft.Plan.plan1d(N, aclfft.C2C, BATCH)
        p.set_precision(mode)
        p.set_stream(stream)
        ms = time_steps(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), steps, 3, stream)
        ref = oracle.c2c(x_cpu[:256].numpy().astype(np.complex128))
        err = oracle.rel_l2(y[:256].cpu().numpy(), ref)
        out[name] = {"ms": round(ms, 5), "GFLOP/s": round(5.0 * N * math.log2(N) * BATCH / ms / 1e6, 1),
                     "hbm_frac": round(bytes_ / ms / 1e6 / HBM_PEAK_GBS, 4), "rel_l2_vs_fp64": err,
                     "err_ratio_vs_fp32": round(err / e32, 3), "sixty_db_rel_l2_vs_fp64": err60(mode)}
        p.destroy()
    return out


def extra_real(torch, aclfft, oracle, np, steps, stream):
    out = {}
    for n, b in ((4096, 8192), (3000, 8192), (8192, 1024)):
        g = torch.Generator().manual_seed(1)
        xr = torch.randn(b, n, generator=g).cuda()
        X = torch.empty(b, n // 2 + 1, dtype=torch.complex64, device="cuda")
        xb = torch.empty_like(xr)
        pr = aclfft.Plan.plan1d(n, aclfft.R2C, b)
        pc = aclfft.Plan.plan1d(n, aclfft.C2R, b)
        pr.set_stream(stream)
        pc.set_stream(stream)
        ms_r = time_steps(torch, lambda: pr.exec_r2c(xr, X), steps, 3, stream)
        ms_c = time_steps(torch, lambda: pc.exec_c2r(X, xb), steps, 3, stream)
        byts = (4.0 * n + 8.0 * (n // 2 + 1)) * b
        rt = float(((xb / n - xr).norm() / xr.norm()).item())
        err = oracle.rel_l2(X[:64].cpu().numpy(), oracle.r2c(xr[:64].cpu().numpy().astype(np.float64)))
        out["n%d_b%d" % (n, b)] = {
            "r2c_ms": round(ms_r, 5), "c2r_ms": round(ms_c, 5),
            "r2c_hbm_frac": round(byts / ms_r / 1e6 / HBM_PEAK_GBS, 4),
            "c2r_hbm_frac": round(byts / ms_c / 1e6 / HBM_PEAK_GBS, 4),
            "r2c_GFLOP/s_2.5nlogn": round(2.5 * n * math.log2(n) * b / ms_r / 1e6, 1),
            "roundtrip_rel_l2": rt, "r2c_rel_l2_vs_fp64": err, "plan_r2c": pr.describe()}
        pr.destroy()
        pc.destroy()
    return out


def extra_double(torch, aclfft, oracle, np, steps, stream):
    """Double precision (aclfftExecZ2Z / D2Z / Z2D): C2C n=1024 batch 32768 (512 MiB of complex128
    each way, the headline's points at twice the bytes) and R2C/C2R n=4096 batch 4096."""
    out = {}
    n, b = 1024, 32768
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.randn(b, n, dtype=torch.complex128, device="cuda", generator=g)
    y = torch.empty_like(x)
    p = aclfft.Plan.plan1d(n, aclfft.Z2Z, b)
    p.set_stream(stream)
    ms = time_steps(torch, lambda: p.exec_z2z(x, y, aclfft.FORWARD), steps, 3, stream)
    err = oracle.rel_l2(y[:64].cpu().numpy(), oracle.c2c(x[:64].cpu().numpy()))
    out["z2z_n1024_b32768"] = {"ms": round(ms, 5), "hbm_frac": round(32.0 * n * b / ms / 1e6 / HBM_PEAK_GBS, 4),
                               "GFLOP/s": round(5.0 * n * math.log2(n) * b / ms / 1e6, 1),
                               "rel_l2_vs_fp64_oracle": err, "plan": p.describe()}
    p.destroy()
    del x, y
    n, b = 4096, 4096
    xr = torch.randn(b, n, dtype=torch.float64, device="cuda", generator=g)
    X = torch.empty(b, n // 2 + 1, dtype=torch.complex128, device="cuda")
    xb = torch.empty_like(xr)
    pr = aclfft.Plan.plan1d(n, aclfft.D2Z, b)
    pc = aclfft.Plan.plan1d(n, aclfft.Z2D, b)
    pr.set_stream(stream)
    pc.set_stream(stream)
    ms_r = time_steps(torch, lambda: pr.exec_d2z(xr, X), steps, 3, stream)
    ms_c = time_steps(torch, lambda: pc.exec_z2d(X, xb), steps, 3, stream)
    byts = (8.0 * n + 16.0 * (n // 2 + 1)) * b
    err = oracle.rel_l2(X[:32].cpu().numpy(), oracle.r2c(xr[:32].cpu().numpy()))
    rt = float(((xb / n - xr).norm() / xr.norm()).item())
    out["d2z_z2d_n4096_b4096"] = {"d2z_ms": round(ms_r, 5), "z2d_ms": round(ms_c, 5),
                                  "d2z_hbm_frac": round(byts / ms_r / 1e6 / HBM_PEAK_GBS, 4),
                                  "z2d_hbm_frac": round(byts / ms_c / 1e6 / HBM_PEAK_GBS, 4),
                                  "d2z_rel_l2_vs_fp64_oracle": err, "roundtrip_rel_l2": rt,
                                  "plan_d2z": pr.describe()}
    pr.destroy()
    pc.destroy()
    return out


def extra_small_latency(torch, aclfft, oracle, np, stream):
    """The launch-bound configuration (C2C N=1024, batch 64, 512 KiB each way): time per call,
    eager and replayed from a hipGraph, measured over many back-to-back calls (device time per
    call with events, host time per call with the wall clock)."""
    x = torch.randn(64, 1024, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    out = {}
    for name, graph in (("eager", False), ("graph", True)):
        p = aclfft.Plan.plan1d(1024, aclfft.C2C, 64)
        p.set_stream(stream)
        p.set_graph_mode(graph)
        fn = lambda: p.exec_c2c(x, y, aclfft.FORWARD)  # noqa: E731
        warm(torch, fn, 20)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        calls = 2000
        e0.record(stream)
        t0 = time.perf_counter()
        for _ in range(calls):
            fn()
        t1 = time.perf_counter()
        e1.record(stream)
        torch.cuda.synchronize()
        out[name] = {"device_us_per_call": round(e0.elapsed_time(e1) * 1e3 / calls, 2),
                     "host_us_per_call": round((t1 - t0) * 1e6 / calls, 2)}
        p.destroy()
    err = oracle.rel_l2(y.cpu().numpy(), oracle.c2c(x.cpu().numpy().astype(np.complex128)))
    out["rel_l2_vs_fp64"] = err
    # device cost without host overhead: 100 eager calls captured into one torch CUDA graph
    # (hipGraph), replayed; each replay's output is checked against the eager result
    p = aclfft.Plan.plan1d(1024, aclfft.C2C, 64)
    p.set_graph_mode(False)
    side = torch.cuda.Stream()
    p.set_stream(side)
    ref = y.clone()
    with torch.cuda.stream(side):
        for _ in range(3):
            p.exec_c2c(x, y, aclfft.FORWARD)
    side.synchronize()
    g = torch.cuda.CUDAGraph()
    per_graph = 100
    with torch.cuda.graph(g, stream=side, capture_error_mode="relaxed"):
        for _ in range(per_graph):
            p.exec_c2c(x, y, aclfft.FORWARD)
    replays = 20
    for _ in range(3):
        y.zero_()
        g.replay()
        torch.cuda.synchronize()
        if not torch.equal(y, ref):
            out["captured_error"] = "replay differs from the eager result"
    cur = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(cur)
    for _ in range(replays):
        g.replay()
    e1.record(cur)
    torch.cuda.synchronize()
    out["captured_graph"] = {"device_us_per_call": round(e0.elapsed_time(e1) * 1e3 / (replays * per_graph), 2),
                             "calls_per_graph": per_graph}
    del g
    p.destroy()
    out["ideal_us_at_8TBs"] = round(2 * x.numel() * 8 / 8e12 * 1e6, 3)
    return out


def extra_2d(torch, aclfft, oracle, np, steps, stream):
    B, NX, NY = 4, 2048, 2048
    g = torch.Generator().manual_seed(2)
    x = torch.randn(B, NX, NY, dtype=torch.complex64, generator=g).cuda()
    y = torch.empty_like(x)
    p = aclfft.Plan.plan2d(B, NX, NY, aclfft.C2C)
    p.set_stream(stream)
    ms = time_steps(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), steps, 3, stream)
    err = oracle.rel_l2(y[:1].cpu().numpy(), oracle.c2c_2d(x[:1].cpu().numpy().astype(np.complex128)))
    pts = B * NX * NY
    out = {"ms": round(ms, 5), "GFLOP/s": round(5.0 * pts * math.log2(NX * NY) / ms / 1e6, 1),
           "hbm_frac_compulsory": round(16.0 * pts / ms / 1e6 / HBM_PEAK_GBS, 4),
           "hbm_frac_two_pass": round(32.0 * pts / ms / 1e6 / HBM_PEAK_GBS, 4),
           "rel_l2_vs_fp64": err, "plan": p.describe()}
    p.destroy()
    return out


def extra_loss_grad(torch, steps, stream):
    """The differentiable configuration: loss = sum |W * FFT(x)|^2 (W ~ U[0,1), fixed) and its
    gradient through torch.autograd over the library (forward transform, then the adjoint kernel),
    for 2D C2C 2048x2048 batch 4 and 1D R2C n=8192 batch 1024.  One step = forward + backward."""
    from aclfft import autograd as AG
    out = {}
    g = torch.Generator(device="cuda").manual_seed(7)
    for name, shape, fn, wshape in (("fft2_2048x2048_b4", (4, 2048, 2048), AG.fft2, (4, 2048, 2048)),
                                    ("rfft_8192_b1024", (1024, 8192), AG.rfft, (1024, 4097))):
        if name.startswith("fft2"):
            x = torch.randn(shape, dtype=torch.complex64, device="cuda", generator=g)
        else:
            x = torch.randn(shape, device="cuda", generator=g)
        x.requires_grad_(True)
        W = torch.rand(wshape, device="cuda", generator=g)

        def step():
            x.grad = None
            loss = (W * fn(x)).abs().pow(2).sum()
            loss.backward()
        ms = time_steps(torch, step, steps, 3, stream)
        fft_ms = time_steps(torch, lambda: fn(x.detach()), steps, 3, stream)
        out[name] = {"fwd_bwd_ms": round(ms, 5), "transform_ms": round(fft_ms, 5),
                     "note": "step = transform + loss (torch elementwise) + adjoint transform"}
    return out


def extra_batched(torch, aclfft, oracle, np, steps, stream):
    """512 C2C transforms, lengths log-uniform over the set, batches U{1..64}; one plan (persistent
    kernel, static schedule; launched directly, and replayed from the plan's hipGraph) vs one plan
    (and launch) per group."""
    lengths = (60, 128, 243, 1000, 1024, 2048, 4096, 6561)
    rng = np.random.default_rng(4)
    logs = np.log(np.array(lengths, float))
    u = rng.uniform(logs.min(), logs.max(), 512)
    ln = [int(lengths[int(np.argmin(np.abs(logs - v)))]) for v in u]
    bt = [int(b) for b in rng.integers(1, 65, 512)]
    total = sum(a * b for a, b in zip(ln, bt))
    x = (torch.rand(total, dtype=torch.complex64, device="cuda") * 2 - (1 + 1j)).to(torch.complex64)
    y = torch.empty_like(x)
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    p.set_stream(stream)
    ms = time_steps(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), steps, 3, stream)
    pg = aclfft.Plan.plan_many_uneven(ln, bt)
    pg.set_stream(stream)
    pg.set_graph_mode(True)
    ms_graph = time_steps(torch, lambda: pg.exec_c2c(x, y, aclfft.FORWARD), steps, 3, stream)
    pg.destroy()
    # correctness on a few groups (output of the directly launched plan)
    p.exec_c2c(x, y, aclfft.FORWARD)
    torch.cuda.synchronize()
    off = 0
    worst = 0.0
    ycpu = y.cpu().numpy()
    xcpu = x.cpu().numpy()
    for gi, (n, b) in enumerate(zip(ln, bt)):
        if gi % 64 == 0:
            seg = xcpu[off:off + n * b].reshape(b, n).astype(np.complex128)
            worst = max(worst, oracle.rel_l2(ycpu[off:off + n * b].reshape(b, n), oracle.c2c(seg)))
        off += n * b
    plans = [aclfft.Plan.plan1d(n, aclfft.C2C, b) for n, b in zip(ln, bt)]
    offs = np.cumsum([0] + [a * b for a, b in zip(ln, bt)])

    def per_group():
        for i, q in enumerate(plans):
            q.exec_c2c(x[offs[i]:offs[i + 1]], y[offs[i]:offs[i + 1]], aclfft.FORWARD)

    for q in plans:
        q.set_stream(stream)
    ms_naive = time_steps(torch, per_group, max(3, steps // 5), 2, stream)
    flops = sum(5.0 * n * math.log2(n) * b for n, b in zip(ln, bt))
    out = {"groups": 512, "points": int(total), "persistent_ms": round(ms, 5), "persistent_hipgraph_ms": round(ms_graph, 5),
           "per_group_launches_ms": round(ms_naive, 5), "speedup": round(ms_naive / ms, 2),
           "GFLOP/s": round(flops / ms / 1e6, 1), "hbm_frac": round(16.0 * total / ms / 1e6 / HBM_PEAK_GBS, 4),
           "worst_rel_l2_vs_fp64": worst, "plan": p.describe()}
    for q in plans:
        q.destroy()
    p.destroy()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--no-extras", action="store_true", help="only the headline C2C configuration")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16", "fp16x2", "bf16x3"],
                    help="arithmetic of the timed loop (profiling runs; the headline value is fp32)")
    ap.add_argument("--with-copy", action="store_true",
                    help="also run one device copy of the 1 GiB input (counter calibration for profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the FFT path has no multi-GPU mode (a batch would be "
              "sharded across GPUs as contiguous ranges of transforms, with no communication)", file=sys.stderr)
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
    if args.precision != "fp32":
        plan.set_precision({"bf16": aclfft.PRECISION_BF16, "fp16x2": aclfft.PRECISION_FP16X2,
                            "bf16x3": aclfft.PRECISION_BF16X3}[args.precision])
    plan.set_stream(stream)
    if args.with_copy:
        ycopy = torch.empty_like(x)
        ycopy.copy_(x)
        torch.cuda.synchronize()
        del ycopy
    # parity on a subset of the output against the FP64 oracle (before the timed region, so that
    # the CPU work does not leave the GPU idle between warmup and timing)
    plan.exec_c2c(x, y, aclfft.FORWARD)
    torch.cuda.synchronize()
    sub = 256
    ref = oracle.c2c(x_cpu[:sub].numpy().astype(np.complex128))
    err = oracle.rel_l2(y[:sub].cpu().numpy(), ref)
    log("bench: rel_l2 vs FP64 oracle on %d transforms = %.3e" % (sub, err))

    log("bench: plan ready, warmup %d" % args.warmup)
    for _ in range(args.warmup):
        plan.exec_c2c(x, y, aclfft.FORWARD)
    torch.cuda.synchronize()

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
                   "n": N, "batch": BATCH, "path": ("FP32 Stockham" if args.precision == "fp32" else
                                                     "MFMA " + args.precision) + " (aclfftExecC2C, device pointers)",
                   "kernel": plan.describe()},
        "rel_l2_vs_fp64": err,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
    }
    if not args.no_extras:
        extras = {}
        for name, fn in (("copy_1GiB", lambda: extra_copy(torch, aclfft, x, 20, stream, ms)),
                         ("mfma_split_n1024_b65536", lambda: extra_mfma(torch, aclfft, oracle, np, x, x_cpu, y,
                                                                        args.steps, stream, err)),
                         ("real_r2c_c2r", lambda: extra_real(torch, aclfft, oracle, np, args.steps, stream)),
                         ("c2c_n1024_b64_latency", lambda: extra_small_latency(torch, aclfft, oracle, np, stream)),
                         ("double_precision", lambda: extra_double(torch, aclfft, oracle, np, args.steps, stream)),
                         ("c2c_2d_2048x2048_b4", lambda: extra_2d(torch, aclfft, oracle, np, 10, stream)),
                         ("uneven_batched_512", lambda: extra_batched(torch, aclfft, oracle, np, 20, stream)),
                         ("loss_grad", lambda: extra_loss_grad(torch, 10, stream))):
            try:
                log("bench: extra %s" % name)
                extras[name] = fn()
            except Exception as ex:  # an extra must not hide the headline line
                extras[name] = {"error": "%s: %s" % (type(ex).__name__, ex)}
        out["extras"] = extras
    if not args.no_cpu:
        sample = x_cpu[:512].numpy().astype(np.complex128)
        out["cpu_baseline"] = cpu_baseline(sample)
        out["cpu_golden_path"] = cpu_golden(x_cpu, torch)
    plan.destroy()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
