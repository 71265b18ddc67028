#!/usr/bin/env python3
"""Benchmark of the aclfft FP32 Stockham C2C path on one MI355X.

Workload (the headline metric's configuration): 1D C2C forward FFT, complex64, N = 1024,
batch = 65536, contiguous, re/im ~ N(0, 1/2) from seed 0 (synthetic), input resident in HBM.
A step is one aclfftExecC2C over the whole batch (one kernel launch).

Procedure of the headline: plan, then untimed calls for at least 0.25 s of device time (the clocks
settle over tens of milliseconds of sustained load), then exactly K timed steps between HIP events
recorded on the plan's stream, then the parity check against the FP64 oracle (after the timed
region, so that host work never leaves the GPU idle just before it).

Prints ONE JSON line:
  value        = C2C GFLOP/s = 5 N log2 N batch / t
  roofline     = algorithmic HBM bytes (16 B per point) / average kernel time vs 8 TB/s
  cpu_baseline = the reference's CPU golden path (torch.fft.fft on the host cores, as its
                 gen_data.py computes the goldens) on the same workload, warm median of >= 10
  extras       = the other workloads (real transforms, MFMA split, uneven batch, 2D, adjoints,
                 double precision, latency), each with its parity figure, and the CPU golden path
                 timed on every workload

--profile-workloads [secondary|reflen|mfma|generic] runs a set of workloads back to back for a profiler
instead (each plan is called until 0.1 s of device time has passed, then --reps more times) and
prints no JSON line: the secondary workloads, the reference's own test-case lengths at batch
>= 1024, or the MFMA split-precision kernels in every mode.
"""
import argparse
import ctypes
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ops-fft_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N = 1024
BATCH = 65536
UNEVEN_LENGTHS = (60, 128, 243, 1000, 1024, 2048, 4096, 6561)


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


def uneven_groups(count=512, seed=4):
    """The uneven-batch workload: lengths log-uniform over the set, batches U{1..64}."""
    import numpy as np
    rng = np.random.default_rng(seed)
    logs = np.log(np.array(UNEVEN_LENGTHS, float))
    u = rng.uniform(logs.min(), logs.max(), count)
    ln = [int(UNEVEN_LENGTHS[int(np.argmin(np.abs(logs - v)))]) for v in u]
    bt = [int(b) for b in rng.integers(1, 65, count)]
    return ln, bt


# ---------------------------------------------------------------------------------------------
# CPU baselines
# ---------------------------------------------------------------------------------------------
def host_threads():
    n = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return n


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_time(fn, reps=10, min_reps_seconds=0.0):
    """Warm median of `reps` calls (one untimed call first), wall clock."""
    fn()
    ts = []
    t_all = time.perf_counter()
    while len(ts) < reps or time.perf_counter() - t_all < min_reps_seconds:
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    return ts[len(ts) // 2], len(ts)


def cpu_golden_headline(torch, x_cpu):
    """The reference's CPU golden path, torch.fft.fft (src/fft1_d/tests/fft1_d_data/gen_data.py:71-86),
    on the whole headline workload with every host core."""
    threads = host_threads()
    torch.set_num_threads(threads)
    t, reps = cpu_time(lambda: torch.fft.fft(x_cpu, dim=1), reps=10)
    n, b = x_cpu.shape[1], x_cpu.shape[0]
    return dict(value=round(5.0 * n * math.log2(n) * b / t / 1e9, 3), unit="GFLOP/s", cores=threads,
                kind="reference",
                sample="the reference's golden path torch.fft.fft (gen_data.py) on the full workload "
                       "(65536 x 1024 complex64), warm median of %d reps, %.4f s per call" % (reps, t),
                cpu_model=cpu_model(), torch=torch.__version__, mkl=bool(torch.backends.mkl.is_available()),
                ms=round(t * 1e3, 3))


def cpu_oracle_port(x_sample_np, budget_s=5.0):
    """The FP64 oracle (a single-thread C restatement of the reference's Stockham) on a bounded
    sample: transforms per second, scaled to GFLOP/s."""
    from oracle import oracle
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


def cpu_golden_configs(torch, np):
    """torch.fft on the host cores (the reference's golden path: fft1_d, rfft1_d, irfft1_d and
    fft2_d gen_data.py) for every other workload, warm median of >= 10 reps each."""
    torch.set_num_threads(host_threads())
    out = {"cores": host_threads(), "cpu_model": cpu_model(), "torch": torch.__version__,
           "mkl": bool(torch.backends.mkl.is_available())}
    g = torch.Generator().manual_seed(0)

    def rec(name, fn, flops, nbytes):
        t, reps = cpu_time(fn, reps=10)
        out[name] = {"ms": round(t * 1e3, 4), "GFLOP/s": round(flops / t / 1e9, 2),
                     "GB/s": round(nbytes / t / 1e9, 2), "reps": reps}

    x = torch.randn(64, 1024, dtype=torch.complex64, generator=g)
    rec("c2c_n1024_b64", lambda: torch.fft.fft(x, dim=1), 5.0 * 1024 * 10 * 64, 16.0 * 1024 * 64)
    for n in (4096, 3000):
        xr = torch.randn(8192, n, generator=g)
        X = torch.fft.rfft(xr, dim=1)
        byts = (4.0 * n + 8.0 * (n // 2 + 1)) * 8192
        fl = 2.5 * n * math.log2(n) * 8192
        rec("r2c_n%d_b8192" % n, lambda: torch.fft.rfft(xr, dim=1), fl, byts)
        rec("c2r_n%d_b8192" % n, lambda: torch.fft.irfft(X, n=n, dim=1) * n, fl, byts)
    ln, bt = uneven_groups()
    segs = [torch.randn(b, n, dtype=torch.complex64, generator=g) for n, b in zip(ln, bt)]
    fl = sum(5.0 * n * math.log2(n) * b for n, b in zip(ln, bt))
    pts = sum(n * b for n, b in zip(ln, bt))
    rec("uneven_512_groups", lambda: [torch.fft.fft(s, dim=1) for s in segs], fl, 16.0 * pts)
    del segs
    x2 = torch.randn(4, 2048, 2048, dtype=torch.complex64, generator=g)
    pts = 4 * 2048 * 2048
    rec("fft2_2048x2048_b4", lambda: torch.fft.fft2(x2), 5.0 * pts * math.log2(2048 * 2048), 16.0 * pts)
    del x2
    xr = torch.randn(1024, 8192, generator=g)
    rec("r2c_n8192_b1024", lambda: torch.fft.rfft(xr, dim=1), 2.5 * 8192 * 13 * 1024,
        (4.0 * 8192 + 8.0 * 4097) * 1024)
    return out


# ---------------------------------------------------------------------------------------------
# GPU timing helpers
# ---------------------------------------------------------------------------------------------
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


def time_steps(torch, fn, steps, warmup, stream, min_seconds=0.1):
    warm(torch, fn, warmup, min_seconds)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def time_cold(torch, fn, stream, reps=10):
    """Median device time of single calls, each issued right after reading 1 GiB (four times the
    256 MB last-level cache) on the same stream, so the call finds none of its data cached.  The
    flush reads rather than writes: a written flush would leave the caches full of dirty lines
    whose write-back the timed call would pay for.  The flush is outside the timed events."""
    buf = torch.zeros(1 << 28, dtype=torch.float32, device="cuda")
    times = []
    for _ in range(reps + 1):
        buf.sum()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    del buf
    times = sorted(times[1:])
    return times[len(times) // 2]


def time_single(torch, fn, stream, reps=10):
    """Median device time of single calls, each after the previous one finished (no flush): the
    cold figure's counterpart with a warm cache, including the same per-call launch cost."""
    times = []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
    times = sorted(times[1:])
    return times[len(times) // 2]


def frac(nbytes, ms):
    return round(nbytes / ms / 1e6 / HBM_PEAK_GBS, 4)


# ---------------------------------------------------------------------------------------------
# extras
# ---------------------------------------------------------------------------------------------
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
    out["torch_copy"] = {"ms": round(ms, 5), "GB/s": round(b / ms / 1e6, 1), "frac_of_peak": frac(b, ms)}
    h = ctypes.c_void_p(stream.cuda_stream)
    nb = x.numel() * x.element_size()
    ms = time_steps(torch, lambda: stream_copy(x.data_ptr(), y.data_ptr(), nb, h), steps, 3, stream)
    out["stream_copy"] = {"ms": round(ms, 5), "GB/s": round(b / ms / 1e6, 1), "frac_of_peak": frac(b, ms)}
    out["headline_vs_stream_copy"] = round(ms / headline_ms, 4)
    if not torch.equal(x, y):
        out["error"] = "stream copy mismatch"
    return out


def sixty_db_signal(np, batch, seed, n=N):
    """The precision workload's 60 dB case: per transform, 8 complex tones at random bins
    (amplitudes 1 x4 and 1e-3 x4, random phases) plus complex Gaussian noise of deviation 1e-6."""
    rng = np.random.default_rng(seed)
    t = np.arange(n)
    x = np.empty((batch, n), np.complex128)
    amp = np.array([1, 1, 1, 1, 1e-3, 1e-3, 1e-3, 1e-3])
    for b in range(batch):
        f = rng.choice(n, 8, replace=False)
        ph = rng.uniform(0, 2 * np.pi, 8)
        x[b] = (amp * np.exp(1j * (2 * np.pi * np.outer(t, f) / n + ph))).sum(1)
    x += 1e-6 * (rng.standard_normal((batch, n)) + 1j * rng.standard_normal((batch, n)))
    return x.astype(np.complex64)


def errs(oracle, np, got, ref):
    """rel-L2 and max-abs error (relative to the reference's largest magnitude) vs FP64."""
    got = np.asarray(got, np.complex128)
    return oracle.rel_l2(got, ref), float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


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
        return errs(oracle, np, o.cpu().numpy(), ref60)

    out["sixty_db_fp32_vs_fp64"] = dict(zip(("rel_l2", "max_abs_rel"), err60(0)))
    ref = oracle.c2c(x_cpu[:256].numpy().astype(np.complex128))
    for mode, name in ((aclfft.PRECISION_BF16, "bf16"), (aclfft.PRECISION_FP16X2, "fp16x2"),
                       (aclfft.PRECISION_BF16X3, "bf16x3")):
        p = aclfft.Plan.plan1d(N, aclfft.C2C, BATCH)
        p.set_precision(mode)
        p.set_stream(stream)
        ms = time_steps(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), steps, 3, stream)
        el2, emax = errs(oracle, np, y[:256].cpu().numpy(), ref)
        e60 = err60(mode)
        out[name] = {"ms": round(ms, 5), "GFLOP/s": round(5.0 * N * math.log2(N) * BATCH / ms / 1e6, 1),
                     "hbm_frac": frac(bytes_, ms), "rel_l2_vs_fp64": el2, "max_abs_rel_vs_fp64": emax,
                     "err_ratio_vs_fp32": round(el2 / e32, 3), "sixty_db_rel_l2_vs_fp64": e60[0],
                     "sixty_db_max_abs_rel_vs_fp64": e60[1], "plan": p.describe()}
        p.destroy()
    return out


def extra_mfma_lengths(torch, aclfft, oracle, np, x, y, steps, stream):
    """The MFMA split path at the other lengths of its registry (16-, 32- and 64-point dense DFT
    stages; n = 16, 32, 64 as one dense stage), each at 64 Mi points (1 GiB moved) against the FP32 Stockham kernel: time, HBM
    fraction and the error of 64 transforms against the FP64 oracle, per operand precision; then
    2D images of 32 and 64 points per axis (the stages as F X F) against the one-pass FP32 2D
    kernel, the same way."""
    out = {}
    pts = x.numel()
    for n in (16, 32, 64, 256, 512, 2048, 4096):
        b = pts // n
        xv, yv = x.view(-1)[:b * n].view(b, n), y.view(-1)[:b * n].view(b, n)
        ref = oracle.c2c(xv[:64].cpu().numpy().astype(np.complex128))
        rec = {}
        for mode, name in ((0, "fp32"), (aclfft.PRECISION_BF16, "bf16"), (aclfft.PRECISION_FP16X2, "fp16x2"),
                           (aclfft.PRECISION_BF16X3, "bf16x3")):
            p = aclfft.Plan.plan1d(n, aclfft.C2C, b)
            if mode:
                p.set_precision(mode)
            p.set_stream(stream)
            ms = time_steps(torch, lambda: p.exec_c2c(xv, yv, aclfft.FORWARD), steps, 3, stream)
            el2, emax = errs(oracle, np, yv[:64].cpu().numpy(), ref)
            rec[name] = {"ms": round(ms, 5), "hbm_frac": frac(16.0 * pts, ms), "rel_l2_vs_fp64": el2,
                         "max_abs_rel_vs_fp64": emax, "plan": p.describe()}
            p.destroy()
        for name in ("bf16", "fp16x2", "bf16x3"):
            rec[name]["vs_fp32"] = round(rec["fp32"]["ms"] / rec[name]["ms"], 3)
        out["n%d_b%d" % (n, b)] = rec
    # 2D images (the reference's 2D sizes with 32- and 64-point axes) as F X F on the same stages
    for nx, ny in ((32, 32), (32, 64), (64, 64)):
        b = pts // (nx * ny)
        xv, yv = x.view(-1)[:b * nx * ny].view(b, nx, ny), y.view(-1)[:b * nx * ny].view(b, nx, ny)
        ref = oracle.c2c_2d(xv[:64].cpu().numpy().astype(np.complex128))
        rec = {}
        for mode, name in ((0, "fp32"), (aclfft.PRECISION_BF16, "bf16"), (aclfft.PRECISION_FP16X2, "fp16x2"),
                           (aclfft.PRECISION_BF16X3, "bf16x3")):
            p = aclfft.Plan.plan2d(b, nx, ny, aclfft.C2C)
            if mode:
                p.set_precision(mode)
            p.set_stream(stream)
            ms = time_steps(torch, lambda: p.exec_c2c(xv, yv, aclfft.FORWARD), steps, 3, stream)
            el2, emax = errs(oracle, np, yv[:64].cpu().numpy(), ref)
            rec[name] = {"ms": round(ms, 5), "hbm_frac": frac(16.0 * pts, ms), "rel_l2_vs_fp64": el2,
                         "max_abs_rel_vs_fp64": emax, "plan": p.describe()}
            p.destroy()
        for name in ("bf16", "fp16x2", "bf16x3"):
            rec[name]["vs_fp32"] = round(rec["fp32"]["ms"] / rec[name]["ms"], 3)
        out["2d_%dx%d_b%d" % (nx, ny, b)] = rec
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
        cold_r = time_cold(torch, lambda: pr.exec_r2c(xr, X), stream)
        cold_c = time_cold(torch, lambda: pc.exec_c2r(X, xb), stream)
        single_r = time_single(torch, lambda: pr.exec_r2c(xr, X), stream)
        byts = (4.0 * n + 8.0 * (n // 2 + 1)) * b
        rt = float(((xb / n - xr).norm() / xr.norm()).item())
        err = oracle.rel_l2(X[:64].cpu().numpy(), oracle.r2c(xr[:64].cpu().numpy().astype(np.float64)))
        rec_mfma = {}
        if n == 8192:
            # the MFMA precisions: the 4096-point MFMA transform with the Hermitian unpack / pre-pack
            for mode, name in ((aclfft.PRECISION_BF16, "bf16"), (aclfft.PRECISION_FP16X2, "fp16x2"),
                               (aclfft.PRECISION_BF16X3, "bf16x3")):
                qr = aclfft.Plan.plan1d(n, aclfft.R2C, b)
                qc = aclfft.Plan.plan1d(n, aclfft.C2R, b)
                qr.set_precision(mode)
                qc.set_precision(mode)
                qr.set_stream(stream)
                qc.set_stream(stream)
                Xm = torch.empty_like(X)
                mr = time_steps(torch, lambda: qr.exec_r2c(xr, Xm), steps, 3, stream)
                mc = time_steps(torch, lambda: qc.exec_c2r(Xm, xb), steps, 3, stream)
                em = oracle.rel_l2(Xm[:64].cpu().numpy(), oracle.r2c(xr[:64].cpu().numpy().astype(np.float64)))
                rec_mfma[name] = {"r2c_ms": round(mr, 5), "c2r_ms": round(mc, 5), "r2c_hbm_frac": frac(byts, mr),
                                  "c2r_hbm_frac": frac(byts, mc), "r2c_vs_fp32": round(ms_r / mr, 3),
                                  "r2c_rel_l2_vs_fp64": em, "plan_r2c": qr.describe()}
                qr.destroy()
                qc.destroy()
        out["n%d_b%d" % (n, b)] = {
            "r2c_ms": round(ms_r, 5), "c2r_ms": round(ms_c, 5),
            "r2c_hbm_frac": frac(byts, ms_r), "c2r_hbm_frac": frac(byts, ms_c),
            "r2c_cold_llc_ms": round(cold_r, 5), "c2r_cold_llc_ms": round(cold_c, 5),
            "r2c_single_call_warm_ms": round(single_r, 5),
            "r2c_cold_llc_hbm_frac": frac(byts, cold_r), "c2r_cold_llc_hbm_frac": frac(byts, cold_c),
            "r2c_GFLOP/s_2.5nlogn": round(2.5 * n * math.log2(n) * b / ms_r / 1e6, 1),
            "roundtrip_rel_l2": rt, "r2c_rel_l2_vs_fp64": err, "plan_r2c": pr.describe(),
            "plan_c2r": pc.describe()}
        if rec_mfma:
            out["n%d_b%d" % (n, b)]["mfma"] = rec_mfma
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
    out["z2z_n1024_b32768"] = {"ms": round(ms, 5), "hbm_frac": frac(32.0 * n * b, ms),
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
                                  "d2z_hbm_frac": frac(byts, ms_r), "z2d_hbm_frac": frac(byts, ms_c),
                                  "d2z_rel_l2_vs_fp64_oracle": err, "roundtrip_rel_l2": rt,
                                  "plan_d2z": pr.describe()}
    pr.destroy()
    pc.destroy()
    return out


def extra_small_latency(torch, aclfft, oracle, np, stream):
    """The launch-bound configuration (C2C N=1024, batch 64, 512 KiB each way): time per call,
    eager and replayed from the plan's hipGraph, measured over many back-to-back calls (device time
    per call with events, host time per call with the wall clock)."""
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
    # the runtime's own cost of replaying a one-kernel graph: the same call captured alone in a torch
    # graph, replayed back to back (what the plan's per-call hipGraph pays on top of the kernel)
    g1 = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g1, stream=side, capture_error_mode="relaxed"):
        p.exec_c2c(x, y, aclfft.FORWARD)
    for _ in range(20):
        g1.replay()
    torch.cuda.synchronize()
    reps1 = 2000
    e0.record(cur)
    for _ in range(reps1):
        g1.replay()
    e1.record(cur)
    torch.cuda.synchronize()
    out["one_call_graph_replay"] = {"device_us_per_replay": round(e0.elapsed_time(e1) * 1e3 / reps1, 2)}
    del g1
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
    cold = time_cold(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), stream)
    single = time_single(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), stream)
    err = oracle.rel_l2(y[:1].cpu().numpy(), oracle.c2c_2d(x[:1].cpu().numpy().astype(np.complex128)))
    pts = B * NX * NY
    out = {"ms": round(ms, 5), "cold_llc_ms": round(cold, 5), "cold_llc_hbm_frac_two_pass": frac(32.0 * pts, cold),
           "single_call_warm_ms": round(single, 5),
           "GFLOP/s": round(5.0 * pts * math.log2(NX * NY) / ms / 1e6, 1),
           "hbm_frac_compulsory": frac(16.0 * pts, ms), "hbm_frac_two_pass": frac(32.0 * pts, ms),
           "rel_l2_vs_fp64": err, "plan": p.describe()}
    p.destroy()
    return out


def extra_2d_small(torch, aclfft, oracle, np, x, y, steps, stream):
    """The reference's 2D sizes ({32, 64, 128}^2, its DD domain) over 2^26 points (the headline's
    1 GiB buffers): the one-pass kernel (image in LDS, 16 B/point) against the row + column kernel
    pair (ACLFFT_2D_FUSED=0, 32 B/point), with parity of the first 16 images against the FP64
    oracle."""
    out = {}
    pts = x.numel()
    for n in (32, 64, 128):
        b = pts // (n * n)
        rec = {}
        for fused in (True, False):
            old = os.environ.get("ACLFFT_2D_FUSED")
            if not fused:
                os.environ["ACLFFT_2D_FUSED"] = "0"
            try:
                p = aclfft.Plan.plan2d(b, n, n, aclfft.C2C)
            finally:
                if old is None:
                    os.environ.pop("ACLFFT_2D_FUSED", None)
                else:
                    os.environ["ACLFFT_2D_FUSED"] = old
            p.set_stream(stream)
            ms = time_steps(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), steps, 3, stream)
            k = 16 * n * n
            err = oracle.rel_l2(y[:k].cpu().numpy().reshape(16, n, n),
                                oracle.c2c_2d(x[:k].cpu().numpy().reshape(16, n, n).astype(np.complex128)))
            rec["one_pass" if fused else "two_pass"] = {
                "ms": round(ms, 5), "GFLOP/s": round(5.0 * pts * math.log2(n * n) / ms / 1e6, 1),
                "hbm_frac_compulsory": frac(16.0 * pts, ms), "rel_l2_vs_fp64": err, "plan": p.describe()}
            p.destroy()
        rec["speedup"] = round(rec["two_pass"]["ms"] / rec["one_pass"]["ms"], 3)
        out["%dx%d_b%d" % (n, n, b)] = rec
    return out


def extra_loss_grad(torch, aclfft, steps, stream):
    """The differentiable configuration: loss = sum |W * FFT(x)|^2 (W ~ U[0,1), fixed) and its
    gradient through torch.autograd over the library (forward transform, then the adjoint kernel),
    for 2D C2C 2048x2048 batch 4 and 1D R2C n=8192 batch 1024.  One step = forward + backward.
    The adjoint launches are also timed alone: the 2D adjoint is the backward 2D transform, the
    R2C adjoint is aclfftExtExecR2CAdjoint (C2R kernels with the Hermitian weights)."""
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
        rec = {"fwd_bwd_ms": round(ms, 5), "transform_ms": round(fft_ms, 5),
               "note": "step = transform + loss (torch elementwise) + adjoint transform"}
        if name.startswith("fft2"):
            p = aclfft.Plan.plan2d(4, 2048, 2048, aclfft.C2C)
            p.set_stream(stream)
            gi = torch.randn(shape, dtype=torch.complex64, device="cuda", generator=g)
            go = torch.empty_like(gi)
            ams = time_steps(torch, lambda: p.exec_c2c(gi, go, aclfft.BACKWARD), steps, 3, stream)
            pts = 4 * 2048 * 2048
            rec.update({"adjoint_ms": round(ams, 5), "adjoint_hbm_frac_two_pass": frac(32.0 * pts, ams),
                        "adjoint_plan": p.describe()})
        else:
            p = aclfft.Plan.plan1d(8192, aclfft.R2C, 1024)
            p.set_stream(stream)
            gi = torch.randn(1024, 4097, dtype=torch.complex64, device="cuda", generator=g)
            go = torch.empty(1024, 8192, device="cuda")
            ams = time_steps(torch, lambda: p.exec_r2c_adjoint(gi, go), steps, 3, stream)
            byts = (4.0 * 8192 + 8.0 * 4097) * 1024
            rec.update({"adjoint_ms": round(ams, 5), "adjoint_hbm_frac": frac(byts, ams)})
        p.destroy()
        out[name] = rec
    return out


def extra_batched(torch, aclfft, oracle, np, steps, stream):
    """512 C2C transforms, lengths log-uniform over the set, batches U{1..64}; one plan (persistent
    kernel, static schedule; launched directly, and replayed from the plan's hipGraph) vs one plan
    (and launch) per group.  Every group's output of both the direct and the replayed plan is
    checked against the FP64 oracle."""
    ln, bt = uneven_groups()
    total = sum(a * b for a, b in zip(ln, bt))
    x = (torch.rand(total, dtype=torch.complex64, device="cuda") * 2 - (1 + 1j)).to(torch.complex64)
    y = torch.empty_like(x)
    yg = torch.empty_like(x)
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    p.set_stream(stream)
    ms = time_steps(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), steps, 3, stream)
    ms_cold = time_cold(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), stream)
    ms_single = time_single(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), stream)
    pg = aclfft.Plan.plan_many_uneven(ln, bt)
    pg.set_stream(stream)
    pg.set_graph_mode(True)
    ms_graph = time_steps(torch, lambda: pg.exec_c2c(x, yg, aclfft.FORWARD), steps, 3, stream)
    pg.destroy()
    p.exec_c2c(x, y, aclfft.FORWARD)
    torch.cuda.synchronize()
    off = 0
    worst = 0.0
    ycpu = y.cpu().numpy()
    ygcpu = yg.cpu().numpy()
    xcpu = x.cpu().numpy()
    for n, b in zip(ln, bt):
        seg = xcpu[off:off + n * b].reshape(b, n).astype(np.complex128)
        ref = oracle.c2c(seg)
        worst = max(worst, oracle.rel_l2(ycpu[off:off + n * b].reshape(b, n), ref),
                    oracle.rel_l2(ygcpu[off:off + n * b].reshape(b, n), ref))
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
    out = {"groups": 512, "points": int(total), "persistent_ms": round(ms, 5),
           "persistent_cold_llc_ms": round(ms_cold, 5), "cold_llc_hbm_frac": frac(16.0 * total, ms_cold),
           "persistent_single_call_warm_ms": round(ms_single, 5),
           "persistent_hipgraph_ms": round(ms_graph, 5),
           "per_group_launches_ms": round(ms_naive, 5), "speedup": round(ms_naive / ms, 2),
           "GFLOP/s": round(flops / ms / 1e6, 1), "hbm_frac": frac(16.0 * total, ms),
           "worst_rel_l2_vs_fp64_all_groups": worst, "plan": p.describe()}
    for q in plans:
        q.destroy()
    p.destroy()

    def worst_err(yy):
        yc = yy.cpu().numpy()
        o, w = 0, 0.0
        for n, b in zip(ln, bt):
            ref = oracle.c2c(xcpu[o:o + n * b].reshape(b, n).astype(np.complex128))
            w = max(w, oracle.rel_l2(yc[o:o + n * b].reshape(b, n), ref))
            o += n * b
        return w

    # the tuned plan (aclfftExtTune: form, schedule, workgroups per CU, class set, step kernels)
    # and the MFMA precisions (MFMA tile launches for 1024 / 2048 / 4096, the FP32 persistent
    # kernel for the other lengths; "auto" = the tuner's choice, MFMA kept only within 1e-6 of the
    # FP32 result); the tuning cache goes to a private temporary file
    import tempfile
    old_cache = os.environ.get("ACLFFT_TUNE_CACHE")
    with tempfile.TemporaryDirectory() as td:
        os.environ["ACLFFT_TUNE_CACHE"] = os.path.join(td, "tune.txt")
        try:
            for name, mode, tune in (("tuned", 0, True), ("mfma_fp16x2", 2, False), ("mfma_bf16x3", 3, False),
                                     ("mfma_bf16", 1, False), ("auto_tuned", 4, True)):
                q = aclfft.Plan.plan_many_uneven(ln, bt)
                q.set_stream(stream)
                if mode:
                    q.set_precision(mode)
                if tune:
                    q.tune()
                mq = time_steps(torch, lambda: q.exec_c2c(x, y, aclfft.FORWARD), steps, 3, stream)
                q.exec_c2c(x, y, aclfft.FORWARD)
                torch.cuda.synchronize()
                out[name] = {"ms": round(mq, 5), "hbm_frac": frac(16.0 * total, mq),
                             "vs_fp32_default": round(ms / mq, 3), "worst_rel_l2_vs_fp64_all_groups": worst_err(y),
                             "plan": q.describe()}
                q.destroy()
        finally:
            if old_cache is None:
                os.environ.pop("ACLFFT_TUNE_CACHE", None)
            else:
                os.environ["ACLFFT_TUNE_CACHE"] = old_cache
    return out


def extra_vendor_fft(torch, aclfft, stream):
    """The vendor FFT on the same GPU, as a yardstick beside this library: torch.fft on device
    tensors runs hipFFT / rocFFT (ROCm 7.2).  Each benchmark workload is timed both ways on the same
    device-resident input, with the same warm-up and event timing; the uneven batch has no rocFFT
    form, so it is one torch.fft call per group there."""
    dev = torch.device("cuda:0")
    g = torch.Generator().manual_seed(11)
    out = {"library": "torch.fft on cuda:0 (hipFFT / rocFFT), torch %s" % torch.__version__}

    def both(name, nbytes, lib_fn, vendor_fn, steps):
        ml = time_steps(torch, lib_fn, steps, 3, stream)
        mv = time_steps(torch, vendor_fn, steps, 3, stream)
        out[name] = {"this_ms": round(ml, 5), "rocfft_ms": round(mv, 5), "this_hbm_frac": frac(nbytes, ml),
                     "rocfft_hbm_frac": frac(nbytes, mv), "speedup_vs_rocfft": round(mv / ml, 2)}

    for n, b, steps in ((1024, 65536, 20), (1024, 64, 100)):
        x = torch.randn(b, n, dtype=torch.complex64, generator=g).to(dev)
        y = torch.empty_like(x)
        p = aclfft.Plan.plan1d(n, aclfft.C2C, b)
        p.set_stream(stream)
        both("c2c_n%d_b%d" % (n, b), 16.0 * n * b, lambda: p.exec_c2c(x, y, aclfft.FORWARD),
             lambda: torch.fft.fft(x), steps)
        p.destroy()
        del x, y
    for n, b in ((4096, 8192), (3000, 8192), (8192, 1024)):
        x = torch.randn(b, n, generator=g).to(dev)
        xf = torch.empty(b, n // 2 + 1, dtype=torch.complex64, device=dev)
        z = torch.empty(b, n, device=dev)
        pr = aclfft.Plan.plan1d(n, aclfft.R2C, b)
        pc = aclfft.Plan.plan1d(n, aclfft.C2R, b)
        pr.set_stream(stream)
        pc.set_stream(stream)
        nb = b * (4.0 * n + 8.0 * (n // 2 + 1))
        both("r2c_n%d_b%d" % (n, b), nb, lambda: pr.exec_r2c(x, xf), lambda: torch.fft.rfft(x), 50)
        both("c2r_n%d_b%d" % (n, b), nb, lambda: pc.exec_c2r(xf, z), lambda: torch.fft.irfft(xf, n=n), 50)
        pr.destroy()
        pc.destroy()
        del x, xf, z
    x = torch.randn(4, 2048, 2048, dtype=torch.complex64, generator=g).to(dev)
    y = torch.empty_like(x)
    p = aclfft.Plan.plan2d(4, 2048, 2048, aclfft.C2C)
    p.set_stream(stream)
    both("fft2_2048x2048_b4", 16.0 * x.numel(), lambda: p.exec_c2c(x, y, aclfft.FORWARD),
         lambda: torch.fft.fft2(x), 20)
    p.destroy()
    del x, y
    ln, bt = uneven_groups()
    total = sum(a * b for a, b in zip(ln, bt))
    x = torch.randn(total, dtype=torch.complex64, generator=g).to(dev)
    y = torch.empty_like(x)
    offs = [0]
    for a, b in zip(ln, bt):
        offs.append(offs[-1] + a * b)
    views = [x[offs[i]:offs[i + 1]].view(bt[i], ln[i]) for i in range(len(ln))]
    p = aclfft.Plan.plan_many_uneven(ln, bt)
    p.set_stream(stream)

    def vendor_groups():
        for v in views:
            torch.fft.fft(v)

    both("uneven_512_groups", 16.0 * total, lambda: p.exec_c2c(x, y, aclfft.FORWARD), vendor_groups, 5)
    p.destroy()
    return out


# ---------------------------------------------------------------------------------------------
# secondary workloads for a profiler (--profile-workloads)
# ---------------------------------------------------------------------------------------------
def _profile_rep(torch, fn, reps):
    """Call fn until 0.1 s of device time has passed (tables built, clocks settled), then `reps`
    more times: a trace's last `reps` dispatches of a kernel are the measured ones."""
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.1:
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()


# the reference's own test-case lengths (its fft1_d / rfft1_d / irfft1_d case lists) that the
# compiled kernels cover at batch >= 1024: C2C lengths, and even real lengths (R2C then C2R)
REFLEN_C2C = (6, 105, 210, 420, 1536, 2100, 2310, 5040)
REFLEN_REAL = (1536, 2100, 5040, 25200)


def profile_reflen(torch, aclfft, reps):
    """C2C at each of REFLEN_C2C and R2C + C2R at each of REFLEN_REAL, 2^24 points per call (batch
    at least 1024), 2D C2C of the reference's sizes 32^2, 64^2, 128^2 over 2^24 points (one-pass
    kernel), then the streaming copy of 512 MiB (counter calibration)."""
    s = torch.cuda.current_stream()
    pts = 1 << 24
    x = torch.randn(2 * pts, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    for n in REFLEN_C2C:
        b = max(1024, pts // n)
        p = aclfft.Plan.plan1d(n, aclfft.C2C, b)
        p.set_stream(s)
        _profile_rep(torch, lambda: p.exec_c2c(x[:n * b], y[:n * b], aclfft.FORWARD), reps)
        log("c2c n=%d b=%d: %s" % (n, b, p.describe()))
        p.destroy()
    for n in (32, 64, 128):
        b = pts // (n * n)
        p = aclfft.Plan.plan2d(b, n, n, aclfft.C2C)
        p.set_stream(s)
        _profile_rep(torch, lambda: p.exec_c2c(x[:pts], y[:pts], aclfft.FORWARD), reps)
        log("2d %dx%d b=%d: %s" % (n, n, b, p.describe()))
        p.destroy()
    del x, y
    for n in REFLEN_REAL:
        b = max(1024, pts // n)
        xr = torch.randn(b, n, device="cuda")
        X = torch.empty(b, n // 2 + 1, dtype=torch.complex64, device="cuda")
        pr = aclfft.Plan.plan1d(n, aclfft.R2C, b)
        pc = aclfft.Plan.plan1d(n, aclfft.C2R, b)
        pr.set_stream(s)
        pc.set_stream(s)
        _profile_rep(torch, lambda: pr.exec_r2c(xr, X), reps)
        _profile_rep(torch, lambda: pc.exec_c2r(X, xr), reps)
        log("real n=%d b=%d: %s | %s" % (n, b, pr.describe(), pc.describe()))
        pr.destroy()
        pc.destroy()
        del xr, X
    _profile_copy(torch, reps)


def profile_generic(torch, aclfft, reps):
    """The runtime-radix kernel (lengths without a compiled kernel): C2C at 7, 100, 360, 1200, 2000,
    4000, 4913 and 9000 points over 2^26 points; then the streaming copy of 512 MiB (counter
    calibration)."""
    s = torch.cuda.current_stream()
    pts = 1 << 26
    x = torch.randn(pts, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    for n in (7, 100, 360, 1200, 2000, 4000, 4913, 9000):
        b = pts // n
        p = aclfft.Plan.plan1d(n, aclfft.C2C, b)
        p.set_stream(s)
        _profile_rep(torch, lambda: p.exec_c2c(x[:n * b], y[:n * b], aclfft.FORWARD), reps)
        log("c2c n=%d b=%d: %s" % (n, b, p.describe()))
        p.destroy()
    del x, y
    _profile_copy(torch, reps)


def profile_pow2(torch, aclfft, reps):
    """C2C at 1024, 2048 and 4096 points over 2^26 points (the default plans), then the streaming
    copy of 512 MiB (counter calibration)."""
    s = torch.cuda.current_stream()
    pts = 1 << 26
    x = torch.randn(pts, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    for n in (1024, 2048, 4096):
        b = pts // n
        p = aclfft.Plan.plan1d(n, aclfft.C2C, b)
        p.set_stream(s)
        _profile_rep(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), reps)
        log("c2c n=%d b=%d: %s" % (n, b, p.describe()))
        p.destroy()
    del x, y
    _profile_copy(torch, reps)


def profile_mfma(torch, aclfft, reps):
    """The MFMA split-precision path in each mode (BF16, FP16 hi/lo, BF16 hi/mid/lo): C2C at
    n = 16, 32, 64 (dense single stage) and 256, 512, 1024, 2048, 4096 over 2^26 points, 2D 32^2
    and 64^2 images over 2^26 points,
    R2C and C2R at n = 8192 batch 1024, and the
    512-group uneven batch; then the streaming copy of 512 MiB."""
    s = torch.cuda.current_stream()
    modes = ((aclfft.PRECISION_BF16, "bf16"), (aclfft.PRECISION_FP16X2, "fp16x2"),
             (aclfft.PRECISION_BF16X3, "bf16x3"))
    pts = 1 << 26
    x = torch.randn(pts, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    for n in (16, 32, 64, 256, 512, 1024, 2048, 4096):
        for mode, name in modes:
            p = aclfft.Plan.plan1d(n, aclfft.C2C, pts // n)
            p.set_precision(mode)
            p.set_stream(s)
            _profile_rep(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), reps)
            log("c2c n=%d %s: %s" % (n, name, p.describe()))
            p.destroy()
    for nx, ny in ((32, 32), (64, 64)):
        for mode, name in modes:
            p = aclfft.Plan.plan2d(pts // (nx * ny), nx, ny, aclfft.C2C)
            p.set_precision(mode)
            p.set_stream(s)
            _profile_rep(torch, lambda: p.exec_c2c(x, y, aclfft.FORWARD), reps)
            log("2d %dx%d %s: %s" % (nx, ny, name, p.describe()))
            p.destroy()
    del x, y
    n, b = 8192, 1024
    xr = torch.randn(b, n, device="cuda")
    X = torch.empty(b, n // 2 + 1, dtype=torch.complex64, device="cuda")
    for mode, name in modes:
        pr = aclfft.Plan.plan1d(n, aclfft.R2C, b)
        pc = aclfft.Plan.plan1d(n, aclfft.C2R, b)
        for p in (pr, pc):
            p.set_precision(mode)
            p.set_stream(s)
        _profile_rep(torch, lambda: pr.exec_r2c(xr, X), reps)
        _profile_rep(torch, lambda: pc.exec_c2r(X, xr), reps)
        log("real n=%d b=%d %s: %s | %s" % (n, b, name, pr.describe(), pc.describe()))
        pr.destroy()
        pc.destroy()
    del xr, X
    ln, bt = uneven_groups()
    total = sum(a * c for a, c in zip(ln, bt))
    xs = torch.randn(total, dtype=torch.complex64, device="cuda")
    ys = torch.empty_like(xs)
    for mode, name in modes:
        q = aclfft.Plan.plan_many_uneven(ln, bt)
        q.set_precision(mode)
        q.set_stream(s)
        _profile_rep(torch, lambda: q.exec_c2c(xs, ys, aclfft.FORWARD), reps)
        log("uneven 512 groups %s: %s" % (name, q.describe()))
        q.destroy()
    del xs, ys
    _profile_copy(torch, reps)


def _profile_copy(torch, reps):
    from aclfft._lib import stream_copy
    s = torch.cuda.current_stream()
    a = torch.empty(1 << 29, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    h = s.cuda_stream
    _profile_rep(torch, lambda: stream_copy(a.data_ptr(), b.data_ptr(), a.numel(), h), reps)
    log("stream copy 512 MiB")


def profile_workloads(torch, aclfft, reps):
    """R2C and C2R at n=4096 b=8192, n=3000 b=8192, n=8192 b=1024; 2D C2C 2048x2048 batch 4; the
    512-group uneven batch; 1D C2C n=2^20 batch 64 (four-step); 3D C2C 256^3; the streaming copy of
    512 MiB.  Every buffer is allocated at the size its plan reads or writes."""
    s = torch.cuda.current_stream()

    def rep(fn):
        _profile_rep(torch, fn, reps)

    for n, b in ((4096, 8192), (3000, 8192), (8192, 1024)):
        xr = torch.randn(b, n, device="cuda")
        X = torch.empty(b, n // 2 + 1, dtype=torch.complex64, device="cuda")
        pr = aclfft.Plan.plan1d(n, aclfft.R2C, b)
        pc = aclfft.Plan.plan1d(n, aclfft.C2R, b)
        pr.set_stream(s)
        pc.set_stream(s)
        rep(lambda: pr.exec_r2c(xr, X))
        rep(lambda: pc.exec_c2r(X, xr))
        log("real n=%d b=%d: %s | %s" % (n, b, pr.describe(), pc.describe()))
        pr.destroy()
        pc.destroy()
        del xr, X
    x = torch.randn(4, 2048, 2048, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    p = aclfft.Plan.plan2d(4, 2048, 2048, aclfft.C2C)
    p.set_stream(s)
    rep(lambda: p.exec_c2c(x, y, aclfft.FORWARD))
    log("2d 2048x2048 b=4: %s" % p.describe())
    p.destroy()
    del x, y
    ln, bt = uneven_groups()
    total = sum(a * b for a, b in zip(ln, bt))
    xs = torch.randn(total, dtype=torch.complex64, device="cuda")
    ys = torch.empty_like(xs)
    q = aclfft.Plan.plan_many_uneven(ln, bt)
    q.set_stream(s)
    rep(lambda: q.exec_c2c(xs, ys, aclfft.FORWARD))
    log("uneven 512 groups, %d points: %s" % (total, q.describe()))
    q.destroy()
    del xs, ys
    n, b = 1 << 20, 64
    x = torch.randn(b, n, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    p = aclfft.Plan.plan1d(n, aclfft.C2C, b)
    p.set_stream(s)
    rep(lambda: p.exec_c2c(x, y, aclfft.FORWARD))
    log("c2c n=2^20 b=64: %s" % p.describe())
    p.destroy()
    del x, y
    x = torch.randn(256, 256, 256, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    p = aclfft.Plan.plan3d(256, 256, 256, aclfft.C2C)
    p.set_stream(s)
    rep(lambda: p.exec_c2c(x, y, aclfft.FORWARD))
    log("3d 256^3: %s" % p.describe())
    p.destroy()
    del x, y
    _profile_copy(torch, reps)


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
    ap.add_argument("--profile-workloads", nargs="?", const="secondary", default=None,
                    choices=["secondary", "reflen", "mfma", "generic", "pow2"],
                    help="run a set of workloads for a profiler, print no JSON line")
    ap.add_argument("--reps", type=int, default=20, help="--profile-workloads: measured calls per plan")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the FFT path has no multi-GPU mode (a batch would be "
              "sharded across GPUs as contiguous ranges of transforms, with no communication)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import aclfft

    if args.profile_workloads:
        {"secondary": profile_workloads, "reflen": profile_reflen,
         "mfma": profile_mfma, "generic": profile_generic, "pow2": profile_pow2}[args.profile_workloads](torch, aclfft, args.reps)
        return

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

    fn = lambda: plan.exec_c2c(x, y, aclfft.FORWARD)  # noqa: E731
    log("bench: plan ready, warmup %d calls + at least 0.25 s" % args.warmup)
    warm(torch, fn, args.warmup, min_seconds=0.25)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    log("bench: %.4f ms per step" % ms)

    # parity after the timed region: 4096 transforms against the FP64 oracle (the full batch is
    # checked by tests/test_gpu_fullsize.py)
    sub = 4096
    ref = oracle.c2c(x_cpu[:sub].numpy().astype(np.complex128))
    err, emax = errs(oracle, np, y[:sub].cpu().numpy(), ref)
    log("bench: rel_l2 vs FP64 oracle on %d transforms = %.3e (max %.3e)" % (sub, err, emax))

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
        "max_abs_rel_vs_fp64": emax,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
    }
    if not args.no_extras:
        extras = {}
        def headline_cold():
            cms = time_cold(torch, lambda: plan.exec_c2c(x, y, aclfft.FORWARD), stream)
            sms = time_single(torch, lambda: plan.exec_c2c(x, y, aclfft.FORWARD), stream)
            return {"ms": round(cms, 5), "hbm_frac": frac(algo_bytes, cms), "single_call_warm_ms": round(sms, 5),
                    "note": "single calls after reading 1 GiB (4x the last-level cache), median of 10"}

        for name, fn_ in (("headline_cold_llc", headline_cold),
                          ("copy_1GiB", lambda: extra_copy(torch, aclfft, x, 20, stream, ms)),
                          ("mfma_split_n1024_b65536", lambda: extra_mfma(torch, aclfft, oracle, np, x, x_cpu, y,
                                                                         args.steps, stream, err)),
                          ("mfma_split_other_lengths", lambda: extra_mfma_lengths(torch, aclfft, oracle, np, x, y,
                                                                                   args.steps, stream)),
                          ("real_r2c_c2r", lambda: extra_real(torch, aclfft, oracle, np, args.steps, stream)),
                          ("c2c_n1024_b64_latency", lambda: extra_small_latency(torch, aclfft, oracle, np, stream)),
                          ("double_precision", lambda: extra_double(torch, aclfft, oracle, np, args.steps, stream)),
                          ("c2c_2d_2048x2048_b4", lambda: extra_2d(torch, aclfft, oracle, np, 10, stream)),
                          ("c2c_2d_reference_sizes", lambda: extra_2d_small(torch, aclfft, oracle, np, x.view(-1),
                                                                            y.view(-1), 20, stream)),
                          ("uneven_batched_512", lambda: extra_batched(torch, aclfft, oracle, np, 20, stream)),
                          ("loss_grad", lambda: extra_loss_grad(torch, aclfft, 10, stream)),
                          ("vendor_fft_same_gpu", lambda: extra_vendor_fft(torch, aclfft, stream))):
            try:
                log("bench: extra %s" % name)
                extras[name] = fn_()
            except Exception as ex:  # an extra must not hide the headline line
                extras[name] = {"error": "%s: %s" % (type(ex).__name__, ex)}
        out["extras"] = extras
    if not args.no_cpu:
        log("bench: CPU baselines")
        out["cpu_baseline"] = cpu_golden_headline(torch, x_cpu)
        out.setdefault("extras", {})["cpu_oracle_port"] = cpu_oracle_port(x_cpu[:512].numpy().astype(np.complex128))
        try:
            out["extras"]["cpu_golden_path_other_workloads"] = cpu_golden_configs(torch, np)
        except Exception as ex:
            out["extras"]["cpu_golden_path_other_workloads"] = {"error": "%s: %s" % (type(ex).__name__, ex)}
    plan.destroy()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
