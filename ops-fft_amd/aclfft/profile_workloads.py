"""Run the library's secondary workloads back to back, for a profiler to trace.

    python3 ops-fft_amd/aclfft/profile_workloads.py [--reps 20]

Workloads (each plan is called until 0.1 s of device time has passed, so that its tables exist and
the clocks have settled, then `reps` more times: a trace's last `reps` dispatches of a kernel are
the measured ones):
  R2C and C2R at n=4096 b=8192, n=3000 b=8192, n=8192 b=1024;
  2D C2C 2048x2048, batch 4;
  the 512-group uneven batch (lengths log-uniform over 60..6561, batches U{1..64}, seed 4);
  1D C2C n=2^20, batch 64 (four-step);
  3D C2C 256^3;
  the streaming copy of 1 GiB (512 MiB read + 512 MiB written), the bandwidth yardstick.
Every buffer is allocated at the size its plan reads or writes.
"""
import argparse
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np  # noqa: E402
import torch  # noqa: E402

import aclfft  # noqa: E402

UNEVEN_LENGTHS = (60, 128, 243, 1000, 1024, 2048, 4096, 6561)


def uneven_groups(count=512, seed=4):
    rng = np.random.default_rng(seed)
    logs = np.log(np.array(UNEVEN_LENGTHS, float))
    u = rng.uniform(logs.min(), logs.max(), count)
    ln = [int(UNEVEN_LENGTHS[int(np.argmin(np.abs(logs - v)))]) for v in u]
    bt = [int(b) for b in rng.integers(1, 65, count)]
    return ln, bt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args()
    s = torch.cuda.current_stream()

    def rep(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < 0.1:
            for _ in range(5):
                fn()
            torch.cuda.synchronize()
        for _ in range(args.reps):
            fn()
        torch.cuda.synchronize()

    for n, b in ((4096, 8192), (3000, 8192), (8192, 1024)):
        xr = torch.randn(b, n, device="cuda")
        X = torch.empty(b, n // 2 + 1, dtype=torch.complex64, device="cuda")
        pr = aclfft.Plan.plan1d(n, aclfft.R2C, b)
        pc = aclfft.Plan.plan1d(n, aclfft.C2R, b)
        pr.set_stream(s)
        pc.set_stream(s)
        rep(lambda: pr.exec_r2c(xr, X))
        rep(lambda: pc.exec_c2r(X, xr))
        print("real n=%d b=%d: %s | %s" % (n, b, pr.describe(), pc.describe()), flush=True)
        pr.destroy()
        pc.destroy()
        del xr, X
    x = torch.randn(4, 2048, 2048, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    p = aclfft.Plan.plan2d(4, 2048, 2048, aclfft.C2C)
    p.set_stream(s)
    rep(lambda: p.exec_c2c(x, y, aclfft.FORWARD))
    print("2d 2048x2048 b=4: %s" % p.describe(), flush=True)
    p.destroy()
    del x, y
    ln, bt = uneven_groups()
    total = sum(a * b for a, b in zip(ln, bt))
    xs = torch.randn(total, dtype=torch.complex64, device="cuda")
    ys = torch.empty_like(xs)
    q = aclfft.Plan.plan_many_uneven(ln, bt)
    q.set_stream(s)
    rep(lambda: q.exec_c2c(xs, ys, aclfft.FORWARD))
    print("uneven 512 groups, %d points: %s" % (total, q.describe()), flush=True)
    q.destroy()
    del xs, ys
    n, b = 1 << 20, 64
    x = torch.randn(b, n, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    p = aclfft.Plan.plan1d(n, aclfft.C2C, b)
    p.set_stream(s)
    rep(lambda: p.exec_c2c(x, y, aclfft.FORWARD))
    print("c2c n=2^20 b=64: %s" % p.describe(), flush=True)
    p.destroy()
    del x, y
    x = torch.randn(256, 256, 256, dtype=torch.complex64, device="cuda")
    y = torch.empty_like(x)
    p = aclfft.Plan.plan3d(256, 256, 256, aclfft.C2C)
    p.set_stream(s)
    rep(lambda: p.exec_c2c(x, y, aclfft.FORWARD))
    print("3d 256^3: %s" % p.describe(), flush=True)
    p.destroy()
    del x, y
    from aclfft._lib import stream_copy
    a = torch.empty(1 << 29, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    h = s.cuda_stream
    rep(lambda: stream_copy(a.data_ptr(), b.data_ptr(), a.numel(), h))
    print("stream copy 512 MiB", flush=True)


if __name__ == "__main__":
    main()
