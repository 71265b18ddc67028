#!/usr/bin/env python3
"""Write the small golden fixtures of tests/golden/.

The reference keeps no golden vectors: its gen_data.py scripts regenerate them at test time with
torch.manual_seed(12345) and torch.fft on the CPU, then delete them
(src/common/host/test_helper.py:194, tests/test_common.h:387-395).  This script restates that
generation (it does not import the reference) for the reference's cases with n <= 2048, keeping
the first rows of each batch (at most 2048 elements per case; a row's input and golden do not
depend on the other rows), and stores inputs and goldens as float32 data, the dtype the reference
writes to its .bin files.  Cases whose rows are coupled (VERTICAL) are kept only when whole.

Run from the repository root:  python3 tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import ref_cases as rc  # noqa: E402

LIMIT = 2048


def rows(n):
    return max(1, LIMIT // n)


def main():
    out = {}
    for name, n, batch, mode, atol, rtol in rc.C2C:
        if n > LIMIT or (mode == "vertical" and n * batch > LIMIT):
            continue
        torch.manual_seed(rc.SEED)
        x = torch.randn(batch, n, dtype=torch.complex64)
        if mode != "vertical":
            x = x[:rows(n)]
        if mode == "vertical":
            g = torch.fft.fft(x.reshape(n, batch), dim=0).reshape(batch, n)
        elif mode == "roundtrip":
            g = torch.fft.ifft(torch.fft.fft(x, dim=1), dim=1) * n
        elif mode == "forward":
            g = torch.fft.fft(x, dim=1)
        else:
            g = torch.fft.ifft(x, dim=1) * n
        out["c2c/%s/in" % name] = x.numpy()
        out["c2c/%s/golden" % name] = g.numpy().astype(np.complex64)
    for name, n, batch, atol, rtol in rc.R2C:
        if n > LIMIT:
            continue
        torch.manual_seed(rc.SEED)
        x = torch.randn(batch * n).float().reshape(batch, n)[:rows(n)]
        out["r2c/%s/in" % name] = x.numpy()
        out["r2c/%s/golden" % name] = torch.fft.rfft(x, dim=1).numpy().astype(np.complex64)
    for name, n, batch, nyq, atol, rtol in rc.C2R:
        if n > LIMIT:
            continue
        torch.manual_seed(rc.SEED)
        if nyq:
            X = torch.zeros(batch, n // 2 + 1, dtype=torch.complex64)
            X[:, -1] = 1.0
        else:
            X = torch.randn(batch, n // 2 + 1, dtype=torch.complex64)
        X = X[:rows(n)]
        out["c2r/%s/in" % name] = X.numpy()
        out["c2r/%s/golden" % name] = (torch.fft.irfft(X, n=n, dim=1) * n).numpy().astype(np.float32)
    for name, nx, ny, batch, fwd, atol, rtol in rc.FFT2:
        if nx * ny > 4 * LIMIT:
            continue
        torch.manual_seed(rc.SEED)
        x = torch.randn(batch, nx, ny, dtype=torch.complex64)[:1]
        g = torch.fft.fft2(x, dim=(1, 2)) if fwd else torch.fft.ifft2(x, dim=(1, 2)) * (nx * ny)
        out["fft2/%s/in" % name] = x.numpy()
        out["fft2/%s/golden" % name] = g.numpy().astype(np.complex64)
    meta = "torch %s, seed %d, <= %d elements per case" % (torch.__version__, rc.SEED, LIMIT)
    np.savez_compressed(os.path.join(HERE, "reference_goldens.npz"), __meta__=np.array(meta), **out)
    print("wrote %d arrays (%s)" % (len(out), meta))


if __name__ == "__main__":
    main()
