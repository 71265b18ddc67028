#!/usr/bin/env python3
"""Generate registry_1d.inc: the compiled Stockham configurations of kernels_1d.hip.

Each line instantiates one kernel (kind, complex length N, threads per transform T, transforms per
block TPB, LDS pad shift, staged flag, radix sequence).  The host chooser picks among the entries
for a length at plan time (heuristic order = order in this file, or autotuned).

Run: python3 gen_registry.py  (rewrites registry_1d.inc next to this script)
"""
import os

KINDS = {"C2C": "OFX_KIND_C2C", "R2C": "OFX_KIND_R2C", "C2R": "OFX_KIND_C2R",
         "R2C_ODD": "OFX_KIND_R2C_ODD", "C2R_ODD": "OFX_KIND_C2R_ODD", "ROWT": "OFX_KIND_ROWT",
         "C2C_TILES": "OFX_KIND_C2C_TILES", "Z2Z": "OFX_KIND_Z2Z", "D2Z": "OFX_KIND_D2Z",
         "Z2D": "OFX_KIND_Z2D", "D2Z_ODD": "OFX_KIND_D2Z_ODD", "Z2D_ODD": "OFX_KIND_Z2D_ODD",
         "ROWT_Z": "OFX_KIND_ROWT_Z"}

# Radix plans for the lengths that get a dedicated (fully unrolled) kernel.  Everything else runs
# on the runtime-radix kernel (kernels_generic.hip).
PLANS = {
    2: [(2,)], 4: [(4,)], 8: [(8,)], 16: [(16,)],
    32: [(4, 8)], 64: [(8, 8)], 128: [(8, 16)], 256: [(16, 16)],
    512: [(8, 8, 8), (16, 16, 2)],
    1024: [(16, 16, 4), (4, 16, 16), (32, 32), (8, 8, 16), (16, 4, 16)],
    2048: [(16, 16, 8), (8, 16, 16)],
    4096: [(16, 16, 16)],
    8192: [(16, 16, 16, 2)],
    # mixed lengths of the uneven-batch workload and the real-transform halves
    60: [(4, 3, 5)], 243: [(27, 9)], 1000: [(10, 10, 10)], 6561: [(27, 27, 9)],
    1500: [(12, 5, 25), (10, 10, 15)],
    3000: [(8, 15, 25)],
    # the reference's own case lengths (its mixed-radix and real-transform tests) and the complex
    # halves of its real lengths 1536, 2100, 5040 and 25200
    6: [(6,)], 105: [(15, 7)], 210: [(14, 15)], 420: [(20, 21)], 768: [(16, 16, 3)], 1050: [(10, 15, 7)],
    1536: [(16, 16, 6)], 2100: [(12, 25, 7)], 2310: [(10, 21, 11)], 2520: [(12, 14, 15)], 5040: [(12, 20, 21)],
    12600: [(8, 9, 25, 7)],
}


def elems(n, t, rad):
    return max(((n // r + t - 1) // t) * r for r in rad)


def choose_t(n, rad):
    """Threads per transform: about 16 elements per thread, wave-contained when possible; lengths
    beyond 8192 take a whole 1024-thread workgroup."""
    if n > 8192:
        return 1024
    best = None
    for t in (1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024):
        e = elems(n, t, rad)
        if e <= 32 and (best is None):
            best = t
        if e <= 16:
            return t
    return best or 1024


def tpb_for(n, t):
    """Transforms per block: aim at 256 threads and at most ~64 KiB of LDS."""
    tpb = max(1, 256 // t)
    while tpb > 1 and tpb * (n + n // 16) * 8 > 65536:
        tpb //= 2
    return tpb


# paired 16-byte I/O variants (IOV = 2): (kind, N, T, TPB, radices)
PAIRED = [
    ("C2C", 1024, 64, 4, (8, 16, 8)), ("C2C", 1024, 64, 8, (8, 16, 8)),
    ("C2C", 1024, 128, 2, (8, 8, 16)),
    ("C2C", 2048, 128, 2, (8, 16, 16)), ("C2C", 2048, 128, 2, (8, 4, 8, 8)),
    ("C2C", 4096, 256, 1, (8, 8, 8, 8)),
    ("C2C", 512, 32, 8, (8, 8, 8)),
    ("R2C", 2048, 128, 2, (8, 16, 16)), ("C2R", 2048, 128, 2, (16, 16, 8)),
    ("R2C", 4096, 256, 1, (8, 8, 8, 8)), ("C2R", 4096, 256, 1, (8, 8, 8, 8)),
]


# further single-access (IOV = 1) configurations: (kind, N, T, TPB, radices)
EXTRA = [
    ("C2C", 2048, 64, 1, (16, 16, 8)), ("C2C", 2048, 64, 2, (16, 16, 8)), ("C2C", 2048, 256, 1, (8, 8, 8, 4)),
    ("R2C", 2048, 64, 1, (16, 16, 8)), ("R2C", 2048, 64, 2, (16, 16, 8)), ("R2C", 2048, 128, 1, (16, 16, 8)),
    ("C2R", 2048, 64, 1, (16, 16, 8)), ("C2R", 2048, 64, 2, (16, 16, 8)), ("C2R", 2048, 128, 1, (16, 16, 8)),
    ("R2C", 1500, 64, 2, (12, 5, 25)), ("R2C", 1500, 64, 1, (12, 5, 25)),
    ("C2R", 1500, 128, 2, (10, 10, 15)), ("C2R", 1500, 64, 2, (12, 5, 25)), ("C2R", 1500, 64, 1, (12, 5, 25)),
    ("R2C", 4096, 128, 1, (16, 16, 16)), ("C2R", 4096, 128, 1, (16, 16, 16)),
    ("R2C", 1500, 128, 2, (10, 15, 10)), ("R2C", 1500, 128, 2, (10, 10, 15)),
]


# split-exchange C2C configurations (real and imaginary parts through LDS in turn: half the LDS, so
# more workgroups per CU for the long transforms): (kind, N, T, TPB, radices)
# SPLIT_FIRST lead their length's candidates: 6561 points take 52 KiB of complex LDS, so only three
# workgroups share a CU; split, four do (C2C 6561 x 1024: 20.6 vs 23.8 us, one round of workgroups
# instead of two; x 8192: 172.4 vs 173.3 us).  At 4096 and 2048 points the split exchange measured
# slower (197.6 vs 195.6 us, 189.7 vs 187.3 us at 2^26 points) and is not built.
SPLIT_FIRST = [("C2C", 6561, 256, 1, (27, 27, 9)), ("C2C_TILES", 6561, 256, 1, (27, 27, 9))]
SPLIT = [("C2C", 6561, 256, 1, (9, 27, 27))]


# tile-list C2C kernels of the uneven-batch lengths (one launch per length): (N, T, TPB, radices)
TILES = [
    (60, 4, 64, (4, 3, 5)), (128, 8, 32, (8, 16)), (243, 16, 16, (27, 9)), (256, 16, 16, (16, 16)),
    (512, 32, 8, (8, 8, 8)), (1000, 128, 2, (10, 10, 10)), (1024, 64, 2, (8, 16, 8)), (1024, 128, 2, (8, 8, 16)),
    (2048, 256, 1, (8, 8, 8, 4)), (2048, 128, 2, (16, 16, 8)), (4096, 256, 1, (16, 16, 16)),
    (4096, 512, 1, (8, 8, 8, 8)), (6561, 256, 1, (27, 27, 9)),
]


# real transforms with the Hermitian partner butterflies in one thread (IOV = 3): the R2C unpack
# runs in the last pass's registers, the C2R pre-pack in the first pass's, so the spectrum makes
# one trip through LDS fewer: (kind, N, T, TPB, radices[, pad shift]).  HPAIR_FIRST lead their length's
# candidates (real 4096 x 8192: R2C 45.5 vs 46.4 us, C2R 45.3 vs 48.4 us for the LDS unpack /
# pre-pack kernels of the same radices).  At n = 8192 (N = 4096) they measured no faster (R2C x 1024
# 15.2 us either way, x 8192 117.9 vs 113.9 us; C2R 16.9 vs 15.6 us) and are not built.  Real 3000
# (N = 1500): the C2R pair kernel at 64 lanes (unit count 63) with streaming stores is the store
# rule's choice (36.0 vs 37.3-38.4 us); the R2C one measured slower (36.3 vs 33.5 us).
HPAIR_FIRST = [("R2C", 2048, 256, 1, (8, 8, 8, 4)), ("C2R", 2048, 256, 1, (4, 8, 8, 8))]
HPAIR = [("R2C", 2048, 128, 1, (16, 16, 8)), ("C2R", 2048, 128, 1, (8, 16, 16)),
         ("C2R", 1500, 64, 2, (12, 5, 25), 30)]

# streaming variants by access (mode 1 loads and stores, 2 stores only, 3 loads only):
# (kind, N, T, TPB, radices, iov, mode[, pad shift]).  The mode-2 twins of the default kernels are
# what single-launch plans take for outputs of >= 16 MiB (the plan's store rule); the rest are
# tuner candidates.
NTVAR = [
    ("R2C", 2048, 256, 1, (8, 8, 8, 4), 3, 2), ("R2C", 2048, 256, 1, (8, 8, 8, 4), 3, 3),
    ("C2R", 2048, 256, 1, (4, 8, 8, 8), 3, 2),
    ("R2C", 4096, 512, 1, (8, 8, 8, 8), 1, 2), ("R2C", 4096, 512, 1, (8, 8, 8, 8), 1, 3),
    ("R2C", 4096, 512, 1, (8, 8, 8, 8), 1, 1), ("C2R", 4096, 512, 1, (8, 8, 8, 8), 1, 2),
    ("R2C", 1500, 256, 1, (10, 15, 10), 1, 2, 30), ("C2R", 1500, 64, 2, (12, 5, 25), 3, 2, 30),
    ("C2R", 1500, 128, 2, (12, 5, 25), 1, 2, 30),
    ("R2C", 1024, 64, 4, (16, 16, 4), 1, 2), ("C2R", 1024, 64, 4, (16, 16, 4), 1, 2),
    ("C2C", 2048, 128, 2, (16, 16, 8), 1, 2), ("C2C", 4096, 256, 1, (16, 16, 16), 1, 2),
    ("C2C", 1024, 64, 2, (8, 16, 8), 2, 2), ("C2C", 1024, 64, 2, (8, 8, 16), 2, 2),
    ("C2C", 256, 16, 16, (16, 16), 1, 2), ("C2C", 512, 32, 8, (8, 8, 8), 1, 2),
    ("C2C", 8192, 512, 1, (16, 16, 16, 2), 1, 2),
]


# streaming-store twins (mode 2) of the default complex128 kernels (the plan's store rule)
ZNT = [
    ("Z2Z", 1024, 128, 2, (8, 8, 16)), ("Z2Z", 2048, 256, 1, (8, 8, 8, 4)), ("Z2Z", 4096, 512, 1, (8, 8, 8, 8)),
    ("D2Z", 1024, 128, 1, (8, 8, 16)), ("Z2D", 1024, 128, 1, (8, 8, 16)),
    ("D2Z", 2048, 256, 1, (8, 8, 8, 4)), ("Z2D", 2048, 256, 1, (8, 8, 8, 4)),
    ("D2Z", 4096, 512, 1, (8, 8, 8, 8)), ("Z2D", 4096, 512, 1, (8, 8, 8, 8)),
]


# LDS pad variants (pad shift other than the default): (kind, N, T, TPB, radices, pad shift)
PADX = [
    ("R2C", 1500, 256, 1, (15, 10, 10), 30), ("R2C", 1500, 256, 1, (25, 15, 4), 30),
    ("C2R", 1500, 256, 1, (15, 10, 10), 30),
]


def tpb_for_z(n, t):
    """complex128: 16-byte elements, so half the transforms per ~64 KiB of LDS."""
    tpb = max(1, 256 // t)
    while tpb > 1 and tpb * (n + n // 16) * 16 > 65536:
        tpb //= 2
    return tpb


# complex128 configurations measured fastest on MI355X (about 8 elements per thread: complex128
# doubles the registers per element); they lead their length's candidate list
ZFIRST = [
    ("Z2Z", 256, 32, 4, (4, 8, 8)), ("Z2Z", 512, 64, 2, (8, 8, 8)), ("Z2Z", 1024, 128, 2, (8, 8, 16)),
    ("Z2Z", 2048, 256, 1, (8, 8, 8, 4)), ("Z2Z", 4096, 512, 1, (8, 8, 8, 8)),
    ("D2Z", 1024, 128, 1, (8, 8, 16)), ("Z2D", 1024, 128, 1, (8, 8, 16)),
    ("D2Z", 2048, 256, 1, (8, 8, 8, 4)), ("Z2D", 2048, 256, 1, (8, 8, 8, 4)),
    ("D2Z", 4096, 512, 1, (8, 8, 8, 8)), ("Z2D", 4096, 512, 1, (8, 8, 8, 8)),
]


def z_lines():
    """complex128 kernels: Z2Z for every length with a radix plan, D2Z/Z2D for the even real
    lengths up to 8192 (their half-length complex kernels) and for the odd lengths 243 and 6561."""
    out = [(kind, n, t, tpb, padsh(n), rad) for kind, n, t, tpb, rad in ZFIRST]
    for n, plans in sorted(PLANS.items()):
        if n > 6561:
            continue
        rad = plans[0]
        t = choose_t(n, rad)
        out.append(("Z2Z", n, t, tpb_for_z(n, t), padsh(n), rad))
    for n in (2, 4, 8, 16, 32, 64, 128, 256, 512, 1000, 1024, 1500, 2048, 3000, 4096):
        rad = PLANS[n][0]
        t = choose_t(n, rad)
        out.append(("D2Z", n, t, tpb_for_z(n, t), padsh(n), rad))
        out.append(("Z2D", n, t, tpb_for_z(n, t), padsh(n), rad))
    for n in (243, 6561):
        rad = PLANS[n][0]
        t = choose_t(n, rad)
        out.append(("D2Z_ODD", n, t, tpb_for_z(n, t), padsh(n), rad))
        out.append(("Z2D_ODD", n, t, tpb_for_z(n, t), padsh(n), rad))
    out.append(("Z2Z", 1024, 128, 1, padsh(1024), (8, 8, 16)))
    # four-step row steps (transposed store), ~8 elements per thread, LDS <= 147456 bytes
    for n, t, rad in ((128, 16, (8, 16)), (256, 32, (4, 8, 8)), (512, 64, (8, 8, 8)), (1024, 128, (8, 8, 16)),
                      (2048, 256, (8, 8, 8, 4))):
        tpb = 16
        while tpb > 1 and (tpb * (n + n // 16 + 1) * 16 > 147456 or tpb * t > 1024):
            tpb //= 2
        out.append(("ROWT_Z", n, t, tpb, 4, rad))
    return out


# streaming variants (non-temporal HBM input/output) of the default C2C / Z2Z kernels, chosen by
# the plan for working sets well beyond the last-level cache: (kind, N, T, TPB, radices, iov)
# A streaming store pays only where each store instruction writes whole 128-byte lines: the
# transform's lanes must span at least 128 bytes (T >= 16 for complex64).  At 128 points (T = 8,
# 64-byte runs) the streaming kernels measured slower than the cached one over 2^26 points (loads
# and stores 208 us, stores 197 us, cached 185 us), at 64 points (T = 4) streaming stores took 4x
# as long (779 vs 216 us), so those lengths have none.
STREAM = [
    ("C2C", 256, 16, 16, (16, 16), 1), ("C2C", 512, 32, 8, (8, 8, 8), 1),
    # 1024: radix 8 x 8 x 16 first (measured 171.8 vs 173.0 us for 8 x 16 x 8 at batch 65536)
    ("C2C", 1024, 64, 2, (8, 8, 16), 2), ("C2C", 1024, 64, 2, (8, 16, 8), 2),
    ("C2C", 2048, 128, 2, (16, 16, 8), 1),
    ("C2C", 4096, 256, 1, (16, 16, 16), 1), ("C2C", 8192, 512, 1, (16, 16, 16, 2), 1),
    ("Z2Z", 256, 32, 4, (4, 8, 8), 1), ("Z2Z", 512, 64, 2, (8, 8, 8), 1), ("Z2Z", 1024, 128, 2, (8, 8, 16), 1),
    ("Z2Z", 2048, 256, 1, (8, 8, 8, 4), 1), ("Z2Z", 4096, 512, 1, (8, 8, 8, 8), 1),
]


def padsh(n):
    """LDS pad shift: odd lengths have no power-of-two strides, so they go unpadded."""
    return 30 if n % 2 else 4


# configurations that lead their length's candidate list (measured fastest on MI355X; the planner
# takes the first candidate unless autotuning or a cached choice says otherwise)
FIRST = [
    ("C2C", 1024, 64, 2, (8, 16, 8), 2),
    ("R2C", 2048, 256, 1, (8, 8, 8, 4), 1), ("C2R", 2048, 256, 1, (8, 8, 8, 4), 1),
    # 1500 points (real 3000) unpadded: the radix-10/12 strides of these orders meet no LDS bank
    # conflicts without the pad that power-of-two strides need (R2C 3000 x 8192: 35.8 vs 39.4-40.7
    # us, C2R 38.9-39.2 vs 39.7-40.8 us on one box, against the padded (10, 10, 15) / (12, 5, 25))
    ("R2C", 1500, 256, 1, (10, 15, 10), 1, 30), ("C2R", 1500, 128, 2, (12, 5, 25), 1, 30),
    ("R2C", 1500, 256, 1, (10, 10, 15), 1), ("C2R", 1500, 128, 2, (12, 5, 25), 1),
    ("R2C", 4096, 512, 1, (8, 8, 8, 8), 1), ("C2R", 4096, 512, 1, (8, 8, 8, 8), 1),
    # four-step row step of 1024 points: two radix-32 passes (2^20 x 64: 472 vs 482 us); pad one
    # element per 32, so the first pass's stride-32 writes spread over all banks (463.5 vs 466.6 us)
    # 1050 points: radix 6 x 25 x 7 at 64 lanes (2^26 points 208 vs 219 us for 10 x 15 x 7 at 128)
    ("C2C", 1050, 64, 4, (6, 25, 7), 1),
    ("ROWT", 1024, 32, 16, (32, 32), 1, 5),
    ("ROWT", 1024, 32, 16, (32, 32), 1),
]


def lines():
    out = []
    for n, plans in sorted(PLANS.items()):
        for pi, rad in enumerate(plans):
            t = choose_t(n, rad)
            if n == 1024 and rad == (32, 32):
                t = 32
            tpb = tpb_for(n, t)
            # up to 64 points (at most 4 lanes per transform) the staged form leads: its loads and
            # stores go through LDS as whole rows of the batch, where the direct form's lanes touch
            # 8-32-byte pieces of many lines (2^26 points: 16 points 191 vs 288 us, 32: 199 vs 266,
            # 64: 199 vs 216 us)
            # (at 1000 and 60 points the staged form is faster at 2^26 points but 20-45% slower at
            # 2^22: 1000 points 197 vs 211 and 16.8 vs 13.7 us, 60 points 224 vs 251 and 20.8 vs
            # 14.4 us; they keep the direct form first)
            order = (True, False) if n <= 64 and n & (n - 1) == 0 else (False, True)
            for staged in (order if pi == 0 else (False,)):
                out.append(("C2C", n, t, tpb, padsh(n), staged, rad))
    # four-step row step (rows in, transposed out): as many rows per workgroup as LDS allows,
    # so that the transposed stores run over >= 32-byte segments
    for n in (128, 256, 512, 1024, 2048, 4096):
        rad = PLANS[n][0]
        t = choose_t(n, rad)
        tpb = 16
        while tpb > 1 and (tpb * (n + n // 16 + 1) * 8 > 147456 or tpb * t > 1024):
            tpb //= 2
        out.append(("ROWT", n, t, tpb, 4, False, rad))
    # real transforms: even n = 2N uses the N-point complex kernel with fused pack/unpack
    for n in (1024, 2048, 1500, 4096, 512, 256, 128, 64, 32, 16, 8, 4, 2, 8192, 3000, 1000,
              768, 1050, 2520, 12600):
        rad = PLANS[n][0]
        t = choose_t(n, rad)
        tpb = tpb_for(n, t)
        if n <= 16:
            # C2R of up to 32 points: the staged form leads (2^26 points: 32 points 102 vs 139 us,
            # 16 points 99 vs 121 us); staged R2C measured no faster at any short length, and from
            # 64 real points on the direct C2R is faster as well
            out.append(("C2R", n, t, tpb, 4, True, rad))
        out.append(("R2C", n, t, tpb, 4, False, rad))
        out.append(("C2R", n, t, tpb, 4, False, rad))
        if n in (2048, 1500, 4096):
            out.append(("R2C", n, t, tpb, 4, True, rad))
            out.append(("C2R", n, t, tpb, 4, True, rad))
    return out


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    rows = lines()
    with open(os.path.join(here, "registry_1d.inc"), "w") as f:
        f.write("// generated by gen_registry.py - do not edit\n")
        f.write("static const Entry kRegistry[] = {\n")
        for kind, n, t, tpb, rad in HPAIR_FIRST:
            f.write("    OFX_EH(%s, %d, %d, %d, %d, %s),\n" % (KINDS[kind], n, t, tpb, padsh(n), ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, rad in SPLIT_FIRST:
            f.write("    OFX_EX(%s, %d, %d, %d, %d, %s),\n" % (KINDS[kind], n, t, tpb, padsh(n), ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, rad, iov, *psh in FIRST:
            assert t * tpb <= 1024
            psh = psh[0] if psh else 4
            if iov == 2:
                f.write("    OFX_EV(%s, %d, %d, %d, %d, %s),\n" % (KINDS[kind], n, t, tpb, psh, ", ".join(str(r) for r in rad)))
            else:
                f.write("    OFX_E(%s, %d, %d, %d, %d, false, %s),\n" % (KINDS[kind], n, t, tpb, psh, ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, psh, staged, rad in rows:
            assert t * tpb <= 1024
            f.write("    OFX_E(%s, %d, %d, %d, %d, %s, %s),\n" % (
                KINDS[kind], n, t, tpb, psh, "true" if staged else "false",
                ", ".join(str(r) for r in rad)))
        for n, t, tpb, rad in TILES:
            assert t * tpb <= 1024
            f.write("    OFX_E(%s, %d, %d, %d, %d, false, %s),\n" % (KINDS["C2C_TILES"], n, t, tpb, padsh(n),
                                                                 ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, rad in EXTRA:
            assert t * tpb <= 1024
            f.write("    OFX_E(%s, %d, %d, %d, 4, false, %s),\n" % (KINDS[kind], n, t, tpb, ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, rad in PAIRED:
            f.write("    OFX_EV(%s, %d, %d, %d, 4, %s),\n" % (KINDS[kind], n, t, tpb, ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, psh, rad in z_lines():
            assert t * tpb <= 1024
            f.write("    OFX_EZ(%s, %d, %d, %d, %d, %s),\n" % (KINDS[kind], n, t, tpb, psh, ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, rad, iov in STREAM:
            assert t * tpb <= 1024
            mac = "OFX_EZS" if kind == "Z2Z" else ("OFX_EVS" if iov == 2 else "OFX_ES")
            f.write("    %s(%s, %d, %d, %d, %d, %s),\n" % (mac, KINDS[kind], n, t, tpb, padsh(n), ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, rad in SPLIT:
            assert t * tpb <= 1024
            f.write("    OFX_EX(%s, %d, %d, %d, %d, %s),\n" % (KINDS[kind], n, t, tpb, padsh(n), ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, rad, psh in PADX:
            assert t * tpb <= 1024
            f.write("    OFX_E(%s, %d, %d, %d, %d, false, %s),\n" % (KINDS[kind], n, t, tpb, psh, ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, rad, *psh in HPAIR:
            assert t * tpb <= 1024
            f.write("    OFX_EH(%s, %d, %d, %d, %d, %s),\n" % (KINDS[kind], n, t, tpb, psh[0] if psh else padsh(n),
                                                       ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, rad in ZNT:
            assert t * tpb <= 1024
            f.write("    OFX_EZM(%s, %d, %d, %d, %d, 2, %s),\n" % (KINDS[kind], n, t, tpb, padsh(n), ", ".join(str(r) for r in rad)))
        for kind, n, t, tpb, rad, iov, mode, *psh in NTVAR:
            assert t * tpb <= 1024
            f.write("    OFX_EM(%s, %d, %d, %d, %d, %d, %d, %s),\n" % (KINDS[kind], n, t, tpb, psh[0] if psh else padsh(n),
                                                                iov, mode, ", ".join(str(r) for r in rad)))
        f.write("};\n")
    print("wrote %d registry entries" % (len(FIRST) + len(rows) + len(TILES) + len(EXTRA) + len(PAIRED) +
                                         len(z_lines()) + len(STREAM) + len(PADX) + len(SPLIT) + len(SPLIT_FIRST) + len(HPAIR) + len(HPAIR_FIRST) + len(NTVAR) + len(ZNT)))


if __name__ == "__main__":
    main()
