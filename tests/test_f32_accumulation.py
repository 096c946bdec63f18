"""CPU emulation of the FP32 accumulation arithmetic (NC_ACCUM_FP32): the
float-float weight product and the TwoSum compensation the kernel applies per
accepted pair (hist_add_f32 in cudancorr_amd/csrc/ncorr_device.hip), replayed in
numpy float32 / float64 in the same order of operations, against exact sums
(math.fsum of the exact FP64 products of FP32 weights).  Also checks the
header's accumulation enum against the Python binding."""
import math
import os
import re

import numpy as np

from cudancorr_amd import NC_ACCUM_EXACT, NC_ACCUM_FP32, NC_ACCUM_FP64

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
f32 = np.float32


def _f32_pair(x):
    """The staged row factor: FP64 value -> (hi, lo) in FP32."""
    hi = f32(x)
    return hi, f32(x - float(hi))


def _product(wr, wl):
    """(hi, lo) of (wr.hi + wr.lo) * wl, as the kernel forms it (fmaf)."""
    ph = f32(wr[0] * wl)
    e = f32(float(wr[0]) * float(wl) - float(ph))  # fmaf(hi, wl, -ph): exact in FP64
    pl = f32(float(wr[1]) * float(wl) + float(e))  # fmaf(lo, wl, e), one rounding
    return ph, pl


def _accumulate(values_row, values_lane, compensate=True):
    s = f32(0.0)
    comp = 0.0
    for wr, wl in zip(values_row, values_lane):
        ph, pl = _product(wr, wl)
        old = s
        s = f32(old + ph)
        if compensate:
            bv = f32(s - old)
            err = f32(f32(old - f32(s - bv)) + f32(ph - bv))
            comp += float(err) + float(pl)
    return float(s) + comp


def test_compensated_sum_matches_exact_sum():
    rng = np.random.default_rng(3)
    n = 20000
    w1 = float(f32(rng.lognormal(0.0, 0.3)))
    wrow = rng.lognormal(0.0, 0.3, n).astype(np.float32)
    wlane = rng.lognormal(0.0, 0.3, n).astype(np.float32)
    c = 6.0
    rows = [_f32_pair(c * w1 * float(w)) for w in wrow]  # exact in FP64: <= 51 bits
    exact = math.fsum(c * w1 * float(a) * float(b) for a, b in zip(wrow, wlane))  # each term rounded once
    got = _accumulate(rows, wlane)
    plain = _accumulate(rows, wlane, compensate=False)
    assert abs(got / exact - 1.0) < 1e-13
    # without the compensation an FP32 running sum is far off: the check can fail
    assert abs(plain / exact - 1.0) > 1e-7


def test_product_error_below_2_to_minus_46():
    rng = np.random.default_rng(4)
    for _ in range(2000):
        w1, a, b = (float(f32(x)) for x in rng.lognormal(0.0, 1.0, 3))
        x = 6.0 * w1 * a  # exact in FP64
        ph, pl = _product(_f32_pair(x), f32(b))
        exact = x * b  # one FP64 rounding, far below the bound
        assert abs((float(ph) + float(pl)) / exact - 1.0) < 2.0 ** -46


def test_accumulation_enum_matches_header():
    text = open(os.path.join(ROOT, "include", "ncorr_hip.h")).read()
    m = re.search(r"enum nc_accum \{([^}]*)\}", text)
    vals = dict((k.strip(), int(v)) for k, v in (e.split("=") for e in m.group(1).split(",")))
    assert vals == {"NC_ACCUM_FP64": NC_ACCUM_FP64, "NC_ACCUM_EXACT": NC_ACCUM_EXACT, "NC_ACCUM_FP32": NC_ACCUM_FP32}
