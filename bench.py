"""Benchmark: binned triplets/s on one MI355X, batched mixed path vs plain path.

Workload (default): 64 clustered periodic boxes (Neyman-Scott, L=1000 Mpc/h,
sigma=5, 100 children per parent, N per box log-uniform in [5e4, 2e6], float32
coordinates, unit weights), rmax=30 with 15 linear bins per side, DDD per
box.  One step = the whole batch: cell-list builds of all 64 boxes plus the
triplet count of all 64 DDD terms.

  plain : 64 nc_count calls (FP64 path), one per box, as a reference user would
          make 64 calculateD1D2D3 calls
  mixed : one nc_count_batched call (FP32 separations + FP64 guard-band
          recheck), a single persistent launch over all 64 terms

Unit of work: one ordered binned triplet (one increment of the reference's
unweighted histogram, counted over all sorted bins).  Counts of both paths are
checked bit-identical every run.  Inputs are uploaded to HBM before the
timed region; the timed region is bracketed by hipDeviceSynchronize.

Prints one JSON line (see the end of main()).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_TRIPLET = 8  # r23^2 = 3 sub + 1 mul + 2 fma (fma = 2 flops)
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def _hip():
    return ctypes.CDLL("libamdhip64.so")


def device_sync(hip):
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def copy_calibration(nbytes):
    """Device-to-device copy of a known size (the runtime's copy kernel), run
    under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE to get the counters' scale."""
    hip = _hip()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for buf in (a, b):
        if hip.hipMalloc(ctypes.byref(buf), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMemset failed")
    device_sync(hip)
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    device_sync(hip)
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": nbytes, "copies": 3}), flush=True)


def from_profiles(key):
    """A figure of the committed rocprofv3 summary of this workload's triplet
    kernel: traffic_bytes_per_launch (FETCH_SIZE + WRITE_SIZE, scaled by a copy
    kernel) or valu_busy_pct (VALUBusy counter), None if absent."""
    path = os.path.join(ROOT, "profiles", "k_triplets_mixed_batched_boxes_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(cats, edges, box, budget_s=15.0):
    """Time the CPU oracle (the reference's 27x27-cell loop, FP64, OpenMP over
    primaries) on a bounded sample: the first primaries of the largest box,
    doubling the sample until the run takes a few seconds."""
    from oracle import oracle

    big = max(cats, key=lambda c: c.shape[0])
    threads = os.cpu_count() or 1
    try:
        threads = min(threads, len(os.sched_getaffinity(0)))
    except AttributeError:
        pass
    threads = min(threads, 16)
    rng = np.random.default_rng(1)
    m = 256
    best = None
    while True:
        idx = rng.choice(big.shape[0], size=min(m, big.shape[0]), replace=False)
        sample = big[np.sort(idx)]
        t0 = time.perf_counter()
        c, _ = oracle.count(sample, big, big, edges, box, nthreads=threads)
        dt = time.perf_counter() - t0
        best = (int(c.sum()), dt, sample.shape[0])
        if dt > budget_s / 4 or m >= big.shape[0]:
            break
        m *= 4
    trip, dt, n = best
    return {"value": trip / dt, "unit": "binned triplets/s", "cores": threads, "kind": "port",
            "sample": f"{n} random primaries of the largest box (N={big.shape[0]}) against the full box, "
                      f"{trip} binned triplets in {dt:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--boxes", type=int, default=64)
    ap.add_argument("--no-plain", action="store_true", help="skip the 64-call plain leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN, Context
    from cudancorr_amd.gen import WORKLOADS, batched_boxes

    wl = WORKLOADS["batched_boxes"]
    edges, L = wl["edges"], wl["L"]
    box = np.array([L, L, L])
    t0 = time.perf_counter()
    cats = batched_boxes()[: args.boxes]
    log(f"generated {len(cats)} boxes, {sum(c.shape[0] for c in cats)} points in {time.perf_counter() - t0:.1f}s")
    hip = _hip()
    ctx = Context(0)
    dev = [ctx.catalog(c) for c in cats]
    terms = [(c, c, c) for c in dev]
    device_sync(hip)

    # ---- mixed, batched (the measured value); the autotuner runs first, untimed
    t0 = time.perf_counter()
    tune_m = ctx.autotune(terms, edges, box, path=NC_PATH_MIXED)
    log(f"autotune mixed: {tune_m} ({time.perf_counter() - t0:.1f}s)")
    for _ in range(args.warmup):
        hist_m, cnt_m = ctx.count_batched(terms, edges, box, path=NC_PATH_MIXED)
    device_sync(hip)
    ker_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        hist_m, cnt_m = ctx.count_batched(terms, edges, box, path=NC_PATH_MIXED)
        t = ctx.timing()
        ker_ms.append((t.ms_build, t.ms_count))
    device_sync(hip)
    wall_m = (time.perf_counter() - t0) / args.steps
    trip = int(cnt_m.sum())
    dev_ms_m = float(np.mean([a + b for a, b in ker_ms]))
    count_ms_m = float(np.mean([b for _, b in ker_ms]))
    log(f"mixed batched: {trip} triplets, wall {wall_m * 1e3:.1f} ms, device {dev_ms_m:.1f} ms "
        f"(count kernel {count_ms_m:.1f} ms)")

    # ---- plain, one call per box
    plain = None
    tune_p = None
    if not args.no_plain:
        tune_p = ctx.autotune(terms, edges, box, path=NC_PATH_PLAIN)
        log(f"autotune plain: {tune_p}")
        for c in dev[:1]:
            ctx.count(c, c, c, edges, box, path=NC_PATH_PLAIN)
        device_sync(hip)
        t0 = time.perf_counter()
        dms, cms = 0.0, 0.0
        cnt_p = []
        for c in dev:
            _, cp = ctx.count(c, c, c, edges, box, path=NC_PATH_PLAIN)
            t = ctx.timing()
            dms += t.ms_build + t.ms_count
            cms += t.ms_count
            cnt_p.append(cp)
        device_sync(hip)
        wall_p = time.perf_counter() - t0
        cnt_p = np.stack(cnt_p)
        if not np.array_equal(cnt_p, cnt_m):
            raise SystemExit("plain and mixed counts differ")
        plain = {"wall_ms": wall_p * 1e3, "device_ms": dms, "count_kernel_ms": cms,
                 "triplets_per_s": trip / (wall_p)}
        log(f"plain x{len(dev)}: wall {wall_p * 1e3:.1f} ms, device {dms:.1f} ms (count kernels {cms:.1f} ms)")

    value = trip / wall_m
    achieved = trip * FLOPS_PER_TRIPLET / (count_ms_m * 1e-3) / 1e12
    cpu = None if args.no_cpu else cpu_baseline(cats, edges, box)
    out = {
        "metric": "binned triplets/s on 1 MI355X, batched mixed vs plain, at stated error",
        "value": value,
        "unit": "binned triplets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_m * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded Neyman-Scott boxes)",
        "config": {"workload": "64 clustered periodic boxes, N in [5e4, 2e6] log-uniform, L=1000, "
                               "rmax=30, 15 bins/side, DDD", "boxes": len(cats),
                   "points": int(sum(c.shape[0] for c in cats)), "binned_triplets": trip},
        "error": {"unweighted": "bit-exact (plain vs mixed, every run)"},
        "plain": plain,
        "tuning": {"mixed": tune_m, "plain": tune_p},
        "speedup_vs_plain": (plain["wall_ms"] / (wall_m * 1e3)) if plain else None,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": from_profiles("traffic_bytes_per_launch"),
                     "valu_busy_pct": from_profiles("valu_busy_pct"),
                     "kernel": "k_triplets (mixed, batched)", "kernel_ms": count_ms_m},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    for c in dev:
        c.close()
    ctx.close()


if __name__ == "__main__":
    main()
