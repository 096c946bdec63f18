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

Survey workloads (--workload survey, survey-small): the four terms DDD, DDR,
DRR, RRR of a cap/shell survey (weighted data, unit-weight randoms) as in
source/main.cu:72-95.  Each term's primaries are split into --slices slices of
their cell-sorted order (16 for survey, 1 for survey-small) (nc_set_primary_slice, the same catalogue handles);
each slice is one nc_count_batched call over the four terms (mixed) or four
nc_count calls (plain), and one step is all slices.  Binned triplets
of the weighted terms come from an untimed unit-weight run of the same terms;
weighted sums of the mixed path are compared with the plain FP64 path.

rppi workload (--workload rppi-box): the reference's rppi binning on the
1M-point clustered box, the key stream against direct global atomics.

Prints one JSON line (see the end of main(), run_survey() and run_rppi()).
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


def atomic_rate():
    """Scattered global atomics per second (torch index_add_: one atomic add
    per element, random bins), u64 and u32, for histograms of 4096 to 2^21
    bins: the ceiling of any path that adds each binned triplet to a global
    histogram."""
    import torch

    dev = torch.device("cuda:0")
    n = 1 << 27
    g = torch.Generator(device=dev)
    g.manual_seed(5)
    res = []
    for bins in (4096, 32768, 262144, 1 << 21):
        idx = torch.randint(0, bins, (n,), device=dev, generator=g)
        for dt in (torch.int64, torch.int32):
            h = torch.zeros(bins, dtype=dt, device=dev)
            one = torch.ones(n, dtype=dt, device=dev)
            h.index_add_(0, idx, one)
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            best = 1e30
            for _ in range(3):
                a.record()
                h.index_add_(0, idx, one)
                b.record()
                b.synchronize()
                best = min(best, a.elapsed_time(b))
            if int(h.sum().item()) != 4 * n:
                raise SystemExit("index_add_ lost updates")
            res.append({"bins": bins, "dtype": str(dt).replace("torch.", ""), "ms": best, "atomics_per_s": n / (best * 1e-3)})
    print(json.dumps({"atomic_rate": res}), flush=True)


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


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(cats, edges, box, budget_s=15.0, survey=False, gpu_count=None):
    """Time the CPU oracle (the reference's 27x27-cell loop, FP64, OpenMP over
    primaries) on a bounded sample: random primaries of the largest catalogue
    against all of it, the sample growing until the run takes a few seconds.
    gpu_count(sample, catalogue) -> counts: the same subset counted on the GPU,
    which must agree bit for bit."""
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
        c, _ = oracle.count(sample, big, big, edges, box, survey=survey, nthreads=threads)
        dt = time.perf_counter() - t0
        best = (int(c.sum()), dt, sample.shape[0], c, sample)
        if dt > budget_s / 4 or m >= big.shape[0]:
            break
        m *= 4
    trip, dt, n, c, sample = best
    checked = None
    if gpu_count is not None:
        checked = bool(np.array_equal(gpu_count(sample, big), c))
        if not checked:
            raise SystemExit("the GPU count of the CPU baseline's sample differs from the oracle")
    return {"value": trip / dt, "unit": "binned triplets/s", "cores": threads, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": f"{n} random primaries of the largest catalogue (N={big.shape[0]}) against all of it, "
                      f"{trip} binned triplets in {dt:.2f} s",
            "gpu_counts_bit_exact_on_sample": checked}


def survey_workload(name):
    """(data, randoms, edges, box, ori, description) of a survey workload,
    shifted by prep_cat_survey (source/survey.cu:60-128)."""
    from cudancorr_amd.gen import SEED, WORKLOADS, prep_cat_survey, survey_cap

    if name == "survey":
        d, r = survey_cap(4_000_000, 20_000_000, seed=SEED + 4000, clustered=True, sigma=4.0)
        edges = WORKLOADS["survey_large"]["edges"]
        desc = ("survey cap RA 0-60, Dec 0-60 deg, comoving 1000-2000 Mpc/h: 4M Neyman-Scott data (sigma=4, "
                "lognormal(0,0.3) f32 weights) + 20M uniform randoms (w=1), rmax=50, 25 bins/side, "
                "DDD+DDR+DRR+RRR")
    else:
        d, r = survey_cap(500_000, 2_500_000, seed=SEED + 2000)
        edges = WORKLOADS["survey_cap"]["edges"]
        desc = ("survey cap RA 0-60, Dec 0-60 deg, comoving 1000-2000 Mpc/h: 500k uniform data "
                "(lognormal(0,0.3) f32 weights) + 2.5M randoms (w=1), rmax=40, 20 bins/side, DDD+DDR+DRR+RRR")
    box, ori = prep_cat_survey(d, r)
    return d, r, edges, box, ori, desc


def run_survey(args):
    from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN, Context

    t0 = time.perf_counter()
    d, r, edges, box, _, desc = survey_workload(args.workload)
    log(f"generated {d.shape[0]} data + {r.shape[0]} randoms in {time.perf_counter() - t0:.1f}s")
    hip = _hip()
    ctx = Context(0)
    K = max(1, args.slices if args.slices is not None else (1 if args.workload == "survey-small" else 16))
    du = d.copy()
    du[:, 3] = 1.0
    D, R, DU = ctx.catalog(d), ctx.catalog(r), ctx.catalog(du)
    names = ["DDD", "DDR", "DRR", "RRR"]

    def terms(k, unit=False):
        # slice k of K of each term's cell-sorted primaries (nc_set_primary_slice),
        # on the same catalogue handles: DDD and RRR keep their six-fold symmetry
        ctx.set_primary_slice(k, K)
        dd = DU if unit else D
        return [(dd, dd, dd), (dd, dd, R), (dd, R, R), (R, R, R)]

    device_sync(hip)

    def run(path):
        hist = None
        cnt = None
        dev = cms = 0.0
        for _ in range(args.warmup):
            ctx.count_batched(terms(0), edges, box, survey=True, path=path)
        device_sync(hip)
        t0 = time.perf_counter()
        for k in range(K):
            if path == NC_PATH_MIXED:
                h, c = ctx.count_batched(terms(k), edges, box, survey=True, path=path)
                t = ctx.timing()
                dev += t.ms_build + t.ms_count
                cms += t.ms_count
            else:
                hc = []
                for tm in terms(k):
                    hc.append(ctx.count(*tm, edges, box, survey=True, path=path))
                    t = ctx.timing()
                    dev += t.ms_build + t.ms_count
                    cms += t.ms_count
                h, c = np.stack([a for a, _ in hc]), np.stack([b for _, b in hc])
            hist = h if hist is None else hist + h
            cnt = c if cnt is None else cnt + c
            log(f"{'mixed' if path == NC_PATH_MIXED else 'plain'} slice {k + 1}/{K}: "
                f"{time.perf_counter() - t0:.1f}s")
        device_sync(hip)
        return hist, cnt, time.perf_counter() - t0, dev, cms

    t0 = time.perf_counter()
    tune_m = ctx.autotune(terms(0), edges, box, survey=True, path=NC_PATH_MIXED)
    log(f"autotune mixed: {tune_m} ({time.perf_counter() - t0:.1f}s)")
    hist_m, cnt_m, wall_m, dev_m, cms_m = run(NC_PATH_MIXED)
    log(f"mixed batched: wall {wall_m:.2f} s, device {dev_m / 1e3:.2f} s (count kernels {cms_m / 1e3:.2f} s)")

    # binned triplets of the weighted terms: the same terms with unit weights (untimed)
    trip_terms = np.zeros(4, dtype=np.int64)
    trip_terms[3] = int(cnt_m[3].sum())
    for k in range(K):
        _, c = ctx.count_batched(terms(k, unit=True)[:3], edges, box, survey=True, path=NC_PATH_MIXED)
        trip_terms[:3] += c.reshape(3, -1).sum(axis=1).astype(np.int64)
        log(f"unit-weight recount slice {k + 1}/{K}")
    trip = int(trip_terms.sum())
    log(f"binned triplets per term: {dict(zip(names, trip_terms.tolist()))}")

    plain = None
    tune_p = None
    err = {}
    if not args.no_plain:
        tune_p = ctx.autotune(terms(0), edges, box, survey=True, path=NC_PATH_PLAIN)
        log(f"autotune plain: {tune_p}")
        hist_p, cnt_p, wall_p, dev_p, cms_p = run(NC_PATH_PLAIN)
        if not np.array_equal(cnt_p[3], cnt_m[3]):
            raise SystemExit("plain and mixed RRR counts differ")
        for i, n in enumerate(names[:3]):
            a, b = hist_m[i].ravel(), hist_p[i].ravel()
            nz = b != 0
            if np.any(a[~nz] != 0):
                raise SystemExit(f"{n}: mixed has weight in bins the plain path leaves empty")
            err[n] = float(np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz]), initial=0.0))
        if max(err.values()) > 1e-9:
            raise SystemExit(f"weighted sums differ by more than 1e-9 relative: {err}")
        plain = {"wall_ms": wall_p * 1e3, "device_ms": dev_p, "count_kernel_ms": cms_p,
                 "triplets_per_s": trip / wall_p}
        log(f"plain: wall {wall_p:.2f} s, device {dev_p / 1e3:.2f} s (count kernels {cms_p / 1e3:.2f} s)")

    ctx.set_primary_slice(0, 1)
    achieved = trip * FLOPS_PER_TRIPLET / (cms_m * 1e-3) / 1e12
    cpu = None if args.no_cpu else cpu_baseline(
        [r], edges, box, survey=True,
        gpu_count=lambda smp, big: ctx.count(smp, R, R, edges, box, survey=True)[1])
    out = {
        "metric": "binned triplets/s on 1 MI355X, batched mixed vs plain, at stated error",
        "value": trip / wall_m,
        "unit": "binned triplets/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": args.warmup,
        "ms_per_step": wall_m * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded survey cap)",
        "config": {"workload": desc, "data_points": int(d.shape[0]), "random_points": int(r.shape[0]),
                   "slices": K, "binned_triplets": trip,
                   "binned_triplets_per_term": dict(zip(names, trip_terms.tolist()))},
        "error": {"unweighted": "RRR bit-exact (plain vs mixed)" if plain else "not compared (no plain leg)",
                  "weighted_rel_max_vs_plain_fp64": err or None},
        "plain": plain,
        "tuning": {"mixed": tune_m, "plain": tune_p},
        "speedup_vs_plain": (plain["wall_ms"] / (wall_m * 1e3)) if plain else None,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None, "valu_busy_pct": None,
                     "kernel": "k_triplets (mixed, batched)", "kernel_ms": cms_m},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    for c in (D, R, DU):
        c.close()
    ctx.close()


def config4_subset(args, hip):
    """The large weighted survey (4M Neyman-Scott data with lognormal weights +
    20M uniform randoms, rmax=50, 25 bins/side, DDD+DDR+DRR+RRR as
    source/main.cu:72-95) on a stated subset: primary slices --config4-slices
    of 32 in cell-sorted order (nc_set_primary_slice), all four terms per slice
    in one nc_count_batched call.  Legs: mixed path with FP64 accumulation (the
    measured one), plain FP64 path (four nc_count calls per slice), and the
    exact fixed-point accumulation (NC_ACCUM_EXACT on the plain path) as the
    FP64 reference of the weighted sums.  Binned triplets of the weighted terms
    from an untimed unit-weight run.  The CPU oracle times a sample of RRR
    primaries against all randoms, checked bit for bit on the GPU."""
    from cudancorr_amd import NC_ACCUM_EXACT, NC_ACCUM_FP32, NC_ACCUM_FP64, NC_PATH_MIXED, NC_PATH_PLAIN, Context

    slices = [int(x) for x in args.config4_slices.split(",")]
    K = 32
    t0 = time.perf_counter()
    d, r, edges, box, _, desc = survey_workload("survey")
    log(f"config 4: generated {d.shape[0]} data + {r.shape[0]} randoms in {time.perf_counter() - t0:.1f}s")
    names = ["DDD", "DDR", "DRR", "RRR"]
    du = d.copy()
    du[:, 3] = 1.0
    out = {"workload": desc + f"; primary slices {slices} of {K} (cell-sorted order) of every term",
           "slices": slices, "nslices": K}
    with Context(0) as ctx:
        D, R, DU = ctx.catalog(d), ctx.catalog(r), ctx.catalog(du)
        terms = [(D, D, D), (D, D, R), (D, R, R), (R, R, R)]
        unit_terms = [(DU, DU, DU), (DU, DU, R), (DU, R, R)]
        tune = ctx.autotune(terms, edges, box, survey=True, path=NC_PATH_MIXED)
        log(f"config 4: autotune mixed {tune}")

        def leg(path, batched=True, exact=False, tl=terms, acc=None):
            ctx.set_accumulation(acc if acc is not None else NC_ACCUM_EXACT if exact else NC_ACCUM_FP64)
            hist = cnt = None
            ms = 0.0
            device_sync(hip)
            t1 = time.perf_counter()
            for k in slices:
                ctx.set_primary_slice(k, K)
                if batched:
                    h, c = ctx.count_batched(tl, edges, box, survey=True, path=path)
                    ms += ctx.timing().ms_count
                else:
                    hc = []
                    for tm in tl:
                        hc.append(ctx.count(*tm, edges, box, survey=True, path=path))
                        ms += ctx.timing().ms_count
                    h, c = np.stack([a for a, _ in hc]), np.stack([b for _, b in hc])
                hist = h if hist is None else hist + h
                cnt = c if cnt is None else cnt + c
            device_sync(hip)
            wall = time.perf_counter() - t1
            ctx.set_primary_slice(0, 1)
            ctx.set_accumulation(NC_ACCUM_FP64)
            return hist, cnt, wall, ms

        leg(NC_PATH_MIXED, tl=[terms[3]])  # warm-up (one term)
        hm_, cm_, wall_m, ms_m = leg(NC_PATH_MIXED)
        tri_terms = ctx.timing().tri_terms
        log(f"config 4: mixed {wall_m:.2f} s (count kernels {ms_m / 1e3:.2f} s)")
        _, cu, _, _ = leg(NC_PATH_MIXED, tl=unit_terms)
        trip_terms = [int(c.sum()) for c in cu] + [int(cm_[3].sum())]
        trip = int(sum(trip_terms))
        log(f"config 4: binned triplets {dict(zip(names, trip_terms))}")
        tune_p = ctx.autotune(terms, edges, box, survey=True, path=NC_PATH_PLAIN)
        log(f"config 4: autotune plain {tune_p}")
        hx, cx, wall_x, ms_x = leg(NC_PATH_PLAIN, exact=True)
        log(f"config 4: exact reference (plain path) {wall_x:.2f} s")
        hf_, cf_, wall_f, ms_f = leg(NC_PATH_MIXED, acc=NC_ACCUM_FP32)
        log(f"config 4: mixed with FP32 accumulation {wall_f:.2f} s")
        hp, cp, wall_p, ms_p = leg(NC_PATH_PLAIN, batched=False)
        log(f"config 4: plain x4 per slice {wall_p:.2f} s")
        if not (np.array_equal(cm_[3], cp[3]) and np.array_equal(cm_[3], cx[3]) and np.array_equal(cm_[3], cf_[3])):
            raise SystemExit("config 4: RRR counts differ between legs")

        def rel(a, b):
            nz = b != 0
            if np.any(a[~nz] != 0):
                raise SystemExit("config 4: weight in bins the exact reference leaves empty")
            return float(np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz]), initial=0.0))

        err_m = {n: rel(hm_[i], hx[i]) for i, n in enumerate(names[:3])}
        err_p = {n: rel(hp[i], hx[i]) for i, n in enumerate(names[:3])}
        err_f = {n: rel(hf_[i], hx[i]) for i, n in enumerate(names[:3])}
        tot_m = {n: float(abs(hm_[i].sum() - hx[i].sum()) / abs(hx[i].sum())) for i, n in enumerate(names[:3])}
        if max(err_m.values()) > 1e-9 or max(err_p.values()) > 1e-9 or max(err_f.values()) > 1e-9:
            raise SystemExit(f"config 4: weighted sums off the exact reference by more than 1e-9: {err_m} {err_p} "
                             f"{err_f}")
        cpu = None
        if not args.no_cpu:
            from oracle import oracle

            threads = min(16, len(os.sched_getaffinity(0)))
            rng = np.random.default_rng(4)
            m = 128
            while True:
                sample = r[np.sort(rng.choice(r.shape[0], size=m, replace=False))]
                t1 = time.perf_counter()
                c, _ = oracle.count(sample, r, r, edges, box, survey=True, nthreads=threads)
                dt = time.perf_counter() - t1
                if dt > 3.0 or m >= 4096:
                    break
                m *= 2
            g = ctx.count(sample, R, R, edges, box, survey=True)[1]
            if not np.array_equal(g, c):
                raise SystemExit("config 4: the GPU count of the CPU sample differs from the oracle")
            cpu = {"value": float(c.sum()) / dt, "unit": "binned triplets/s", "cores": threads, "kind": "port",
                   "sample": f"{m} random RRR primaries against all {r.shape[0]} randoms, {int(c.sum())} "
                             f"triplets in {dt:.2f} s", "gpu_counts_bit_exact_on_sample": True}
        for c in (D, R, DU):
            c.close()
    out.update({
        "binned_triplets": trip, "binned_triplets_per_term": dict(zip(names, trip_terms)),
        "mixed": {"wall_s": wall_m, "count_kernel_ms": ms_m, "triplets_per_s": trip / wall_m},
        "plain": {"wall_s": wall_p, "count_kernel_ms": ms_p, "triplets_per_s": trip / wall_p},
        "speedup_vs_plain": wall_p / wall_m,
        "mixed_fp32_accumulation": {
            "wall_s": wall_f, "count_kernel_ms": ms_f, "triplets_per_s": trip / wall_f,
            "weighted_rel_max_vs_exact": err_f,
            "what": "NC_ACCUM_FP32: FP32 float-float weight products, FP32 running sums with exact TwoSum "
                    "compensation in FP64, folded into FP64 totals (the default accumulates FP64 products)"},
        "exact_reference": {"wall_s": wall_x, "count_kernel_ms": ms_x,
                            "what": "plain FP64 path with NC_ACCUM_EXACT: the FP64 products (w1*w2)*w3 summed "
                                    "exactly in 128-bit fixed point (each truncated by < 2^-K, K from the "
                                    "largest weights), rounded once"},
        "error": {"unweighted": "RRR counts bit-exact (mixed = plain = exact reference)",
                  "weighted_rel_max_vs_exact": err_m, "weighted_rel_max_plain_vs_exact": err_p,
                  "weighted_total_rel_vs_exact": tot_m},
        "six_fold_terms": tri_terms, "tuning": {"mixed": tune, "plain": tune_p}, "cpu_baseline": cpu,
    })
    return out


def run_rppi_survey(args):
    """rppi binning of the survey workload (500k weighted data + 2.5M randoms,
    cap/shell), rp edges 0..40 in 20 bins, 20 unit pi bins: DDD, DDR, DRR, RRR
    in one batched launch with the key stream, against the reference's direct
    global atomics (NC_RPPI_STREAM=0); binned triplets of the weighted terms
    from a unit-weight run; the CPU oracle on a sample of RRR primaries, checked
    against the GPU on the same sample."""
    from cudancorr_amd import Context

    from cudancorr_amd.gen import edges_linear

    d, r, _, box, ori, _ = survey_workload("survey-small")
    edges, npi = edges_linear(0.0, 40.0, 20), 20
    names = ["DDD", "DDR", "DRR", "RRR"]
    hip = _hip()

    def run(ctx, unit=False, steps=None, warmup=None):
        """(hist, counts, wall s/step, count ms, tuning); the side legs (unit
        weights, direct atomics) run one step after one warm-up."""
        steps = args.steps if steps is None else steps
        warmup = args.warmup if warmup is None else warmup
        dd = d.copy() if unit else d
        if unit:
            dd[:, 3] = 1.0
        D, R = ctx.catalog(dd), ctx.catalog(r)
        terms = [(D, D, D), (D, D, R), (D, R, R), (R, R, R)]
        tune = ctx.autotune(terms, edges, box, survey=True, npi=npi, ori=ori)
        for _ in range(warmup):
            ctx.count_batched(terms, edges, box, survey=True, npi=npi, ori=ori)
        device_sync(hip)
        t0 = time.perf_counter()
        for _ in range(steps):
            hist, cnt = ctx.count_batched(terms, edges, box, survey=True, npi=npi, ori=ori)
        device_sync(hip)
        wall = (time.perf_counter() - t0) / steps
        ms = ctx.timing().ms_count
        D.close()
        R.close()
        return hist, cnt, wall, ms, tune

    os.environ["NC_RPPI_STREAM"] = "1"  # surveys default to direct atomics: force the key stream here
    with Context(0) as ctx:
        hist_s, cnt_s, wall_s, ms_s, tune_s = run(ctx)
        _, cnt_u, _, _, _ = run(ctx, unit=True, steps=1, warmup=0)
        trip_terms = [int(c.sum()) for c in cnt_u[:3]] + [int(cnt_s[3].sum())]
        trip = int(sum(trip_terms))
        log(f"rppi survey key stream: {trip} triplets {dict(zip(names, trip_terms))}, wall {wall_s * 1e3:.1f} ms, "
            f"count {ms_s:.1f} ms, tuning {tune_s}")
        cpu = None
        if not args.no_cpu:
            from oracle import oracle

            threads = min(16, len(os.sched_getaffinity(0)))
            rng = np.random.default_rng(1)
            sample = r[np.sort(rng.choice(r.shape[0], size=2048, replace=False))]
            t0 = time.perf_counter()
            c, _, _ = oracle.count_rppi(sample, r, r, edges, npi, box, survey=True, ori=ori, nthreads=threads)
            dt = time.perf_counter() - t0
            R = ctx.catalog(r)
            _, g = ctx.count(sample, R, R, edges, box, survey=True, npi=npi, ori=ori)
            R.close()
            if not np.array_equal(g, c):
                raise SystemExit("the GPU count of the CPU baseline's sample differs from the oracle")
            cpu = {"value": float(c.sum()) / dt, "unit": "binned triplets/s", "cores": threads, "kind": "port",
                   "nproc": os.cpu_count(), "cpu_model": cpu_model(),
                   "sample": f"2048 random RRR primaries against all 2.5M randoms, {int(c.sum())} triplets "
                             f"in {dt:.2f} s", "gpu_counts_bit_exact_on_sample": True}
    os.environ["NC_RPPI_STREAM"] = "0"
    try:
        with Context(0) as ctx:
            hist_d, cnt_d, wall_d, ms_d, _ = run(ctx, steps=1, warmup=1)
    finally:
        del os.environ["NC_RPPI_STREAM"]
    if not np.array_equal(cnt_s[3], cnt_d[3]):
        raise SystemExit("RRR: key stream and direct atomics differ")
    err = {}
    for i, n in enumerate(names[:3]):
        a, b = hist_s[i].ravel(), hist_d[i].ravel()
        nz = b != 0
        err[n] = float(np.max(np.abs(a[nz] - b[nz]) / np.abs(b[nz]), initial=0.0))
    out = {
        "metric": "rppi binned triplets/s on 1 MI355X (survey, key stream vs direct atomics)",
        "value": trip / wall_s,
        "unit": "binned triplets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded survey cap)",
        "config": {"workload": "rppi survey: 500k weighted data + 2.5M randoms in the cap/shell, rp edges "
                               "0..40 x20, 20 unit pi bins, DDD+DDR+DRR+RRR in one batched launch",
                   "binned_triplets": trip, "binned_triplets_per_term": dict(zip(names, trip_terms))},
        "error": {"unweighted": "RRR bit-exact (key stream vs direct atomics)",
                  "weighted_rel_max_vs_direct": err},
        "default_output": "direct atomics (surveys; NC_RPPI_STREAM=1 selects the key stream)",
        "direct_atomics": {"wall_ms": wall_d * 1e3, "count_ms": ms_d, "triplets_per_s": trip / wall_d},
        "count_ms": ms_s,
        "tuning": tune_s,
        "speedup_vs_direct": wall_d / wall_s,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def run_rppi(args):
    """rppi binning (the reference's binType 1) on the 1M-point clustered box:
    rp edges 0..30 in 15 bins, 20 unit pi bins, DDD.  The key stream (default)
    against the reference's direct global atomics (NC_RPPI_STREAM=0, a second
    context), counts compared bit for bit; the CPU oracle on a sample."""
    from cudancorr_amd import Context
    from cudancorr_amd.gen import SEED, edges_linear, neyman_scott_box

    p = neyman_scott_box(10000, 100, 5.0, 1000.0, seed=SEED + 1000)
    edges, npi, box = edges_linear(0.0, 30.0, 15), 20, np.array([1000.0] * 3)
    hip = _hip()

    def timed(ctx):
        c = ctx.catalog(p)
        for _ in range(max(1, args.warmup)):
            ctx.count(c, c, c, edges, box, npi=npi)
        device_sync(hip)
        ms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            _, cnt = ctx.count(c, c, c, edges, box, npi=npi)
            ms.append(ctx.timing().ms_count)
        device_sync(hip)
        wall = (time.perf_counter() - t0) / args.steps
        layout = ctx.timing().hist_layout
        c.close()
        return cnt, wall, float(np.mean(ms)), layout

    with Context(0) as ctx:
        cnt_s, wall_s, ms_s, lay_s = timed(ctx)
    os.environ["NC_RPPI_STREAM"] = "0"
    try:
        with Context(0) as ctx:
            cnt_d, wall_d, ms_d, lay_d = timed(ctx)
    finally:
        del os.environ["NC_RPPI_STREAM"]
    if not np.array_equal(cnt_s, cnt_d):
        raise SystemExit("key stream and direct atomics differ")
    trip = int(cnt_s.sum())
    log(f"rppi key stream (layout {lay_s}): {trip} triplets, count {ms_s:.1f} ms; "
        f"direct atomics (layout {lay_d}): {ms_d:.1f} ms")
    cpu = None
    if not args.no_cpu:
        from oracle import oracle

        threads = min(16, len(os.sched_getaffinity(0)))
        rng = np.random.default_rng(1)
        sample = p[np.sort(rng.choice(p.shape[0], size=131072, replace=False))]
        t0 = time.perf_counter()
        c, _, _ = oracle.count_rppi(sample, p, p, edges, npi, box, nthreads=threads)
        dt = time.perf_counter() - t0
        cpu = {"value": float(c.sum()) / dt, "unit": "binned triplets/s", "cores": threads, "kind": "port",
               "sample": f"131072 random primaries against the full box, {int(c.sum())} triplets in {dt:.2f} s"}
    out = {
        "metric": "rppi binned triplets/s on 1 MI355X (key stream vs direct atomics)",
        "value": trip / wall_s,
        "unit": "binned triplets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Neyman-Scott box)",
        "config": {"workload": "rppi, 1M-point clustered box (10,000 parents x 100, sigma=5, L=1000), "
                               "rp edges 0..30 x15, 20 unit pi bins, DDD", "binned_triplets": trip},
        "error": {"unweighted": "bit-exact (key stream vs direct atomics)"},
        "direct_atomics": {"wall_ms": wall_d * 1e3, "count_ms": ms_d, "triplets_per_s": trip / wall_d},
        "count_ms": ms_s,
        "speedup_vs_direct": wall_d / wall_s,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["boxes", "survey", "survey-small", "rppi-box", "rppi-survey"],
                    default="boxes",
                    help="boxes: 64 clustered periodic boxes (default); survey: 4M data + 20M randoms, "
                         "four terms; survey-small: 500k data + 2.5M randoms; rppi-box: rppi binning of "
                         "the 1M-point box; rppi-survey: rppi binning of the survey-small catalogues")
    ap.add_argument("--slices", type=int, default=None,
                    help="survey: primaries split into this many launches (default 16 for survey, about 6 s "
                         "each; 1 for survey-small, whose whole count takes about 60 ms)")
    ap.add_argument("--boxes", type=int, default=64)
    ap.add_argument("--no-plain", action="store_true", help="skip the 64-call plain leg")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-config4", action="store_true",
                    help="boxes: skip the block on a subset of the large weighted survey")
    ap.add_argument("--config4-slices", default="8,24",
                    help="boxes: primary slices (of 32) of the large weighted survey block")
    ap.add_argument("--atomic-rate", action="store_true",
                    help="only measure scattered global atomic adds per second and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device copy of BYTES (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.atomic_rate:
        atomic_rate()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    if args.workload == "rppi-box":
        run_rppi(args)
        return
    if args.workload == "rppi-survey":
        run_rppi_survey(args)
        return
    if args.workload != "boxes":
        run_survey(args)
        return

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
    count_ms_med = float(np.median([b for _, b in ker_ms]))
    log(f"mixed batched: {trip} triplets, wall {wall_m * 1e3:.1f} ms, device {dev_ms_m:.1f} ms "
        f"(count kernel {count_ms_m:.1f} ms)")

    # ---- plain, one call per box
    plain = None
    tune_p = None
    if not args.no_plain:
        # tuned on a launch of the kind this leg runs: one box, the largest
        big = dev[int(np.argmax([c.shape[0] for c in cats]))]
        tune_p = ctx.autotune([(big, big, big)], edges, box, path=NC_PATH_PLAIN)
        log(f"autotune plain (largest box): {tune_p}")
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
    big_dev = dev[int(np.argmax([c.shape[0] for c in cats]))]
    cpu = None if args.no_cpu else cpu_baseline(
        cats, edges, box, gpu_count=lambda smp, big: ctx.count(smp, big_dev, big_dev, edges, box)[1])
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
        "error": {"unweighted": "bit-exact (plain vs mixed, every run)" if plain else "not compared (no plain leg)"},
        "plain": plain,
        "tuning": {"mixed": tune_m, "plain": tune_p},
        "speedup_vs_plain": (plain["wall_ms"] / (wall_m * 1e3)) if plain else None,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": from_profiles("traffic_bytes_per_launch"),
                     "valu_busy_pct": from_profiles("valu_busy_pct"),
                     "kernel": "k_triplets (mixed, batched)", "kernel_ms": count_ms_m,
                     "kernel_ms_median": count_ms_med},
        "cpu_baseline": cpu,
    }
    for c in dev:
        c.close()
    ctx.close()
    if not args.no_config4:
        out["config4"] = config4_subset(args, hip)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
