"""nc_count_batched: one persistent launch over many terms, captured as HIP
graphs after the first call and replayed while the same terms come back.

Every replay is checked against the oracle (and the eager per-term counts),
with host work and device synchronisations between replays; the four survey
terms DDD/DDR/DRR/RRR in one batch are checked against the oracle, also on
clustered data at 25 bins up to rmax 50.
"""
import ctypes

import numpy as np
import pytest

from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN
from cudancorr_amd.gen import edges_linear, neyman_scott_box, prep_cat_survey, survey_cap
from oracle import oracle

pytestmark = pytest.mark.gpu


def _sync():
    hip = ctypes.CDLL("libamdhip64.so")
    assert hip.hipDeviceSynchronize() == 0


def test_batched_boxes_replays(ctx):
    print("batched boxes", flush=True)
    L = 300.0
    edges = edges_linear(0.0, 30.0, 15)
    box = [L, L, L]
    pts = [neyman_scott_box(n, 40, 4.0, L, seed=100 + n) for n in (150, 400, 900)]
    cats = [ctx.catalog(p) for p in pts]
    want = np.stack([oracle.count(p, p, p, edges, box)[0] for p in pts])  # the oracle, per box
    eager = np.stack([ctx.count(c, c, c, edges, box, path=NC_PATH_PLAIN)[1] for c in cats])
    assert np.array_equal(eager, want)
    terms = [(c, c, c) for c in cats]
    for rep in range(4):
        hist, cnt = ctx.count_batched(terms, edges, box, path=NC_PATH_MIXED)
        assert np.array_equal(cnt, want), f"call {rep}"
        assert np.array_equal(hist, want.astype(np.float64))
        _ = np.sort(np.random.default_rng(rep).random(100000))  # host work between replays
        _sync()
    # a different term list (reordered, with a repeat) gets its own plan
    terms2 = [terms[2], terms[0], terms[2]]
    for rep in range(2):
        _, cnt2 = ctx.count_batched(terms2, edges, box, path=NC_PATH_MIXED)
        assert np.array_equal(cnt2, want[[2, 0, 2]])
    # plain batched equals mixed batched
    _, cntp = ctx.count_batched(terms, edges, box, path=NC_PATH_PLAIN)
    assert np.array_equal(cntp, want)
    # destroying a catalogue of the cached plan drops the plan
    cats[1].close()
    terms3 = [terms[0], terms[2]]
    _, cnt3 = ctx.count_batched(terms3, edges, box, path=NC_PATH_MIXED)
    assert np.array_equal(cnt3, want[[0, 2]])
    for c in cats:
        c.close()


def test_batched_survey_four_terms(ctx):
    print("batched survey", flush=True)
    d, r = survey_cap(3000, 4500, seed=21)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.12
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 40.0, 20)
    D, R = ctx.catalog(d), ctx.catalog(r)
    terms = [(D, D, D), (D, D, R), (D, R, R), (R, R, R)]
    arrs = {id(D): d, id(R): r}
    for rep in range(3):
        hist, cnt = ctx.count_batched(terms, edges, box, survey=True, path=NC_PATH_MIXED)
        for i, (a, b, c) in enumerate(terms):
            wc, ws = oracle.count(arrs[id(a)], arrs[id(b)], arrs[id(c)], edges, box, survey=True)
            if i == 3:
                assert np.array_equal(cnt[i], wc)
            else:
                nz = ws != 0
                assert np.array_equal(hist[i][~nz], ws[~nz])
                assert np.max(np.abs(hist[i][nz] - ws[nz]) / ws[nz]) <= 1e-9
        _sync()
    D.close()
    R.close()


def test_autotune_keeps_results(ctx, tmp_path, monkeypatch):
    print("autotune", flush=True)
    monkeypatch.setenv("NC_TUNE_CACHE", str(tmp_path / "tune.txt"))
    L = 250.0
    edges = edges_linear(0.0, 25.0, 10)
    box = [L, L, L]
    cats = [ctx.catalog(neyman_scott_box(n, 30, 4.0, L, seed=7 + n)) for n in (200, 500)]
    terms = [(c, c, c) for c in cats]
    before = ctx.get_tuning()
    _, want = ctx.count_batched(terms, edges, box, path=NC_PATH_MIXED)
    best = ctx.autotune(terms, edges, box, path=NC_PATH_MIXED)
    assert best["waves_per_group"] in (4, 8, 16) and best["list_chunk"] in (64, 128, 256) and best["from_cache"] == 0
    _, got = ctx.count_batched(terms, edges, box, path=NC_PATH_MIXED)
    assert np.array_equal(got, want)
    again = ctx.autotune(terms, edges, box, path=NC_PATH_MIXED)
    assert again["from_cache"] == 1 and again["waves_per_group"] == best["waves_per_group"]
    ctx.set_tuning(*before)
    for c in cats:
        c.close()


def test_catalogue_outlives_its_context():
    """A catalogue destroyed after its context (as at interpreter exit) only
    frees its own buffers."""
    from cudancorr_amd import Context

    print("catalogue after context", flush=True)
    c = Context(0)
    p = neyman_scott_box(50, 20, 2.0, 100.0, seed=3)
    a, b = c.catalog(p), c.catalog(p)
    b.close()
    c.close()
    a.close()
    with Context(0) as c2:  # the device still works
        q = c2.catalog(p)
        _, cnt = c2.count(q, q, q, edges_linear(0.0, 10.0, 5), [100.0] * 3)
        assert cnt.sum() > 0
        q.close()


def test_wrapped_device_catalogue_changes_under_replay(ctx):
    """A catalogue wrapping device memory (nc_catalog_wrap_device): the first
    batched count plans the dense layout from the exact longest list (chunk
    16384, so the list-bound pass runs) and captures a graph; the points are
    then overwritten in place with clustered ones whose lists are far longer.
    The replay must notice the overflow and recount eagerly: both counts equal
    the oracle's."""
    from cudancorr_amd.gen import uniform_box

    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    H2D = 1  # hipMemcpyHostToDevice

    print("batched wrapped device catalogue", flush=True)
    L = 216.0
    edges = edges_linear(0.0, 24.0, 12)
    box = np.array([L] * 3)
    a = uniform_box(20000, L, seed=41)
    b = neyman_scott_box(200, 100, 2.0, L, seed=43)
    want_a, _ = oracle.count(a, a, a, edges, box)
    want_b, _ = oracle.count(b, b, b, edges, box)
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    assert a.shape == b.shape
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), a.nbytes) == 0
    before = ctx.get_tuning()
    try:
        assert hip.hipMemcpy(dev, a.ctypes.data, a.nbytes, H2D) == 0
        ctx.set_tuning(-1, 16384, -1, 0)
        C = ctx.catalog_device(dev.value, a.shape[0])
        _, got = ctx.count_batched([(C, C, C)], edges, box, path=NC_PATH_MIXED)
        assert ctx.timing().hist_layout == 2
        assert np.array_equal(got[0], want_a)
        assert hip.hipMemcpy(dev, b.ctypes.data, b.nbytes, H2D) == 0
        _, got = ctx.count_batched([(C, C, C)], edges, box, path=NC_PATH_MIXED)
        assert np.array_equal(got[0], want_b)
        C.close()
    finally:
        ctx.set_tuning(*before)
        hip.hipFree(dev)


def test_wrapped_device_catalogue_weights_change(ctx):
    """The weights of a wrapped device catalogue change in place between
    counts: unit -> lognormal -> unit.  The unit-weight status decides the
    histogram layout (dense integer counts), the batch split and whether
    counts are returned, so every call must follow the current contents:
    weighted sums within 1e-9 of the oracle's compensated FP64 sums while
    the weights are not 1, exact counts when they are, for nc_count and for
    nc_count_batched (whose captured plan must not be replayed across the
    change)."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    H2D = 1

    print("wrapped device catalogue, weights changed in place", flush=True)
    L = 120.0
    edges = edges_linear(0.0, 12.0, 6)
    box = np.array([L] * 3)
    unit = np.ascontiguousarray(neyman_scott_box(60, 50, 2.0, L, seed=51), dtype=np.float64)
    wtd = unit.copy()
    rng = np.random.default_rng(52)
    wtd[:, 3] = rng.lognormal(0.0, 0.3, unit.shape[0]).astype(np.float32)
    want_c, _ = oracle.count(unit, unit, unit, edges, box)
    _, want_w = oracle.count(wtd, wtd, wtd, edges, box)
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), unit.nbytes) == 0
    try:
        C = ctx.catalog_device(dev.value, unit.shape[0])
        for step, pts in enumerate((unit, wtd, unit, wtd)):
            assert hip.hipMemcpy(dev, pts.ctypes.data, pts.nbytes, H2D) == 0
            weighted = pts is wtd
            for path in (NC_PATH_PLAIN, NC_PATH_MIXED):
                print("step", step, "weighted", weighted, "path", path, flush=True)
                for rep in range(2):  # the second batched call may replay a plan
                    h, c = ctx.count_batched([(C, C, C)], edges, box, path=path)
                    h1, c1 = ctx.count(C, C, C, edges, box, path=path)
                    for hh, cc in ((h[0], c[0]), (h1, c1)):
                        if weighted:
                            nz = want_w != 0
                            assert np.array_equal(hh[~nz], want_w[~nz])
                            assert np.max(np.abs(hh[nz] / want_w[nz] - 1.0)) <= 1e-9
                            assert not cc.any()  # no integer counts for weighted terms
                        else:
                            assert np.array_equal(cc, want_c)
                            assert np.array_equal(hh, want_c.astype(np.float64))
        C.close()
    finally:
        hip.hipFree(dev)


def test_batched_survey_four_terms_clustered_25_bins(ctx):
    """The large survey's shape at a size the oracle covers: clustered data
    (Neyman-Scott, sigma 4) with lognormal weights and uniform randoms in a
    shrunk cap, rmax 50 in 25 bins, the four terms in one batch on both paths,
    replayed: counts bit-exact, weighted sums within 1e-9 of the oracle."""
    print("batched survey clustered 25 bins", flush=True)
    d, r = survey_cap(2500, 6000, seed=29, clustered=True, sigma=4.0)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.15
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 50.0, 25)
    want = [oracle.count(a, b, c, edges, box, survey=True) for a, b, c in ((d, d, d), (d, d, r), (d, r, r), (r, r, r))]
    assert want[3][0].sum() > 10000
    D, R = ctx.catalog(d), ctx.catalog(r)
    terms = [(D, D, D), (D, D, R), (D, R, R), (R, R, R)]
    for path in (NC_PATH_PLAIN, NC_PATH_MIXED):
        for rep in range(2):
            hist, cnt = ctx.count_batched(terms, edges, box, survey=True, path=path)
            assert np.array_equal(cnt[3], want[3][0]), f"path {path} replay {rep}"
            for i in range(3):
                ws = want[i][1]
                nz = ws != 0
                assert np.array_equal(hist[i][~nz], ws[~nz])
                assert np.max(np.abs(hist[i][nz] / ws[nz] - 1.0)) <= 1e-9
    D.close()
    R.close()
