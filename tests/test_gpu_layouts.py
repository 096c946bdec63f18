"""Workgroup histogram layouts of the triplet kernel, each against the oracle.

Unit-weight launches count into dense nb^3 32-bit words (layout 2), with work
items cut so that none can add 2^31 pairs before a merge; weighted terms count
into the sorted slots through the slot table (layout 1).  Every layout must give the
same, bit-exact counts; the dense counts must survive being merged after every
chunk (NC_DENSE_FLUSH_PAIRS set low in a fresh context).
"""
import numpy as np
import pytest

from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN, Context
from cudancorr_amd.gen import edges_linear, neyman_scott_box, prep_cat_survey, survey_cap
from oracle import oracle

pytestmark = pytest.mark.gpu

PATHS = [NC_PATH_PLAIN, NC_PATH_MIXED]


def _log(msg):
    print(msg, flush=True)


def _box_case():
    p = neyman_scott_box(400, 40, 3.0, 220.0, seed=31)
    edges = edges_linear(0.0, 24.0, 12)
    box = np.array([220.0] * 3)
    want, _ = oracle.count(p, p, p, edges, box)
    return p, edges, box, want


@pytest.mark.parametrize("path", PATHS)
def test_dense_and_slot_table_layouts(ctx, path):
    _log(f"layouts path={path}")
    p, edges, box, want = _box_case()
    assert want.sum() > 100_000
    before = ctx.get_tuning()
    try:
        _, got = ctx.count(p, p, p, edges, box, path=path)
        assert ctx.timing().hist_layout == 2
        assert np.array_equal(got, want)
        # chunks of 2^20 primaries could add more than 2^31 pairs between
        # merges: the work items are cut instead, and the dense layout stays
        ctx.set_tuning(-1, 1 << 20, -1, 0)
        _, got = ctx.count(p, p, p, edges, box, path=path)
        assert ctx.timing().hist_layout == 2
        assert np.array_equal(got, want)
    finally:
        ctx.set_tuning(*before)


def test_weighted_terms_use_sorted_slots(ctx):
    _log("layouts weighted")
    d, r = survey_cap(1500, 2500, seed=8)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.1
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 40.0, 16)
    _, ws = oracle.count(d, d, r, edges, box, survey=True)
    hist, _ = ctx.count(d, d, r, edges, box, survey=True, path=NC_PATH_MIXED)
    assert ctx.timing().hist_layout == 1
    nz = ws != 0
    assert np.array_equal(hist[~nz], ws[~nz])
    assert np.max(np.abs(hist[nz] - ws[nz]) / ws[nz]) <= 1e-9


@pytest.mark.parametrize("path", PATHS)
def test_dense_counts_merged_every_chunk(monkeypatch, path):
    _log(f"layouts dense flush path={path}")
    p, edges, box, want = _box_case()
    monkeypatch.setenv("NC_DENSE_FLUSH_PAIRS", "1")
    with Context(0) as c2:
        c2.set_tuning(-1, 4, -1, 0)  # many small chunks, each followed by a merge
        _, got = c2.count(p, p, p, edges, box, path=path)
        assert c2.timing().hist_layout == 2
        assert np.array_equal(got, want)
        # batched (graph capture and replays) with the same merges
        cat = c2.catalog(p)
        for _ in range(3):
            _, cnt = c2.count_batched([(cat, cat, cat), (cat, cat, cat)], edges, box, path=path)
            assert np.array_equal(cnt[0], want) and np.array_equal(cnt[1], want)
        cat.close()


HALF_CU_LDS = 81920 - 256


@pytest.mark.parametrize("path", PATHS)
def test_wide_binning_uses_the_whole_cu_lds(ctx, path):
    """36 bins over rmax = 50: the weighted sorted-bin histogram alone
    (C(38,3) slots, 67.5 KB) leaves too little of half the CU's LDS for the
    waves' staging, so the launch runs one workgroup per CU."""
    _log(f"layouts wide binning path={path}")
    d, r = survey_cap(1500, 2500, seed=21)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.1
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 50.0, 36)
    cats = {"D": ctx.catalog(d), "R": ctx.catalog(r)}
    want = {}
    for term in ("DDR", "RRR"):
        a, b, c = ({"D": d, "R": r}[ch] for ch in term)
        want[term] = oracle.count(a, b, c, edges, box, survey=True)
    terms = [tuple(cats[ch] for ch in t) for t in ("DDR", "RRR")]
    hist, cnt = ctx.count_batched(terms, edges, box, survey=True, path=path)
    if path == NC_PATH_MIXED:
        assert ctx.timing().lds_bytes > HALF_CU_LDS
    assert want["RRR"][0].sum() > 100_000
    assert np.array_equal(cnt[1], want["RRR"][0])
    ws = want["DDR"][1]
    nz = ws != 0
    assert np.array_equal(hist[0][~nz], ws[~nz])
    assert np.max(np.abs(hist[0][nz] - ws[nz]) / ws[nz]) <= 1e-9
    # the unit-weight term alone: dense counts
    _, got = ctx.count(cats["R"], cats["R"], cats["R"], edges, box, survey=True, path=path)
    assert np.array_equal(got, want["RRR"][0])
    for c in cats.values():
        c.close()


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("waves", [4, 16])
def test_workgroup_sizes(ctx, path, waves):
    """4 and 16 waves per workgroup (1024 threads) give the same counts."""
    _log(f"layouts waves={waves} path={path}")
    p, edges, box, want = _box_case()
    before = ctx.get_tuning()
    try:
        ctx.set_tuning(waves, 16, 0, 128)
        _, got = ctx.count(p, p, p, edges, box, path=path)
        assert np.array_equal(got, want)
        cat = ctx.catalog(p)
        _, cnt = ctx.count_batched([(cat, cat, cat)], edges, box, path=path)
        assert np.array_equal(cnt[0], want)
        cat.close()
    finally:
        ctx.set_tuning(*before)


def test_dense_admitted_by_exact_list_bound(ctx):
    """chunk * (27-cell candidate bound)^2 >= 2^31 but chunk * (longest list)^2
    < 2^31: the exact-length pass (k_list_bound) admits the dense layout, and
    the counts stay bit-exact.  20,000 uniform points in 9^3 cells: about 740
    candidates but about 115 list entries per primary."""
    _log("layouts exact list bound")
    from cudancorr_amd.gen import uniform_box
    L = 216.0
    p = uniform_box(20000, L, seed=5)
    edges = edges_linear(0.0, 24.0, 12)
    box = np.array([L] * 3)
    want, _ = oracle.count(p, p, p, edges, box)
    before = ctx.get_tuning()
    try:
        ctx.set_tuning(-1, 16384, -1, 0)
        for path in PATHS:
            _, got = ctx.count(p, p, p, edges, box, path=path)
            assert ctx.timing().hist_layout == 2
            assert np.array_equal(got, want)
    finally:
        ctx.set_tuning(*before)


def test_batched_unit_terms_split_to_dense(ctx):
    """A weighted term beside a unit-weight one at 25 bins: the unit-weight
    term runs as its own launch in the dense layout (whole-CU LDS), the
    weighted one in the sorted slots; both match the oracle."""
    _log("layouts batched split")
    d, r = survey_cap(1200, 2000, seed=17)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.12
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 50.0, 25)
    _, wsw = oracle.count(d, d, r, edges, box, survey=True)
    cr, _ = oracle.count(r, r, r, edges, box, survey=True)
    D, R = ctx.catalog(d), ctx.catalog(r)
    hist, cnt = ctx.count_batched([(D, D, R), (R, R, R)], edges, box, survey=True, path=NC_PATH_MIXED)
    t = ctx.timing()
    assert t.hist_layout == 2 and t.path == NC_PATH_MIXED
    assert cr.sum() > 10_000
    assert np.array_equal(cnt[1], cr)
    nz = wsw != 0
    assert np.array_equal(hist[0][~nz], wsw[~nz])
    assert np.max(np.abs(hist[0][nz] - wsw[nz]) / wsw[nz]) <= 1e-9
    D.close()
    R.close()


def test_autotune_split_batch(ctx):
    """The autotuner times a mixed-weight batch as it runs (two launches) and
    the tuned shape counts it bit-exact / within tolerance."""
    _log("layouts autotune split batch")
    d, r = survey_cap(1200, 2000, seed=23)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.12
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 50.0, 25)
    _, wsw = oracle.count(d, d, r, edges, box, survey=True)
    cr, _ = oracle.count(r, r, r, edges, box, survey=True)
    before = ctx.get_tuning()
    D, R = ctx.catalog(d), ctx.catalog(r)
    try:
        best = ctx.autotune([(D, D, R), (R, R, R)], edges, box, survey=True, path=NC_PATH_MIXED)
        assert best["waves_per_group"] in (4, 8, 16) and best["ms"] > 0
        hist, cnt = ctx.count_batched([(D, D, R), (R, R, R)], edges, box, survey=True, path=NC_PATH_MIXED)
        assert np.array_equal(cnt[1], cr)
        nz = wsw != 0
        assert np.array_equal(hist[0][~nz], wsw[~nz])
        assert np.max(np.abs(hist[0][nz] - wsw[nz]) / wsw[nz]) <= 1e-9
    finally:
        D.close()
        R.close()
        ctx.set_tuning(*before)


def test_batch_split_switched_off(monkeypatch):
    """NC_BATCH_SPLIT=0: the mixed-weight batch stays one launch in a sorted-slot
    layout, with the same results."""
    _log("layouts batch split off")
    d, r = survey_cap(1200, 2000, seed=17)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.12
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 50.0, 25)
    _, wsw = oracle.count(d, d, r, edges, box, survey=True)
    cr, _ = oracle.count(r, r, r, edges, box, survey=True)
    monkeypatch.setenv("NC_BATCH_SPLIT", "0")
    with Context(0) as c2:
        D, R = c2.catalog(d), c2.catalog(r)
        hist, cnt = c2.count_batched([(D, D, R), (R, R, R)], edges, box, survey=True, path=NC_PATH_MIXED)
        assert c2.timing().hist_layout in (0, 1)
        assert np.array_equal(cnt[1], cr)
        nz = wsw != 0
        assert np.array_equal(hist[0][~nz], wsw[~nz])
        assert np.max(np.abs(hist[0][nz] - wsw[nz]) / wsw[nz]) <= 1e-9
        D.close()
        R.close()
