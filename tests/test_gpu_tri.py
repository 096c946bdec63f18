"""The six-fold path for terms over one catalogue (c1 == c2 == c3): each
unordered triple away from the last cells is evaluated once, from its
lowest-index point, and counted six times; primaries in a last cell, and
pairs with a last-cell point, keep their own ordered counts.  Every result is
checked against the oracle (the reference's ordered loop restated), on both
arithmetic paths, and against the same library with NC_TRI=0."""
import numpy as np
import pytest

from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN
from cudancorr_amd.gen import edges_linear, neyman_scott_box, prep_cat_survey, survey_cap, uniform_box
from oracle import oracle

pytestmark = pytest.mark.gpu

PATHS = (NC_PATH_PLAIN, NC_PATH_MIXED)


def _log(msg):
    print(msg, flush=True)


@pytest.fixture
def ctx_notri(monkeypatch):
    from cudancorr_amd import Context

    monkeypatch.setenv("NC_TRI", "0")
    c = Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("L", [250.0, 110.0])
def test_partial_last_cell_box(ctx, ctx_notri, L):
    """L = 250 and 110 with rmax 20: 13 and 6 cells, the last one 10 wide, so
    the reference misses wrap pairs and some triples are not seen in all six
    orders.  Clustered points crowd the wrap corner."""
    _log(f"tri partial last cell L={L}")
    rng = np.random.default_rng(int(L))
    p = neyman_scott_box(120, 40, 4.0, L, seed=int(L) + 1)
    q = rng.uniform(-25.0, 25.0, (1500, 3)) % L  # the corner across the wrap
    p = np.concatenate([p, np.c_[q, np.ones(len(q))]])
    p[:, :3] = p[:, :3].astype(np.float32)
    edges = edges_linear(2.0, 20.0, 10)
    box = np.array([L, L, L])
    want, _ = oracle.count(p, p, p, edges, box)
    assert want.sum() > 0 and np.any(want % 6 != 0)  # the quirk is present
    for path in PATHS:
        _, got = ctx.count(p, p, p, edges, box, path=path)
        assert ctx.timing().tri_terms == 1
        assert np.array_equal(got, want), f"path {path}"
        _, got0 = ctx_notri.count(p, p, p, edges, box, path=path)
        assert ctx_notri.timing().tri_terms == 0
        assert np.array_equal(got0, want)


@pytest.mark.parametrize("L", [15.0, 30.0, 45.0, 60.0])
def test_small_grids(ctx, L):
    """1, 2, 3 and 3 cells per side (rmax 20): one cell is all last cells; two
    cells list the other cell twice; the six-fold part only runs on cells
    before the last."""
    _log(f"tri small grid L={L}")
    p = uniform_box(300, L, seed=int(L) + 7)
    p[:, :3] = p[:, :3].astype(np.float32)
    edges = edges_linear(0.0, 20.0, 8)
    box = np.array([L, L, L])
    want, _ = oracle.count(p, p, p, edges, box)
    for path in PATHS:
        _, got = ctx.count(p, p, p, edges, box, path=path)
        assert np.array_equal(got, want), f"path {path}"


def test_commensurate_box_with_wrap_images(ctx):
    """L = 4 rmax: wraps join cell 3 (last) and cell 0; FP64 coordinates whose
    image x + L is not exact; counts must still follow the reference."""
    _log("tri commensurate")
    L = 80.0
    p = uniform_box(6000, L, seed=17)
    edges = edges_linear(1.0, 20.0, 9)
    box = np.array([L, L, L])
    want, _ = oracle.count(p, p, p, edges, box)
    for path in PATHS:
        _, got = ctx.count(p, p, p, edges, box, path=path)
        assert ctx.timing().tri_terms == 1
        assert np.array_equal(got, want), f"path {path}"


def test_inexact_gridding_falls_back(ctx):
    """Coordinates just below zero are put in cell 0 by the reference's
    truncation (not the floor): the six-fold path is not taken, and the count
    still equals the oracle's."""
    _log("tri inexact gridding")
    L = 120.0
    p = uniform_box(4000, L, seed=21)
    p[:50, 0] = -1e-9
    edges = edges_linear(0.0, 12.0, 6)
    box = np.array([L, L, L])
    want, _ = oracle.count(p, p, p, edges, box)
    for path in PATHS:
        _, got = ctx.count(p, p, p, edges, box, path=path)
        assert ctx.timing().tri_terms == 0
        assert np.array_equal(got, want), f"path {path}"


def test_survey_terms_weighted(ctx):
    """DDD (weighted) and RRR take the six-fold path, DDR the pair-swap path;
    DRR neither.  Counts exact, weighted sums within 1e-9."""
    _log("tri survey terms")
    d, r = survey_cap(3000, 5000, seed=23)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.12
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 40.0, 12)
    for path in PATHS:
        for name, (a, b, c) in {"DDD": (d, d, d), "RRR": (r, r, r), "DDR": (d, d, r)}.items():
            h, cnt = ctx.count(a, b, c, edges, box, survey=True, path=path)
            assert ctx.timing().tri_terms == (1 if name != "DDR" else 0)
            assert ctx.timing().swap_terms == (1 if name == "DDR" else 0)
            wc, ws = oracle.count(a, b, c, edges, box, survey=True)
            if name == "RRR":
                assert np.array_equal(cnt, wc)
            else:
                nz = ws != 0
                assert np.array_equal(h[~nz], ws[~nz])
                assert np.max(np.abs(h[nz] / ws[nz] - 1.0)) <= 1e-9, name


def test_mixed_weight_launch_six_fold_for_weighted_terms(monkeypatch):
    """NC_BATCH_SPLIT=0: DDD (weighted) and RRR (unit weights) in one launch.
    Its kernel variant carries the six-fold path for weighted terms only, so
    DDD takes it and RRR the two-fold path; both equal the oracle."""
    from cudancorr_amd import Context

    _log("tri mixed-weight launch")
    monkeypatch.setenv("NC_BATCH_SPLIT", "0")
    d, r = survey_cap(2500, 4000, seed=41)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.12
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 40.0, 10)
    wc_d, ws_d = oracle.count(d, d, d, edges, box, survey=True)
    wc_r, _ = oracle.count(r, r, r, edges, box, survey=True)
    with Context(0) as c:
        D, R = c.catalog(d), c.catalog(r)
        try:
            for path in PATHS:
                h, cnt = c.count_batched([(D, D, D), (R, R, R)], edges, box, survey=True, path=path)
                t = c.timing()
                assert t.hist_layout in (0, 1) and t.tri_terms == 1, path  # one sorted-slot launch
                assert np.array_equal(cnt[1], wc_r), path
                nz = ws_d != 0
                assert np.array_equal(h[0][~nz], ws_d[~nz])
                assert np.max(np.abs(h[0][nz] / ws_d[nz] - 1.0)) <= 1e-9, path
        finally:
            D.close()
            R.close()


@pytest.mark.parametrize("unit", [False, True])
def test_survey_pair_swap_terms(ctx, ctx_notri, unit):
    """Survey terms with c1 == c2 != c3 (DDR): the triplets (p1, p2, p3) and
    (p2, p1, p3) are evaluated once, from the lower index, and counted twice
    (weighted: twice the product (w1 w2) w3).  Weighted sums within 1e-9,
    counts bit-exact, against the oracle and against NC_TRI=0; a clustered
    catalogue so that many pairs of D lie in one cell."""
    _log(f"tri pair swap unit={unit}")
    d, r = survey_cap(2500, 4000, seed=43, clustered=True, sigma=3.0, n_children=50)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.12
    if unit:
        d[:, 3] = 1.0
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 40.0, 10)
    wc, ws = oracle.count(d, d, r, edges, box, survey=True)
    assert wc.sum() > 0 and np.all(wc % 2 == 0)
    for path in PATHS:
        h, cnt = ctx.count(d, d, r, edges, box, survey=True, path=path)
        t = ctx.timing()
        assert t.swap_terms == 1 and t.tri_terms == 0, path
        h0, cnt0 = ctx_notri.count(d, d, r, edges, box, survey=True, path=path)
        assert ctx_notri.timing().swap_terms == 0
        if unit:
            assert np.array_equal(cnt, wc) and np.array_equal(cnt0, wc), path
        else:
            nz = ws != 0
            for got in (h, h0):
                assert np.array_equal(got[~nz], ws[~nz])
                assert np.max(np.abs(got[nz] / ws[nz] - 1.0)) <= 1e-9, path


def test_pair_swap_off_in_boxes(ctx):
    """A box term with c1 == c2 != c3 keeps the ordered path (a wrap or a last
    cell can break the p1/p2 symmetry) and equals the oracle."""
    _log("tri pair swap box")
    L = 110.0
    a = neyman_scott_box(60, 30, 3.0, L, seed=47)
    b = uniform_box(2500, L, seed=48)
    edges = edges_linear(1.0, 20.0, 8)
    box = np.array([L, L, L])
    want, _ = oracle.count(a, a, b, edges, box)
    for path in PATHS:
        _, got = ctx.count(a, a, b, edges, box, path=path)
        assert ctx.timing().swap_terms == 0
        assert np.array_equal(got, want), path


def test_dense_merges_under_six_fold_counts(monkeypatch):
    """Dense 32-bit counts with increments of 6, merged after every work item
    (NC_DENSE_FLUSH_PAIRS=1): the same counts."""
    from cudancorr_amd import Context

    _log("tri dense merges")
    monkeypatch.setenv("NC_DENSE_FLUSH_PAIRS", "1")
    L = 200.0
    p = neyman_scott_box(200, 50, 3.0, L, seed=29)
    edges = edges_linear(0.0, 18.0, 9)
    box = np.array([L, L, L])
    want, _ = oracle.count(p, p, p, edges, box)
    with Context(0) as c:
        _, got = c.count(p, p, p, edges, box, path=NC_PATH_MIXED)
        assert c.timing().hist_layout == 2 and c.timing().tri_terms == 1
        assert np.array_equal(got, want)


def test_batched_boxes_replayed(ctx):
    """Several boxes in one batched launch, replayed from the captured graph:
    each equals the oracle on every replay."""
    _log("tri batched replay")
    L = 150.0
    boxes = [neyman_scott_box(n, 30, 3.0, L, seed=30 + i) for i, n in enumerate((40, 90, 150))]
    edges = edges_linear(0.0, 15.0, 5)
    box = np.array([L, L, L])
    cats = [ctx.catalog(b) for b in boxes]
    want = [oracle.count(b, b, b, edges, box)[0] for b in boxes]
    for rep in range(3):
        _, got = ctx.count_batched([(c, c, c) for c in cats], edges, box, path=NC_PATH_MIXED)
        assert ctx.timing().tri_terms == 3
        for g, w in zip(got, want):
            assert np.array_equal(g, w), f"replay {rep}"
    for c in cats:
        c.close()


@pytest.mark.parametrize("nslices", [3, 7])
def test_primary_slices_add_up(ctx, nslices):
    """nc_set_primary_slice: the slices of every term's cell-sorted primaries,
    on the same catalogue handles, add up to the full count (exactly for
    counts, within 1e-12 for weighted sums), and the six-fold path stays on."""
    _log(f"tri primary slices {nslices}")
    d, r = survey_cap(2500, 4000, seed=37)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.12
    box, _ = prep_cat_survey(d, r)
    edges = edges_linear(0.0, 40.0, 10)
    D, R = ctx.catalog(d), ctx.catalog(r)
    terms = [(D, D, D), (D, D, R), (D, R, R), (R, R, R)]
    try:
        for path in PATHS:
            h_full, c_full = ctx.count_batched(terms, edges, box, survey=True, path=path)
            h_sum = np.zeros_like(h_full)
            c_sum = np.zeros_like(c_full)
            for k in range(nslices):
                ctx.set_primary_slice(k, nslices)
                h, c = ctx.count_batched(terms, edges, box, survey=True, path=path)
                assert ctx.timing().tri_terms == 2  # DDD and RRR (its own launch on both paths)
                assert ctx.timing().swap_terms == 1  # DDR
                h_sum += h
                c_sum += c
            ctx.set_primary_slice(0, 1)
            assert np.array_equal(c_sum[3], c_full[3])
            want_rrr, _ = oracle.count(r, r, r, edges, box, survey=True)
            assert np.array_equal(c_full[3], want_rrr)
            for t in range(3):
                nz = h_full[t] != 0
                assert np.array_equal(h_sum[t][~nz], h_full[t][~nz])
                assert np.max(np.abs(h_sum[t][nz] / h_full[t][nz] - 1.0)) <= 1e-12
    finally:
        ctx.set_primary_slice(0, 1)
        D.close()
        R.close()
