"""FP32 accumulation of weighted sums (nc_set_accumulation NC_ACCUM_FP32):
weight products formed in FP32 as float-float pairs, FP32 running sums in the
workgroup's LDS slots whose rounding errors (TwoSum on the value a returning
LDS atomic gives back) go to FP64 compensations, folded into FP64 totals at
every merge.  Checked against the oracle's compensated FP64 sums at the 1e-9
bar, and against the exact fixed-point accumulation at a much tighter bound on
inputs that pile many additions into few slots: an uncompensated FP32 sum (or
a compensation that missed the atomic's rounding) would be off by ~1e-7."""
import numpy as np
import pytest

from cudancorr_amd import NC_ACCUM_EXACT, NC_ACCUM_FP32, NC_ACCUM_FP64, NC_PATH_MIXED, NC_PATH_PLAIN, NcError
from cudancorr_amd.gen import edges_linear, neyman_scott_box, prep_cat_survey, survey_cap
from oracle import oracle

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


@pytest.fixture
def f32_ctx(ctx):
    ctx.set_accumulation(NC_ACCUM_FP32)
    yield ctx
    ctx.set_accumulation(NC_ACCUM_FP64)


def _survey():
    d, r = survey_cap(3000, 5000, seed=71)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.12
    box, _ = prep_cat_survey(d, r)
    return d, r, box


@pytest.mark.parametrize("path", [NC_PATH_PLAIN, NC_PATH_MIXED])
def test_f32_survey_terms_match_oracle(f32_ctx, path):
    ctx = f32_ctx
    _log(f"f32 survey terms path={path}")
    d, r, box = _survey()
    edges = edges_linear(0.0, 40.0, 12)
    D, R = ctx.catalog(d), ctx.catalog(r)
    try:
        h, c = ctx.count_batched([(D, D, D), (D, D, R), (D, R, R), (R, R, R)], edges, box, survey=True, path=path)
        for i, (a, b, cc) in enumerate(((d, d, d), (d, d, r), (d, r, r), (r, r, r))):
            wc, ws = oracle.count(a, b, cc, edges, box, survey=True)
            if i == 3:
                assert np.array_equal(c[3], wc)
                continue
            nz = ws != 0
            assert nz.sum() > 20
            assert np.all(h[i][~nz] == 0.0)
            assert np.max(np.abs(h[i][nz] / ws[nz] - 1.0)) <= 1e-9, i
    finally:
        D.close()
        R.close()


def test_f32_box_six_fold_matches_oracle(f32_ctx):
    """A weighted DDD box term on the six-fold path (rows carry 6 w1 w_row)."""
    ctx = f32_ctx
    _log("f32 box six-fold")
    L = 120.0
    p = neyman_scott_box(80, 40, 2.0, L, seed=73)
    p[:, 3] = np.random.default_rng(74).lognormal(0.0, 0.3, p.shape[0]).astype(np.float32)
    edges = edges_linear(0.0, 12.0, 6)
    box = np.array([L, L, L])
    wc, ws = oracle.count(p, p, p, edges, box)
    for path in (NC_PATH_PLAIN, NC_PATH_MIXED):
        h, _ = ctx.count(p, p, p, edges, box, path=path)
        assert ctx.timing().tri_terms == 1
        assert ctx.timing().hist_layout == 6
        nz = ws != 0
        assert np.all(h[~nz] == 0.0)
        assert np.max(np.abs(h[nz] / ws[nz] - 1.0)) <= 1e-9


def test_f32_many_additions_per_slot_against_exact(ctx):
    """Three bins over a dense clustered box: millions of additions land in a
    few dozen slots between merges.  The FP32 mode must agree with the exact
    fixed-point sums far below FP32 resolution (2^-24 = 6e-8)."""
    _log("f32 many additions")
    L = 60.0
    p = neyman_scott_box(30, 300, 1.5, L, seed=75)
    p[:, 3] = np.random.default_rng(76).lognormal(0.0, 0.3, p.shape[0]).astype(np.float32)
    edges = np.array([0.0, 2.0, 4.0, 6.0])
    box = np.array([L, L, L])
    q = p.copy()
    q[:, 3] = np.random.default_rng(77).lognormal(0.0, 0.3, p.shape[0]).astype(np.float32)
    try:
        ctx.set_accumulation(NC_ACCUM_EXACT)
        hx, _ = ctx.count(p, q, q, edges, box, path=NC_PATH_MIXED)
        ctx.set_accumulation(NC_ACCUM_FP32)
        hf, _ = ctx.count(p, q, q, edges, box, path=NC_PATH_MIXED)
        assert ctx.timing().hist_layout == 6
    finally:
        ctx.set_accumulation(NC_ACCUM_FP64)
    pu, qu = p.copy(), q.copy()
    pu[:, 3] = qu[:, 3] = 1.0
    _, cu = ctx.count(pu, qu, qu, edges, box, path=NC_PATH_MIXED)
    nz = hx != 0
    assert np.array_equal(nz, cu != 0)
    assert cu.max() > 1e6  # many additions into one slot
    assert np.all(hf[~nz] == 0.0)
    assert np.max(np.abs(hf[nz] / hx[nz] - 1.0)) <= 1e-12


def test_f32_rejects_weights_fp32_cannot_hold(f32_ctx):
    ctx = f32_ctx
    p = neyman_scott_box(20, 20, 2.0, 60.0, seed=78)
    p[:, 3] = 1.0
    p[5, 3] = 1.0 + 2.0 ** -30  # not an FP32 value
    with pytest.raises(NcError) as e:
        ctx.count(p, p, p, edges_linear(0.0, 6.0, 3), [60.0] * 3)
    assert e.value.code == 1
    # unit-weight terms need no weights at all
    p[5, 3] = 1.0
    _, c = ctx.count(p, p, p, edges_linear(0.0, 6.0, 3), [60.0] * 3)
    wc, _ = oracle.count(p, p, p, edges_linear(0.0, 6.0, 3), [60.0] * 3)
    assert np.array_equal(c, wc)
