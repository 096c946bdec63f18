"""Exact accumulation of weighted sums (nc_set_accumulation NC_ACCUM_EXACT):
every FP64 weight product is truncated to a fixed-point multiple of 2^-K and
summed as a 128-bit integer.  The sums do not depend on the order of the
additions, so two runs, and the plain and mixed paths (which bin every triplet
identically), give bit-identical results; against the oracle's compensated
FP64 sums they agree to the truncation bound."""
import numpy as np
import pytest

from cudancorr_amd import NC_ACCUM_EXACT, NC_ACCUM_FP64, NC_PATH_MIXED, NC_PATH_PLAIN, NcError
from cudancorr_amd.gen import edges_linear, neyman_scott_box, prep_cat_survey, survey_cap
from oracle import oracle

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


@pytest.fixture
def exact_ctx(ctx):
    ctx.set_accumulation(NC_ACCUM_EXACT)
    yield ctx
    ctx.set_accumulation(NC_ACCUM_FP64)


def _survey():
    d, r = survey_cap(3000, 5000, seed=61)
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.12
    box, _ = prep_cat_survey(d, r)
    return d, r, box


def test_exact_survey_terms_reproducible_and_path_independent(exact_ctx):
    ctx = exact_ctx
    _log("exact survey terms")
    d, r, box = _survey()
    edges = edges_linear(0.0, 40.0, 12)
    D, R = ctx.catalog(d), ctx.catalog(r)
    terms = [(D, D, D), (D, D, R), (D, R, R), (R, R, R)]
    try:
        runs = {}
        for path in (NC_PATH_PLAIN, NC_PATH_MIXED, NC_PATH_MIXED):
            h, c = ctx.count_batched(terms, edges, box, survey=True, path=path)
            runs.setdefault(path, []).append((h, c))
        h0, c0 = runs[NC_PATH_PLAIN][0]
        for h, c in runs[NC_PATH_MIXED]:
            assert np.array_equal(h, h0) and np.array_equal(c, c0)  # bit for bit
        for i, (a, b, cc) in enumerate(((d, d, d), (d, d, r), (d, r, r), (r, r, r))):
            wc, ws = oracle.count(a, b, cc, edges, box, survey=True)
            if i == 3:
                assert np.array_equal(c0[3], wc)
                continue
            nz = ws != 0
            assert nz.sum() > 20
            assert np.array_equal(h0[i][~nz], ws[~nz])
            # lognormal(0, 0.3) weights: products below 2^e with e <= 12 here, each
            # off by < 2^(e-60) against products >= 1e-3: relative < 1e-12
            assert np.max(np.abs(h0[i][nz] / ws[nz] - 1.0)) <= 1e-12, i
        # the FP64 accumulation agrees to rounding
        ctx.set_accumulation(NC_ACCUM_FP64)
        hf, cf = ctx.count_batched(terms, edges, box, survey=True, path=NC_PATH_MIXED)
        assert np.array_equal(cf[3], c0[3])
        for i in range(3):
            nz = h0[i] != 0
            assert np.max(np.abs(hf[i][nz] / h0[i][nz] - 1.0)) <= 1e-12
    finally:
        D.close()
        R.close()


def test_exact_box_six_fold_and_negative_weights(exact_ctx):
    """Weights in [-1, 1] (the fixed-point sums carry signs through the high
    word) on a box DDD term, which takes the six-fold path: agreement with the
    oracle to K's truncation bound, and bit identity across paths."""
    ctx = exact_ctx
    _log("exact box negative weights")
    L = 120.0
    p = neyman_scott_box(80, 40, 2.0, L, seed=63)
    p[:, 3] = np.random.default_rng(64).uniform(-1.0, 1.0, p.shape[0])
    edges = edges_linear(0.0, 12.0, 6)
    box = np.array([L, L, L])
    wc, ws = oracle.count(p, p, p, edges, box)
    got = []
    for path in (NC_PATH_PLAIN, NC_PATH_MIXED):
        h, _ = ctx.count(p, p, p, edges, box, path=path)
        assert ctx.timing().tri_terms == 1
        got.append(h)
    assert np.array_equal(got[0], got[1])
    # |w| <= 1: products below 2^1, each truncated by < 2^(1-60); the oracle's
    # compensated sum is within a few ulp of the exact one
    tol = wc.astype(np.float64) * 2.0 ** -55 + 4e-16 * np.abs(ws) + 1e-300
    assert np.all(np.abs(got[0] - ws) <= tol)


def test_exact_rejects_non_finite_weights(exact_ctx):
    ctx = exact_ctx
    p = neyman_scott_box(20, 20, 2.0, 60.0, seed=65)
    p[3, 3] = np.inf
    with pytest.raises(NcError) as e:
        ctx.count(p, p, p, edges_linear(0.0, 6.0, 3), [60.0] * 3)
    assert e.value.code == 1


def test_exact_batched_replay(exact_ctx):
    """A captured exact-mode plan replays to the same bits."""
    ctx = exact_ctx
    _log("exact batched replay")
    d, r, box = _survey()
    edges = edges_linear(0.0, 40.0, 8)
    D, R = ctx.catalog(d), ctx.catalog(r)
    try:
        first = None
        for rep in range(3):
            h, c = ctx.count_batched([(D, D, R), (D, R, R)], edges, box, survey=True, path=NC_PATH_MIXED)
            if first is None:
                first = h
            assert np.array_equal(h, first), rep
    finally:
        D.close()
        R.close()
