"""Parity at the full sizes of the benchmark workloads.  The oracle cannot
count a full catalogue in seconds, so each test checks
  * a seeded random sample of primaries against the FULL secondary catalogues
    (the sample as c1, the catalogues as c2 and c3), GPU against oracle, bit
    for bit for counts and within 1e-9 for weighted sums;
  * the full count on both arithmetic paths, which must agree bit for bit for
    counts (and within 1e-9 for weighted sums), with the total of the full
    count checked against the sum over primary slices.
Workloads: the 1M-point clustered box (rmax 30, 15 bins), the 500k + 2.5M
survey (rmax 40, 20 bins, four terms) and a sample of the 4M + 20M survey
(rmax 50, 25 bins)."""
import numpy as np
import pytest

from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN
from cudancorr_amd.gen import SEED, WORKLOADS, neyman_scott_box, prep_cat_survey, survey_cap
from oracle import oracle

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)


def _sample(rng, p, m):
    return p[np.sort(rng.choice(p.shape[0], size=m, replace=False))]


def _wcheck(got, want, rtol=1e-9):
    nz = want != 0
    assert np.all(got[~nz] == 0.0)
    assert np.max(np.abs(got[nz] / want[nz] - 1.0), initial=0.0) <= rtol


def test_box_1m_points(ctx):
    _log("full size: 1M-point clustered box")
    p = neyman_scott_box(10000, 100, 5.0, 1000.0, seed=SEED + 1000)
    edges = WORKLOADS["box_clustered"]["edges"]
    box = np.array([1000.0] * 3)
    P = ctx.catalog(p)
    try:
        _, plain = ctx.count(P, P, P, edges, box, path=NC_PATH_PLAIN)
        _, mixed = ctx.count(P, P, P, edges, box, path=NC_PATH_MIXED)
        assert ctx.timing().path == NC_PATH_MIXED and ctx.timing().tri_terms == 1
        assert plain.sum() > 3e10
        assert np.array_equal(plain, mixed)
        # four primary slices add up to the full count
        parts = np.zeros_like(mixed)
        for k in range(4):
            ctx.set_primary_slice(k, 4)
            parts += ctx.count(P, P, P, edges, box, path=NC_PATH_MIXED)[1]
        ctx.set_primary_slice(0, 1)
        assert np.array_equal(parts, mixed)
        s = _sample(np.random.default_rng(101), p, 1024)
        want, _ = oracle.count(s, p, p, edges, box)
        for path in (NC_PATH_PLAIN, NC_PATH_MIXED):
            _, got = ctx.count(s, P, P, edges, box, path=path)
            assert np.array_equal(got, want), f"path {path}"
    finally:
        ctx.set_primary_slice(0, 1)
        P.close()


def test_survey_500k_and_2_5m(ctx):
    _log("full size: 500k data + 2.5M randoms survey")
    d, r = survey_cap(500_000, 2_500_000, seed=SEED + 2000)
    box, _ = prep_cat_survey(d, r)
    edges = WORKLOADS["survey_cap"]["edges"]
    D, R = ctx.catalog(d), ctx.catalog(r)
    try:
        terms = [(D, D, D), (D, D, R), (D, R, R), (R, R, R)]
        hp, cp = ctx.count_batched(terms, edges, box, survey=True, path=NC_PATH_PLAIN)
        hm, cm = ctx.count_batched(terms, edges, box, survey=True, path=NC_PATH_MIXED)
        assert np.array_equal(cp[3], cm[3]) and cm[3].sum() > 1e11
        for i in range(3):
            _wcheck(hm[i], hp[i])
        rng = np.random.default_rng(102)
        sd, sr = _sample(rng, d, 256), _sample(rng, r, 256)
        for s, S2, S3, a, b in ((sd, D, R, d, r), (sd, D, D, d, d), (sr, R, R, r, r)):
            wc, ws = oracle.count(s, a, b, edges, box, survey=True)
            for path in (NC_PATH_PLAIN, NC_PATH_MIXED):
                h, c = ctx.count(s, S2, S3, edges, box, survey=True, path=path)
                if b is r and a is r:
                    assert np.array_equal(c, wc), f"RRR sample, path {path}"
                else:
                    _wcheck(h, ws)
    finally:
        D.close()
        R.close()


def test_survey_large_sample(ctx):
    _log("full size: 4M + 20M survey, sampled primaries")
    d, r = survey_cap(4_000_000, 20_000_000, seed=SEED + 4000, clustered=True, sigma=4.0)
    box, _ = prep_cat_survey(d, r)
    edges = WORKLOADS["survey_large"]["edges"]
    D, R = ctx.catalog(d), ctx.catalog(r)
    try:
        rng = np.random.default_rng(103)
        sd, sr = _sample(rng, d, 48), _sample(rng, r, 48)
        for s, S2, S3, a, b, weighted in ((sr, R, R, r, r, False), (sd, R, R, r, r, True), (sd, D, D, d, d, True)):
            wc, ws = oracle.count(s, a, b, edges, box, survey=True)
            assert wc.sum() > 0
            h, c = ctx.count(s, S2, S3, edges, box, survey=True, path=NC_PATH_MIXED)
            if weighted:
                _wcheck(h, ws)
            else:
                assert np.array_equal(c, wc)
    finally:
        D.close()
        R.close()
