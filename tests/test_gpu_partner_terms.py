"""Partner terms: a term (c1, c3, c3) counted beside a term (c1, c2, c3) of the
same call, c2 != c3 (the survey's DRR beside DDR), from the first term's
primaries and its lists of c3, into a second workgroup histogram.  Checked
against the oracle on both paths (counts bit for bit, weighted sums within
1e-9), against a context with the pairing turned off (NC_FUSE_TERMS=0), with
primaries whose list of c2 is empty (their partner pairs must still count),
over three catalogues, in a periodic box with partial last cells, and under
graph replay."""
import os

import numpy as np
import pytest

from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN, Context
from cudancorr_amd.gen import edges_linear, prep_cat_survey, survey_cap
from oracle import oracle

pytestmark = pytest.mark.gpu

PATHS = [NC_PATH_PLAIN, NC_PATH_MIXED]


def _log(msg):
    print(msg, flush=True)


def _wcheck(got, want, rtol=1e-9):
    nz = want != 0
    assert nz.sum() > 10
    assert np.all(got[~nz] == 0.0)
    assert np.max(np.abs(got[nz] / want[nz] - 1.0)) <= rtol


@pytest.fixture(scope="module")
def cats():
    d, r = survey_cap(1500, 5000, seed=91)
    r2 = survey_cap(10, 3000, seed=92)[1]
    sparse = survey_cap(60, 10, seed=93)[0]  # a few data points: most have no data neighbour
    for c in (d, r, r2, sparse):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.1
    box, _ = prep_cat_survey(d, r, r2, sparse)  # (shifts the catalogues in place)
    rw = r.copy()
    rw[:, 3] = np.random.default_rng(94).lognormal(0.0, 0.3, r.shape[0]).astype(np.float32)
    return dict(d=d, r=r, r2=r2, rw=rw, s=sparse, box=box, edges=edges_linear(0.0, 30.0, 12))


def _oracle(c, a, b, e):
    return oracle.count(c[a], c[b], c[e], c["edges"], c["box"], survey=True)


@pytest.mark.parametrize("path", PATHS)
def test_drr_beside_ddr(ctx, cats, path):
    """Weighted data, unit-weight randoms: DDR (pair-swap path) with DRR as
    its partner (the unit-secondary weighted pair loop)."""
    _log(f"partner DDR+DRR path={path}")
    c = cats
    D, R = ctx.catalog(c["d"]), ctx.catalog(c["r"])
    try:
        h, _ = ctx.count_batched([(D, D, R), (D, R, R)], c["edges"], c["box"], survey=True, path=path)
        assert ctx.timing().partner_terms == 1
        _wcheck(h[0], _oracle(c, "d", "d", "r")[1])
        _wcheck(h[1], _oracle(c, "d", "r", "r")[1])
    finally:
        D.close()
        R.close()


@pytest.mark.parametrize("path", PATHS)
def test_weighted_partner_over_three_catalogues(ctx, cats, path):
    """(D, R2, Rw) with (D, Rw, Rw): the first term on the plain two-list path,
    its partner with every weight staged; and the same with the catalogue
    order of the call reversed."""
    _log(f"partner three catalogues path={path}")
    c = cats
    D, R2, RW = ctx.catalog(c["d"]), ctx.catalog(c["r2"]), ctx.catalog(c["rw"])
    try:
        want0 = _oracle(c, "d", "r2", "rw")[1]
        want1 = _oracle(c, "d", "rw", "rw")[1]
        for terms, (i0, i1) in (([(D, R2, RW), (D, RW, RW)], (0, 1)), ([(D, RW, RW), (D, R2, RW)], (1, 0))):
            h, _ = ctx.count_batched(terms, c["edges"], c["box"], survey=True, path=path)
            assert ctx.timing().partner_terms == 1
            _wcheck(h[i0], want0)
            _wcheck(h[i1], want1)
    finally:
        for x in (D, R2, RW):
            x.close()


def test_partner_of_primaries_without_c2_neighbours(ctx, cats):
    """60 sparse data points as c1 and c2: nearly every primary has an empty
    list of c2, and its partner pairs over the randoms must still count."""
    _log("partner sparse")
    c = cats
    S, R = ctx.catalog(c["s"]), ctx.catalog(c["r"])
    try:
        want_ssr, want_srr = _oracle(c, "s", "s", "r")[1], _oracle(c, "s", "r", "r")[1]
        assert want_srr.sum() > 0
        for path in PATHS:
            h, _ = ctx.count_batched([(S, S, R), (S, R, R)], c["edges"], c["box"], survey=True, path=path)
            assert ctx.timing().partner_terms == 1
            nz = want_ssr != 0
            assert np.all(h[0][~nz] == 0.0)
            if nz.any():
                assert np.max(np.abs(h[0][nz] / want_ssr[nz] - 1.0)) <= 1e-9
            _wcheck(h[1], want_srr)
    finally:
        S.close()
        R.close()


def test_paired_equals_unpaired_and_replays(ctx, cats):
    """The four survey terms against a context with the pairing off (sums to
    rounding, unit-weight counts identical), and graph replays of the paired
    batch against the eager result."""
    _log("partner vs NC_FUSE_TERMS=0")
    c = cats
    old = os.environ.get("NC_FUSE_TERMS")
    os.environ["NC_FUSE_TERMS"] = "0"
    try:
        off = Context(0)
    finally:
        if old is None:
            del os.environ["NC_FUSE_TERMS"]
        else:
            os.environ["NC_FUSE_TERMS"] = old
    D, R = ctx.catalog(c["d"]), ctx.catalog(c["r"])
    D2, R2 = off.catalog(c["d"]), off.catalog(c["r"])
    try:
        terms = [(D, D, D), (D, D, R), (D, R, R), (R, R, R)]
        h, n = ctx.count_batched(terms, c["edges"], c["box"], survey=True, path=NC_PATH_MIXED)
        assert ctx.timing().partner_terms == 1
        h0, n0 = off.count_batched([(D2, D2, D2), (D2, D2, R2), (D2, R2, R2), (R2, R2, R2)], c["edges"], c["box"],
                                   survey=True, path=NC_PATH_MIXED)
        assert off.timing().partner_terms == 0
        assert np.array_equal(n[3], n0[3])
        for i in range(3):
            _wcheck(h[i], h0[i], rtol=1e-12)
        for rep in range(3):
            hr, nr = ctx.count_batched(terms, c["edges"], c["box"], survey=True, path=NC_PATH_MIXED)
            assert np.array_equal(nr[3], n[3]), rep
            for i in range(3):
                _wcheck(hr[i], h[i], rtol=1e-12)
    finally:
        for x in (D, R, D2, R2):
            x.close()
        off.close()


@pytest.mark.parametrize("path", PATHS)
def test_partner_in_a_box_with_partial_last_cells(ctx, path):
    """(D, D, R) with (D, R, R) in a periodic box of 4.7 cells per axis: the
    partner's two-fold loop over the first term's lists of R, last-cell
    primaries included."""
    _log(f"partner box path={path}")
    from cudancorr_amd.gen import neyman_scott_box

    L = 47.0
    d = neyman_scott_box(40, 25, 2.0, L, seed=95)
    d[:, 3] = np.random.default_rng(96).lognormal(0.0, 0.3, d.shape[0]).astype(np.float32)
    r = np.ones((3000, 4))
    r[:, :3] = np.random.default_rng(98).uniform(0.0, L, size=(3000, 3)).astype(np.float32)
    edges = edges_linear(0.0, 10.0, 8)
    box = np.array([L, L, L])
    D, R = ctx.catalog(d), ctx.catalog(r)
    try:
        h, _ = ctx.count_batched([(D, D, R), (D, R, R)], edges, box, path=path)
        assert ctx.timing().partner_terms == 1
        _wcheck(h[0], oracle.count(d, d, r, edges, box)[1])
        _wcheck(h[1], oracle.count(d, r, r, edges, box)[1])
    finally:
        D.close()
        R.close()
