"""Mixed lists around the list chunk C (the rows a wave stages in LDS at a
time): clustered inputs give list lengths from a few entries to several times
C, so lists fit one row chunk, end just past one, or span several with a
partial tail block.  Each case runs at list chunks 64, 128 and 256 and is
checked against the oracle (counts bit for bit, weighted sums within 1e-9):
the six-fold path (with its last-cell lists and the rows its low pass
appends), the two-fold path of last-cell primaries, and weighted survey
terms (DRR over weighted and over unit-weight randoms, and D R R' over two
random catalogues)."""
import numpy as np
import pytest

from cudancorr_amd import NC_PATH_MIXED
from cudancorr_amd.gen import edges_linear, neyman_scott_box, prep_cat_survey, survey_cap
from oracle import oracle

pytestmark = pytest.mark.gpu

CHUNKS = [64, 128, 256]


def _log(msg):
    print(msg, flush=True)


def _wcheck(got, want, rtol=1e-9):
    nz = want != 0
    assert nz.sum() > 20
    assert np.all(got[~nz] == 0.0)
    assert np.max(np.abs(got[nz] / want[nz] - 1.0)) <= rtol


def _box():
    """Dense clusters over a sparse background in a periodic box of 12 cells
    per side: lists of ~5 to ~600 entries."""
    L = 120.0
    c = neyman_scott_box(25, 300, 2.0, L, seed=81)
    u = np.random.default_rng(82).uniform(0.0, L, size=(3000, 3))
    p = np.ones((c.shape[0] + u.shape[0], 4))
    p[: c.shape[0]] = c
    p[c.shape[0]:, :3] = u.astype(np.float32)
    return p, edges_linear(0.0, 10.0, 10), np.array([L, L, L])


@pytest.fixture(scope="module")
def box_case():
    p, edges, box = _box()
    pw = p.copy()
    pw[:, 3] = np.random.default_rng(83).lognormal(0.0, 0.3, p.shape[0]).astype(np.float32)
    return p, pw, edges, box, oracle.count(p, p, p, edges, box)[0], oracle.count(pw, pw, pw, edges, box)[1]


@pytest.mark.parametrize("chunk", CHUNKS)
def test_box_lists_around_the_chunk(ctx, box_case, chunk):
    _log(f"list head box chunk={chunk}")
    p, pw, edges, box, want, wsum = box_case
    before = ctx.get_tuning()
    try:
        ctx.set_tuning(-1, 16, -1, chunk)
        _, got = ctx.count(p, p, p, edges, box, path=NC_PATH_MIXED)
        assert ctx.timing().tri_terms == 1
        assert np.array_equal(got, want)
        h, _ = ctx.count(pw, pw, pw, edges, box, path=NC_PATH_MIXED)
        _wcheck(h, wsum)
    finally:
        ctx.set_tuning(*before)


@pytest.fixture(scope="module")
def survey_case():
    d, r = survey_cap(2000, 6000, seed=84, clustered=True, sigma=2.0)
    r2 = survey_cap(10, 4000, seed=86)[1]
    for c in (d, r, r2):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.1
    box, _ = prep_cat_survey(d, r, r2)  # (shifts the catalogues in place)
    rw = r.copy()
    rw[:, 3] = np.random.default_rng(85).lognormal(0.0, 0.3, r.shape[0]).astype(np.float32)
    edges = edges_linear(0.0, 30.0, 12)
    want = {
        "DRwRw": oracle.count(d, rw, rw, edges, box, survey=True)[1],
        "DRR": oracle.count(d, r, r, edges, box, survey=True)[1],
        "DRR2": oracle.count(d, r, r2, edges, box, survey=True)[1],
    }
    return d, r, rw, r2, box, edges, want


@pytest.mark.parametrize("chunk", CHUNKS)
def test_survey_weighted_lists_around_the_chunk(ctx, survey_case, chunk):
    """DRR with weighted randoms (every weight staged), and with unit-weight
    randoms over one and two random catalogues (only the primary's weight:
    an accepted pair adds the number of its orders times w1)."""
    _log(f"list chunks survey chunk={chunk}")
    d, r, rw, r2, box, edges, want = survey_case
    before = ctx.get_tuning()
    try:
        ctx.set_tuning(-1, 16, -1, chunk)
        for key, (a, b) in (("DRwRw", (rw, rw)), ("DRR", (r, r)), ("DRR2", (r, r2))):
            h, _ = ctx.count(d, a, b, edges, box, survey=True, path=NC_PATH_MIXED)
            _wcheck(h, want[key])
    finally:
        ctx.set_tuning(*before)
