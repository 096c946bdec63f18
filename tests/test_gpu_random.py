"""Seeded random configurations against the oracle (threeD and rppi, box and
survey, linear and log-spaced edges as source/main.cu:21 uses, unit and
lognormal weights, both arithmetic paths).  Counts bit-exact, weighted sums
within 1e-9 relative."""
import numpy as np
import pytest

from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN, NcError
from cudancorr_amd._lib import mixed_classify
from cudancorr_amd.gen import edges_linear, logspace, neyman_scott_box, prep_cat_survey, survey_cap, uniform_box
from oracle import oracle

pytestmark = pytest.mark.gpu

WEIGHT_RTOL = 1e-9


def _log(msg):
    print(msg, flush=True)


def _check(hist, cnt, want_c, want_w, weighted):
    if weighted:
        nz = want_w != 0
        assert np.array_equal(hist[~nz], want_w[~nz])
        assert np.max(np.abs(hist[nz] - want_w[nz]) / np.abs(want_w[nz]), initial=0.0) <= WEIGHT_RTOL
    else:
        assert np.array_equal(cnt, want_c)


def _mixed_runs(edges):
    """Whether the mixed path can take these edges (its lookup tables exist)."""
    try:
        mixed_classify(edges, np.zeros(1, np.float32))
        return True
    except NcError:
        return False


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    survey = bool(rng.integers(2))
    rmax = float(rng.uniform(4.0, 12.0))
    nb = int(rng.integers(3, 16))
    if rng.integers(2):
        edges = logspace(rmax * float(rng.uniform(0.005, 0.05)), rmax, nb + 1)
    else:
        edges = edges_linear(float(rng.choice([0.0, rmax * 0.1])), rmax, nb)
    weighted = bool(rng.integers(2))
    if survey:
        d, r = survey_cap(int(rng.integers(800, 2000)), int(rng.integers(800, 2000)), seed=seed,
                          clustered=bool(rng.integers(2)))
        for c in (d, r):
            c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * float(rng.uniform(0.04, 0.1))
        box, ori = prep_cat_survey(d, r)
        if not weighted:
            d[:, 3] = 1.0
        cats = (d, d, r) if rng.integers(2) else (d, d, d)
    else:
        L = float(rng.uniform(3.2, 6.0)) * rmax
        if rng.integers(2):
            p = uniform_box(int(rng.integers(1500, 3500)), L, seed=seed)
        else:
            p = neyman_scott_box(int(rng.integers(30, 80)), 30, rmax / 6.0, L, seed=seed)
        if weighted:
            p[:, 3] = rng.lognormal(0.0, 0.3, p.shape[0])
        box, ori = np.array([L, L, L]), (0.0, 0.0, 0.0)
        cats = (p, p, p)
    return survey, edges, weighted, cats, box, ori, rng


@pytest.mark.parametrize("seed", range(12))
def test_random_threed(ctx, seed):
    survey, edges, weighted, cats, box, _, _ = _case(seed)
    _log(f"random threeD seed={seed} survey={survey} nb={edges.size - 1} edges {edges[0]:.3g}..{edges[-1]:.3g} "
         f"weighted={weighted}")
    want_c, want_w = oracle.count(*cats, edges, box, survey=survey)
    assert want_c.sum() > 0
    for path in (NC_PATH_PLAIN, NC_PATH_MIXED):
        hist, cnt = ctx.count(*cats, edges, box, survey=survey, path=path)
        assert ctx.timing().path == (NC_PATH_MIXED if path == NC_PATH_MIXED and _mixed_runs(edges) else NC_PATH_PLAIN)
        _check(hist, cnt, want_c, want_w, weighted)


@pytest.mark.parametrize("seed", range(8))
def test_random_rppi(ctx, seed):
    survey, edges, weighted, cats, box, ori, rng = _case(100 + seed)
    npi = int(rng.integers(1, 9))
    if not survey:  # box cells (rpMax, rpMax, npi) must fit the box
        box = box.copy()
        box[2] = max(box[2], 3.2 * npi)
        cats = tuple(np.c_[c[:, 0], c[:, 1], np.mod(c[:, 2], box[2]), c[:, 3]] for c in cats)
    _log(f"random rppi seed={seed} survey={survey} nb={edges.size - 1} npi={npi} weighted={weighted}")
    want_c, want_w, _ = oracle.count_rppi(*cats, edges, npi, box, survey=survey, ori=ori)
    hist, cnt = ctx.count(*cats, edges, box, survey=survey, npi=npi, ori=ori)
    _check(hist, cnt, want_c, want_w, weighted)


def test_reference_driver_bins(ctx):
    """The bins of source/main.cu:21 (logspace(0.1, 20, 13)) on a clustered box."""
    _log("reference driver bins")
    edges = logspace(0.1, 20.0, 13)
    p = neyman_scott_box(200, 40, 2.0, 120.0, seed=77)
    box = np.array([120.0] * 3)
    want, _ = oracle.count(p, p, p, edges, box)
    assert want.sum() > 100_000
    for path in (NC_PATH_PLAIN, NC_PATH_MIXED):
        _, got = ctx.count(p, p, p, edges, box, path=path)
        assert ctx.timing().path == path  # the log-spaced lookup takes these bins
        assert np.array_equal(got, want)
