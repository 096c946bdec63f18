"""Parity of the HIP triplet count (through the C-ABI) with the CPU oracle.

Unweighted counts must be bit-exact; weighted sums (FP64) within the stated
relative tolerance of the oracle's compensated FP64 sums.  Sizes are chosen so
the oracle finishes in seconds.
"""
import numpy as np
import pytest

from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN
from cudancorr_amd.gen import edges_linear, neyman_scott_box, prep_cat_survey, survey_cap, uniform_box
from oracle import oracle

pytestmark = pytest.mark.gpu

PATHS = [NC_PATH_PLAIN, NC_PATH_MIXED]
WEIGHT_RTOL = 1e-9


def _log(msg):
    print(msg, flush=True)


def _wcheck(got, want, rtol=WEIGHT_RTOL):
    nz = want != 0
    assert np.array_equal(got[~nz], want[~nz]), "weighted bins that must be empty are not"
    rel = np.abs(got[nz] - want[nz]) / np.abs(want[nz])
    assert rel.max(initial=0.0) <= rtol, f"max relative error {rel.max():.3e}"
    return rel.max(initial=0.0)


@pytest.mark.parametrize("path", PATHS)
def test_box_uniform_full(ctx, path):
    """20,000 uniform points, L=250 (13 cells, last one partial), 10 bins on [2,20]."""
    _log(f"box_uniform path={path}")
    p = uniform_box(20000, 250.0, seed=11)
    edges = edges_linear(2.0, 20.0, 10)
    box = np.array([250.0] * 3)
    want, _ = oracle.count(p, p, p, edges, box)
    hist, got = ctx.count(p, p, p, edges, box, path=path)
    assert want.sum() > 1_000_000
    assert np.array_equal(got, want)
    assert np.array_equal(hist, want.astype(np.float64))


@pytest.mark.parametrize("path", PATHS)
def test_box_clustered(ctx, path):
    """Neyman-Scott box, float32 coordinates, bins from 0 (duplicates possible)."""
    _log(f"box_clustered path={path}")
    p = neyman_scott_box(300, 50, 3.0, 200.0, seed=5)
    edges = edges_linear(0.0, 30.0, 15)
    box = np.array([200.0] * 3)
    want, _ = oracle.count(p, p, p, edges, box)
    _, got = ctx.count(p, p, p, edges, box, path=path)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("L", [10.0, 25.0, 47.0, 60.0])
def test_box_small_grids(ctx, L):
    """1, 2 and 3 cells per side: repeated neighbour entries with different
    image shifts, and the partial last cell."""
    _log(f"small_grid L={L}")
    p = uniform_box(400, L, seed=int(L))
    edges = edges_linear(0.0, 20.0, 8)
    box = np.array([L, L, L])
    want, _ = oracle.count(p, p, p, edges, box)
    for path in PATHS:
        _, got = ctx.count(p, p, p, edges, box, path=path)
        assert np.array_equal(got, want), f"path {path}"


def test_box_asymmetric_three_catalogues(ctx):
    _log("box three catalogues")
    L = 120.0
    a = uniform_box(3000, L, seed=1)
    b = uniform_box(2500, L, seed=2)
    c = uniform_box(3500, L, seed=3)
    edges = edges_linear(1.0, 15.0, 7)
    box = np.array([L, 1.2 * L, L])
    b[:, 1] *= 1.2
    a[:, 1] *= 1.2
    c[:, 1] *= 1.2
    want, _ = oracle.count(a, b, c, edges, box)
    for path in PATHS:
        _, got = ctx.count(a, b, c, edges, box, path=path)
        assert np.array_equal(got, want), f"path {path}"


def _small_survey(seed=3):
    d, r = survey_cap(6000, 9000, seed=seed)
    # shrink the shell so the sample is dense enough to have many triplets
    for c in (d, r):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.15
    box, ori = prep_cat_survey(d, r)
    return d, r, box


@pytest.mark.parametrize("term", ["DDD", "DDR", "DRR", "RRR"])
def test_survey_terms(ctx, term):
    _log(f"survey term={term}")
    d, r, box = _small_survey()
    cats = {"D": d, "R": r}
    c1, c2, c3 = (cats[t] for t in term)
    edges = edges_linear(0.0, 40.0, 20)
    wc, ws = oracle.count(c1, c2, c3, edges, box, survey=True)
    assert wc.sum() > 10000
    for path in PATHS:
        hist, got = ctx.count(c1, c2, c3, edges, box, survey=True, path=path)
        if term == "RRR":
            assert np.array_equal(got, wc)
        else:
            _wcheck(hist, ws)


def test_survey_symmetry_invariants(ctx):
    """Survey DDD/RRR bins are multiples of 6, DDR/DRR multiples of 2."""
    _log("survey invariants")
    d, r, box = _small_survey(seed=9)
    one = d.copy()
    one[:, 3] = 1.0
    edges = edges_linear(0.0, 40.0, 20)
    _, ddd = ctx.count(one, one, one, edges, box, survey=True)
    _, ddr = ctx.count(one, one, r, edges, box, survey=True)
    assert ddd.sum() > 0 and np.all(ddd % 6 == 0)
    assert ddr.sum() > 0 and np.all(ddr % 2 == 0)


def test_weighted_box_scaling(ctx):
    """w -> c*w scales every weighted bin by c^3 (c a power of two: exact)."""
    _log("weight scaling")
    p = uniform_box(5000, 100.0, seed=4)
    rng = np.random.default_rng(4)
    p[:, 3] = rng.lognormal(0.0, 0.3, p.shape[0]).astype(np.float32)
    edges = edges_linear(2.0, 20.0, 10)
    box = np.array([100.0] * 3)
    _, ws = oracle.count(p, p, p, edges, box)
    h1, _ = ctx.count(p, p, p, edges, box)
    q = p.copy()
    q[:, 3] *= 2.0
    h2, _ = ctx.count(q, q, q, edges, box)
    _wcheck(h1, ws)
    _wcheck(h2 / 8.0, ws)


def test_empty_and_tiny(ctx):
    _log("empty/tiny")
    edges = edges_linear(0.0, 10.0, 5)
    box = np.array([50.0] * 3)
    empty = np.zeros((0, 4))
    p = uniform_box(10, 50.0, seed=1)
    h, c = ctx.count(empty, p, p, edges, box)
    assert c.sum() == 0 and h.sum() == 0
    h, c = ctx.count(p, empty, p, edges, box)
    assert c.sum() == 0
    # equilateral triangle: 6 ordered triplets in bin (i,i,i)
    s = 3.0
    tri = np.array([[10, 10, 10, 1], [10 + s, 10, 10, 1], [10 + s / 2, 10 + s * np.sqrt(3) / 2, 10, 1]], float)
    want, _ = oracle.count(tri, tri, tri, edges, box)
    _, got = ctx.count(tri, tri, tri, edges, box)
    assert np.array_equal(got, want)
    assert got[1, 1, 1] == 6 and got.sum() == 6


def test_host_buffer_entry(ctx):
    """nc_count_host (the calculateD1D2D3 form) equals nc_count."""
    _log("host entry")
    p = uniform_box(4000, 80.0, seed=21)
    r = uniform_box(4000, 80.0, seed=22)
    edges = edges_linear(0.0, 16.0, 8)
    box = np.array([80.0] * 3)
    want, _ = oracle.count(p, p, r, edges, box)
    _, got = ctx.count_host(p, p, r, edges, box)
    assert np.array_equal(got, want)


def test_domain_error(ctx):
    from cudancorr_amd import NcError

    edges = edges_linear(0.0, 10.0, 5)
    p = uniform_box(100, 50.0, seed=1)
    p[5, 0] = 500.0  # far outside the 50-wide box
    with pytest.raises(NcError) as e:
        ctx.count(p, p, p, edges, np.array([50.0] * 3))
    assert e.value.code == 3


def test_repeat_bit_identical(ctx):
    _log("repeat")
    p = neyman_scott_box(200, 40, 3.0, 150.0, seed=8)
    edges = edges_linear(0.0, 30.0, 15)
    box = np.array([150.0] * 3)
    _, a = ctx.count(p, p, p, edges, box)
    _, b = ctx.count(p, p, p, edges, box)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("path", PATHS)
def test_lattice_edges_and_collinear_triples(ctx, path):
    """An integer lattice with edges sqrt(k): nearly every squared separation
    lies exactly on a squared edge and many triples are exactly collinear, so
    on the mixed path most pairs go through the FP64 recheck queue, which fills
    and drains over and over.  Unit weights (dense layout, 16-bit queue
    entries) and weighted (sorted slots, two-word entries): the counts
    bit-exact and the sums within tolerance."""
    _log(f"lattice path={path}")
    n = 12
    g = np.arange(n, dtype=np.float64)
    xyz = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    p = np.c_[xyz, np.ones(xyz.shape[0])]
    edges = np.sqrt(np.arange(10, dtype=np.float64))  # 0, 1, sqrt2, ..., 3
    box = np.array([float(n)] * 3)
    want, _ = oracle.count(p, p, p, edges, box)
    assert want.sum() > 1_000_000
    _, got = ctx.count(p, p, p, edges, box, path=path)
    assert np.array_equal(got, want)
    w = p.copy()
    w[:, 3] = 1.0 + (np.arange(w.shape[0]) % 7) / 8.0
    _, ws = oracle.count(w, w, w, edges, box)
    hw, _ = ctx.count(w, w, w, edges, box, path=path)
    _wcheck(hw, ws)
