"""Slab pruning of the neighbour lists: every cell's points are sorted into
z slabs, and a list skips the slabs of a neighbour cell that lie beyond rmax
of the primary when the catalogue is exactly gridded.  The skipped points
would fail the range filter, so the counts must equal the oracle's on inputs
that put points on slab and cell boundaries, at exactly rmax, and on inputs
whose gridding is not exact (where the lists take whole cells)."""
import numpy as np
import pytest

from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN
from cudancorr_amd.gen import edges_linear, uniform_box
from oracle import oracle

pytestmark = pytest.mark.gpu

PATHS = (NC_PATH_PLAIN, NC_PATH_MIXED)


def _log(msg):
    print(msg, flush=True)


def _lattice(n, L, step, seed):
    """n distinct points of the lattice step * Z^3 inside [0, L)^3 (every
    coordinate on a slab boundary when step divides the slab width)."""
    rng = np.random.default_rng(seed)
    m = int(round(L / step))
    idx = rng.choice(m ** 3, size=n, replace=False)
    p = np.ones((n, 4))
    p[:, 0] = (idx // (m * m)) * step
    p[:, 1] = ((idx // m) % m) * step
    p[:, 2] = (idx % m) * step
    return p


@pytest.mark.parametrize("survey", [False, True])
def test_points_on_slab_boundaries(ctx, survey):
    """rmax 8 (cells of 8, slabs of 1) and a lattice of spacing 0.5: points on
    every slab boundary, and many pairs at exactly rmax (rejected by the strict
    filter) or at the edges of bins."""
    _log(f"slabs lattice survey={survey}")
    L = 32.0
    p = _lattice(2500, L, 0.5, seed=3 + survey)
    edges = edges_linear(0.0, 8.0, 8)
    box = np.array([L, L, L])
    want, _ = oracle.count(p, p, p, edges, box, survey=survey)
    assert want.sum() > 0
    for path in PATHS:
        _, got = ctx.count(p, p, p, edges, box, survey=survey, path=path)
        assert np.array_equal(got, want), f"path {path}"


def test_three_catalogues_on_slab_boundaries(ctx):
    """Three different lattice catalogues (c2 != c3): both lists cut slabs."""
    _log("slabs lattice three catalogues")
    L = 24.0
    a, b, c = (_lattice(1200, L, 0.25, seed=s) for s in (11, 12, 13))
    edges = edges_linear(0.5, 8.0, 6)
    box = np.array([L, L, L])
    want, _ = oracle.count(a, b, c, edges, box)
    assert want.sum() > 0
    for path in PATHS:
        _, got = ctx.count(a, b, c, edges, box, path=path)
        assert np.array_equal(got, want), f"path {path}"


@pytest.mark.parametrize("where", ["negative", "upper"])
def test_inexact_gridding_takes_whole_cells(ctx, where):
    """Points the reference accepts but whose cell index is not the floor of
    coordinate / cell size: -cs < x < 0 (truncated into cell 0) or x == L
    (clamped into the last cell).  Such a point can lie outside its cell's
    slabs, so that catalogue's lists must take whole cells; the other
    catalogue keeps its slab cut."""
    _log(f"slabs inexact gridding {where}")
    L = 96.0  # 8 cells of 12: x == L is clamped into the last cell
    p = uniform_box(1500, L, seed=21)
    rng = np.random.default_rng(22)
    k = 40
    if where == "negative":
        p[:k, 2] = -rng.uniform(0.0, 5.0, k)  # z in (-5, 0): cell 0, slab 0
        p[k:2 * k, 0] = -rng.uniform(0.0, 5.0, k)
    else:
        p[:k, 2] = L
        p[k:2 * k, 1] = L
    q = uniform_box(1500, L, seed=23)  # exactly gridded
    edges = edges_linear(1.0, 12.0, 8)
    box = np.array([L, L, L])
    for cats in ((p, p, p), (q, p, p), (p, q, q), (q, q, p)):
        want, _ = oracle.count(*cats, edges, box)
        assert want.sum() > 0
        for path in PATHS:
            _, got = ctx.count(*cats, edges, box, path=path)
            assert np.array_equal(got, want), f"path {path}"
