"""Last-cell triples of one-catalogue box terms counted six-fold
(KArgs.tri_last6): a primary outside the last cells counts every triple of
its lists once, with six orders, when it is the triple's lowest-index point
outside the last cells; a last-cell primary counts a pair of its list only if
the two entries' cells are not adjacent, or are both last cells.  Checked
against the oracle (counts bit for bit, weighted sums within 1e-9) on boxes of
3 to 13 cells per axis with partial last cells of several widths, points
crowded around the periodic seam, both paths, unit and weighted catalogues
and all three accumulation modes, and against a context that keeps the
last-cell triples two-fold (NC_TRI_LAST=0)."""
import os

import numpy as np
import pytest

from cudancorr_amd import NC_ACCUM_EXACT, NC_ACCUM_FP32, NC_ACCUM_FP64, NC_PATH_MIXED, NC_PATH_PLAIN, Context
from cudancorr_amd.gen import edges_linear
from oracle import oracle

pytestmark = pytest.mark.gpu

PATHS = [NC_PATH_PLAIN, NC_PATH_MIXED]
RMAX = 10.0


def _log(msg):
    print(msg, flush=True)


def _seam_box(L, n, seed, weighted=False):
    """Uniform points plus clumps astride the periodic seam (x, y or z near 0
    and L) and inside the partial last cells."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, L, size=(n, 3))
    k = n // 2
    c = rng.uniform(0.0, L, size=(k, 3))
    ax = rng.integers(0, 3, size=k)
    c[np.arange(k), ax] = np.mod(rng.normal(0.0, 0.6 * RMAX, size=k), L)  # around 0 == L
    xyz = np.concatenate([u, c]).astype(np.float32).astype(np.float64)
    xyz = np.minimum(xyz, np.nextafter(L, 0.0))
    p = np.ones((xyz.shape[0], 4))
    p[:, :3] = xyz
    if weighted:
        p[:, 3] = rng.lognormal(0.0, 0.3, p.shape[0]).astype(np.float32)
    return p


GEOMS = [(2.5, 800), (3.2, 1500), (3.5, 1600), (4.0, 2000), (4.7, 3000), (12.5, 12000)]


@pytest.mark.parametrize("ratio,n", GEOMS)
def test_unit_weights_against_oracle(ctx, ratio, n):
    L = ratio * RMAX
    _log(f"tri_last unit L/rmax={ratio}")
    p = _seam_box(L, n, seed=int(ratio * 10))
    edges = edges_linear(0.0, RMAX, 8)
    box = np.array([L, L, L])
    want, _ = oracle.count(p, p, p, edges, box)
    assert want.sum() > 0
    for path in PATHS:
        _, got = ctx.count(p, p, p, edges, box, path=path)
        assert ctx.timing().tri_terms == 1
        assert np.array_equal(got, want), f"path {path}"


@pytest.mark.parametrize("ratio,n", [(2.5, 800), (3.2, 1500), (12.5, 8000)])
@pytest.mark.parametrize("accum", [NC_ACCUM_FP64, NC_ACCUM_EXACT, NC_ACCUM_FP32])
def test_weighted_against_oracle(ctx, ratio, n, accum):
    L = ratio * RMAX
    _log(f"tri_last weighted L/rmax={ratio} accum={accum}")
    p = _seam_box(L, n, seed=int(ratio * 10) + 1, weighted=True)
    edges = edges_linear(0.0, RMAX, 8)
    box = np.array([L, L, L])
    _, ws = oracle.count(p, p, p, edges, box)
    nz = ws != 0
    assert nz.sum() > 20
    try:
        ctx.set_accumulation(accum)
        for path in PATHS:
            h, _ = ctx.count(p, p, p, edges, box, path=path)
            assert np.all(h[~nz] == 0.0)
            assert np.max(np.abs(h[nz] / ws[nz] - 1.0)) <= 1e-9, f"path {path}"
    finally:
        ctx.set_accumulation(NC_ACCUM_FP64)


def test_same_counts_as_two_fold_last_cells(ctx):
    """A batch of boxes of several grids against a context with NC_TRI_LAST=0,
    and graph replays of the batch."""
    _log("tri_last vs NC_TRI_LAST=0")
    old = os.environ.get("NC_TRI_LAST")
    os.environ["NC_TRI_LAST"] = "0"
    try:
        off = Context(0)
    finally:
        if old is None:
            del os.environ["NC_TRI_LAST"]
        else:
            os.environ["NC_TRI_LAST"] = old
    L = 3.3 * RMAX
    edges = edges_linear(0.0, RMAX, 8)
    box = np.array([L, L, L])
    pts = [_seam_box(L, 1200 + 300 * i, seed=40 + i) for i in range(3)]
    try:
        cats = [ctx.catalog(p) for p in pts]
        cats0 = [off.catalog(p) for p in pts]
        _, n0 = off.count_batched([(c, c, c) for c in cats0], edges, box, path=NC_PATH_MIXED)
        for rep in range(3):
            _, n = ctx.count_batched([(c, c, c) for c in cats], edges, box, path=NC_PATH_MIXED)
            assert np.array_equal(n, n0), rep
        for c in cats + cats0:
            c.close()
    finally:
        off.close()
