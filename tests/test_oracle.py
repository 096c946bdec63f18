"""Pin the CPU oracle (oracle/oracle.c) -- CPU only.

Checks: golden fixtures from the brute force, oracle == brute force on fresh
random inputs, closed-form cases, the reference's quirks (missed wrap pair,
FP32 triangle test, strict range filter, half-open bins), invariants, the
evaluation order of sqrSep, output compaction sizes and input validation.
A planted mistake in the oracle (wrong sign, dropped term, swapped bins) fails
at least one of these.
"""
import glob
import os
from fractions import Fraction

import numpy as np
import pytest

from cudancorr_amd.gen import edges_linear, neyman_scott_box, uniform_box
from oracle import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def pts(rows, w=1.0):
    a = np.array(rows, dtype=np.float64).reshape(-1, 3)
    out = np.full((a.shape[0], 4), w)
    out[:, :3] = a
    return out


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))), ids=os.path.basename)
def test_golden_fixtures(path):
    d = np.load(path)  # allow_pickle is False by default
    c, w = oracle.count(d["p1"], d["p2"], d["p3"], d["edges"], d["box"], survey=bool(d["survey"]))
    assert np.array_equal(c, d["counts"])
    assert np.array_equal(w, d["wsum"])


def test_random_survey_matches_brute_force():
    rng = np.random.default_rng(5)
    p = np.ones((70, 4))
    p[:, :3] = np.round(rng.uniform(0, 25, (70, 3)) * 128) / 128
    q = np.ones((60, 4))
    q[:, :3] = np.round(rng.uniform(0, 25, (60, 3)) * 128) / 128
    edges = np.array([0.25, 2.0, 4.0, 5.0, 8.0])
    box = np.array([26.0, 26.0, 26.0])
    for a, b, c in ((p, p, p), (p, p, q), (p, q, q), (q, p, q)):
        got, _ = oracle.count(a, b, c, edges, box, survey=True)
        want, _ = oracle.brute_force(a, b, c, edges, None)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("m", [3, 4, 5])
def test_commensurate_box_matches_minimal_image(m):
    rng = np.random.default_rng(m)
    rmax = 5.0
    L = m * rmax
    p = np.ones((150, 4))
    p[:, :3] = np.round(rng.uniform(0, L, (150, 3)) * 64) / 64 % L
    edges = np.linspace(0.0, rmax, 6)
    got, _ = oracle.count(p, p, p, edges, [L, L, L])
    want, _ = oracle.brute_force(p, p, p, edges, [L, L, L])
    assert want.sum() > 0
    assert np.array_equal(got, want)


def test_missed_wrap_pair_quirk():
    """L=250, rmax=20: 13 cells, the last one 10 wide.  A triangle across the
    wrap whose points are two cells apart is never found by the one-wrap
    27-cell search (device.cu:96-119), although its periodic sides are < rmax."""
    L = 250.0
    edges = edges_linear(2.0, 20.0, 9)
    tri = pts([[236.0, 100.0, 100.0], [2.0, 100.0, 100.0], [245.0, 108.0, 100.0]])
    got, _ = oracle.count(tri, tri, tri, edges, [L, L, L])
    want_mi, _ = oracle.brute_force(tri, tri, tri, edges, [L, L, L])
    assert want_mi.sum() == 6  # minimal image: one triangle, 6 orders
    # the pair (236, 2) sits in cells 11 and 0, never neighbours: only p1 = 245
    # (cell 12, which wraps to cell 0) sees both others -> 2 ordered triplets
    assert got.sum() == 2
    # shifting everything by +5 cells away from the wrap restores the count
    tri2 = tri.copy()
    tri2[:, 0] = (tri2[:, 0] + 100.0) % L
    got2, _ = oracle.count(tri2, tri2, tri2, edges, [L, L, L])
    assert got2.sum() == 6


def test_equilateral_and_collinear():
    edges = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    s = 2.5
    tri = pts([[5, 5, 5], [5 + s, 5, 5], [5 + s / 2, 5 + s * np.sqrt(3) / 2, 5]])
    c, w = oracle.count(tri, tri, tri, edges, [20, 20, 20], survey=True)
    assert c[2, 2, 2] == 6 and c.sum() == 6
    # exactly collinear, sides 1, 2, 3 (all exact): sqrtf(1)+sqrtf(4) <= sqrtf(9) -> rejected
    col = pts([[5, 5, 5], [6, 5, 5], [8, 5, 5]])
    c, _ = oracle.count(col, col, col, np.array([0.0, 1.5, 2.5, 3.5]), [20, 20, 20], survey=True)
    assert c.sum() == 0


def test_fp32_triangle_test():
    """Nearly collinear triple: FP64 says triangle, FP32 sqrtf says degenerate
    (device.cu:173 is evaluated in float) -> rejected."""
    a, b = 1.0, 2.0
    h = 1e-5  # height of the apex: slack ~ h^2, far below float resolution
    tri = pts([[0, 0, 0], [a + b, 0, 0], [a, h, 0]])
    tri[:, :3] += 5.0
    d = tri[:, :3]
    s = sorted(float(np.sum((d[i] - d[j]) ** 2)) for i, j in ((0, 1), (0, 2), (1, 2)))
    assert np.sqrt(s[0]) + np.sqrt(s[1]) > np.sqrt(s[2])  # a triangle in FP64
    f = np.sqrt(np.array(s, dtype=np.float32))
    assert f[0] + f[1] <= f[2]  # degenerate in FP32
    c, _ = oracle.count(tri, tri, tri, np.array([0.0, 1.5, 2.5, 3.5]), [20, 20, 20], survey=True)
    assert c.sum() == 0


def test_edges_half_open_and_strict_range():
    """A side exactly on an interior edge goes to the upper bin
    (device.cu:182); sides exactly at rmin or rmax are excluded (device.cu:287)."""
    edges = np.array([1.0, 2.0, 3.0, 4.0])
    # right triangle 3-4-5 scaled by 0.5: sides 1.5, 2, 2.5 -> bins 0, 1 (2 is the edge), 1
    tri = pts([[5, 5, 5], [6.5, 5, 5], [5, 7, 5]])
    c, _ = oracle.count(tri, tri, tri, edges, [20, 20, 20], survey=True)
    assert c[0, 1, 1] == 6 and c.sum() == 6
    # side exactly rmax (4) -> triangle 3-4-5 with sides 3, 4, 5: excluded (4 not < rmax and 5 > rmax)
    tri = pts([[5, 5, 5], [8, 5, 5], [5, 9, 5]])
    c, _ = oracle.count(tri, tri, tri, np.array([1.0, 2.0, 3.0, 4.0, 4.5]), [20, 20, 20], survey=True)
    assert c.sum() == 0
    # side exactly rmin: 1-1-1 equilateral with rmin = 1 -> excluded
    s = 1.0
    tri = pts([[5, 5, 5], [5 + s, 5, 5], [5, 5, 5 + s]])
    c, _ = oracle.count(tri, tri, tri, np.array([1.0, 1.5, 2.0]), [20, 20, 20], survey=True)
    assert c.sum() == 0


def test_duplicates_excluded_and_weights_scale():
    rng = np.random.default_rng(3)
    p = uniform_box(200, 20.0, seed=4)
    p[:, 3] = np.round(rng.uniform(0.5, 2.0, 200) * 256) / 256
    edges = np.array([0.0, 1.0, 2.0, 3.0, 4.0])
    c1, w1 = oracle.count(p, p, p, edges, [20, 20, 20])
    dup = np.concatenate([p, p[:1]])  # duplicate one point: its zero separation never counts
    q = p.copy()
    q[:, 3] *= 2.0
    c2, w2 = oracle.count(q, q, q, edges, [20, 20, 20])
    assert np.array_equal(c1, c2)
    np.testing.assert_allclose(w2, 8.0 * w1, rtol=1e-15, atol=0)
    cd, _ = oracle.count(dup, dup, dup, edges, [20, 20, 20])
    assert cd.sum() >= c1.sum()
    single = p[:1]
    cs, _ = oracle.count(dup, single, single, edges, [20, 20, 20])
    assert cs.sum() == 0  # p2 == p3 == the same point: r23 = 0 is never in range


def test_survey_symmetry_multiples():
    d = uniform_box(300, 30.0, seed=9)
    r = uniform_box(300, 30.0, seed=10)
    edges = edges_linear(0.0, 6.0, 6)
    box = [31.0, 31.0, 31.0]
    ddd, _ = oracle.count(d, d, d, edges, box, survey=True)
    ddr, _ = oracle.count(d, d, r, edges, box, survey=True)
    drr, _ = oracle.count(d, r, r, edges, box, survey=True)
    assert ddd.sum() > 0 and np.all(ddd % 6 == 0)
    assert np.all(ddr % 2 == 0) and np.all(drr % 2 == 0)


def test_sqrsep_evaluation_order():
    """The oracle's sqrSep is fma(dz,dz,fma(dy,dy,dx*dx)): find a pair where
    the uncontracted order lands on the other side of an edge and check that
    the oracle's bin follows the exact FMA emulation."""
    rng = np.random.default_rng(11)
    found = None
    for _ in range(20000):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        dx, dy, dz = (float(t) for t in v)
        exact = oracle.sqrsep_exact(dx, dy, dz)
        plain = (dx * dx + dy * dy) + dz * dz
        if (exact < 1.0) != (plain < 1.0):
            found = (dx, dy, dz, exact)
            break
    assert found is not None
    dx, dy, dz, exact = found
    p1 = pts([[0.0, 0.0, 0.0]])
    p2 = pts([[-dx, -dy, -dz]])  # p1 - p2 = (dx, dy, dz) exactly
    u = np.cross([dx, dy, dz], [0.0, 0.0, 1.0])
    u = 1.25 * u / np.linalg.norm(u)  # p3 perpendicular to p1-p2: r13 = 1.25, r23 ~ 1.6
    p3 = pts([u])
    edges = np.array([0.5, 1.0, 1.5, 2.5])
    box = [10.0, 10.0, 10.0]
    c0, _ = oracle.count(p1, p2, p3, edges, box, survey=True, contract=0)
    c1, _ = oracle.count(p1, p2, p3, edges, box, survey=True, contract=1)
    want_bin = 1 if exact >= 1.0 else 0
    assert c0.sum() == 1 and c1.sum() == 1
    assert np.argwhere(c0)[0].tolist().count(want_bin) >= 1
    assert not np.array_equal(c0, c1)
    assert oracle.sensitivity(p1, p2, edges, box, survey=True) == 1


def test_float32_data_is_contraction_insensitive():
    p = neyman_scott_box(100, 30, 3.0, 150.0, seed=2)
    edges = edges_linear(0.0, 30.0, 15)
    box = [150.0] * 3
    assert oracle.sensitivity(p, p, edges, box) == 0
    c0, _ = oracle.count(p, p, p, edges, box, contract=0)
    c1, _ = oracle.count(p, p, p, edges, box, contract=1)
    c2, _ = oracle.count(p, p, p, edges, box, contract=2)
    assert np.array_equal(c0, c1) and np.array_equal(c0, c2)


@pytest.mark.parametrize("edges,n_out", [(edges_linear(2.0, 20.0, 10), 150), (edges_linear(0.0, 30.0, 15), 372),
                                         (edges_linear(0.0, 40.0, 20), 825), (edges_linear(0.0, 50.0, 25), 1547)])
def test_output_compaction_sizes(edges, n_out):
    nb = len(edges) - 1
    dense = np.arange(nb ** 3, dtype=np.float64)
    out = oracle.compact(edges, dense)
    assert out.size == n_out == oracle.output_bins(edges).shape[0]
    mids = oracle.output_bins(edges)
    assert np.all(mids[:, 2] <= mids[:, 0] + mids[:, 1])


def test_input_validation():
    p = uniform_box(10, 10.0, seed=1)
    with pytest.raises(ValueError):
        oracle.count(p, p, p, np.array([0.0, 2.0, 1.0]), [10, 10, 10])
    with pytest.raises(ValueError):
        oracle.count(p, p, p, np.array([-1.0, 2.0]), [10, 10, 10])
    q = p.copy()
    q[0, 0] = 1000.0
    with pytest.raises(ValueError):
        oracle.count(q, q, q, np.array([0.0, 2.0]), [10, 10, 10])


def test_exact_fma_helper():
    """_fma_exact rounds once: a case where fma differs from mul-then-add."""
    a = 1.0 + 2.0 ** -30
    b = 1.0 - 2.0 ** -30
    c = -1.0
    want = float(Fraction(a) * Fraction(b) + Fraction(c))
    assert oracle._fma_exact(a, b, c) == want == -(2.0 ** -60)
    assert a * b + c == 0.0


# ---------------------------------------------------------------------------
# Known answers for the output compaction (cudaNcorr.cu:72-86) and the
# one-wrap image rule of non-commensurate boxes (device.cu:89-122).


def _compact_transposed(edges, dense):
    """The compaction with the dense index transposed (i_largest major): a
    planted mistake that the known answer must reject."""
    e = np.asarray(edges, dtype=np.float64)
    mid = (e[1:] + e[:-1]) / 2.0
    nb = mid.size
    d = np.asarray(dense, dtype=np.float64).reshape(-1)
    return np.array([d[i + j * nb + k * nb * nb] for i in range(nb) for j in range(i, nb) for k in range(j, nb)
                     if mid[k] <= mid[i] + mid[j]])


def test_compaction_known_answer():
    """edges 0,1,2,3 (mids 0.5, 1.5, 2.5): the sorted bin triples with
    mid_k <= mid_i + mid_j, worked out by hand, are (0,0,0) (0,1,1) (0,2,2)
    (1,1,1) (1,1,2) (1,2,2) (2,2,2); their dense indices k + 3j + 9i are
    0, 4, 8, 13, 14, 17, 26."""
    edges = np.array([0.0, 1.0, 2.0, 3.0])
    dense = np.arange(27, dtype=np.float64)
    want = [0, 4, 8, 13, 14, 17, 26]
    assert oracle.compact(edges, dense).tolist() == want
    mids = oracle.output_bins(edges).tolist()
    assert mids == [[0.5, 0.5, 0.5], [0.5, 1.5, 1.5], [0.5, 2.5, 2.5], [1.5, 1.5, 1.5], [1.5, 1.5, 2.5],
                    [1.5, 2.5, 2.5], [2.5, 2.5, 2.5]]
    # the planted transposition passes the size check but not the known answer
    bad = _compact_transposed(edges, dense)
    assert bad.size == len(want) and bad.tolist() != want


def test_compaction_known_answer_four_bins():
    """edges 0..4 (mids 0.5 .. 3.5): 13 of the 20 sorted triples pass the
    filter; spot values by hand: (0,0,0) -> 0, (0,3,3) -> 15, (1,2,3) -> 27,
    (3,3,3) -> 63, and (0,1,2) (2.5 > 2.0) is dropped."""
    edges = np.arange(5, dtype=np.float64)
    dense = np.arange(64, dtype=np.float64)
    got = oracle.compact(edges, dense).tolist()
    assert got == [0, 5, 10, 15, 21, 22, 26, 27, 31, 42, 43, 47, 63]
    assert 6 not in got  # (0,1,2)


def _grid_points(rng, n, L, span):
    """Points within `span` of the corner of the periodic box, on a 1/64 grid
    (exact differences and squares: every sqrSep order agrees)."""
    x = rng.uniform(-span, span, (n, 3))
    x = np.round(x * 64.0) / 64.0
    x = np.where(x < 0.0, x + L, x)
    p = np.ones((n, 4))
    p[:, :3] = x
    p[:, 3] = np.round(rng.uniform(0.5, 2.0, n) * 256.0) / 256.0
    return p


@pytest.mark.parametrize("ratio", [0.8, 1.5, 2.5, 3.5, 12.5])
def test_non_commensurate_box_matches_wrap_brute_force(ratio):
    """Boxes of L = ratio * rmax: 1, 2, 3, 4 and 13 cells, the last cell
    partial.  The oracle's 27-entry search equals the per-axis restatement of
    the one-wrap rule bin for bin, including the pairs the minimal image
    would count but the reference misses, and the repeated images of grids
    of one and two cells."""
    rng = np.random.default_rng(int(ratio * 10))
    rmax = 2.0
    L = ratio * rmax
    span = min(2.6, L / 2.0)
    # one or two cells hold every point several times over: fewer points
    n = {0.8: 14, 1.5: 40}.get(ratio, 90)
    p = _grid_points(rng, n, L, span)
    q = _grid_points(rng, n - n // 5, L, span)
    edges = np.array([0.25, 0.75, 1.25, 1.6, 2.0])
    box = [L, L, L]
    for a, b, c in ((p, p, p), (p, q, q), (q, p, q)):
        got, gw = oracle.count(a, b, c, edges, box, contract=1)
        want, ww = oracle.brute_force_wrap(a, b, c, edges, box)
        assert want.sum() > 50
        assert np.array_equal(got, want)
        np.testing.assert_allclose(gw, ww, rtol=1e-12, atol=0)


def test_wrap_brute_force_detects_planted_shift_errors():
    """A planted wrong shift makes the independent restatement disagree with
    the oracle: (a) the minimal image (what a correct periodic count would
    give) on a 13-cell box with a partial last cell, (b) the wrap shift with
    the wrong sign."""
    rng = np.random.default_rng(7)
    rmax, L = 2.0, 25.0
    p = _grid_points(rng, 120, L, 2.6)
    edges = np.array([0.25, 0.75, 1.25, 1.6, 2.0])
    got, _ = oracle.count(p, p, p, edges, [L, L, L], contract=1)
    mi, _ = oracle.brute_force(p, p, p, edges, [L, L, L])
    assert not np.array_equal(got, mi)  # (a): the reference misses pairs across the partial cell
    good = oracle.wrap_shifts
    try:
        oracle.wrap_shifts = lambda c1, cq, n, Lx: [-s for s in good(c1, cq, n, Lx)]
        flipped, _ = oracle.brute_force_wrap(p, p, p, edges, [L, L, L])
    finally:
        oracle.wrap_shifts = good
    assert not np.array_equal(got, flipped)  # (b)
    right, _ = oracle.brute_force_wrap(p, p, p, edges, [L, L, L])
    assert np.array_equal(got, right)


def test_wrap_shifts_small_grids():
    """The per-axis rule by hand: one cell is its own neighbour three times
    (shifts -L, 0, +L); in two cells, cell 1 is reached from cell 0 directly
    and across the wrap (shifts -L and 0)."""
    assert oracle.wrap_shifts(0, 0, 1, 10.0) == [-10.0, 0.0, 10.0]
    assert oracle.wrap_shifts(0, 1, 2, 10.0) == [-10.0, 0.0]
    assert oracle.wrap_shifts(1, 0, 2, 10.0) == [0.0, 10.0]
    assert oracle.wrap_shifts(0, 12, 13, 10.0) == [-10.0]
    assert oracle.wrap_shifts(12, 0, 13, 10.0) == [10.0]
    assert oracle.wrap_shifts(5, 8, 13, 10.0) == []
