"""CPU check of the error bounds behind the mixed (FP32) path -- no GPU.

The mixed kernel (cudancorr_amd/csrc/ncorr_device.hip, MixedPol; its bounds are
set by make_mixed_tables in cudancorr_amd/csrc/ncorr_geometry.hip) decides a pair without FP64 arithmetic when
  * its FP32 r23^2 is farther than g32 = 80uR^2 from every squared edge, and
  * s_a*s_b - dot^2 > tri_t = 80uR^2 * 2R^2 (1 + 1e-5) + 64uR^4, the largest value of
    80uR^2 (s_a + s_b) + 64uR^4 over list entries (triangle far from degenerate),
with u = 2^-24, R = rmax, r = fl32(q - p1), s = fl32(|r|^2),
dot = fl32(r_a . r_b), d32 = fl32(s_a + s_b - 2 dot).  Every other pair is
re-evaluated in FP64.  The kernel works in units scaled by a power of 4
(coordinates times sigma = 2^kk, thresholds and edges times sigma^2 or
sigma^4); the emulation does the same and the decisions are checked to be
identical for every scale.  This test emulates that FP32 arithmetic with numpy and
checks, on random and adversarial pairs (near bin edges, near-collinear
triples, survey-sized coordinates), that every decided pair gets the bin and
the triangle outcome of the reference's FP64 evaluation (device.cu:154-201).
"""
import numpy as np
import pytest

U = 2.0 ** -24


def f32(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32)


def fma32(a, b, c):
    # one rounding of the exact a*b (exact in FP64 for FP32 inputs) plus c;
    # the FP64 intermediate adds a rounding far below the 2x margins
    return f32(np.float64(a) * np.float64(b) + np.float64(c))


def kernel_scale_exponent(edges):
    """kk of the kernel's scale: the power of 4, s32 = 4^kk, with
    rmax^2*s32 in [1, 4)."""
    return -int(np.floor(np.log2(float(edges[-1]) ** 2) / 2.0))


def mixed_decisions(P, A, B, edges, kk=0, guard=80.0):
    R = float(edges[-1])
    sg = 2.0 ** kk
    s32 = sg * sg
    sq = edges * edges
    E32 = f32(f32(sq) * s32)
    g = guard * U * R * R * s32
    R2s = R * R * s32
    tri_t = np.nextafter(np.float32(80.0 * U * R2s * 2.0 * R2s * (1.0 + 1e-5) + 64.0 * U * R2s * R2s),
                         np.float32(np.inf))
    ra = f32((A - P) * sg)
    rb = f32((B - P) * sg)
    sa = fma32(ra[:, 2], ra[:, 2], fma32(ra[:, 1], ra[:, 1], f32(ra[:, 0] * ra[:, 0])))
    sb = fma32(rb[:, 2], rb[:, 2], fma32(rb[:, 1], rb[:, 1], f32(rb[:, 0] * rb[:, 0])))
    dot = fma32(ra[:, 2], rb[:, 2], fma32(ra[:, 1], rb[:, 1], f32(ra[:, 0] * rb[:, 0])))
    ss = f32(sa + sb)
    d32 = fma32(np.float32(-2.0), dot, ss)
    near_edge = np.min(np.abs(d32[:, None].astype(np.float64) - E32[None, :].astype(np.float64)), axis=1) <= g
    k32 = np.searchsorted(E32, d32, side="right") - 1
    valid32 = (k32 >= 0) & (k32 < edges.size - 1)
    delta = fma32(-dot, dot, f32(sa * sb))
    amb_t = delta <= tri_t
    decided = ~near_edge & ~(valid32 & amb_t)
    accepted = decided & valid32
    return decided, accepted, k32


def reference(P, A, B, edges):
    sq = edges * edges

    def sep(X, Y):
        d = X - Y
        return (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]

    d12, d13, d23 = sep(P, A), sep(P, B), sep(A, B)
    inr = (d23 > sq[0]) & (d23 < sq[-1])
    k = np.searchsorted(sq, d23, side="right") - 1
    s = np.sort(np.stack([d12, d13, d23], -1), axis=-1)
    f = np.sqrt(s.astype(np.float32))
    tri = (f[:, 0] + f[:, 1]) > f[:, 2]
    return inr, k, tri


def check(P, A, B, edges, guard=80.0):
    sq = edges * edges
    # only list entries are paired: both within (rmin, rmax) of the primary
    da = np.sum((A - P) ** 2, 1)
    db = np.sum((B - P) ** 2, 1)
    keep = (da > sq[0]) & (da < sq[-1]) & (db > sq[0]) & (db < sq[-1])
    P, A, B = P[keep], A[keep], B[keep]
    decided, accepted, k32 = mixed_decisions(P, A, B, edges, kernel_scale_exponent(edges), guard)
    # the scaling is exact: unscaled and other scales decide identically
    for kk in (0, -7, 9):
        d2, a2, k2 = mixed_decisions(P, A, B, edges, kk, guard)
        assert np.array_equal(d2, decided) and np.array_equal(a2, accepted) and np.array_equal(k2, k32)
    inr, k, tri = reference(P, A, B, edges)
    # decided pairs: out of range iff the reference's value is out of range
    assert np.array_equal(accepted, decided & inr)
    # accepted pairs: same bin and the reference's triangle test passes
    assert np.array_equal(k32[accepted], k[accepted])
    assert np.all(tri[accepted])
    return decided.mean(), accepted.sum()


def around(rng, n, center, R, lo=0.0):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    r = R * np.cbrt(rng.uniform((lo / R) ** 3, 1.0, n))
    return center + v * r[:, None]


@pytest.mark.parametrize("scale", [1.0, 1500.0])
@pytest.mark.parametrize("edges", [np.linspace(0.0, 30.0, 16), np.linspace(0.0, 50.0, 26),
                                   np.linspace(2.0, 20.0, 11)])
def test_random_pairs(scale, edges):
    rng = np.random.default_rng(1)
    n = 200_000
    R = edges[-1]
    P = rng.uniform(0, 3 * R, (n, 3)) + scale
    A = around(rng, n, P, R)
    B = around(rng, n, P, R)
    frac, acc = check(P, A, B, edges)
    assert acc > n // 4
    assert frac > 0.999  # rechecks are rare


@pytest.mark.parametrize("scale", [7.0, 1800.0])
def test_pairs_on_bin_edges(scale):
    """|a - b| placed within a few ulp of every edge."""
    rng = np.random.default_rng(2)
    edges = np.linspace(0.0, 30.0, 16)
    n = 50_000
    P = rng.uniform(0, 60, (n, 3)) + scale
    A = around(rng, n, P, 25.0, lo=5.0)
    e = edges[rng.integers(1, edges.size - 1, n)]
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    B = A + v * (e * (1.0 + rng.normal(0, 1e-7, n)))[:, None]
    frac, _ = check(P, A, B, edges)
    assert frac < 1.0  # some pairs land in the guard band and are rechecked


@pytest.mark.parametrize("scale", [3.0, 1200.0])
def test_near_collinear(scale):
    """a, b and the primary nearly on one line, on the same side (c ~ |a-b|)
    and on opposite sides (c ~ a+b), at offsets across the FP32 threshold."""
    rng = np.random.default_rng(3)
    edges = np.linspace(0.0, 40.0, 21)
    n = 100_000
    P = rng.uniform(0, 80, (n, 3)) + scale
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    w = np.cross(v, rng.normal(size=(n, 3)))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    ta = rng.uniform(1.0, 19.0, n)
    tb = rng.uniform(1.0, 19.0, n) * np.where(rng.uniform(size=n) < 0.5, -1.0, 1.0)
    h = 10.0 ** rng.uniform(-8, 0.5, n)
    A = P + v * ta[:, None]
    B = P + v * tb[:, None] + w * h[:, None]
    frac, acc = check(P, A, B, edges)
    assert acc > 0


@pytest.mark.parametrize("scale", [5.0, 1500.0])
def test_pairs_on_log_edges(scale):
    """Log-spaced edges and |a - b| within ~1e-6 of an edge while a and b sit
    far from the primary, where r23^2 = s_a + s_b - 2 dot loses the most to
    cancellation (errors far above the spread of the pairs around the edge)."""
    rng = np.random.default_rng(4)
    edges = np.logspace(0.0, np.log10(30.0), 16)
    n = 400_000
    P = rng.uniform(0, 60, (n, 3)) + scale
    A = around(rng, n, P, 29.0, lo=15.0)
    e = edges[rng.integers(1, edges.size - 1, n)]
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    B = A + v * (e * (1.0 + rng.normal(0, 1e-6, n)))[:, None]
    frac, _ = check(P, A, B, edges)
    assert frac < 1.0
    # the same pairs without the guard band: some get a wrong bin
    with pytest.raises(AssertionError):
        check(P, A, B, edges, guard=0.0)
