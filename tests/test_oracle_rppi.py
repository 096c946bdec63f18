"""The rppi oracle (oracle.count_rppi, a restatement of device.cu:203-255,
317-379, 434-505 and cudaNcorr.cu:159-216) pinned on the CPU: against an
independent O(n^3) enumeration with no cell list, closed-form triangles, the
reference's quirks, and the output compaction sizes."""
import numpy as np
import pytest

from oracle import oracle


def _box_points(rng, n, box, w=True):
    p = np.c_[rng.uniform(0, box[0], n), rng.uniform(0, box[1], n), rng.uniform(0, box[2], n), np.ones(n)]
    if w:
        p[:, 3] = rng.uniform(0.5, 1.5, n)
    return p


@pytest.mark.parametrize("seed", [1, 2])
def test_box_matches_brute_force(seed):
    """Box a whole number (5) of cells (rpmax, rpmax, npi) per axis: the 27-cell
    search with one wrap finds the nearest periodic image."""
    rng = np.random.default_rng(seed)
    edges = np.array([0.5, 1.5, 3.0, 4.5, 6.0])
    npi = 5
    box = np.array([30.0, 30.0, 25.0])
    p = _box_points(rng, 300, box)
    q = _box_points(rng, 200, box)
    for p2, p3 in ((p, p), (p, q), (q, p)):
        c, w, lost = oracle.count_rppi(p, p2, p3, edges, npi, box, contract=1)
        cb, wb, lostb = oracle.brute_force_rppi(p, p2, p3, edges, npi, box=box)
        assert c.sum() > 1000
        assert np.array_equal(c, cb) and lost == lostb == 0
        assert np.allclose(w, wb, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("ori_x", [0.0, -700.0])
def test_survey_matches_brute_force(ori_x):
    """Cell size sqrtf(3^2 + 4^2) = 5 exactly, so the 27-cell search is complete."""
    rng = np.random.default_rng(3)
    edges = np.array([0.0, 1.0, 2.0, 3.0])
    npi = 4
    box = np.array([21.0, 21.0, 21.0])
    d = np.c_[rng.uniform(0, 20, (260, 3)), rng.uniform(0.5, 1.5, 260)]
    r = np.c_[rng.uniform(0, 20, (200, 3)), np.ones(200)]
    for p2, p3 in ((d, d), (d, r), (r, r)):
        c, w, lost = oracle.count_rppi(d, p2, p3, edges, npi, box, survey=True, ori=(ori_x, 9.0, -4.0), contract=1)
        cb, wb, lostb = oracle.brute_force_rppi(d, p2, p3, edges, npi, box=None, ori_x=ori_x)
        assert c.sum() > 500
        assert np.array_equal(c, cb) and lost == lostb == 0
        assert np.allclose(w, wb, rtol=1e-12, atol=1e-12)


def test_equilateral_in_plane():
    """Three points on an equilateral triangle of side 2 in the plane z = const:
    rp = 2 on every side, pi = 0; 6 ordered triplets in (k,k,k, pi 0, pi 0)."""
    s = 2.0
    pts = np.array([[50.0, 50.0, 40.0, 1.0], [50.0 + s, 50.0, 40.0, 1.0],
                    [50.0 + s / 2, 50.0 + s * np.sqrt(3) / 2, 40.0, 1.0]])
    edges = np.array([1.0, 1.9, 2.1, 3.0])
    c, w, _ = oracle.count_rppi(pts, pts, pts, edges, 4, np.array([100.0, 100.0, 80.0]))
    assert c.sum() == 6 and c[1, 1, 1, 0, 0] == 6


def test_pi_bins_follow_the_sorted_rp_sides():
    """p1 at the origin, p2 at rp 1 with dz 2.5, p3 at rp 2 with dz 0.5:
    rp sides 1 (pi 2.5), 2 (pi 0.5), |p2-p3| (pi 2.0); the pi bins belong to
    the two smallest rp sides."""
    p1 = [10.0, 10.0, 10.0, 1.0]
    p2 = [11.0, 10.0, 12.5, 1.0]
    p3 = [10.0, 12.0, 10.5, 1.0]
    pts = np.array([p1, p2, p3])
    edges = np.array([0.5, 1.5, 2.1, 2.5])
    c, _, _ = oracle.count_rppi(pts[:1], pts[1:], pts[1:], edges, 4, np.array([40.0, 40.0, 40.0]))
    # |p2 - p3| in the plane: sqrt(1 + 4) = 2.236 -> bin 2; pi 2.0 -> bin 2
    assert c.sum() == 2
    assert c[0, 1, 2, 2, 0] == 2  # sorted rp: 1 (pi bin 2), 2 (pi bin 0), 2.236


def test_survey_line_of_sight():
    """Three points far along z from the origin: the line of sight is ~z, so pi
    is ~|dz| and rp the transverse separation."""
    base = np.array([0.0, 0.0, 1000.0])
    pts = np.array([list(base) + [1.0], list(base + [1.0, 0.0, 2.5]) + [1.0], list(base + [0.0, 1.2, 0.3]) + [1.0]])
    box = np.array([50.0, 50.0, 1100.0])
    sh = np.array([20.0, 20.0, 0.0, 0.0])
    q = pts + sh  # shift inside the grid; the LOS uses x - ori.x, y, z
    c, _, _ = oracle.count_rppi(q, q, q, np.array([0.5, 1.1, 1.3, 2.0]), 3, box, survey=True, ori=(20.0, 0.0, 0.0))
    # sides (rp, pi): p1-p2 (1.0, ~2.5), p1-p3 (1.2, ~0.3), p2-p3 (~1.56, ~2.2),
    # whichever point is the primary: the 6 ordered triplets share one bin
    assert c.sum() == 6 and c[0, 1, 2, 2, 0] == 6


def test_only_ori_x_reaches_the_line_of_sight():
    rng = np.random.default_rng(5)
    d = np.c_[rng.uniform(0, 20, (150, 3)) + 100.0, np.ones(150)]
    edges = np.array([0.0, 1.5, 3.0])
    box = np.array([130.0, 130.0, 130.0])
    a, _, _ = oracle.count_rppi(d, d, d, edges, 4, box, survey=True, ori=(-50.0, 0.0, 0.0))
    b, _, _ = oracle.count_rppi(d, d, d, edges, 4, box, survey=True, ori=(-50.0, 77.0, -31.0))
    c, _, _ = oracle.count_rppi(d, d, d, edges, 4, box, survey=True, ori=(-350.0, 0.0, 0.0))
    assert a.sum() > 0
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)


def test_pi_index_alias():
    """pi^2 just below npi^2 rounds to npi^2 in FP32: floor(sqrtf) = npi and the
    flat index moves into the next bin (device.cu:248-252), as the reference's
    arithmetic does."""
    npi = 3
    z2 = np.nextafter(3.5, 0.0)
    dz = z2 - 0.5  # exact
    assert dz * dz < 9.0 and np.sqrt(np.float32(dz * dz)) == 3.0
    p1 = [10.0, 10.0, 0.5, 1.0]
    p2 = [11.0, 10.0, z2, 1.0]   # rp 1, pi^2 < 9 with pi index 3
    p3 = [10.0, 11.9, 0.5, 1.0]  # rp 1.9, pi 0
    edges = np.array([0.5, 1.5, 2.5])
    pts = np.array([p1, p2, p3])
    c, _, lost = oracle.count_rppi(pts[:1], pts[1:], pts[1:], edges, npi, np.array([40.0, 40.0, 42.0]))
    # sorted rp (1, 1.9, |p2-p3| = 2.15): bins (0, 1, 1); pi indices (3, 0): the
    # flat index m + l*npi + (i3 + i2*nb + i1*nb^2)*npi^2 with l = npi is the
    # first pi bin of the next rp triple, (1, 0, 0) with nb = 2
    flat = 0 + 3 * npi + (1 + 1 * 2 + 0 * 4) * npi * npi
    got = np.nonzero(c.reshape(-1))[0]
    assert list(got) == [flat] and lost == 0
    assert c[1, 0, 0, 0, 0] == 2


@pytest.mark.parametrize("lo,nb,npi,n3d", [(2.0, 10, 5, 150), (0.0, 15, 20, 372)])
def test_compaction_sizes(lo, nb, npi, n3d):
    e = np.linspace(lo, 2.0 * nb if lo == 0.0 else 20.0, nb + 1)
    dense = np.arange(nb ** 3 * npi * npi, dtype=np.float64)
    out = oracle.compact_rppi(e, npi, dense)
    assert out.size == n3d * npi * npi
    # first selected bin is (0,0,0): its npi^2 pi bins come first, in (l, m) order
    assert np.array_equal(out[: npi * npi], np.arange(npi * npi))


def test_compaction_known_answer():
    """cudaNcorr.cu:88-105 with edges 0,1,2,3 and 2 pi bins: the selected rp
    triples (by hand, as the threeD compaction) have dense triple indices
    0, 4, 8, 13, 14, 17, 26; each is followed by its pi bins (l, m) in the
    order (0,0) (0,1) (1,0) (1,1), i.e. dense index 4T + 2l + m."""
    edges = np.array([0.0, 1.0, 2.0, 3.0])
    npi = 2
    dense = np.arange(27 * npi * npi, dtype=np.float64)
    want = [4 * t + u for t in (0, 4, 8, 13, 14, 17, 26) for u in range(4)]
    got = oracle.compact_rppi(edges, npi, dense).tolist()
    assert got == want
    assert got[4:8] == [16, 17, 18, 19]  # the second rp triple (0,1,1), past the first rp bin
    # planted mistake: pi bins transposed (m major) -> a different order
    bad = [4 * t + 2 * m + l for t in (0, 4, 8, 13, 14, 17, 26) for l in range(2) for m in range(2)]
    assert bad != got
