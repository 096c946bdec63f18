"""CPU check of the mixed path's FP32 bin lookup tables -- no GPU.

nc_mixed_classify runs the kernel's lookup (clamp, log-spaced cell, table
entry, one edge comparison and the guard-band test) on the host.  The truth
is brute force over all squared edges: a value farther than the guard band
g = 80 * 2^-24 * rmax^2 from every FP32 squared edge must get the bin that
counting the edges below it gives (-1 outside (E_0, E_nb)), and the guard
band of the one nearest edge is the only place where -2 (FP64 recheck) may
appear.  Values are random over the whole range and packed around every
edge at multiples of g, plus zero, negatives, denormals and huge values.
"""
import numpy as np
import pytest

from cudancorr_amd._lib import NcError, mixed_classify
from cudancorr_amd.gen import edges_linear, logspace

U = 2.0 ** -24


def _probe_values(E, g, rng):
    r2 = float(E[-1])
    parts = [rng.uniform(-0.01 * r2, 1.6 * r2, 200_000),
             np.exp(rng.uniform(np.log(1e-12 * r2), np.log(4.0 * r2), 200_000)),
             np.array([0.0, -0.0, 1e-45, 1e-40, 1.17e-38, -1e-30, 1e30, 3e38, -3e38, r2, 4.0 * r2])]
    for e in E:
        mult = np.r_[-np.logspace(3, -3, 60), 0.0, np.logspace(-3, 3, 60)]
        parts.append(e + g * mult)
        parts.append(e * (1.0 + np.linspace(-0.05, 0.05, 201)))
        x = np.float32(e)
        parts.append(np.array([np.nextafter(x, np.float32(-np.inf)), x, np.nextafter(x, np.float32(np.inf))],
                              dtype=np.float64))
    return np.concatenate(parts).astype(np.float32)


def _truth(E, d):
    """(bin by counting edges, distance to the nearest edge), in FP64 on the FP32 values."""
    x = d.astype(np.float64)
    below = np.searchsorted(E, x, side="left")  # edges strictly below x
    k = below - 1
    k[(k < 0) | (k >= E.size - 1)] = -1
    dist = np.min(np.abs(x[:, None] - E[None, :]), axis=1)
    return k, dist


def _check(edges, seed=0):
    E = np.asarray(edges, dtype=np.float64) ** 2
    E = E.astype(np.float32).astype(np.float64)  # the kernel's edges are FP32
    g = 80.0 * U * float(edges[-1]) ** 2
    d = _probe_values(E, g, np.random.default_rng(seed))
    got, entries = mixed_classify(edges, d)
    k, dist = _truth(E, d)
    # float rounding of |d - E| and of the scaled band: a relative 1e-6 of g
    clear = dist > g * (1 + 1e-6)
    inside = dist < g * (1 - 1e-6)
    assert np.all(got[inside] == -2), "values in a guard band must be rechecked"
    assert np.all(got[clear] != -2), "values clear of every band must be decided in FP32"
    decided = got != -2
    assert np.all(dist[decided] > g * (1 - 1e-6))
    np.testing.assert_array_equal(got[decided], k[decided])
    assert 2 <= entries <= 4096
    return entries


@pytest.mark.parametrize("edges", [
    edges_linear(0.0, 10.0, 20),
    edges_linear(0.0, 1.0, 40),
    edges_linear(1.0, 10.0, 9),
    edges_linear(0.0, 150.0, 30),
    logspace(0.1, 20.0, 13),        # source/main.cu:21
    logspace(0.05, 10.0, 30),       # lowest squared bin just wider than two guard bands
    logspace(0.005, 1.0, 10),       # 18 octaves of squared edges
    edges_linear(0.0, 1e-3, 10),
    edges_linear(0.0, 3e4, 10),
], ids=["lin20", "lin40", "lin-offset", "lin150", "driver-log", "log30", "log-wide", "tiny", "huge"])
def test_lookup_matches_brute_force(edges):
    _check(np.asarray(edges, dtype=np.float64))


def test_lookup_random_edges():
    rng = np.random.default_rng(5)
    for t in range(20):
        nb = int(rng.integers(1, 41))
        rmax = float(10.0 ** rng.uniform(-2, 3))
        lo = float(rng.choice([0.0, rmax * rng.uniform(1e-3, 0.5)]))
        inner = np.sort(rng.uniform(lo, rmax, nb - 1)) if nb > 1 else np.array([])
        edges = np.r_[lo, inner, rmax]
        if np.any(np.diff(edges) <= 0):
            continue
        try:
            _check(edges, seed=t)
        except NcError as e:  # bins narrower than the guard bands: the path is off, as the kernel's is
            assert e.code == 6  # NC_ERR_LIMIT
            E = edges.astype(np.float64) ** 2
            g = 80.0 * U * rmax ** 2
            assert np.any(np.diff(E.astype(np.float32).astype(np.float64)) <= 2 * g)


def test_narrow_bins_disable_the_path():
    edges = np.array([0.0, 5.0, 5.0 + 1e-7, 10.0])
    with pytest.raises(NcError):
        mixed_classify(edges, np.zeros(1, np.float32))


def test_driver_bins_table_is_small():
    """The log-spaced cells keep the reference driver's bins on the mixed path in a few KB of LDS."""
    assert _check(logspace(0.1, 20.0, 13)) <= 512


def test_non_finite_inputs():
    """NaN is outside every bin; +-inf clamp to the ends of the table."""
    edges = edges_linear(0.0, 10.0, 20)
    got, _ = mixed_classify(edges, np.array([np.nan, np.inf, -np.inf, -np.nan], dtype=np.float32))
    np.testing.assert_array_equal(got, [-1, -1, -1, -1])
