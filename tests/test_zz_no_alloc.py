"""nc_count_noalloc (the allocation-free form of nc_count) against nc_count on
the same device-resident points, through the C-ABI: two calls in a row, the
second with a smaller input, on both arithmetic paths, unit weights (counts
bit-exact) and lognormal weights (sums within 1e-12 of nc_count's and 1e-9 of
the oracle's).  Device memory free before and after a call must not change,
and a count larger than the scratch must fail with NC_ERR_LIMIT."""
import ctypes

import numpy as np
import pytest

from cudancorr_amd import NC_PATH_MIXED, NC_PATH_PLAIN, NcError
from cudancorr_amd.gen import edges_linear, neyman_scott_box, prep_cat_survey, survey_cap
from oracle import oracle

pytestmark = pytest.mark.gpu

H2D = 1


def _hip():
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    return hip


class DevPoints:
    """Points copied to a hipMalloc'ed buffer owned by the test."""

    def __init__(self, hip, pts):
        self.hip = hip
        self.a = np.ascontiguousarray(pts, dtype=np.float64)
        self.n = self.a.shape[0]
        self.p = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.p), max(self.a.nbytes, 32)) == 0
        if self.n:
            assert hip.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, H2D) == 0

    def free(self):
        self.hip.hipFree(self.p)


def _free_mem(hip):
    f, t = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(f), ctypes.byref(t)) == 0
    return f.value


@pytest.mark.parametrize("path", [NC_PATH_PLAIN, NC_PATH_MIXED])
def test_noalloc_box_matches_nc_count(ctx, path):
    hip = _hip()
    L = 200.0
    edges = edges_linear(0.0, 16.0, 8)
    box = [L] * 3
    big = neyman_scott_box(300, 60, 2.0, L, seed=71)     # 18000 points
    small = neyman_scott_box(100, 40, 2.0, L, seed=72)   # 4000 points
    for step, pts in enumerate((big, small)):
        print("noalloc box, path", path, "call", step, "n", pts.shape[0], flush=True)
        d = DevPoints(hip, pts)
        try:
            C = ctx.catalog_device(d.p.value, d.n)
            want_h, want_c = ctx.count(C, C, C, edges, box, path=path)
            C.close()
            free0 = _free_mem(hip)
            h, c = ctx.count_noalloc(d.p.value, d.n, d.p.value, d.n, d.p.value, d.n, edges, box, path=path)
            assert _free_mem(hip) == free0  # nothing allocated on the device
            assert c.sum() > 0
            assert np.array_equal(c, want_c)
            assert np.array_equal(h, want_h)
            if step == 1:
                oc, _ = oracle.count(pts, pts, pts, edges, box)
                assert np.array_equal(c, oc)
        finally:
            d.free()


@pytest.mark.parametrize("path", [NC_PATH_PLAIN, NC_PATH_MIXED])
def test_noalloc_survey_weighted_matches_nc_count(ctx, path):
    """DDR on a weighted survey sample (three distinct buffers of which two
    are the same data), then a smaller DRR."""
    hip = _hip()
    d_, r_ = survey_cap(2500, 4000, seed=73)
    for c in (d_, r_):
        c[:, :3] = 1000.0 + (c[:, :3] - 1000.0) * 0.1
    box, _ = prep_cat_survey(d_, r_)
    edges = edges_linear(0.0, 40.0, 10)
    D = DevPoints(hip, d_)
    R = DevPoints(hip, r_)
    Rs = DevPoints(hip, r_[:1500])
    try:
        for step, (a, b, c) in enumerate(((D, D, R), (D, Rs, Rs))):
            print("noalloc survey, path", path, "call", step, flush=True)
            cats = [ctx.catalog_device(x.p.value, x.n) for x in (a, b, c)]
            cc = {id(x): k for x, k in zip((a, b, c), cats)}
            want_h, _ = ctx.count(cc[id(a)], cc[id(b)], cc[id(c)], edges, box, survey=True, path=path)
            for k in set(cats):
                k.close()
            free0 = _free_mem(hip)
            h, cnt = ctx.count_noalloc(a.p.value, a.n, b.p.value, b.n, c.p.value, c.n, edges, box, survey=True,
                                       path=path)
            assert _free_mem(hip) == free0
            assert not cnt.any()  # weighted terms return no integer counts
            _, ow = oracle.count(a.a, b.a, c.a, edges, box, survey=True)
            nz = ow != 0
            assert nz.sum() > 10
            assert np.array_equal(h[~nz], ow[~nz]) and np.array_equal(want_h[~nz], ow[~nz])
            assert np.max(np.abs(h[nz] / want_h[nz] - 1.0)) <= 1e-12
            assert np.max(np.abs(h[nz] / ow[nz] - 1.0)) <= 1e-9
    finally:
        for x in (D, R, Rs):
            x.free()


def test_noalloc_rejects_counts_past_the_scratch(ctx):
    hip = _hip()
    from cudancorr_amd._lib import lib

    cap = lib().nc_noalloc_capacity()
    assert cap == 256 << 20
    n = cap // 32 + 1024  # more points than the scratch can hold sorted copies of
    pts = np.zeros((n, 4))
    pts[:, :3] = np.random.default_rng(5).uniform(0.0, 100.0, (n, 3))
    pts[:, 3] = 1.0
    print("noalloc oversize, n", n, flush=True)
    d = DevPoints(hip, pts)
    try:
        with pytest.raises(NcError) as e:
            ctx.count_noalloc(d.p.value, d.n, d.p.value, d.n, d.p.value, d.n, edges_linear(0.0, 5.0, 5), [100.0] * 3)
        assert e.value.code == 6  # NC_ERR_LIMIT
        assert "no-alloc scratch" in str(e.value)
    finally:
        d.free()
    # rppi is not offered without allocation
    p, _ = ctx.params(edges_linear(0.0, 5.0, 5), [100.0] * 3, npi=4)
    h = np.zeros(5 ** 3 * 16)
    rc = lib().nc_count_noalloc(ctx.h, ctypes.byref(p), None, 0, None, 0, None, 0, NC_PATH_PLAIN,
                                h.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None)
    assert rc == 1  # NC_ERR_ARG
