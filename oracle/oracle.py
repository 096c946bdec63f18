"""CPU oracle for the three-point triangle count -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this module, and only as the checker / the timed CPU
baseline.  The product library never calls it.

Two independent restatements of the reference's threeD count
(``source/device.cu:154-201``, ``259-315``, ``381-431``;
``source/cudaNcorr.cu:37-86``, ``118-157``):

* ``count``: ctypes wrapper of ``oracle.c`` -- the cell-list loop structure
  of the reference (27 neighbour entries for p2, 27 for p3), OpenMP over
  primaries.  This is the oracle the HIP library is compared with.
* ``brute_force``: an O(n1*n2*n3) enumeration with no cell list at all, used to
  pin ``count`` on small inputs.  Survey mode enumerates all triples directly;
  box mode uses the nearest periodic image of p2 and p3 relative to p1, which
  is what the reference's one-wrap 27-cell search finds when the box is a
  whole number (>= 3) of cells.  With ``exact=True`` every squared separation
  is evaluated as fma(dz,dz,fma(dy,dy,dx*dx)) with exact rational arithmetic
  and a single rounding per operation, so the canonical evaluation order is
  checked independently of any FMA hardware.
"""
from __future__ import annotations

import ctypes
import os
from fractions import Fraction

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

ERRORS = {1: "invalid argument", 2: "invalid bin edges", 3: "point outside the cell grid", 4: "out of memory"}


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run __graft_entry__.build())")
        lib = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        lib.nc_oracle_count.argtypes = [dp, ctypes.c_long, dp, ctypes.c_long, dp, ctypes.c_long, dp, ctypes.c_int,
                                        dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_uint64), dp]
        lib.nc_oracle_count.restype = ctypes.c_int
        lib.nc_oracle_sensitivity.argtypes = [dp, ctypes.c_long, dp, ctypes.c_long, dp, ctypes.c_int, dp, ctypes.c_int]
        lib.nc_oracle_sensitivity.restype = ctypes.c_long
        lib.nc_oracle_compact.argtypes = [dp, ctypes.c_int, dp, dp, ctypes.c_int]
        lib.nc_oracle_compact.restype = ctypes.c_int
        lib.nc_oracle_count_rppi.argtypes = [dp, ctypes.c_long, dp, ctypes.c_long, dp, ctypes.c_long, dp,
                                             ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int,
                                             ctypes.c_int, ctypes.POINTER(ctypes.c_uint64), dp,
                                             ctypes.POINTER(ctypes.c_long)]
        lib.nc_oracle_count_rppi.restype = ctypes.c_int
        lib.nc_oracle_compact_rppi.argtypes = [dp, ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int]
        lib.nc_oracle_compact_rppi.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _pts(p):
    a = np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(-1, 4))
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.shape[0]


def count(p1, p2, p3, edges, box, survey=False, contract=0, nthreads=0):
    """Ordered-triplet histogram of the reference semantics.

    p1, p2, p3: (n,4) float64 arrays of x,y,z,w.  edges: nb+1 bin edges in r.
    box: (3,) box size (survey mode: the padded bounding box of
    ``survey::prep_cat_survey``).  Returns (counts uint64 (nb,nb,nb),
    wsum float64 (nb,nb,nb)) indexed [i_smallest, i_middle, i_largest].
    """
    lib = _lib()
    a1, q1, n1 = _pts(p1)
    a2, q2, n2 = _pts(p2)
    a3, q3, n3 = _pts(p3)
    e = np.ascontiguousarray(edges, dtype=np.float64)
    b = np.ascontiguousarray(box, dtype=np.float64)
    nb = e.size - 1
    nh = max(nb, 0) ** 3
    cnt = np.zeros(max(nh, 1), dtype=np.uint64)
    ws = np.zeros(max(nh, 1), dtype=np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.nc_oracle_count(q1, n1, q2, n2, q3, n3, e.ctypes.data_as(dp), e.size, b.ctypes.data_as(dp),
                             int(bool(survey)), int(contract), int(nthreads),
                             cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), ws.ctypes.data_as(dp))
    del a1, a2, a3
    if rc != 0:
        raise ValueError(f"oracle: {ERRORS.get(rc, rc)}")
    return cnt[:nh].reshape(nb, nb, nb), ws[:nh].reshape(nb, nb, nb)


def sensitivity(p1, p2, edges, box, survey=False):
    """Number of (p1,p2) pairs whose bin could depend on the FMA contraction order."""
    lib = _lib()
    a1, q1, n1 = _pts(p1)
    a2, q2, n2 = _pts(p2)
    e = np.ascontiguousarray(edges, dtype=np.float64)
    b = np.ascontiguousarray(box, dtype=np.float64)
    dp = ctypes.POINTER(ctypes.c_double)
    r = lib.nc_oracle_sensitivity(q1, n1, q2, n2, e.ctypes.data_as(dp), e.size, b.ctypes.data_as(dp), int(bool(survey)))
    del a1, a2
    if r < 0:
        raise ValueError(f"oracle: {ERRORS.get(-r, r)}")
    return int(r)


def compact(edges, dense):
    """Output compaction of cudaNcorr.cu:72-86 (sorted bins with mid_k <= mid_i + mid_j)."""
    lib = _lib()
    e = np.ascontiguousarray(edges, dtype=np.float64)
    d = np.ascontiguousarray(np.asarray(dense, dtype=np.float64).reshape(-1))
    dp = ctypes.POINTER(ctypes.c_double)
    n = lib.nc_oracle_compact(e.ctypes.data_as(dp), e.size, None, None, 0)
    out = np.zeros(max(n, 1), dtype=np.float64)
    lib.nc_oracle_compact(e.ctypes.data_as(dp), e.size, d.ctypes.data_as(dp), out.ctypes.data_as(dp), n)
    return out[:n]


def count_rppi(p1, p2, p3, edges, npi, box, survey=False, ori=(0.0, 0.0, 0.0), contract=0, nthreads=0):
    """Ordered-triplet rppi histogram of the reference semantics
    (device.cu:203-255, 317-379, 434-505; cudaNcorr.cu:159-216).

    edges: nb+1 rp edges; npi: number of unit-width pi bins.  Returns (counts
    uint64 (nb,nb,nb,npi,npi), wsum float64 (same), dropped) indexed
    [rp bin of the smallest, middle, largest rp side, pi bin of the smallest,
    of the middle]; dropped = triplets whose flat index falls past the
    histogram (the reference writes out of bounds there).
    """
    lib = _lib()
    a1, q1, n1 = _pts(p1)
    a2, q2, n2 = _pts(p2)
    a3, q3, n3 = _pts(p3)
    e = np.ascontiguousarray(edges, dtype=np.float64)
    b = np.ascontiguousarray(box, dtype=np.float64)
    o = np.ascontiguousarray(ori, dtype=np.float64)
    nb = e.size - 1
    nh = max(nb, 0) ** 3 * npi * npi
    cnt = np.zeros(max(nh, 1), dtype=np.uint64)
    ws = np.zeros(max(nh, 1), dtype=np.float64)
    lost = ctypes.c_long(0)
    dp = ctypes.POINTER(ctypes.c_double)
    rc = lib.nc_oracle_count_rppi(q1, n1, q2, n2, q3, n3, e.ctypes.data_as(dp), e.size, int(npi),
                                  b.ctypes.data_as(dp), int(bool(survey)), o.ctypes.data_as(dp), int(contract),
                                  int(nthreads), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                  ws.ctypes.data_as(dp), ctypes.byref(lost))
    del a1, a2, a3
    if rc != 0:
        raise ValueError(f"oracle: {ERRORS.get(rc, rc)}")
    shape = (nb, nb, nb, npi, npi)
    return cnt[:nh].reshape(shape), ws[:nh].reshape(shape), int(lost.value)


def compact_rppi(edges, npi, dense):
    """Output compaction of cudaNcorr.cu:88-105."""
    lib = _lib()
    e = np.ascontiguousarray(edges, dtype=np.float64)
    d = np.ascontiguousarray(np.asarray(dense, dtype=np.float64).reshape(-1))
    dp = ctypes.POINTER(ctypes.c_double)
    n = lib.nc_oracle_compact_rppi(e.ctypes.data_as(dp), e.size, int(npi), None, None, 0)
    out = np.zeros(max(n, 1), dtype=np.float64)
    lib.nc_oracle_compact_rppi(e.ctypes.data_as(dp), e.size, int(npi), d.ctypes.data_as(dp),
                               out.ctypes.data_as(dp), n)
    return out[:n]


# ---------------------------------------------------------------------------
# brute force


def _fma_exact(a: float, b: float, c: float) -> float:
    """fma with one rounding: exact rational a*b+c rounded to nearest double."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def sqrsep_exact(dx: float, dy: float, dz: float) -> float:
    """fma(dz,dz, fma(dy,dy, dx*dx)) with exact intermediate rounding."""
    return _fma_exact(dz, dz, _fma_exact(dy, dy, dx * dx))


def _sep2(a, b, exact):
    d = a[..., None, :3] - b[None, :, :3]
    if exact:
        out = np.empty(d.shape[:-1])
        for idx in np.ndindex(out.shape):
            out[idx] = sqrsep_exact(*map(float, d[idx]))
        return out
    return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]


def _images(p, center, box):
    """Per-axis nearest periodic image of every point of p relative to center."""
    q = p.copy()
    for ax in range(3):
        L = box[ax]
        best = q[:, ax].copy()
        for s in (-L, L):
            cand = p[:, ax] + s
            better = np.abs(center[ax] - cand) < np.abs(center[ax] - best)
            best = np.where(better, cand, best)
        q[:, ax] = best
    return q


def brute_force(p1, p2, p3, edges, box=None, exact=False):
    """Independent O(n^3) restatement (see module docstring).

    box=None: survey mode (no periodic images).  Otherwise the periodic box,
    which must be a whole number >= 3 of cells for the reference to agree.
    Returns (counts uint64 (nb,nb,nb), wsum float64 (nb,nb,nb)).
    """
    p1 = np.asarray(p1, dtype=np.float64).reshape(-1, 4)
    p2 = np.asarray(p2, dtype=np.float64).reshape(-1, 4)
    p3 = np.asarray(p3, dtype=np.float64).reshape(-1, 4)
    e = np.asarray(edges, dtype=np.float64)
    sq = e * e
    nb = e.size - 1
    cnt = np.zeros((nb, nb, nb), dtype=np.uint64)
    ws = np.zeros((nb, nb, nb), dtype=np.float64)
    lo, hi = sq[0], sq[-1]
    for a in p1:
        q2 = p2 if box is None else _images(p2, a, box)
        q3 = p3 if box is None else _images(p3, a, box)
        r12 = _sep2(a[None], q2, exact)[0]
        r13 = _sep2(a[None], q3, exact)[0]
        j = np.nonzero((r12 > lo) & (r12 < hi))[0]
        k = np.nonzero((r13 > lo) & (r13 < hi))[0]
        if j.size == 0 or k.size == 0:
            continue
        r23 = _sep2(q2[j], q3[k], exact)
        A = np.broadcast_to(r12[j][:, None], r23.shape)
        B = np.broadcast_to(r13[k][None, :], r23.shape)
        ok = (r23 > lo) & (r23 < hi)
        s = np.sort(np.stack([A, B, r23], -1), axis=-1)
        f = np.sqrt(s.astype(np.float32))
        ok &= (f[..., 0] + f[..., 1]) > f[..., 2]
        bins = np.searchsorted(sq, s, side="right") - 1
        w = (a[3] * q2[j, 3])[:, None] * q3[k, 3][None, :]
        for jj, kk in zip(*np.nonzero(ok)):
            i1, i2, i3 = bins[jj, kk]
            cnt[i1, i2, i3] += 1
            ws[i1, i2, i3] += w[jj, kk]
    return cnt, ws


def wrap_shifts(c1, cq, n, L):
    """Image shifts under which a point of cell cq is a neighbour of cell c1
    along one axis of n cells: the reference's one-wrap rule
    (device.cu:89-122, cudaNcorr.cu:37-48) restated per axis.  Entry c1 + d
    (d = -1, 0, 1) is cell c1 + d, except that n wraps to 0 with shift +L and
    -1 to n - 1 with shift -L.  With n = 1 or 2 a cell is reached more than
    once (three or two images)."""
    out = []
    for d in (-1, 0, 1):
        e, s = c1 + d, 0.0
        if e == n:
            e, s = 0, L
        if e == -1:
            e, s = n - 1, -L
        if e == cq:
            out.append(s)
    return out


def _axis_cells(v, cs, n):
    c = np.trunc(v / cs).astype(np.int64)  # cudaNcorr.cu:50-58: int(x / cellSize)
    c[c == n] = n - 1
    return c


def brute_force_wrap(p1, p2, p3, edges, box):
    """Independent restatement of the box count for ANY box, commensurate or
    not, with no 27-entry loop: for each primary, every image q + s of every
    secondary whose per-axis shift s is one of wrap_shifts(cell(p1), cell(q)),
    filtered on r12 / r13, then all (p2 image, p3 image) pairs filtered on
    r23 (from the image positions), the FP32 triangle test and half-open bins
    on the sorted squared sides.  Cells are ceil(L / rmax) per axis of size
    rmax (cudaNcorr.cu:137-140).  Squared separations in the uncontracted
    order (the oracle's contract=1).  Returns (counts, wsum) as ``count``."""
    p1 = np.asarray(p1, dtype=np.float64).reshape(-1, 4)
    p2 = np.asarray(p2, dtype=np.float64).reshape(-1, 4)
    p3 = np.asarray(p3, dtype=np.float64).reshape(-1, 4)
    e = np.asarray(edges, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64)
    sq = e * e
    nb = e.size - 1
    cs = e[-1]
    n = [int(np.ceil(box[a] / cs)) for a in range(3)]
    cells = [np.stack([_axis_cells(p[:, a], cs, n[a]) for a in range(3)], 1) for p in (p1, p2, p3)]
    lo, hi = sq[0], sq[-1]
    cnt = np.zeros((nb, nb, nb), dtype=np.uint64)
    ws = np.zeros((nb, nb, nb), dtype=np.float64)

    def images(a, ca, q, cq):
        """(index, image position) of every neighbour image of q around a, r in range."""
        idx, pos = [], []
        for j in range(q.shape[0]):
            per_axis = [wrap_shifts(int(ca[ax]), int(cq[j, ax]), n[ax], box[ax]) for ax in range(3)]
            for sx in per_axis[0]:
                for sy in per_axis[1]:
                    for sz in per_axis[2]:
                        idx.append(j)
                        pos.append((q[j, 0] + sx, q[j, 1] + sy, q[j, 2] + sz))
        if not idx:
            return np.zeros(0, dtype=np.int64), np.zeros((0, 3)), np.zeros(0)
        pos = np.array(pos)
        d = a[None, :3] - pos
        r = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        keep = (r > lo) & (r < hi)
        return np.array(idx)[keep], pos[keep], r[keep]

    for i, a in enumerate(p1):
        j2, q2, r12 = images(a, cells[0][i], p2, cells[1])
        j3, q3, r13 = images(a, cells[0][i], p3, cells[2])
        if j2.size == 0 or j3.size == 0:
            continue
        d = q2[:, None, :] - q3[None, :, :]
        r23 = (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]
        s = np.sort(np.stack([np.broadcast_to(r12[:, None], r23.shape), np.broadcast_to(r13[None, :], r23.shape),
                              r23], -1), axis=-1)
        f = np.sqrt(s.astype(np.float32))
        ok = (r23 > lo) & (r23 < hi) & ((f[..., 0] + f[..., 1]) > f[..., 2])
        bins = np.searchsorted(sq, s, side="right") - 1
        w = (a[3] * p2[j2, 3])[:, None] * p3[j3, 3][None, :]
        for jj, kk in zip(*np.nonzero(ok)):
            b1, b2, b3 = bins[jj, kk]
            cnt[b1, b2, b3] += 1
            ws[b1, b2, b3] += w[jj, kk]
    return cnt, ws


def output_bins(edges):
    """Bin mids of the compact output (cudaNcorr.cu:145-154), as an (m,3) array."""
    e = np.asarray(edges, dtype=np.float64)
    mid = [(e[i + 1] + e[i]) / 2.0 for i in range(e.size - 1)]
    nb = len(mid)
    out = []
    for i in range(nb):
        for j in range(i, nb):
            for k in range(j, nb):
                if mid[k] <= mid[i] + mid[j]:
                    out.append((mid[i], mid[j], mid[k]))
    return np.array(out, dtype=np.float64).reshape(-1, 3)


def _sort3_carry(r, q):
    """The compare-and-swap network of device.cu:215-229 on the last axis of r
    (3 values), applying the same swaps to q."""
    r = r.copy()
    q = q.copy()
    for i, j in ((0, 1), (0, 2), (1, 2)):
        sw = r[..., i] > r[..., j]
        ri, rj = r[..., i].copy(), r[..., j].copy()
        qi, qj = q[..., i].copy(), q[..., j].copy()
        r[..., i] = np.where(sw, rj, ri)
        r[..., j] = np.where(sw, ri, rj)
        q[..., i] = np.where(sw, qj, qi)
        q[..., j] = np.where(sw, qi, qj)
    return r, q


def brute_force_rppi(p1, p2, p3, edges, npi, box=None, ori_x=0.0):
    """Independent O(n^3) restatement of the rppi count with no cell list, in
    the no-contraction evaluation order (the oracle's contract=1).

    box=None: survey mode (LOS of each triplet from (ori_x, 0, 0)); the cell
    size sqrtf(rpmax^2 + npi^2) must not round below the true value.  Box
    mode: nearest periodic image relative to p1; the box must be a whole
    number >= 3 of cells (rpmax, rpmax, npi).  Returns (counts, wsum,
    dropped) as count_rppi.
    """
    p1 = np.asarray(p1, dtype=np.float64).reshape(-1, 4)
    p2 = np.asarray(p2, dtype=np.float64).reshape(-1, 4)
    p3 = np.asarray(p3, dtype=np.float64).reshape(-1, 4)
    e = np.asarray(edges, dtype=np.float64)
    sq = e * e
    nb = e.size - 1
    nh = nb ** 3 * npi * npi
    cnt = np.zeros(nh, dtype=np.uint64)
    ws = np.zeros(nh, dtype=np.float64)
    lost = 0
    rlo, rhi = sq[0], sq[-1]
    pimax2 = float(npi) * float(npi)
    smax2 = e[-1] * e[-1] + pimax2

    def rp2(d):
        return d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]

    def s2(d):
        return (d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]

    def dot(l, d):
        return (l[..., 0] * d[..., 0] + l[..., 1] * d[..., 1]) + l[..., 2] * d[..., 2]

    for a in p1:
        q2 = p2 if box is None else _images(p2, a, box)
        q3 = p3 if box is None else _images(p3, a, box)
        d12 = a[None, :3] - q2[:, :3]
        d13 = a[None, :3] - q3[:, :3]
        if box is not None:
            z12 = d12[:, 2] * d12[:, 2]
            z13 = d13[:, 2] * d13[:, 2]
            r12 = rp2(d12)
            r13 = rp2(d13)
            j = np.nonzero((z12 < pimax2) & (r12 < rhi) & (r12 > rlo))[0]
            k = np.nonzero((z13 < pimax2) & (r13 < rhi) & (r13 > rlo))[0]
            if j.size == 0 or k.size == 0:
                continue
            d23 = q2[j][:, None, :3] - q3[k][None, :, :3]
            z23 = d23[..., 2] * d23[..., 2]
            r23 = rp2(d23)
            ok = (z23 < pimax2) & (r23 < rhi) & (r23 > rlo)
            R = np.stack([np.broadcast_to(r12[j][:, None], r23.shape), np.broadcast_to(r13[k][None, :], r23.shape),
                          r23], -1)
            Q = np.stack([np.broadcast_to(z12[j][:, None], r23.shape), np.broadcast_to(z13[k][None, :], r23.shape),
                          z23], -1)
        else:
            s12 = s2(d12)
            s13 = s2(d13)
            j = np.nonzero((s12 < smax2) & (s12 > 0.0))[0]
            k = np.nonzero((s13 < smax2) & (s13 > 0.0))[0]
            if j.size == 0 or k.size == 0:
                continue
            b2 = q2[j][:, None, :3]
            b3 = q3[k][None, :, :3]
            d23 = b2 - b3
            s23 = s2(d23)
            ok = (s23 < smax2) & (s23 > 0.0)
            o = np.array([ori_x, 0.0, 0.0])
            los = ((a[:3] - o) + (b2 - o)) + (b3 - o)
            sl = dot(los, los)
            lim = pimax2 * sl
            t1 = dot(los, a[:3] - b2) ** 2
            t2 = dot(los, a[:3] - b3) ** 2
            t3 = dot(los, d23) ** 2
            ok &= ~(lim < t1) & ~(lim < t2) & ~(lim < t3)
            with np.errstate(divide="ignore", invalid="ignore"):
                Q = np.stack([t1 / sl, t2 / sl, t3 / sl], -1)
            S = np.stack([np.broadcast_to(s12[j][:, None], s23.shape), np.broadcast_to(s13[k][None, :], s23.shape),
                          s23], -1)
            R = S - Q
            ok &= np.all((R < rhi) & (R > rlo), axis=-1) & (Q[..., 0] < pimax2) & (Q[..., 1] < pimax2)
        Rs, Qs = _sort3_carry(R, Q)
        f = np.sqrt(Rs.astype(np.float32))
        ok &= (f[..., 0] + f[..., 1]) > f[..., 2]
        bins = np.searchsorted(sq, Rs, side="right") - 1
        pib = np.floor(np.sqrt(Qs.astype(np.float32))).astype(np.int64)
        idx = (pib[..., 1] + pib[..., 0] * npi + bins[..., 2] * npi * npi + bins[..., 1] * nb * npi * npi +
               bins[..., 0] * nb * nb * npi * npi)
        w = (a[3] * q2[j, 3])[:, None] * q3[k, 3][None, :]
        for jj, kk in zip(*np.nonzero(ok)):
            t = int(idx[jj, kk])
            if t >= nh:
                lost += 1
                continue
            cnt[t] += 1
            ws[t] += w[jj, kk]
    shape = (nb, nb, nb, npi, npi)
    return cnt.reshape(shape), ws.reshape(shape), lost
