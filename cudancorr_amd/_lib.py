"""ctypes binding of the C-ABI in include/ncorr_hip.h (libncorr_hip.so).

The library is loaded from ``cudancorr_amd/lib``; there is no fallback of any
kind -- if the shared object is missing or a call fails, an exception is
raised.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libncorr_hip.so")

NC_PATH_PLAIN = 0
NC_PATH_MIXED = 1
NC_BIN_THREED = 0
NC_BIN_RPPI = 1
NC_ACCUM_FP64 = 0
NC_ACCUM_EXACT = 1
NC_ACCUM_FP32 = 2

STATUS = {0: "ok", 1: "invalid argument", 2: "invalid bin edges", 3: "point outside the cell grid",
          4: "out of memory", 5: "HIP error", 6: "size limit"}

# Every entry point declared in include/ncorr_hip.h.
EXPORTS = (
    "nc_create", "nc_destroy", "nc_last_error", "nc_catalog_create", "nc_catalog_wrap_device",
    "nc_catalog_destroy", "nc_catalog_size", "nc_hist_size", "nc_compact_size", "nc_hist_len", "nc_count",
    "nc_count_host", "nc_count_batched", "nc_last_timing", "nc_set_tuning", "nc_get_tuning", "nc_autotune",
    "nc_mixed_classify", "nc_count_noalloc", "nc_noalloc_capacity", "nc_set_primary_slice", "nc_set_accumulation",
)


class NcParams(ctypes.Structure):
    _fields_ = [("box", ctypes.c_double * 3), ("survey", ctypes.c_int), ("nedges", ctypes.c_int),
                ("edges", ctypes.POINTER(ctypes.c_double)), ("bin_type", ctypes.c_int),
                ("num_pi_bins", ctypes.c_int), ("ori", ctypes.c_double * 3)]


class NcTiming(ctypes.Structure):
    _fields_ = [("ms_build", ctypes.c_float), ("ms_count", ctypes.c_float), ("ms_total", ctypes.c_float),
                ("n_kernel", ctypes.c_int64),
                ("hist_layout", ctypes.c_int32),
                ("lds_bytes", ctypes.c_int32), ("path", ctypes.c_int32), ("tri_terms", ctypes.c_int32),
                ("swap_terms", ctypes.c_int32), ("partner_terms", ctypes.c_int32)]


class NcTuning(ctypes.Structure):
    _fields_ = [("waves_per_group", ctypes.c_int), ("chunk", ctypes.c_int), ("wgs_per_cu", ctypes.c_int),
                ("list_chunk", ctypes.c_int), ("ms", ctypes.c_float), ("from_cache", ctypes.c_int)]


class NcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_LIB = None


def lib():
    """Load libncorr_hip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    up = ctypes.POINTER(ctypes.c_uint64)
    i64 = ctypes.c_int64
    L.nc_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    L.nc_destroy.argtypes = [vp]
    L.nc_last_error.argtypes = [vp]
    L.nc_last_error.restype = ctypes.c_char_p
    L.nc_catalog_create.argtypes = [vp, dp, i64, ctypes.POINTER(vp)]
    L.nc_catalog_wrap_device.argtypes = [vp, vp, i64, ctypes.POINTER(vp)]
    L.nc_catalog_destroy.argtypes = [vp]
    L.nc_catalog_size.argtypes = [vp]
    L.nc_catalog_size.restype = i64
    L.nc_hist_size.argtypes = [ctypes.c_int]
    L.nc_hist_size.restype = i64
    L.nc_hist_len.argtypes = [ctypes.POINTER(NcParams)]
    L.nc_hist_len.restype = i64
    L.nc_compact_size.argtypes = [ctypes.c_int]
    L.nc_compact_size.restype = i64
    L.nc_count.argtypes = [vp, ctypes.POINTER(NcParams), vp, vp, vp, ctypes.c_int, dp, up]
    L.nc_count_host.argtypes = [vp, ctypes.POINTER(NcParams), dp, i64, dp, i64, dp, i64, ctypes.c_int, dp, up]
    L.nc_count_batched.argtypes = [vp, ctypes.POINTER(NcParams), ctypes.c_int, ctypes.POINTER(vp),
                                   ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.c_int, dp, up]
    L.nc_last_timing.argtypes = [vp, ctypes.POINTER(NcTiming)]
    L.nc_set_tuning.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.nc_get_tuning.argtypes = [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.nc_autotune.argtypes = [vp, ctypes.POINTER(NcParams), ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(vp),
                              ctypes.POINTER(vp), ctypes.c_int, ctypes.POINTER(NcTuning)]
    L.nc_mixed_classify.argtypes = [dp, ctypes.c_int, ctypes.POINTER(ctypes.c_float), i64,
                                    ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
    L.nc_count_noalloc.argtypes = [vp, ctypes.POINTER(NcParams), vp, i64, vp, i64, vp, i64, ctypes.c_int, dp, up]
    L.nc_noalloc_capacity.argtypes = []
    L.nc_noalloc_capacity.restype = i64
    L.nc_set_primary_slice.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.nc_set_accumulation.argtypes = [vp, ctypes.c_int]
    _LIB = L
    return L


def mixed_classify(edges, d):
    """The mixed path's FP32 bin lookup evaluated on the host (no GPU): bin of
    each FP32 squared separation in d, -1 out of range, -2 in an edge's guard
    band (FP64 recheck).  Returns (bins, table entries); NcError when the mixed
    path is disabled for these edges."""
    e = np.ascontiguousarray(edges, dtype=np.float64)
    x = np.ascontiguousarray(d, dtype=np.float32)
    out = np.empty(x.size, dtype=np.int32)
    n = ctypes.c_int32(0)
    rc = lib().nc_mixed_classify(_dp(e), e.size, x.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), x.size,
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ctypes.byref(n))
    if rc:
        raise NcError(rc, "nc_mixed_classify")
    return out, n.value


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _points(p):
    a = np.ascontiguousarray(np.asarray(p, dtype=np.float64).reshape(-1, 4))
    return a


class Catalog:
    """Device-resident catalogue of double4 points (nc_catalog)."""

    def __init__(self, ctx: "Context", points=None, device_ptr=None, n=0):
        """Upload host points, or wrap n double4 points at device address
        device_ptr (not copied; the memory must outlive the catalogue)."""
        self.ctx = ctx
        h = ctypes.c_void_p()
        if device_ptr is not None:
            self.n = int(n)
            ctx._check(lib().nc_catalog_wrap_device(ctx.h, ctypes.c_void_p(device_ptr), self.n, ctypes.byref(h)))
        else:
            a = _points(points)
            self.n = a.shape[0]
            ctx._check(lib().nc_catalog_create(ctx.h, _dp(a), self.n, ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().nc_catalog_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """An nc_context on one HIP device."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        rc = lib().nc_create(ctypes.byref(h), device)
        if rc:
            raise NcError(rc, "nc_create failed")
        self.h = h

    def _check(self, rc):
        if rc:
            raise NcError(rc, lib().nc_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().nc_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def params(edges, box, survey=False, npi=0, ori=(0.0, 0.0, 0.0)):
        """nc_params for threeD binning, or rppi binning when npi > 0 (npi
        unit-width pi bins; survey line of sight from ori)."""
        e = np.ascontiguousarray(edges, dtype=np.float64)
        p = NcParams()
        for i in range(3):
            p.box[i] = float(box[i])
            p.ori[i] = float(ori[i])
        p.survey = int(bool(survey))
        p.nedges = e.size
        p.edges = _dp(e)
        p.bin_type = NC_BIN_RPPI if npi > 0 else NC_BIN_THREED
        p.num_pi_bins = int(npi)
        return p, e

    @staticmethod
    def _shape(nb, npi):
        return (nb, nb, nb, npi, npi) if npi > 0 else (nb, nb, nb)

    def catalog(self, points) -> Catalog:
        return Catalog(self, points)

    def catalog_device(self, device_ptr, n) -> Catalog:
        """nc_catalog_wrap_device: n double4 points already in device memory."""
        return Catalog(self, device_ptr=device_ptr, n=n)

    def count(self, c1, c2, c3, edges, box, survey=False, path=NC_PATH_PLAIN, npi=0, ori=(0.0, 0.0, 0.0)):
        """nc_count on catalogues (Catalog or point arrays).  Returns
        (hist float64, counts uint64), shaped (nb,nb,nb) or, for rppi (npi >
        0), (nb,nb,nb,npi,npi)."""
        owned = []
        cats = []
        seen = {}
        for c in (c1, c2, c3):
            if isinstance(c, Catalog):
                cats.append(c)
            else:
                key = id(c)
                if key not in seen:
                    seen[key] = Catalog(self, c)
                    owned.append(seen[key])
                cats.append(seen[key])
        p, e = self.params(edges, box, survey, npi, ori)
        shape = self._shape(e.size - 1, npi)
        hist = np.zeros(max(int(np.prod(shape)), 1), dtype=np.float64)
        cnt = np.zeros(hist.size, dtype=np.uint64)
        try:
            self._check(lib().nc_count(self.h, ctypes.byref(p), cats[0].h, cats[1].h, cats[2].h, int(path),
                                       _dp(hist), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        finally:
            for c in owned:
                c.close()
        n = int(np.prod(shape))
        return hist[:n].reshape(shape), cnt[:n].reshape(shape)

    def count_host(self, p1, p2, p3, edges, box, survey=False, path=NC_PATH_PLAIN, npi=0, ori=(0.0, 0.0, 0.0)):
        """nc_count_host: the drop-in host-buffer form."""
        a1, a2, a3 = _points(p1), None, None
        a2 = a1 if p2 is p1 else _points(p2)
        a3 = a2 if p3 is p2 else (a1 if p3 is p1 else _points(p3))
        p, e = self.params(edges, box, survey, npi, ori)
        shape = self._shape(e.size - 1, npi)
        n = int(np.prod(shape))
        hist = np.zeros(max(n, 1), dtype=np.float64)
        cnt = np.zeros(hist.size, dtype=np.uint64)
        self._check(lib().nc_count_host(self.h, ctypes.byref(p), _dp(a1), a1.shape[0], _dp(a2), a2.shape[0],
                                        _dp(a3), a3.shape[0], int(path), _dp(hist),
                                        cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return hist[:n].reshape(shape), cnt[:n].reshape(shape)

    def count_noalloc(self, d1, n1, d2, n2, d3, n3, edges, box, survey=False, path=NC_PATH_PLAIN):
        """nc_count_noalloc on device buffers of double4 points (integer
        device addresses; equal address and size = one catalogue).  Returns
        (hist, counts) shaped (nb,nb,nb)."""
        p, e = self.params(edges, box, survey)
        shape = self._shape(e.size - 1, 0)
        n = int(np.prod(shape))
        hist = np.zeros(max(n, 1), dtype=np.float64)
        cnt = np.zeros(hist.size, dtype=np.uint64)
        self._check(lib().nc_count_noalloc(self.h, ctypes.byref(p), d1, int(n1), d2, int(n2), d3, int(n3), int(path),
                                           _dp(hist), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return hist[:n].reshape(shape), cnt[:n].reshape(shape)

    def count_batched(self, terms, edges, box, survey=False, path=NC_PATH_MIXED, npi=0, ori=(0.0, 0.0, 0.0)):
        """nc_count_batched over a list of (Catalog, Catalog, Catalog) terms.
        Returns (hist (T,...), counts (T,...)) with the shape of count()."""
        T = len(terms)
        arr = ctypes.c_void_p * T
        c1 = arr(*[t[0].h for t in terms])
        c2 = arr(*[t[1].h for t in terms])
        c3 = arr(*[t[2].h for t in terms])
        p, e = self.params(edges, box, survey, npi, ori)
        shape = self._shape(e.size - 1, npi)
        hist = np.zeros(T * int(np.prod(shape)), dtype=np.float64)
        cnt = np.zeros(hist.size, dtype=np.uint64)
        self._check(lib().nc_count_batched(self.h, ctypes.byref(p), T, c1, c2, c3, int(path), _dp(hist),
                                           cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))))
        return hist.reshape((T,) + shape), cnt.reshape((T,) + shape)

    def autotune(self, terms, edges, box, survey=False, path=NC_PATH_MIXED, npi=0, ori=(0.0, 0.0, 0.0)):
        """nc_autotune over (Catalog, Catalog, Catalog) terms; returns a dict."""
        T = len(terms)
        arr = ctypes.c_void_p * T
        c1 = arr(*[t[0].h for t in terms])
        c2 = arr(*[t[1].h for t in terms])
        c3 = arr(*[t[2].h for t in terms])
        p, e = self.params(edges, box, survey, npi, ori)
        best = NcTuning()
        self._check(lib().nc_autotune(self.h, ctypes.byref(p), T, c1, c2, c3, int(path), ctypes.byref(best)))
        return {f: getattr(best, f) for f, _ in NcTuning._fields_}

    def timing(self) -> NcTiming:
        t = NcTiming()
        self._check(lib().nc_last_timing(self.h, ctypes.byref(t)))
        return t

    def set_tuning(self, waves_per_group=0, chunk=0, wgs_per_cu=-1, list_chunk=0):
        self._check(lib().nc_set_tuning(self.h, waves_per_group, chunk, wgs_per_cu, list_chunk))

    def set_accumulation(self, mode):
        """nc_set_accumulation: NC_ACCUM_FP64 (0, default), NC_ACCUM_EXACT (1) or NC_ACCUM_FP32 (2)."""
        self._check(lib().nc_set_accumulation(self.h, int(mode)))

    def set_primary_slice(self, slice_, nslices):
        """nc_set_primary_slice: later counts take slice `slice_` of `nslices`
        of each term's cell-sorted primaries ((0, 1): all)."""
        self._check(lib().nc_set_primary_slice(self.h, int(slice_), int(nslices)))

    def get_tuning(self):
        a, b, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(lib().nc_get_tuning(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c), ctypes.byref(d)))
        return a.value, b.value, c.value, d.value
