// libncorr_hip.so: MI355X (gfx950) three-point triangle counting.
//
// Hot path of the reference (ahnyu/cudaNcorr): ncorr::calculateD1D2D3
// (source/cudaNcorr.cu:220-430) and its kernels countTrianglesThreeDBox /
// countTrianglesThreeDSurvey (source/device.cu:259-315, 381-431).
//
// Design (not a port): instead of one thread per unsorted primary walking
// per-cell device allocations, the library
//   1. builds a cell list on the device: FP64 cell key per point
//      (cudaNcorr.cu:50-58), a stable counting sort by key (per-cell counts,
//      exclusive scan, then each point placed after the points of its cell
//      with a smaller index) and an SoA scatter, so neighbour cells are
//      contiguous, coalesced ranges in an order that is the same on every run;
//   2. runs one persistent triplet kernel: each workgroup pulls
//      (term, primary tile) items from a device work queue, each wave takes one
//      primary at a time, compacts the neighbours within (rmin, rmax) of its
//      27 neighbour entries with wave ballots into a scratch list (storing the
//      exact bin and FP32 sqrt of the r12 / r13 sides once), then loops over
//      pairs of list entries -- one entry broadcast from LDS, one per lane --
//      binning (r12, r13, r23) into a workgroup-private LDS histogram over the
//      C(nb+2,3) sorted bins, merged once per term into global memory.
//   When p2 and p3 are the same catalogue the (p2,p3) exchange symmetry of the
//   reference's ordered loop is used: every unordered pair is evaluated once
//   and counted twice (its squared separation is bitwise symmetric).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/ncorr_hip.h"
#include "ncorr_kernels.hpp"

using namespace nck;

// ---------------------------------------------------------------------------
// device helpers

// Squared separation in the canonical evaluation order of the oracle:
// fma(dz,dz, fma(dy,dy, dx*dx)), the left-to-right contraction of
// dx*dx + dy*dy + dz*dz (device.cu:30-36).  Built with -ffp-contract=off so
// no other contraction is introduced.
__device__ __forceinline__ double sqrsep(double dx, double dy, double dz) {
    return __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
}

// sqrRpBox (device.cu:38-43): dx*dx + dy*dy, contracted in the same
// left-to-right order as sqrsep.
__device__ __forceinline__ double rp2(double dx, double dy) { return __builtin_fma(dy, dy, dx * dx); }

// dotProduct_d3 (device.cu:23-25) in the same order.
__device__ __forceinline__ double dot3(double ax, double ay, double az, double bx, double by, double bz) {
    return __builtin_fma(az, bz, __builtin_fma(ay, by, ax * bx));
}

// Cell index along one axis: int(v / cs) with truncation (cudaNcorr.cu:51-56,
// device.cu:76), optional clamp of ==n to n-1.  Out of grid => bad.
__device__ __forceinline__ int axis_cell(double v, double cs, int n, bool clamp, bool &bad) {
    const double q = v / cs;
    if (!(q > -1.0 && q < (double)n + 1.0)) {
        bad = true;
        return 0;
    }
    int c = (int)q;
    if (clamp && c == n) c = n - 1;
    if (c < 0 || c >= n) {
        bad = true;
        return 0;
    }
    return c;
}

__device__ __forceinline__ int lane_rank(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// Order LDS / global accesses of the lanes of one wave (wave-private data).
// Wave-uniform copies of values the compiler cannot prove uniform (loads
// through generic pointers may read per-lane private memory, so their results
// count as divergent even at a uniform address).  Keeps the loops that depend
// on them under scalar control.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double uniform(double v) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// A wave-uniform 32-bit value held in a VGPR (the compiler would keep it in a
// scalar register).  Inside the pair loops the scalar registers run out, and
// a constant spilled to a VGPR lane costs a v_readlane_b32 (a VALU issue slot)
// at every use; a VGPR operand costs nothing.
template <typename T>
__device__ __forceinline__ T in_vgpr(T x) {
    static_assert(sizeof(T) == 4, "32-bit values");
    unsigned r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(__builtin_bit_cast(unsigned, x)));
    return __builtin_bit_cast(T, r);
}

// A catalogue array as a global-memory pointer.  The catalogue pointers reach
// the kernels inside TermDesc, so the compiler cannot tell their address
// space and would load through flat instructions, which count against the
// LDS wait counter too: every LDS wait after them (the list build's lane
// searches, the pair loops' table reads) would also wait for those loads.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gptr(const T *p) {
    return (const __attribute__((address_space(1))) T *)p;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Exact bin of a squared separation d with r2min < d < r2max: lut[j] is a
// lower bound of the bin of every d with (int)(d*scale) == j, and the loop
// steps up over the squared edges (half-open bins, device.cu:181-191).
__device__ __forceinline__ int bin_of(double d, double scale, const uint16_t *lut, const double *sq, int nb) {
    int k = lut[(int)(d * scale)];
    while (k < nb - 1 && d >= sq[k + 1]) ++k;
    return k;
}

// One of the 27 neighbour entries (x-major shift order, cudaNcorr.cu:37-48).
// Box: one periodic wrap with image shift +/-L (device.cu:89-122).
// Survey: entries outside the grid are skipped (device.cu:123-153) => -1.
__device__ __forceinline__ int neighbour(const KArgs &A, int cx, int cy, int cz, int s, double &shx, double &shy,
                                         double &shz) {
    int ex = cx + s / 9 - 1, ey = cy + (s / 3) % 3 - 1, ez = cz + s % 3 - 1;
    shx = shy = shz = 0.0;
    if (A.survey) {
        if (ex < 0 || ex >= A.nc[0] || ey < 0 || ey >= A.nc[1] || ez < 0 || ez >= A.nc[2]) return -1;
    } else {
        if (ex == A.nc[0]) { ex = 0; shx = A.box[0]; }
        if (ey == A.nc[1]) { ey = 0; shy = A.box[1]; }
        if (ez == A.nc[2]) { ez = 0; shz = A.box[2]; }
        if (ex == -1) { ex = A.nc[0] - 1; shx = -A.box[0]; }
        if (ey == -1) { ey = A.nc[1] - 1; shy = -A.box[1]; }
        if (ez == -1) { ez = A.nc[2] - 1; shz = -A.box[2]; }
    }
    return ez + A.nc[2] * (ey + A.nc[1] * ex);
}

// ---------------------------------------------------------------------------
// cell-list build


__global__ void k_fill_u32(uint32_t *p, uint32_t v, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

// c is the exact floor of v / cs: c*cs <= v < (c+1)*cs in exact arithmetic
// (each FMA rounds v - c*cs once, which keeps its sign).  Fails for the
// clamped index of a point at or past n*cs and for -cs < v < 0, which
// truncation puts in cell 0.
__device__ __forceinline__ bool exact_floor(double v, double cs, int c) {
    return __builtin_fma(-(double)c, cs, v) >= 0.0 && __builtin_fma(-(double)(c + 1), cs, v) < 0.0;
}

// z slab of a point of cell c along z: trunc(kSlabs * v / cs) - kSlabs * c
// (the scaling by kSlabs, a power of 2, is exact), clamped to [0, kSlabs):
// the clamp of a point at n*cs into cell n-1 and the truncation of
// -cs < v < 0 into cell 0 leave such points outside their cell's range.
__device__ __forceinline__ int slab_of(double v, double cs, int c) {
    const int s = (int)((v / cs) * (double)kSlabs) - kSlabs * c;
    return min(max(s, 0), kSlabs - 1);
}

// Slab key of every point (cell by the host clamp rule, cudaNcorr.cu:50-58,
// then its z slab) and its slot in the slab.  Survey primaries are also checked without the clamp, as the
// reference's device index does not clamp them (device.cu:84-87).
// *inexact gets 1 when some point's cell index is not the exact floor of its
// coordinate over the cell size (then no pair closer than a cell size is
// guaranteed to lie in adjacent cells); with expect_exact the plan relied on
// it, and err gets kErrReplan.
__global__ void k_cell_keys(const double4 *__restrict__ raw, int n, GridGeo g, int check_unclamped,
                            uint32_t *__restrict__ key, uint32_t *__restrict__ rank, uint32_t *__restrict__ cnt,
                            unsigned *__restrict__ err, unsigned *__restrict__ inexact, int expect_exact) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = raw[i];
    bool bad = false;
    int cx = axis_cell(p.x, g.cs[0], g.nc[0], true, bad);
    int cy = axis_cell(p.y, g.cs[1], g.nc[1], true, bad);
    int cz = axis_cell(p.z, g.cs[2], g.nc[2], true, bad);
    const bool exact =
        exact_floor(p.x, g.cs[0], cx) && exact_floor(p.y, g.cs[1], cy) && exact_floor(p.z, g.cs[2], cz);
    if (__ballot(!exact) != 0ull && (threadIdx.x & (kWave - 1)) == 0) {  // lane 0 is active whenever any lane is
        atomicOr(inexact, 1u);
        if (expect_exact) atomicOr(err, kErrReplan);
    }
    if (check_unclamped) {
        axis_cell(p.x, g.cs[0], g.nc[0], false, bad);
        axis_cell(p.y, g.cs[1], g.nc[1], false, bad);
        axis_cell(p.z, g.cs[2], g.nc[2], false, bad);
    }
    if (bad) {
        atomicOr(err, kErrDomain);
        cx = cy = cz = 0;
    }
    const uint32_t k = (uint32_t)((cz + g.nc[2] * (cy + g.nc[1] * cx)) * kSlabs + (bad ? 0 : slab_of(p.z, g.cs[2], cz)));
    key[i] = k;
    rank[i] = atomicAdd(&cnt[k], 1u);
}

// Exclusive scan of n counts in one workgroup of 1024 threads (start[n] gets
// the total): the tile sums of the multi-block scan below.
__global__ __launch_bounds__(1024) void k_scan(const uint32_t *__restrict__ cnt, uint32_t *__restrict__ start,
                                               int ncells) {
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const int per = (ncells + 1023) / 1024;
    const int b = min(t * per, ncells), e = min(b + per, ncells);
    uint32_t s = 0;
    for (int i = b; i < e; ++i) s += cnt[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = t ? part[t - 1] : 0u;
    for (int i = b; i < e; ++i) {
        start[i] = run;
        run += cnt[i];
    }
    if (t == 1023) start[ncells] = part[1023];
}

// Multi-block exclusive scan of the slab counts, ntiles tiles of kScanTile
// counts (zero past the last slab): k_scan_tiles sums every tile, k_scan
// scans the sums into pstart, k_scan_apply scans each tile from its offset.
// 256 threads per tile, 16 consecutive counts per thread.
__global__ __launch_bounds__(256) void k_scan_tiles(const uint32_t *__restrict__ cnt, uint32_t *__restrict__ part) {
    const uint4 *c = reinterpret_cast<const uint4 *>(cnt + (size_t)blockIdx.x * kScanTile);
    uint32_t s = 0;
    for (int j = threadIdx.x; j < kScanTile / 4; j += 256) {
        const uint4 v = c[j];
        s += v.x + v.y + v.z + v.w;
    }
    for (int o = kWave / 2; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    __shared__ uint32_t ws[256 / kWave];
    if ((threadIdx.x & (kWave - 1)) == 0) ws[threadIdx.x / kWave] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < 256 / kWave; ++w) t += ws[w];
        part[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(256) void k_scan_apply(const uint32_t *__restrict__ cnt,
                                                    const uint32_t *__restrict__ pstart, uint32_t *__restrict__ start,
                                                    int ntiles) {
    static_assert(kScanTile == 256 * 16, "16 counts per thread");
    const int t = threadIdx.x, lane = t & (kWave - 1), wave = t / kWave;
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)t * 16;
    const uint4 *c = reinterpret_cast<const uint4 *>(cnt + base);
    uint32_t v[16];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint4 q = c[k];
        v[4 * k] = q.x;
        v[4 * k + 1] = q.y;
        v[4 * k + 2] = q.z;
        v[4 * k + 3] = q.w;
    }
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) tot += v[k];
    uint32_t inc = tot;
    for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    __shared__ uint32_t ws[256 / kWave];
    if (lane == kWave - 1) ws[wave] = inc;
    __syncthreads();
    uint32_t run = pstart[blockIdx.x] + inc - tot;
    for (int w = 0; w < wave; ++w) run += ws[w];
    uint4 *o = reinterpret_cast<uint4 *>(start + base);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        uint4 q;
        q.x = run;
        run += v[4 * k];
        q.y = run;
        run += v[4 * k + 1];
        q.z = run;
        run += v[4 * k + 2];
        q.w = run;
        run += v[4 * k + 3];
        o[k] = q;
    }
    if (blockIdx.x == 0 && t == 0) start[(size_t)ntiles * kScanTile] = pstart[ntiles];
}

// The counting sort by slab key, made stable: k_place puts every point index
// into its slab's range at the slot its atomic rank gave it (an order that
// varies from run to run); k_stable_scatter then moves each point to position
// start + (number of points of its slab with a smaller index), i.e. the slab's
// points in input order, and writes the SoA arrays there.  The order within a
// slab is thus the same on every run, and so is every list, work split and
// floating-point summation order that follows from it.  Each position scans
// its slab once: O(sum of squared slab occupancies), parallel over positions.
__global__ void k_place(int n, const uint32_t *__restrict__ key, const uint32_t *__restrict__ rank,
                        const uint32_t *__restrict__ start, uint32_t *__restrict__ perm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    perm[start[key[i]] + rank[i]] = (uint32_t)i;
}

__global__ void k_stable_scatter(const double4 *__restrict__ raw, int n, const uint32_t *__restrict__ key,
                                 const uint32_t *__restrict__ start, const uint32_t *__restrict__ perm,
                                 double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
                                 double *__restrict__ w) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint32_t me = perm[q];
    const uint32_t c = key[me];
    const int b = (int)start[c], e = (int)start[c + 1];
    int r = 0;
    for (int j = b; j < e; ++j) r += perm[j] < me ? 1 : 0;
    const int d = b + r;
    const double4 p = raw[me];
    x[d] = p.x;
    y[d] = p.y;
    z[d] = p.z;
    w[d] = p.w;
}

// Upper bound of the neighbour-list length: over cells holding primaries, the
// number of secondaries in the 27 neighbour entries.
__global__ void k_nbr_bound(const uint32_t *__restrict__ st1, const uint32_t *__restrict__ st2, GridGeo g,
                            unsigned *__restrict__ out) {
    const int ncells = g.nc[0] * g.nc[1] * g.nc[2];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells || st1[(c + 1) * kSlabs] == st1[c * kSlabs]) return;
    const int cz = c % g.nc[2], cy = (c / g.nc[2]) % g.nc[1], cx = c / (g.nc[2] * g.nc[1]);
    unsigned tot = 0;
    for (int s = 0; s < 27; ++s) {
        int ex = cx + s / 9 - 1, ey = cy + (s / 3) % 3 - 1, ez = cz + s % 3 - 1;
        if (g.survey) {
            if (ex < 0 || ex >= g.nc[0] || ey < 0 || ey >= g.nc[1] || ez < 0 || ez >= g.nc[2]) continue;
        } else {
            ex = ex == g.nc[0] ? 0 : (ex == -1 ? g.nc[0] - 1 : ex);
            ey = ey == g.nc[1] ? 0 : (ey == -1 ? g.nc[1] - 1 : ey);
            ez = ez == g.nc[2] ? 0 : (ez == -1 ? g.nc[2] - 1 : ez);
        }
        const int k = ez + g.nc[2] * (ey + g.nc[1] * ex);
        tot += st2[(k + 1) * kSlabs] - st2[k * kSlabs];
    }
    atomicMax(out, tot);
}


// Weight statistics of a catalogue into s[0..3]: s[0] |= 1 when some weight
// is not 1, s[1] |= 1 when some weight is not finite and |= 2 when some weight
// is not an FP32 weight (fp32_weight), and the 64-bit word at s + 2 gets the
// largest |w| as its bit pattern (non-negative doubles order as their bits).
__global__ void k_weight_stats(const double4 *__restrict__ raw, long long n, unsigned *__restrict__ s) {
    unsigned nonunit = 0u, nonfinite = 0u, nonf32 = 0u;
    unsigned long long mx = 0ull;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double w = raw[i].w;
        nonunit |= w != 1.0 ? 1u : 0u;
        nonf32 |= fp32_weight(w) ? 0u : 2u;
        if (__builtin_isfinite(w))
            mx = max(mx, (unsigned long long)__double_as_longlong(fabs(w)));
        else
            nonfinite = 1u;
    }
    if (__ballot(nonunit != 0u) != 0ull && (threadIdx.x & (kWave - 1)) == 0) atomicOr(s, 1u);
    if (__ballot(nonfinite != 0u) != 0ull && (threadIdx.x & (kWave - 1)) == 0) atomicOr(s + 1, 1u);
    if (__ballot(nonf32 != 0u) != 0ull && (threadIdx.x & (kWave - 1)) == 0) atomicOr(s + 1, 2u);
    atomicMax(reinterpret_cast<unsigned long long *>(s + 2), mx);
}

// ---------------------------------------------------------------------------
// triplet kernel
//
// Per wave and primary p1, the neighbour lists S2 (catalogue c2) and S3 (c3)
// hold every image point q of the 27 neighbour entries with
// rmin^2 < |p1-q|^2 < rmax^2 -- exactly the p2 candidates the reference
// accepts at device.cu:287 and the p3 candidates passing the r13 half of
// device.cu:301.  The (p2,p3) loop of device.cu:288-308 is then a loop over
// S2 x S3 that tests r23 only.
//
// Plain list entry (FP64, 40 B):   Elem {x,y,z image position, k, f} + weight
// Mixed list entry (20 B):         float4 {x,y,z relative to p1, k | neighbour entry << 8}
//                                  + catalogue index
// (the FP64 image position of an entry, and f, are rebuilt from the catalogue
// index and its neighbour entry's shift, exactly as the list build computed them)
// k = exact bin of |p1-q|^2, f = sqrtf(float(|p1-q|^2)) as in device.cu:173.


struct LdsTabs {
    unsigned long long *hist;
    const double *sq;
    const uint16_t *lut;
    const int *ctab;
    const uint2 *lut32;
    const uint16_t *t3;
    const uint16_t *rk;  // rppi key stream: rank of a dense rp triple
    double xscale;       // kHistExact: 2^K, the fixed-point scale of the weight products
    int xhi;             // kHistExact: byte offset of the high words after the low words
};


__device__ __forceinline__ double rp_side_pack(float f, int k, int p) {
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)(k | (p << 8)) << 32) |
                                            (unsigned long long)__float_as_uint(f)));
}
__device__ __forceinline__ void rp_side_unpack(double q, float &f, int &k, int &p) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(q);
    f = __uint_as_float((unsigned)u);
    const unsigned kp = (unsigned)(u >> 32);
    k = (int)(kp & 0xffu);
    p = (int)(kp >> 8);
}


// Arrays of a mixed-path list inside a wave's scratch slot (20 B per entry;
// the exact rechecks rebuild the entry's FP64 side, and with it its FP32
// root, from the catalogue).  The fourth word of R holds the entry's key,
// not s = |r|^2: the pair loops recompute s from r when they stage rows or
// load lanes (the same FP32 operations as the list build, so the same
// value), which keeps one 16 B load per entry and pass.
struct MList {
    float4 *R;  // x, y, z relative to p1 (FP32); .w: the key's bits
    int *I;     // index of the point in the cell-sorted catalogue
};
__device__ __forceinline__ MList mlist(const char *base, int cap) {
    MList m;
    m.R = reinterpret_cast<float4 *>(const_cast<char *>(base));
    m.I = reinterpret_cast<int *>(const_cast<char *>(base) + (size_t)cap * sizeof(float4));
    return m;
}
// An entry's key: exact bin of |p1-q|^2 | neighbour entry << 8.
__device__ __forceinline__ int mkey(const MList &m, int i) {
    return reinterpret_cast<const int *>(m.R + i)[3];
}
// s = |r|^2 of an entry's relative position (FP32, as the list build makes it)
__device__ __forceinline__ float rsq(float4 r) {
    return __builtin_fmaf(r.z, r.z, __builtin_fmaf(r.y, r.y, r.x * r.x));
}

// The primary of a pair loop: its cell (for the image shifts of the exact
// rechecks) and its position.
struct PrimRef {
    int cx, cy, cz;
    double px, py, pz;
    unsigned lastmask;  // bit e: neighbour entry e of the primary is a last cell (box)
};

// Neighbour entries e = (dx+1)*9 + (dy+1)*3 + (dz+1) of a primary's cell
// whose cells are adjacent to the cell of entry s (every axis offset within
// one): as a 27-bit mask.  With at least three cells per axis, two entries
// whose offsets differ by two on an axis are images on either side of the
// primary's cell.
__device__ __forceinline__ unsigned entry_adjacent(int s) {
    const int ix = s / 9, iy = (s / 3) % 3, iz = s % 3;
    auto axis = [](int i) { return i == 0 ? 3u : i == 1 ? 7u : 6u; };
    const unsigned mz = axis(iz), my = axis(iy), mx = axis(ix);
    const unsigned yz = ((my & 1u) ? mz : 0u) | ((my & 2u) ? mz << 3 : 0u) | ((my & 4u) ? mz << 6 : 0u);
    return ((mx & 1u) ? yz : 0u) | ((mx & 2u) ? yz << 9 : 0u) | ((mx & 4u) ? yz << 18 : 0u);
}

// Catalogue arrays behind a mixed list (for weights and exact rechecks).
struct CatRef {
    const double *x, *y, *z, *w;
    int unit;  // every weight is 1: the pair loops read none
};

// Compact the neighbours of primary (px,py,pz) in catalogue (X,Y,Z,Wt,st)
// whose squared separation lies in (r2min, r2max) into the list at `base`.
// Returns the list length (entries beyond cap are not stored).
//
// The 27 neighbour entries follow the reference's x-major order
// (cudaNcorr.cu:37-48), and so does the list.  Their cell ranges are laid end
// to end and scanned as one sequence, 128 candidates per step (two 64-point
// blocks with their loads in flight together), so that short cells leave no
// lanes idle and a primary waits for few memory round trips.  Each list entry
// records its neighbour entry s (0..26), which fixes its image shift.
//
// Selections (SEL) for the six-fold path of terms over one catalogue, where
// pidx is the primary's index in the sorted catalogue and an entry is "last"
// when its cell is the last one along some axis (box mode; surveys have
// none):
//   kSelAll    every entry into base;
//   kSelSplit  entries that are not last and have index > pidx into base,
//              last entries into base_b (*nb_out of them);
//   kSelLow    the remaining ones (not last, index < pidx), into base from
//              position `first` on.
// Each selection scans only the neighbour rows that can hold its entries.
//
// cut (threeD, exactly gridded catalogue): a neighbour cell's range is cut
// to its z slabs within reach of the primary.  With (ax, ay) the distance from
// the primary to the cell's x-y extent (image position), a point of the cell
// can lie within R of the primary only if |z - pz| <= h = sqrt(R^2 - ax^2 -
// ay^2), and the slabs holding such z are a contiguous run.  R = sqrt(prune_r2)
// exceeds rmax by far more than the rounding of the cell and slab bounds, so
// every skipped point fails the range filter below; the lists are the same as
// without pruning.
constexpr int kSelAll = 0, kSelSplit = 1, kSelLow = 2;

// Exact bin of a squared separation d with r2min < d < r2max for a mixed
// list entry: the FP32 lookup of the pair loop (LDS) on d scaled by sigma^2
// (exact) and rounded once to FP32, an error far inside the guard band g32;
// within the band of the edge, the FP64 lookup (global tables, rare).  KS:
// the shift of the table's .x.
template <int KS>
__device__ __forceinline__ int mixed_entry_bin(const KArgs &A, const uint2 *lb32, double d, const double *sq,
                                               const uint16_t *lut, float lo32, int sh32, float g32) {
    const float ds = (float)(d * (A.sigma32 * A.sigma32));
    const uint2 e = lb32[__float_as_uint(__builtin_amdgcn_fmed3f(ds, lo32, 4.0f)) >> sh32];
    const float dE = ds - __uint_as_float(e.y);
    if (fabsf(dE) <= g32) return bin_of(d, A.lut_scale, lut, sq, A.nb);
    return (int)(e.x >> KS) - (dE < 0.0f ? 1 : 0);
}

template <int LM, int SEL = kSelAll, int KS = 0>
__device__ __forceinline__ int build_list(const KArgs &A, const double *__restrict__ X,
                                          const double *__restrict__ Y, const double *__restrict__ Z,
                                          const double *__restrict__ Wt, const uint32_t *__restrict__ st, bool cut,
                                          double px, double py, double pz, int cx, int cy, int cz,
                                          char *__restrict__ base,
                                          int lane, const double *sq, const uint16_t *lut, int pidx = 0,
                                          char *__restrict__ base_b = nullptr, int *nb_out = nullptr, int first = 0,
                                          unsigned *lastmask_out = nullptr) {
    constexpr bool MIXED = LM == kListMixed;
    int n = SEL == kSelLow ? first : 0;
    int nbl = 0;
    // rppi survey: every triplet's line of sight L = 3a + d2 + d3 (a = p1 - ori,
    // |d2|, |d3| < smax) is within the angle th = asin(2 smax / 3|a|) of a, so
    // a side d of the list has, in any triplet, pi >= |ah.d| - th|d| and
    // rp^2 >= |d|^2 - (|ah.d| + th|d|)^2.  A side failing pi < piMax or
    // rp^2 < rpMax^2 by these bounds (with 1e-9 relative margins, far above
    // the FP64 rounding of either evaluation) is rejected in every triplet
    // (device.cu:470-481), so it is left out of the list.
    double ahx = 0.0, ahy = 0.0, ahz = 0.0, th = 0.0;
    bool prune = false;
    if constexpr (LM == kListRppiSurvey) {
        const double ax = px - A.ori_x, na = sqrt(dot3(ax, py, pz, ax, py, pz));
        const double x = 2.0 * sqrt(A.smax2) / (3.0 * na) * (1.0 + 1e-9);
        prune = na > 0.0 && x < 0.5;
        if (prune) {
            ahx = ax / na;
            ahy = py / na;
            ahz = pz / na;
            th = asin(x) * (1.0 + 1e-9) + 1e-15;
        }
    }
    const double pim = sqrt(A.pimax2) * (1.0 + 1e-9), r2m = A.r2max * (1.0 + 1e-9);
    // the mixed path's FP32 lookup in LDS (entry c - c0 at lb32[c])
    extern __shared__ __attribute__((aligned(16))) char lds_tables[];
    const uint2 *lb32 = MIXED ? reinterpret_cast<const uint2 *>(lds_tables + A.off_lut32) - A.lut32_c0 : nullptr;
    // constants of the per-candidate code as VGPR operands (the scalar
    // registers run out here, and a spilled one costs a v_readlane per use)
    const int capv = in_vgpr(A.cap);
    const float lo32v = MIXED ? in_vgpr(A.lut32_lo) : 0.0f, g32v = MIXED ? in_vgpr(A.g32) : 0.0f;
    const int sh32v = MIXED ? in_vgpr(A.lut32_shift) : 0;
    // one threeD list entry (exact bin k and sqrtf of its side) at position pos of the list at b
    auto store = [&](char *b, int pos, int j, double d, double qx, double qy, double qz, double qw, int s) {
        const int k = MIXED ? mixed_entry_bin<KS>(A, lb32, d, sq, lut, lo32v, sh32v, g32v)
                            : bin_of(d, A.lut_scale, lut, sq, A.nb);
        const float f = __builtin_sqrtf((float)d);
        if constexpr (MIXED) {
            const MList ml = mlist(b, A.cap);
            const double sg = A.sigma32;  // power of 2: the scaling is exact
            const float rx = (float)((qx - px) * sg), ry = (float)((qy - py) * sg), rz = (float)((qz - pz) * sg);
            ml.R[pos] = make_float4(rx, ry, rz, __int_as_float(k | (s << 8)));
            ml.I[pos] = j;
        } else {
            Elem e;
            e.x = qx;
            e.y = qy;
            e.z = qz;
            e.k = k | (s << 8);  // exact bin | neighbour entry
            e.f = f;
            reinterpret_cast<Elem *>(b)[pos] = e;
            reinterpret_cast<double *>(b + (size_t)A.cap * sizeof(Elem))[pos] = qw;
        }
    };
    auto keep = [&](int j, bool v, double qx, double qy, double qz, double qw, int s, bool last) {
        if constexpr (LM == kListRppiBox || LM == kListRppiSurvey) {
            // rppi: box keeps pi^2 < piMax^2 and rpMin^2 < rp^2 < rpMax^2
            // (device.cu:343-346), survey 0 < s^2 < sqrRMax (device.cu:454)
            double r, q;
            bool pass;
            if constexpr (LM == kListRppiBox) {
                const double dz = pz - qz;
                q = dz * dz;
                r = rp2(px - qx, py - qy);
                pass = v & (q < A.pimax2) & (r < A.r2max) & (r > A.r2min);
            } else {
                r = sqrsep(px - qx, py - qy, pz - qz);
                q = qx - A.ori_x;
                pass = v & (r < A.smax2) & (r > 0.0);
                if (prune) {
                    const double u = fabs(dot3(ahx, ahy, ahz, qx - px, qy - py, qz - pz));
                    const double dl = th * (double)(__builtin_sqrtf((float)r) * 1.000001f);  // >= th |d|
                    const double hi = u + dl;
                    pass = pass & !(u - dl > pim) & !(r - hi * hi > r2m + 1e-9 * r);
                }
            }
            const unsigned long long m = __ballot(pass);
            const int pos = n + lane_rank(m);
            if (pass & (pos < capv)) {
                if constexpr (LM == kListRppiBox)  // the side's root, rp bin and pi bin (device.cu:214-250)
                    q = rp_side_pack(__builtin_sqrtf((float)r), bin_of(r, A.lut_scale, lut, sq, A.nb),
                                     (int)floorf(__builtin_sqrtf((float)q)));
                reinterpret_cast<RElem *>(base)[pos] = RElem{qx, qy, qz, r, q, qw};
            }
            n += __popcll(m);
            return;
        }
        const double d = sqrsep(px - qx, py - qy, pz - qz);
        const bool in = v & (d > A.r2min) & (d < A.r2max);
        if constexpr (SEL == kSelAll) {
            const unsigned long long m = __ballot(in);
            const int pos = n + lane_rank(m);
            if (in & (pos < capv)) store(base, pos, j, d, qx, qy, qz, qw, s);
            n += __popcll(m);
        } else if constexpr (SEL == kSelSplit) {
            const bool ph = in & !last & (j > pidx), pb = in & last;
            const unsigned long long mh = __ballot(ph), mb = __ballot(pb);
            const int posh = n + lane_rank(mh), posb = nbl + lane_rank(mb);
            // the two lists take disjoint entries: one store path
            const bool sh = ph & (posh < capv), sb = pb & (posb < capv);
            if (sh | sb) store(sh ? base : base_b, sh ? posh : posb, j, d, qx, qy, qz, qw, s);
            n += __popcll(mh);
            nbl += __popcll(mb);
        } else {
            const bool pl = in & !last & (j < pidx);
            const unsigned long long m = __ballot(pl);
            const int pos = n + lane_rank(m);
            if (pl & (pos < capv)) store(base, pos, j, d, qx, qy, qz, qw, s);
            n += __popcll(m);
        }
    };
    // The 27 neighbour entries, one per lane (lane e: entry e of the
    // reference's x-major order, cudaNcorr.cu:37-48): the entry's cell after
    // the one periodic wrap (box: image shift of +-L per wrapped axis,
    // device.cu:89-122) or nothing (survey: outside the grid), its range of
    // the sorted catalogue cut to the selection, and whether it is a last
    // cell.  All 27 ranges load at once.
    int cbeg = 0, clen = 0, cpk = 0;  // cpk: shift signs (2 bits per axis) | last << 6
    if (lane < 27) {
        const int d[3] = {lane / 9 - 1, (lane / 3) % 3 - 1, lane % 3 - 1};
        const int c0[3] = {cx, cy, cz};
        int e[3];
        double sh[3] = {0.0, 0.0, 0.0};
        bool in = true, last = false;
        for (int a = 0; a < 3; ++a) {
            e[a] = c0[a] + d[a];
            if (A.survey) {
                in = in && e[a] >= 0 && e[a] < A.nc[a];
            } else {
                if (e[a] == A.nc[a]) {
                    e[a] = 0;
                    cpk |= 1 << (2 * a);  // +L
                    sh[a] = A.box[a];
                } else if (e[a] == -1) {
                    e[a] = A.nc[a] - 1;
                    cpk |= 2 << (2 * a);  // -L
                    sh[a] = -A.box[a];
                }
                last = last || e[a] == A.nc[a] - 1;
            }
        }
        if (in) {
            const int cell = e[2] + A.nc[2] * (e[1] + A.nc[1] * e[0]);
            int s0 = 0, s1 = kSlabs;  // the slabs [s0, s1) of the cell
            if constexpr (LM == kListPlain || LM == kListMixed) {
                if (cut) {
                    const double pc[2] = {px, py};
                    double dxy2 = 0.0;
                    for (int a = 0; a < 2; ++a) {
                        const double lo = (double)e[a] * A.cs[a] + sh[a], hi = lo + A.cs[a];
                        const double dd = fmax(fmax(lo - pc[a], pc[a] - hi), 0.0);
                        dxy2 = __builtin_fma(dd, dd, dxy2);
                    }
                    if (!(dxy2 <= A.prune_r2)) {
                        s0 = s1 = 0;
                    } else {
                        const double h = sqrt(A.prune_r2 - dxy2);
                        const double z0 = (double)e[2] * A.cs[2] + sh[2], iw = (double)kSlabs / A.cs[2];
                        const double slo = floor((pz - h - z0) * iw), shi = floor((pz + h - z0) * iw);
                        s0 = (int)fmin(fmax(slo, 0.0), (double)kSlabs);
                        s1 = (int)fmin(fmax(shi + 1.0, 0.0), (double)kSlabs);
                        s1 = max(s1, s0);
                    }
                }
            }
            int b = (int)gptr(st)[cell * kSlabs + s0], end = (int)gptr(st)[cell * kSlabs + s1];
            if constexpr (SEL == kSelSplit)
                if (!last) b = max(b, pidx + 1);  // last cells whole (into base_b), others above pidx
            if constexpr (SEL == kSelLow) end = last ? b : min(end, pidx);  // not last, below pidx
            cbeg = b;
            clen = max(end - b, 0);
            cpk |= (int)last << 6;
        }
    }
    // the entries' ranges laid end to end: exclusive offsets (lanes 27..31
    // hold the total, which no candidate reaches)
    int incl = clen;
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) {
        const int t = __shfl_up(incl, o);
        if (lane >= o) incl += t;
    }
    const int cexcl = incl - clen;
    const int total = uniform(__shfl(incl, 26));
    // most primaries have no wrapped (shifted) and no last neighbour cell:
    // then the per-candidate shift and flag are skipped (wave-uniform tests;
    // q + 0.0 would be q in every later use)
    const bool any_shift = __ballot((cpk & 63) != 0) != 0ull;
    if (lastmask_out) *lastmask_out = (unsigned)__ballot((cpk & 64) != 0);
    // Candidates v of [0, total), two 64-lane blocks per step with all their
    // loads in flight.  A candidate belongs to the last non-empty entry whose
    // offset is <= v (offsets of non-empty entries increase strictly).  The
    // non-empty entries are walked once per primary under scalar control, in
    // order: a block [w0, w0 + 64) takes over the entry current at w0 and
    // then each entry starting before w0 + 64, lanes at or past its offset
    // switching to it.  Lane e < 27 holds its entry's catalogue index offset
    // cbeg - cexcl (candidate v of the entry is point v + that) and the word
    // cpk | e << 8 the list entries record.
    const int cdel = cbeg - cexcl, cword = cpk | (lane << 8);
    unsigned pend = (unsigned)__ballot((lane < 27) & (clen > 0));  // entries not yet reached
    int e_nx = pend ? __builtin_ctz(pend) : 0;
    int s_nx = pend ? __builtin_amdgcn_readlane(cexcl, e_nx) : INT_MAX;  // offset of the next entry
    int cur_del = 0, cur_word = 0;  // the last entry reached
    // a point of the first non-empty range: the loads of lanes past the end
    // read it (always in bounds, never kept), so they need no branch
    const int jsafe = pend ? __builtin_amdgcn_readlane(cbeg, e_nx) : 0;
    constexpr int kB = 2;
    for (int v0 = 0; v0 < total; v0 += kB * kWave) {
        int jj[kB], pk[kB], ent[kB];
        bool ok[kB];
        double qx[kB], qy[kB], qz[kB], qw[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int w0 = v0 + u * kWave, v = w0 + lane;
            ok[u] = v < total;
            int dl = cur_del, wd = cur_word;
            while (s_nx < w0 + kWave) {  // wave-uniform
                cur_del = __builtin_amdgcn_readlane(cdel, e_nx);
                cur_word = __builtin_amdgcn_readlane(cword, e_nx);
                if (v >= s_nx) {
                    dl = cur_del;
                    wd = cur_word;
                }
                pend &= pend - 1u;
                e_nx = pend ? __builtin_ctz(pend) : 0;
                s_nx = pend ? __builtin_amdgcn_readlane(cexcl, e_nx) : INT_MAX;
            }
            // the values of an invalid lane are never used
            ent[u] = wd >> 8;
            jj[u] = dl + v;
            pk[u] = wd & 127;
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            if (v0 + u * kWave >= total) break;  // wave-uniform, as below
            // image position: point + shift (device.cu:282-284); weights
            // only for the lists that carry them (the mixed list reads them
            // from the catalogue where they are needed)
            const int jl = ok[u] ? jj[u] : jsafe;
            qx[u] = gptr(X)[jl];
            qy[u] = gptr(Y)[jl];
            qz[u] = gptr(Z)[jl];
            qw[u] = MIXED ? 0.0 : gptr(Wt)[jl];
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            if (v0 + u * kWave >= total) break;  // wave-uniform
            const int p = pk[u];
            if (any_shift) {
                qx[u] += (p & 1) ? A.box[0] : (p & 2) ? -A.box[0] : 0.0;
                qy[u] += (p & 4) ? A.box[1] : (p & 8) ? -A.box[1] : 0.0;
                qz[u] += (p & 16) ? A.box[2] : (p & 32) ? -A.box[2] : 0.0;
            }
            keep(jj[u], ok[u], qx[u], qy[u], qz[u], qw[u], ent[u], (p & 64) != 0);
        }
    }
    if constexpr (SEL == kSelSplit) *nb_out = nbl;
    return n;
}

// Longest neighbour list (S2 or S3) over the primaries of one term: the exact
// lengths build_list returns, which stores nothing when A.cap is 0.  The
// dense layout's no-wrap rule needs chunk * (list length)^2 < 2^31; the
// 27-cell candidate bound is several times the list length, too loose for
// dense catalogues.  One wave per primary, grid-stride.
// Primaries [p0, p1) of the sorted catalogue (the primary slice counted).
__global__ __launch_bounds__(256) void k_list_bound(KArgs A, TermDesc T, int p0, int p1, unsigned *__restrict__ out) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) / kWave);
    const int nwaves = (int)(gridDim.x * blockDim.x / kWave);
    int best = 0;
    for (int p = p0 + wave; p < p1; p += nwaves) {
        const double px = T.x1[p], py = T.y1[p], pz = T.z1[p];
        bool bad = false;
        const bool clamp = !A.survey;
        const int cx = axis_cell(px, A.cs[0], A.nc[0], clamp, bad);
        const int cy = axis_cell(py, A.cs[1], A.nc[1], clamp, bad);
        const int cz = axis_cell(pz, A.cs[2], A.nc[2], clamp, bad);
        if (bad) continue;  // the triplet kernel reports it
        const int n2 = build_list<kListPlain>(A, T.x2, T.y2, T.z2, T.w2, T.st2, T.prune2, px, py, pz, cx, cy, cz, nullptr, lane,
                                              nullptr, nullptr);
        const int n3 = T.sym ? n2
                             : build_list<kListPlain>(A, T.x3, T.y3, T.z3, T.w3, T.st3, T.prune3, px, py, pz, cx, cy, cz,
                                                      nullptr, lane, nullptr, nullptr);
        best = max(best, max(n2, n3));
    }
    if (lane == 0) atomicMax(out, (unsigned)best);
}


// Byte offset in the LDS histogram of the slot of bins (ka, kb, k), k in
// [-1, nb] (slots of k outside [0, nb) are never added).
template <int HM>
__device__ __forceinline__ int sorted_slot(const KArgs &A, const LdsTabs &t, int ka, int kb, int k) {
    if constexpr (t3_like(HM)) return t.t3[__mul24(__mul24(ka, A.nb) + kb, A.nb + 2) + k + 1];
    if constexpr (HM == kHistDense) return 4 * (__mul24(__mul24(ka, A.nb) + kb, A.nb) + k);
    const int klo = min(min(ka, kb), k);
    const int khi = max(max(ka, kb), k);
    const int kmid = ka + kb + k - klo - khi;
    return (t.ctab[klo * A.nb + kmid] + (khi - kmid)) * (int)sizeof(unsigned long long);
}

// Bins as held by the pair loops, pre-multiplied so that the slot of
// (ka, kb, k) costs one add: kHistT3, row 2*(ka*nb*(nb+2) + 1) and lane
// 2*kb*(nb+2) bytes into t3 (symmetric in ka and kb), slot = the u16 at byte
// row + lane + 2k; kHistDense, row 4*ka*nb^2 and lane 4*kb*nb, slot byte
// offset row + lane + 4k; kHistCtab, the plain bins.
template <int HM>
__device__ __forceinline__ int row_bin(const KArgs &A, int k) {
    if constexpr (t3_like(HM)) return 2 * (__mul24(k, A.nb * (A.nb + 2)) + 1);
    if constexpr (HM == kHistDense) return 4 * __mul24(k, A.nb * A.nb);
    return k;
}
template <int HM>
__device__ __forceinline__ int lane_bin(const KArgs &A, int k) {
    if constexpr (t3_like(HM)) return 2 * __mul24(k, A.nb + 2);
    if constexpr (HM == kHistDense) return 4 * __mul24(k, A.nb);
    return k;
}
// (a << SH) + b as one v_lshl_add_u32.  For a in {0, -1} the compiler would
// otherwise rewrite the shift as a mask and spend a third instruction.
template <int SH>
__device__ __forceinline__ int shl_add(int a, int b) {
    if constexpr (SH == 0) {
        return a + b;
    } else {
        int r;
        asm("v_lshl_add_u32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "i"(SH), "v"(b));
        return r;
    }
}


// The same with the third bin in layout units: 4k (kHistDense), 2k (kHistT3)
// or k (kHistCtab).
template <int HM>
__device__ __forceinline__ int pair_slot_x(const KArgs &A, const LdsTabs &t, int kr, int kl, int kx) {
    if constexpr (t3_like(HM))
        return *reinterpret_cast<const uint16_t *>(reinterpret_cast<const char *>(t.t3) + kr + kl + kx);
    if constexpr (HM == kHistDense) return kr + kl + kx;
    return sorted_slot<kHistCtab>(A, t, kr, kl, kx);
}
template <int HM>
__device__ __forceinline__ int pair_slot(const KArgs &A, const LdsTabs &t, int kr, int kl, int k) {
    if constexpr (t3_like(HM))
        return *reinterpret_cast<const uint16_t *>(reinterpret_cast<const char *>(t.t3) + kr + kl + 2 * k);
    if constexpr (HM == kHistDense) return kr + kl + 4 * k;
    return sorted_slot<kHistCtab>(A, t, kr, kl, k);
}

// Triangle test inputs are non-negative floats: their bit patterns order as
// unsigned integers, so min/med/max run on the integer ALU (v_min3_u32,
// v_med3_u32, v_max3_u32) without NaN canonicalisation.
struct Sides {
    float lo, mid, hi;
};
__device__ __forceinline__ Sides sort3(float a, float b, float c) {
    const unsigned ua = __float_as_uint(a), ub = __float_as_uint(b), uc = __float_as_uint(c);
    const unsigned lo = min(min(ua, ub), uc);
    const unsigned hi = max(max(ua, ub), uc);
    const unsigned mid = max(min(ua, ub), min(max(ua, ub), uc));
    return {__uint_as_float(lo), __uint_as_float(mid), __uint_as_float(hi)};
}

// Histogram update of one accepted pair at a precomputed slot (byte offset).
// ORD: the ordered triplets the pair stands for.  With the primary p1 (weight
// w1), the pair's list-2 entry a and list-3 entry b (w1a = w1*wa, w1b =
// w1*wb), the reference adds one product (p1*p2)*p3 per order (device.cu:305):
//   kOrdOne   (p1; a, b)                          (w1 wa) wb
//   kOrdSym   and (p1; b, a)                      + (w1 wb) wa
//   kOrdSix   and the four orders led by a or b:  2 [(w1 wa) wb + (w1 wb) wa + (wa wb) w1]
//   kOrdSwap  (p1; a, b) and (a; p1, b)           2 (w1 wa) wb
// (the products of orders that differ by swapping the first two factors are
// equal).  Counts add the number of orders.
constexpr int kOrdOne = 1, kOrdSym = 2, kOrdSix = 6, kOrdSwap = 12;
constexpr int ord_count(int ord) { return ord == kOrdSwap ? 2 : ord; }
__device__ __forceinline__ long long fixed_point(double p, double scale) { return (long long)(p * scale); }
template <int ORD, bool WEIGHTED, int HM>
__device__ __forceinline__ void hist_add(const LdsTabs &t, int off, double w1, double w1a, double wa, double w1b,
                                         double wb) {
    static_assert(ORD == kOrdOne || ORD == kOrdSym || ORD == kOrdSix || ORD == kOrdSwap, "orders per pair");
    char *h = reinterpret_cast<char *>(t.hist) + off;
    if constexpr (HM == kHistDense) {
        static_assert(!WEIGHTED, "dense counts are for unit weights");
        atomicAdd(reinterpret_cast<unsigned *>(h), (unsigned)ord_count(ORD));
    } else if constexpr (WEIGHTED && HM == kHistExact) {
        // 128-bit sum: the low word's wrap (from the value it had) carries
        // into the high word, whatever the order of the additions
        long long f = fixed_point(w1a * wb, t.xscale);
        if constexpr (ORD == kOrdSym || ORD == kOrdSix) f += fixed_point(w1b * wa, t.xscale);
        if constexpr (ORD == kOrdSix) f = 2 * (f + fixed_point((wa * wb) * w1, t.xscale));
        if constexpr (ORD == kOrdSwap) f *= 2;
        const unsigned long long u = (unsigned long long)f;
        const unsigned long long old = atomicAdd(reinterpret_cast<unsigned long long *>(h), u);
        const long long carry = (f < 0 ? -1 : 0) + (old + u < old ? 1 : 0);
        if (carry != 0) atomicAdd(reinterpret_cast<unsigned long long *>(h + t.xhi), (unsigned long long)carry);
    } else if constexpr (WEIGHTED) {
        double v = w1a * wb;
        if constexpr (ORD == kOrdSym || ORD == kOrdSix) v += w1b * wa;
        if constexpr (ORD == kOrdSix) v = 2.0 * (v + (wa * wb) * w1);
        if constexpr (ORD == kOrdSwap) v *= 2.0;
        atomicAdd(reinterpret_cast<double *>(h), v);
    } else {
        atomicAdd(reinterpret_cast<unsigned long long *>(h), (unsigned long long)ord_count(ORD));
    }
}

// FP64 accumulation (not the exact layout): the pair loops stage each row's
// weight already multiplied by w1 and the number of orders the pair stands
// for, so an accepted pair adds one product (c w1 w_row) w_lane.  For FP32
// weights every order's product (w1 w2) w3 is the correctly rounded w1 w2 w3,
// and c w1 w_row is exact (c = 1, 2, 6), so the one rounding of this product
// matches them; for wider weights the two differ by an ulp or two.
template <bool WEIGHTED, int HM>
constexpr bool pre_weighted() { return WEIGHTED && HM != kHistExact; }
__device__ __forceinline__ void hist_add_value(const LdsTabs &t, int off, double v) {
    atomicAdd(reinterpret_cast<double *>(reinterpret_cast<char *>(t.hist) + off), v);
}

// kHistF32: the row's weight factor c w1 w_row as a float-float pair (its
// bits in the FP64 word the rows are staged in) times the lane's FP32 weight:
// the product (hi, lo) from an FMA, hi into the FP32 sum, the exact rounding
// error of that sum (TwoSum on the value the atomic returns) and lo into the
// FP64 compensation.  -ffp-contract=off keeps the TwoSum's operations apart.
__device__ __forceinline__ double f32_pair(double x) {
    const float hi = (float)x;
    return __builtin_bit_cast(double, make_float2(hi, (float)(x - (double)hi)));
}

// The staged weight of a row: c w1 w_row (pre-weighted layouts, wrow = c w1),
// the catalogue weight otherwise; FP32 accumulation keeps it as a pair.
template <bool WEIGHTED, int HM>
__device__ __forceinline__ double staged_weight(double wrow, double w) {
    if (pre_weighted<WEIGHTED, HM>()) w = wrow * w;
    if (is_f32(HM)) w = f32_pair(w);
    return w;
}
// The U pairs of a pair-loop group together: every returning atomic is issued
// before the first sum it returns is used, so their latencies overlap.
template <int U>
__device__ __forceinline__ void hist_add_f32(const LdsTabs &t, const int (&off)[U], const double (&wrow)[U], float wl,
                                             const bool (&acc)[U]) {
    char *h = reinterpret_cast<char *>(t.hist);
    float ph[U], pl[U], old[U] = {};
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const float2 wr = __builtin_bit_cast(float2, wrow[u]);
        ph[u] = wr.x * wl;
        pl[u] = __builtin_fmaf(wr.y, wl, __builtin_fmaf(wr.x, wl, -ph[u]));
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (acc[u]) old[u] = atomicAdd(reinterpret_cast<float *>(h + t.xhi + (off[u] >> 1)), ph[u]);
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (acc[u]) {
            const float sum = old[u] + ph[u];
            const float bv = sum - old[u];
            const float err = (old[u] - (sum - bv)) + (ph[u] - bv);
            atomicAdd(reinterpret_cast<double *>(h + off[u]), (double)err + (double)pl[u]);
        }
}

// Bin of a squared separation (exact for r2min < d < r2max; any value in
// [0, nb] otherwise).  STEP1: the lookup table is fine enough that one step
// over the squared edges always suffices (checked on the host).
template <bool STEP1>
__device__ __forceinline__ int bin_any(const KArgs &A, const LdsTabs &t, double d) {
    const int j = (int)fmin(d * A.lut_scale, (double)A.lut_m);
    int k = t.lut[j];
    if constexpr (STEP1) {
        k += d >= t.sq[k + 1] ? 1 : 0;
    } else {
        while (k < A.nb - 1 && d >= t.sq[k + 1]) ++k;
    }
    return k;
}

// Exact FP64 evaluation of one pair (device.cu:299-307 restated): r23 from
// the image positions, strict range filter, exact bin, FP32 triangle test
// with the correctly rounded sqrtf.
template <int ORD, bool WEIGHTED, int HM>
__device__ __forceinline__ void exact_pair(const KArgs &A, const LdsTabs &t, double ax, double ay, double az, int ka,
                                           float fa, double w1, double w1a, double wa, double bx, double by, double bz,
                                           int kb, float fb, double w1b, double wb, bool ok) {
    const double d = sqrsep(ax - bx, ay - by, az - bz);
    if (ok && d > A.r2min && d < A.r2max) {
        const int k = bin_any<false>(A, t, d);
        const float f = __builtin_sqrtf((float)d);
        const float lo = fminf(fminf(fa, fb), f);
        const float hi = fmaxf(fmaxf(fa, fb), f);
        const float mid = __builtin_amdgcn_fmed3f(fa, fb, f);
        if (lo + mid > hi)  // the reference rejects sqrtf(r1)+sqrtf(r2) <= sqrtf(r3), device.cu:173
            hist_add<ORD, WEIGHTED, HM>(t, sorted_slot<HM>(A, t, ka, kb, k), w1, w1a, wa, w1b, wb);
    }
}

// ---------------------------------------------------------------------------
// Pair traversal (shared by both paths).
//
// Lanes hold one entry each of a 64-entry block of one list; "rows" are
// entries staged in the wave's LDS chunk, broadcast to all lanes
// (rectangular passes) or gathered per lane (circulant passes: lane l of a
// block of m entries pairs with entry (l + s) mod m, s = 1 .. m/2, so every
// unordered pair of the block is visited once with no lane idle except in the
// half step s = m/2 of an even block).
//   SYM (p2 and p3 the same catalogue, one list of n entries):
//     full region F = the first 64*floor(n/64) entries, tail T = the rest;
//     (1) chunks of F as rows x full blocks of F as lanes, pairs a < b: the
//         rows before the block (rectangular) and the block itself
//         (circulant);
//     (2) T as rows x the full blocks as lanes (rectangular), then T's own
//         pairs (circulant).
//   ASYM (S2 x S3): chunks of S2 as rows x full blocks of S3 as lanes; then
//     the S3 tail as rows x blocks of S2 as lanes (roles swapped, so only the
//     last S2 block has idle lanes).
// Every pair (SYM: unordered pair, counted twice) is visited exactly once.
template <class Pol>
__device__ __forceinline__ void traverse(Pol &P, int n2, int n3) {
    const int C = P.C;
    if constexpr (Pol::kSym) {
        // row segments: chunks of F, then T; lane blocks: the full blocks
        // from the segment on (F) or all of them and T itself (T)
        const int n = n2;
        const int nf = n & ~(kWave - 1);
        for (int a0 = 0, a1; a0 < n; a0 = a1) {
            a1 = a0 < nf ? min(a0 + C, nf) : n;
            P.stage(false, a0, a1);
            const int bend = a0 < nf ? nf : n;
            for (int b0 = a0 < nf ? a0 : 0; b0 < bend; b0 += kWave) {
                P.load_lanes(true, b0, n);
                P.template rect<false>(a0, b0 < a0 ? a1 : min(b0, a1));
                if (b0 >= a0 && b0 < a1) P.circ(b0, min(kWave, n - b0));
            }
        }
    } else {
        const int nf3 = n3 & ~(kWave - 1);
        if (nf3 > 0)
            for (int a0 = 0; a0 < n2; a0 += C) {
                const int a1 = min(a0 + C, n2);
                P.stage(false, a0, a1);
                for (int b0 = 0; b0 < nf3; b0 += kWave) {
                    P.load_lanes(true, b0, nf3);
                    P.template rect<false>(a0, a1);
                }
            }
        if (nf3 < n3) {
            P.stage(true, nf3, n3);
            for (int b0 = 0; b0 < n2; b0 += kWave) {
                P.load_lanes(false, b0, n2);
                P.template rect<true>(nf3, n3);
            }
        }
    }
}

// Plain path (FP64 everywhere).  Rows: Elem + weight in LDS; lanes: Elem +
// weight in registers.  Per group of U pairs: separations, bin lookups, the
// FP32 triangle test with the hardware sqrt (<= 1 ulp) -- redone with the
// correctly rounded sqrtf when the slack is within 8 ulp of the largest side
// -- and the histogram updates.
// LASTP: the two-fold loop of a last-cell primary when one-catalogue terms
// count their last-cell triples six-fold (KArgs.tri_last6): a pair of list
// entries whose cells are adjacent counts only if both cells are last cells
// (the other triples it closes are counted six-fold by a primary outside the
// last cells).  Rows carry their neighbour entry in the top bits of the
// staged bin; lanes hold the mask of row entries they skip.
template <bool SYM, int ORD, bool WEIGHTED, int HM, bool LASTP = false>
struct PlainPol {
    static constexpr bool kSym = SYM;
    static constexpr int U = 2;
    const KArgs &A;
    const LdsTabs &t;
    const Elem *L2, *L3;
    const double *W2, *W3;
    Elem *la;
    double *lw;
    double w1, wrow;  // wrow: the staged rows' weight factor (pre_weighted)
    int C, lane, row_first;
    Elem eb;
    double wb, w1b;
    float wbf;  // the lane's weight in FP32 (kHistF32: FP32 weights)
    bool vb;
    unsigned lastmask, skipb;  // LASTP

    __device__ PlainPol(const KArgs &A_, const LdsTabs &t_, double w1_, const char *l2, const char *l3, char *wlds,
                        int lane_, unsigned lastmask_)
        : A(A_), t(t_), w1(w1_), lane(lane_), lastmask(lastmask_) {
        wrow = w1 * (double)ord_count(ORD);
        C = A.chunk_elems;
        L2 = reinterpret_cast<const Elem *>(l2);
        W2 = reinterpret_cast<const double *>(l2 + (size_t)A.cap * sizeof(Elem));
        L3 = reinterpret_cast<const Elem *>(l3);
        W3 = reinterpret_cast<const double *>(l3 + (size_t)A.cap * sizeof(Elem));
        la = reinterpret_cast<Elem *>(wlds);
        lw = reinterpret_cast<double *>(wlds + (size_t)(C + 4) * sizeof(Elem));
        row_first = 0;
    }

    __device__ void stage(bool from3, int first, int last) {
        const Elem *L = from3 ? L3 : L2;
        const double *W = from3 ? W3 : W2;
        wave_sync();
        for (int i = lane; i < last - first + 4; i += kWave) {
            Elem e;
            double w = 0.0;
            if (i < last - first) {
                e = L[first + i];
                e.k = row_bin<HM>(A, e.k & 0xff) | (LASTP ? (e.k >> 8) << 24 : 0);
                if (WEIGHTED) w = pre_weighted<WEIGHTED, HM>() ? wrow * W[first + i] : W[first + i];
                if (is_f32(HM)) w = f32_pair(w);
            } else {  // padding: far away, never in range
                e.x = e.y = e.z = 1e150;
                e.k = 0;
                e.f = 0.0f;
            }
            la[i] = e;
            if (WEIGHTED) lw[i] = w;
        }
        row_first = first;
        wave_sync();
    }

    __device__ void load_lanes(bool from3, int b0, int end) {
        const int b = b0 + lane;
        vb = b < end;
        wb = 0.0;
        if (vb) {
            eb = (from3 ? L3 : L2)[b];
            if constexpr (LASTP) {
                const int sb = eb.k >> 8;
                skipb = entry_adjacent(sb) & ~(((lastmask >> sb) & 1u) ? lastmask : 0u);
            }
            eb.k = lane_bin<HM>(A, eb.k & 0xff);
            if (WEIGHTED) wb = (from3 ? W3 : W2)[b];
        } else {
            eb.x = eb.y = eb.z = -1e150;
            eb.k = 0;
            eb.f = 0.0f;
        }
        if constexpr (!pre_weighted<WEIGHTED, HM>()) w1b = w1 * wb;
        wbf = (float)wb;
    }

    // U pairs: rows ea[u] (weights wa[u]) with the lane entry.  SWAP: the rows
    // are the p3 entries, so the weight product is (w1*w_lane)*w_row.
    template <bool SWAP>
    __device__ __forceinline__ void group(const Elem (&ea)[U], const double (&wa)[U], const bool (&ok)[U]) {
        const double r2min = A.r2min, r2max = A.r2max, scale = A.lut_scale;
        const float near_rel = 8.0f * 1.1920929e-7f;
        double d[U];
        bool inr[U];
        int k[U], kra[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            bool live = ok[u];
            kra[u] = ea[u].k;
            if constexpr (LASTP) {
                live = live & (((skipb >> ((unsigned)kra[u] >> 24)) & 1u) == 0u);
                kra[u] &= 0xffffff;
            }
            d[u] = sqrsep(ea[u].x - eb.x, ea[u].y - eb.y, ea[u].z - eb.z);
            inr[u] = live & (d[u] > r2min) & (d[u] < r2max);
            const double dj = inr[u] ? d[u] : 0.0;
            k[u] = t.lut[(int)(dj * scale)];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if constexpr (!ctab_like(HM)) {
                k[u] += d[u] >= t.sq[k[u] + 1] ? 1 : 0;
            } else {
                while (k[u] < A.nb - 1 && d[u] >= t.sq[k[u] + 1]) ++k[u];
            }
        }
        bool acc[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float fd = (float)d[u];
            const Sides s = sort3(ea[u].f, eb.f, __builtin_amdgcn_sqrtf(fd));
            const float sum = s.lo + s.mid;
            const bool near = (fabsf(sum - s.hi) <= near_rel * s.hi) | (fd < 1e-30f);
            acc[u] = inr[u] & !near & (sum > s.hi);
            if (inr[u] & near) {  // rare: redo the test with the correctly rounded sqrtf
                const Sides e = sort3(ea[u].f, eb.f, __builtin_sqrtf(fd));
                acc[u] = e.lo + e.mid > e.hi;
            }
        }
        if constexpr (WEIGHTED && is_f32(HM)) {
            int slot[U];
#pragma unroll
            for (int u = 0; u < U; ++u) slot[u] = acc[u] ? pair_slot<HM>(A, t, kra[u], eb.k, k[u]) : 0;
            hist_add_f32<U>(t, slot, wa, wbf, acc);
            return;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (acc[u]) {
                const int slot = pair_slot<HM>(A, t, kra[u], eb.k, k[u]);
                if constexpr (pre_weighted<WEIGHTED, HM>())
                    hist_add_value(t, slot, wa[u] * wb);
                else if (SWAP)
                    hist_add<ORD, WEIGHTED, HM>(t, slot, w1, w1b, wb, w1 * wa[u], wa[u]);
                else
                    hist_add<ORD, WEIGHTED, HM>(t, slot, w1, w1 * wa[u], wa[u], w1b, wb);
            }
    }

    template <bool SWAP>
    __device__ void rect(int r0, int r1) {
        for (int a = r0; a < r1; a += U) {
            Elem ea[U];
            double wa[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ea[u] = la[a - row_first + u];
                wa[u] = WEIGHTED ? lw[a - row_first + u] : 0.0;
                ok[u] = vb && a + u < r1;
            }
            group<SWAP>(ea, wa, ok);
        }
    }

    __device__ void circ(int b0, int m) {
        const int smax = m / 2;
        for (int s = 1; s <= smax; s += U) {
            Elem ea[U];
            double wa[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int st = s + u;
                int off = lane + st;
                off -= off >= m ? m : 0;
                ok[u] = vb && st <= smax && (2 * st != m || lane < st);
                const int li = b0 - row_first + (ok[u] ? off : 0);
                ea[u] = la[li];
                wa[u] = WEIGHTED ? lw[li] : 0.0;
            }
            group<false>(ea, wa, ok);
        }
    }
};

// Recheck queue entries: (S2 index, S3 index).  The unit-weight layout packs
// them into 16 bits each (its lists hold fewer than 46341 entries: the host
// uses kHistDense only when chunk * cap^2 < 2^31); the others keep two words.
template <int HM>
struct QueueEntry {
    using T = uint2;
    static __device__ __forceinline__ T pack(unsigned a, unsigned b) { return make_uint2(a, b); }
    static __device__ __forceinline__ uint2 unpack(T e) { return e; }
};
template <>
struct QueueEntry<kHistDense> {
    using T = uint32_t;
    static __device__ __forceinline__ T pack(unsigned a, unsigned b) { return a | (b << 16); }
    static __device__ __forceinline__ uint2 unpack(T e) { return make_uint2(e & 0xffffu, e >> 16); }
};

// Mixed path, deferred exact rechecks: every queued (S2 entry, S3 entry) pair
// is evaluated with the FP64 image positions (catalogue point + periodic
// shift of its neighbour entry, rebuilt as in build_list), one pair per lane.
// PRIMW: the term's secondaries carry unit weights (wa = wb = 1).
template <int ORD, bool WEIGHTED, int HM, bool PRIMW = false>
__device__ void drain_queue(const KArgs &A, const LdsTabs &t, double w1, const char *l2, const char *l3,
                            const CatRef &c2, const CatRef &c3, const PrimRef &pr,
                            const typename QueueEntry<HM>::T *q, int qn, int lane) {
    const MList m2 = mlist(l2, A.cap), m3 = mlist(l3, A.cap);
    wave_sync();
    for (int i = lane; i < qn; i += kWave) {
        const uint2 e = QueueEntry<HM>::unpack(q[i]);
        const int ka = mkey(m2, e.x), kb = mkey(m3, e.y);
        const int ja = m2.I[e.x], jb = m3.I[e.y];
        double sax, say, saz, sbx, sby, sbz;
        neighbour(A, pr.cx, pr.cy, pr.cz, ka >> 8, sax, say, saz);
        neighbour(A, pr.cx, pr.cy, pr.cz, kb >> 8, sbx, sby, sbz);
        // the image positions and FP32 roots of the two sides, as the list build made them
        const double ax = c2.x[ja] + sax, ay = c2.y[ja] + say, az = c2.z[ja] + saz;
        const double bx = c3.x[jb] + sbx, by = c3.y[jb] + sby, bz = c3.z[jb] + sbz;
        const float fa = __builtin_sqrtf((float)sqrsep(pr.px - ax, pr.py - ay, pr.pz - az));
        const float fb = __builtin_sqrtf((float)sqrsep(pr.px - bx, pr.py - by, pr.pz - bz));
        const double wa = PRIMW ? 1.0 : WEIGHTED ? c2.w[ja] : 0.0, wb = PRIMW ? 1.0 : WEIGHTED ? c3.w[jb] : 0.0;
        exact_pair<ORD, WEIGHTED || PRIMW, HM>(A, t, ax, ay, az, ka & 0xff, fa, w1, w1 * wa, wa, bx, by, bz,
                                                 kb & 0xff, fb, w1 * wb, wb, true);
    }
    wave_sync();
}

// Mixed path.  With r = q - p1 re-centred in FP32 and s = |r|^2:
// r23^2 = s_a + s_b - 2 r_a.r_b, and the triangle (p1, a, b) is degenerate
// only when r_a and r_b are nearly collinear, i.e. when s_a*s_b - (r_a.r_b)^2
// is small.  Per pair: one dot product, one lookup of the log-spaced cell of
// r23^2 that returns the number of squared edges below the cell and the (at
// most one) edge inside its guard-band reach, the bin from the sign of
// r23^2 - edge, and the collinearity bound.  A pair whose r23^2 lies within
// the guard band g32 of that edge, or whose collinearity measure is within
// the error bound (make_mixed_tables), is queued for the exact FP64
// evaluation.
// Rows: float4 {r, s} + bin + weight in LDS; lanes in registers.
// PRIMW (with WEIGHTED false): a weighted term whose secondaries carry unit
// weights; every ordered triplet adds (w1 * 1) * 1 = w1, so an accepted pair
// adds the wave-uniform ord * w1 and no weights are staged or loaded.
// LASTP: as PlainPol's.
template <bool SYM, int ORD, bool WEIGHTED, int HM, bool PRIMW = false, bool LASTP = false>
struct MixedPol {
    static constexpr bool kSym = SYM;
    static constexpr int U = 4;
    static constexpr int kShift = HM == kHistDense ? 2 : t3_like(HM) ? 1 : 0;  // bin units of the layout
    static __device__ __forceinline__ float E_(uint2 e) { return __uint_as_float(e.y); }
    const KArgs &A;
    const LdsTabs &t;
    MList m2, m3;
    CatRef c2, c3;
    const char *l2, *l3;
    PrimRef pr;
    double w1;
    int lane;
    float4 *la;
    int *lk;
    double *lw;
    typename QueueEntry<HM>::T *q;
    int qn, C, row_first;
    const uint2 *lb;  // lookup table shifted down by its first cell: read at cell indices >= lut32_c0 only
    float lo32, g32, tri_t;
    int sh32;
    unsigned nbx;  // nb in the layout's bin units
    float4 rb;
    int kb, bidx;
    double wb, w1b;
    double wrow;  // the staged rows' weight factor (pre_weighted; PRIMW: the value of an accepted pair)
    float wbf;    // the lane's weight in FP32 (kHistF32: FP32 weights)
    bool vb;
    unsigned skipb;  // LASTP: row entries this lane skips
    int rot_b0;      // circ_full: the diagonal block's first list index

    __device__ MixedPol(const KArgs &A_, const LdsTabs &t_, double w1_, const char *l2_, const char *l3_, char *wlds,
                        int lane_, const CatRef &cat2, const CatRef &cat3, const PrimRef &pr_)
        : A(A_), t(t_), c2(cat2), c3(cat3), l2(l2_), l3(l3_), pr(pr_), w1(w1_), lane(lane_) {
        C = A.chunk_elems;
        const int Cp = C + 4;
        m2 = mlist(l2, A.cap);
        m3 = mlist(l3, A.cap);
        la = reinterpret_cast<float4 *>(wlds);
        lk = reinterpret_cast<int *>(wlds + (size_t)Cp * sizeof(float4));
        lw = reinterpret_cast<double *>(wlds + (size_t)Cp * (sizeof(float4) + sizeof(int)));
        // unit-weight layout (kHistDense): no weights staged, the queue follows the bins
        q = reinterpret_cast<typename QueueEntry<HM>::T *>(wlds + (size_t)Cp * (sizeof(float4) + sizeof(int) +
                                                           (HM == kHistDense ? 0 : sizeof(double))));
        qn = 0;
        row_first = 0;
        lb = t.lut32 - A.lut32_c0;
        // the pair loop's constants as VGPR operands
        lo32 = in_vgpr(A.lut32_lo);
        sh32 = in_vgpr(A.lut32_shift);
        g32 = in_vgpr(A.g32);
        tri_t = in_vgpr(A.tri_t);
        nbx = in_vgpr((unsigned)A.nb << kShift);
        wrow = w1 * (double)ord_count(ORD);
    }

    __device__ void stage(bool from3, int first, int last) {
        const MList &m = from3 ? m3 : m2;
        const CatRef &c = from3 ? c3 : c2;
        wave_sync();
        for (int i = lane; i < last - first + 4; i += kWave) {
            float4 r = make_float4(3e18f, 3e18f, 3e18f, 2.7e37f);  // padding: never in range
            int k = 0;
            double w = 0.0;
            if (i < last - first) {
                r = m.R[first + i];
                const int kk = __float_as_int(r.w);
                r.w = rsq(r);
                k = row_bin<HM>(A, kk & 0xff) | (LASTP ? (kk >> 8) << 24 : 0);
                if (WEIGHTED) w = staged_weight<WEIGHTED, HM>(wrow, c.unit ? 1.0 : gptr(c.w)[m.I[first + i]]);
            }
            la[i] = r;
            lk[i] = k;
            if (WEIGHTED) lw[i] = w;
        }
        row_first = first;
        wave_sync();
    }

    __device__ void load_lanes(bool from3, int b0, int end) {
        const MList &m = from3 ? m3 : m2;
        bidx = b0 + lane;
        vb = bidx < end;
        rb = make_float4(-3e18f, -3e18f, -3e18f, 2.7e37f);
        kb = 0;
        wb = 0.0;
        if (vb) {
            rb = m.R[bidx];
            const int kk = __float_as_int(rb.w);
            rb.w = rsq(rb);
            kb = lane_bin<HM>(A, kk & 0xff);
            if constexpr (LASTP) {
                const int sb = kk >> 8;
                skipb = entry_adjacent(sb) & ~(((pr.lastmask >> sb) & 1u) ? pr.lastmask : 0u);
            }
            const CatRef &c = from3 ? c3 : c2;
            if (WEIGHTED) wb = c.unit ? 1.0 : gptr(c.w)[m.I[bidx]];
        }
        if constexpr (!pre_weighted<WEIGHTED, HM>()) w1b = w1 * wb;
        wbf = (float)wb;
    }

    __device__ void drain() {
        drain_queue<ORD, WEIGHTED, HM, PRIMW>(A, t, w1, l2, l3, c2, c3, pr, q, qn, lane);
        qn = 0;
    }

    // U pairs: rows (ra, ka, wa, list index ai) with the lane entry.  SWAP:
    // the rows are S3 entries and the lanes S2 entries.
    // ROT: the rows are the lanes' own block rotated (circ_full): ai holds
    // the step st, and the row's list index is rot_b0 + ((lane + st) & 63).
    template <bool SWAP, bool ROT = false>
    __device__ __forceinline__ void group(const float4 (&ra)[U], const int (&ka)[U], const double (&wa)[U],
                                          const int (&ai)[U], const bool (&ok)[U]) {
        float d[U], dot[U], ss[U];
        uint2 e[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            dot[u] = __builtin_fmaf(ra[u].z, rb.z, __builtin_fmaf(ra[u].y, rb.y, ra[u].x * rb.x));
            ss[u] = ra[u].w + rb.w;
            d[u] = __builtin_fmaf(-2.0f, dot[u], ss[u]);
            e[u] = lb[__float_as_uint(__builtin_amdgcn_fmed3f(d[u], lo32, 4.0f)) >> sh32];
        }
        bool acc[U], amb[U];
        int slot[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // bin in layout units (the table's .x is pre-scaled): edges below
            // the cell, minus one when d < E (sign of d - E; d == E lies in
            // the guard band and is rechecked)
            const float dE = d[u] - E_(e[u]);
            const int kx = shl_add<kShift>(__float_as_int(dE) >> 31, (int)e[u].x);
            const bool amb_b = fabsf(dE) <= g32;
            const bool valid = (unsigned)kx < nbx;
            const float delta = __builtin_fmaf(-dot[u], dot[u], ra[u].w * rb.w);
            const bool amb_t = delta <= tri_t;
            // bitwise, so the predicates stay lane masks in scalar registers
            amb[u] = ok[u] & (amb_b | (valid & amb_t));
            acc[u] = ok[u] & valid & !amb_b & !amb_t;
            // the row bin (an LDS read) in a register here, ahead of the
            // branches below, so that no read waits inside them
            int kr = ka[u];
            asm volatile("" : "+v"(kr));
            if constexpr (LASTP) {
                const bool live = ((skipb >> ((unsigned)kr >> 24)) & 1u) == 0u;
                amb[u] = amb[u] & live;
                acc[u] = acc[u] & live;
                kr &= 0xffffff;
            }
            slot[u] = pair_slot_x<HM>(A, t, kr, kb, kx);
        }
        static_assert(!PRIMW || (!WEIGHTED && (HM == kHistT3 || HM == kHistCtab)), "FP64 slot sums");
        if constexpr (WEIGHTED && is_f32(HM)) hist_add_f32<U>(t, slot, wa, wbf, acc);
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (acc[u] && !(WEIGHTED && is_f32(HM))) {
                if constexpr (PRIMW)
                    hist_add_value(t, slot[u], wrow);
                else if constexpr (pre_weighted<WEIGHTED, HM>())
                    hist_add_value(t, slot[u], wa[u] * wb);
                else if (SWAP)
                    hist_add<ORD, WEIGHTED, HM>(t, slot[u], w1, w1b, wb, w1 * wa[u], wa[u]);
                else
                    hist_add<ORD, WEIGHTED, HM>(t, slot[u], w1, w1 * wa[u], wa[u], w1b, wb);
            }
        bool any = false;
#pragma unroll
        for (int u = 0; u < U; ++u) any = any | amb[u];
        if (__builtin_amdgcn_ballot_w64(any)) {  // rare
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const unsigned long long m = __ballot(amb[u]);
                if (amb[u]) {
                    const unsigned ar = ROT ? (unsigned)(rot_b0 + ((lane + ai[u]) & (kWave - 1))) : (unsigned)ai[u];
                    q[qn + lane_rank(m)] = SWAP ? QueueEntry<HM>::pack((unsigned)bidx, ar)
                                                : QueueEntry<HM>::pack(ar, (unsigned)bidx);
                }
                qn += __popcll(m);
            }
            // one group adds at most U*64 entries: drain once the queue could overflow
            if (qn > kQueueCap - U * kWave) drain();
        }
    }

    template <bool SWAP>
    __device__ void rect(int r0, int r1) {
        for (int a = r0; a < r1; a += U) {
            float4 ra[U];
            int ka[U], ai[U];
            double wa[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                ra[u] = la[a - row_first + u];
                ka[u] = lk[a - row_first + u];
                wa[u] = WEIGHTED ? lw[a - row_first + u] : 0.0;
                ai[u] = a + u;
                ok[u] = vb && a + u < r1;
            }
            group<SWAP>(ra, ka, wa, ai, ok);
        }
    }

    // The diagonal block [b0, b0 + m) against itself: lane l meets the row
    // (l + st) mod m for st = 1 .. m/2 (the last step of an even m on the
    // lower half of the lanes only), each unordered pair once.  For l < m,
    // l + st < 2m, so (l + st) mod m is the smaller of l + st and l + st - m
    // as unsigned values.  A lane past m (m < 64) gets l + st - m <= 63 + m/2
    // - m < 64 from the same formula: its read stays inside the staged chunk
    // (b0 - row_first + 63 < C), and vb masks its pairs.
    __device__ void circ(int b0, int m) {
        if (m == kWave) {
            circ_full(b0);
            return;
        }
        const int smax = m / 2;
        const bool lower = lane < smax;  // lanes of the last step when m is even
        const float4 *lab = la + (b0 - row_first);
        const int *lkb = lk + (b0 - row_first);
        const double *lwb = lw + (b0 - row_first);
        const unsigned mu = (unsigned)m;
        for (int s = 1; s <= smax; s += U) {
            float4 ra[U];
            int ka[U], ai[U];
            double wa[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int st = s + u;
                const unsigned x = (unsigned)(lane + st), xm = x - mu;
                const int off = (int)(xm < x ? xm : x);
                ok[u] = vb & (st <= smax) & ((2 * st != m) | lower);
                ra[u] = lab[off];
                ka[u] = lkb[off];
                wa[u] = WEIGHTED ? lwb[off] : 0.0;
                ai[u] = b0 + off;
            }
            group<false>(ra, ka, wa, ai, ok);
        }
    }

    // A full diagonal block (m = 64) without row reads: each lane loads the
    // staged row form of its own entry once, and the copies move one lane
    // down per step (DPP wave_rol:1, lane l taking lane l + 1 mod 64), so at
    // step st lane l holds entry (l + st) mod 64 -- the circulant pairing of
    // circ, with no LDS gather per pair.
    static __device__ __forceinline__ int rol1(int x) { return __builtin_amdgcn_mov_dpp(x, 0x134, 0xF, 0xF, false); }
    static __device__ __forceinline__ float rol1(float x) { return __int_as_float(rol1(__float_as_int(x))); }
    static __device__ __forceinline__ double rol1(double x) {
        const long long b = __double_as_longlong(x);
        const unsigned lo = (unsigned)rol1((int)b), hi = (unsigned)rol1((int)(b >> 32));
        return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
    }
    __device__ void circ_full(int b0) {
        const int i0 = b0 - row_first + lane;
        float4 r = la[i0];
        int k = lk[i0];
        double w = WEIGHTED ? lw[i0] : 0.0;
        const bool lower = lane < kWave / 2;  // lanes of the last step
        rot_b0 = b0;
        static_assert((kWave / 2) % U == 0, "whole groups of steps");
        for (int s = 1; s <= kWave / 2; s += U) {
            float4 ra[U];
            int ka[U], ai[U];
            double wa[U];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                r.x = rol1(r.x);
                r.y = rol1(r.y);
                r.z = rol1(r.z);
                r.w = rol1(r.w);
                k = rol1(k);
                if (WEIGHTED) w = rol1(w);
                ra[u] = r;
                ka[u] = k;
                wa[u] = w;
                ai[u] = s + u;
                ok[u] = vb & ((s + u != kWave / 2) | lower);
            }
            group<false, true>(ra, ka, wa, ai, ok);
        }
    }
};

// Merge the workgroup's LDS histogram into the term's global compact
// histogram and clear it.  Dense counts go to the sorted slot of their bins.
template <int HM>
__device__ __forceinline__ void flush_hist(const KArgs &A, int term, int weighted, unsigned long long *hist,
                                           unsigned long long *pairs) {
    __syncthreads();
    unsigned long long *out = A.out + (long long)term * A.out_stride;
    if constexpr (HM == kHistExact) {
        // 128-bit merge: low words with the carry of their wrap into the high words
        for (int i = threadIdx.x; i < A.nbc; i += blockDim.x) {
            const unsigned long long lo = hist[i], hi = hist[A.nbc + i];
            if ((lo | hi) != 0ull) {
                const unsigned long long old = atomicAdd(&out[i], lo);
                const unsigned long long up = hi + (old + lo < old ? 1ull : 0ull);
                if (up != 0ull) atomicAdd(&out[A.nbc + i], up);
                hist[i] = 0ull;
                hist[A.nbc + i] = 0ull;
            }
        }
    } else if constexpr (is_f32(HM)) {
        // FP32 sum + FP64 compensation, folded into the FP64 total
        // (a unit-weight term adds u64 counts in the compensation words)
        double *comp = reinterpret_cast<double *>(hist);
        float *sum = reinterpret_cast<float *>(reinterpret_cast<char *>(hist) + (size_t)A.nbc * 8);
        for (int i = threadIdx.x; i < A.nbc; i += blockDim.x) {
            const float s = sum[i];
            const unsigned long long c = hist[i];
            if (s != 0.0f || c != 0ull) {
                if (weighted)
                    unsafeAtomicAdd(reinterpret_cast<double *>(&out[i]), (double)s + comp[i]);
                else
                    atomicAdd(&out[i], c);
                sum[i] = 0.0f;
                hist[i] = 0ull;
            }
        }
        if (threadIdx.x == 0) *pairs = 0ull;
    } else if constexpr (HM == kHistDense) {
        unsigned *h32 = reinterpret_cast<unsigned *>(hist);
        const int nb = A.nb, nb3 = nb * nb * nb;
        for (int i = threadIdx.x; i < nb3; i += blockDim.x) {
            const unsigned v = h32[i];
            if (v != 0u) {
                const int ka = i / (nb * nb), kb = (i / nb) % nb, k = i % nb;
                const int klo = min(min(ka, kb), k), khi = max(max(ka, kb), k), kmid = ka + kb + k - klo - khi;
                atomicAdd(&out[A.ctab[klo * nb + kmid] + (khi - kmid)], (unsigned long long)v);
                h32[i] = 0u;
            }
        }
        if (threadIdx.x == 0) *pairs = 0ull;
    } else {
        for (int i = threadIdx.x; i < A.nslots; i += blockDim.x) {
            const unsigned long long v = hist[i];
            if (v != 0ull) {
                if (weighted)
                    unsafeAtomicAdd(reinterpret_cast<double *>(&out[i]), __longlong_as_double((long long)v));
                else
                    atomicAdd(&out[i], v);
                hist[i] = 0ull;
            }
        }
    }
    __syncthreads();
}

// ORD: ordered triplets per evaluated pair (hist_add); by default kOrdSym for
// a symmetric traversal (p2 and p3 from one list) and kOrdOne otherwise.
// PRIMW: a weighted term with unit-weight secondaries (MixedPol; the plain
// path takes it as any weighted term).
template <bool MIXED, int HM, bool SYM, bool WEIGHTED, int ORD = SYM ? kOrdSym : kOrdOne, bool PRIMW = false,
          bool LASTP = false>
__device__ __forceinline__ void pair_loop(const KArgs &A, const LdsTabs &t, double w1, const char *l2, int n2,
                                          const char *l3, int n3, char *wlds, int lane, const CatRef &c2,
                                          const CatRef &c3, const PrimRef &pr) {
    if constexpr (MIXED) {
        MixedPol<SYM, ORD, WEIGHTED && !PRIMW, HM, PRIMW, LASTP> P(A, t, w1, l2, l3, wlds, lane, c2, c3, pr);
        traverse(P, n2, n3);
        if (P.qn > 0) P.drain();
    } else {
        PlainPol<SYM, ORD, WEIGHTED, HM, LASTP> P(A, t, w1, l2, l3, wlds, lane, pr.lastmask);
        traverse(P, n2, n3);
    }
}

// One primary of a term over one catalogue on the six-fold path (TermDesc.tri).
//
// With every cell index the exact floor of coordinate / cell size, two points
// closer than the cell size (rmax) lie in adjacent cells, and a wrap across
// the periodic boundary only ever joins the last cell of an axis to the
// first.  So for three points none of which lies in a last cell, each sees
// the other two as unshifted neighbours whenever one of them sees the pair,
// and all three compute the same three squared separations bit for bit
// (sqrSep(a - b) == sqrSep(b - a)), the same sorted bins and the same FP32
// triangle test: the reference counts such a triple in all six orders or in
// none.  It is evaluated once, from its lowest-index point, and counted six
// times.  A triple with a point in a last cell may be seen from one primary
// only (the partial last cell of a box that is not a whole number of cells:
// device.cu:89-122), so each of its primaries counts its own two orders, as
// the reference does.  A primary in a last cell therefore takes the ordinary
// path (the caller); any other primary counts
//   pairs of H (not last, index > p)                 x 6
//   pairs of (H + L) x B  (L: not last, index < p;
//                          B: entries in a last cell) x 2
//   pairs of B                                        x 2
// which is its own share of the reference's ordered count (hist_add forms the
// weight products of those orders).  Returns false on a list overflow.
template <bool MIXED, int HM, bool WEIGHTED>
__device__ __forceinline__ bool tri_primary(const KArgs &A, const LdsTabs &t, const TermDesc &T, int p,
                                            double px, double py, double pz, double pw, int cx, int cy, int cz,
                                            char *l2, char *l3, char *wlds, int lane, unsigned long long &wpairs) {
    constexpr int LM = MIXED ? kListMixed : kListPlain;
    int nb = 0;
    const int nh = uniform(build_list<LM, kSelSplit, lut32_units(HM)>(A, T.x2, T.y2, T.z2, T.w2, T.st2, T.prune2, px, py, pz, cx, cy, cz, l2,
                                                     lane, t.sq, t.lut, p, l3, &nb));
    const int nbu = uniform(nb);
    if (nh + nbu > A.nmax) return false;
    wave_sync();
    const CatRef c = {T.x2, T.y2, T.z2, T.w2, T.wunit2};
    const PrimRef pr = {cx, cy, cz, px, py, pz};
    if (nh > 1) pair_loop<MIXED, HM, true, WEIGHTED, kOrdSix>(A, t, pw, l2, nh, l2, nh, wlds, lane, c, c, pr);
    if (A.tri_last6) {
        // Triples with last-cell points, six-fold too: every triple of this
        // primary's lists that has its sides in range is closed under the
        // reference's neighbour rule (A's cell is a full cell on every axis,
        // so entries on either side of it are at least rmax apart), and it is
        // counted here when this primary is its lowest-index point outside
        // the last cells: pairs of the higher entries of list 2 with the
        // last-cell entries, and pairs of last-cell entries.  Last-cell
        // primaries skip such triples (LASTP).
        if (nbu > 0) {
            if (nh > 0)
                pair_loop<MIXED, HM, false, WEIGHTED, kOrdSix>(A, t, pw, l2, nh, l3, nbu, wlds, lane, c, c, pr);
            if (nbu > 1) pair_loop<MIXED, HM, true, WEIGHTED, kOrdSix>(A, t, pw, l3, nbu, l3, nbu, wlds, lane, c, c, pr);
        }
        if constexpr (counts_pairs(HM)) {
            const unsigned long long h = (unsigned long long)nh, b = (unsigned long long)nbu;
            wpairs += 3ull * h * (h > 0 ? h - 1 : 0) + 6ull * h * b + 3ull * b * (b > 0 ? b - 1 : 0);
        }
        return true;
    }
    int nrow = nh;
    if (nbu > 0) {
        nrow = uniform(build_list<LM, kSelLow, lut32_units(HM)>(A, T.x2, T.y2, T.z2, T.w2, T.st2, T.prune2, px, py, pz, cx, cy, cz, l2, lane, t.sq,
                                               t.lut, p, nullptr, nullptr, nh));
        if (nrow + nbu > A.nmax) return false;
        wave_sync();
        if (nrow > 0)
            pair_loop<MIXED, HM, false, WEIGHTED, kOrdSym>(A, t, pw, l2, nrow, l3, nbu, wlds, lane, c, c, pr);
        if (nbu > 1) pair_loop<MIXED, HM, true, WEIGHTED>(A, t, pw, l3, nbu, l3, nbu, wlds, lane, c, c, pr);
    }
    if constexpr (counts_pairs(HM))
        wpairs += 3ull * (unsigned long long)nh * (unsigned long long)max(nh - 1, 0) +
                  2ull * (unsigned long long)nrow * (unsigned long long)nbu +
                  (unsigned long long)nbu * (unsigned long long)max(nbu - 1, 0);
    return true;
}

// Persistent triplet kernel.  Workgroups pull (term, primary range) chunks
// from the queue A.chunks with one atomic per chunk; the waves of a workgroup
// take the primaries of the chunk one at a time.  The LDS histogram of the
// workgroup is merged into the term's global histogram whenever the workgroup
// moves to another term and at exit.  Dense u32 counts are also merged once
// the pairs added since the last merge reach A.flush_pairs <= 2^31: one chunk
// adds at most chunk * cap^2 < 2^31 (checked on the host), so no count wraps.
// The unit-weight mixed variant is held to 96 VGPRs (five waves per SIMD):
// its pair loop is latency bound, and with four-wave workgroups of about
// 31 KB five of them share a CU (20 waves instead of 16).
// WT: the launch holds terms with non-unit weights.  Such a variant carries
// the six-fold path for those terms only (its unit-weight terms take the
// two-fold path); a WT = false variant carries no weighted code.  Both
// six-fold variants in one kernel would not fit its 128 VGPRs.
// The mixed exact-accumulation variant (128-bit slot adds) gets 168 VGPRs
// (three waves per SIMD, workgroups of at most 12 waves) rather than spill.
template <bool MIXED, int HM, bool WT>
__global__ __launch_bounds__(MIXED && HM == kHistExact ? 768 : 1024) __attribute__((
    amdgpu_waves_per_eu(MIXED && HM == kHistDense ? 5 : MIXED && HM == kHistExact ? 3 : 4))) void
k_triplets(KArgs A) {
    static_assert(!(WT && HM == kHistDense), "dense counts are for unit weights");
    static_assert(WT || (HM != kHistExact && !is_f32(HM)), "exact and FP32 accumulation are for weighted terms");
    extern __shared__ __attribute__((aligned(16))) char lds[];
    __shared__ int s_chunk;
    __shared__ int s_next;
    __shared__ int s_seg;  // segments drained so far, counted from this workgroup's own
    __shared__ unsigned long long s_pairs;
    LdsTabs t;
    t.hist = reinterpret_cast<unsigned long long *>(lds);
    double *sq = reinterpret_cast<double *>(lds + A.off_sq);
    uint16_t *lut = reinterpret_cast<uint16_t *>(lds + A.off_lut);
    int *ctab = reinterpret_cast<int *>(lds + A.off_ctab);
    uint2 *lut32 = reinterpret_cast<uint2 *>(lds + A.off_lut32);
    uint16_t *t3 = reinterpret_cast<uint16_t *>(lds + A.off_t3);
    t.sq = sq;
    t.lut = lut;
    t.ctab = ctab;
    t.lut32 = lut32;
    t.t3 = t3;
    t.xscale = A.xscale;
    t.xhi = A.nbc * 8;
    const int tid = threadIdx.x, nthr = blockDim.x;
    // the partner terms' histogram (TermDesc.partner) follows the first
    unsigned long long *hist2 = reinterpret_cast<unsigned long long *>(lds + A.off_hist2);
    for (int i = tid; i < A.nslots; i += nthr) t.hist[i] = 0ull;
    if (A.off_hist2 > 0)
        for (int i = tid; i < A.nslots; i += nthr) hist2[i] = 0ull;
    if (tid == 0) {
        s_pairs = 0ull;
        s_seg = 0;
    }
    if constexpr (MIXED) {
        // FP64 tables are only read by the rare exact rechecks: global memory
        t.sq = A.sq;
        t.lut = A.lut;
        // entry .x (edges below the cell) in the layout's bin units
        constexpr int kshift = lut32_units(HM);
        for (int i = tid; i < A.lut32_n; i += nthr) {
            uint2 v = A.lut32[i];
            v.x <<= kshift;
            lut32[i] = v;
        }
    } else {
        for (int i = tid; i <= A.nb; i += nthr) sq[i] = A.sq[i];
        for (int i = tid; i <= A.lut_m; i += nthr) lut[i] = A.lut[i];
    }
    if constexpr (t3_like(HM))
        for (int i = tid; i < A.nb * A.nb * (A.nb + 2); i += nthr) t3[i] = A.t3[i];
    else if constexpr (ctab_like(HM))
        for (int i = tid; i < A.nb * A.nb; i += nthr) ctab[i] = A.ctab[i];
    const int lane = tid & (kWave - 1), wave = tid / kWave, nw = nthr / kWave;
    char *wlds = lds + A.off_wave + wave * A.wave_bytes;
    char *slot = A.scratch + (long long)(blockIdx.x * nw + wave) * A.slot_bytes;
    char *l2 = slot;
    char *l3 = slot + (size_t)A.cap * list_entry_bytes(MIXED ? kListMixed : kListPlain);
    int cur = -1, cur_weighted = 0, cur_partner = -1, partner_weighted = 0;
    unsigned long long wpairs = 0ull;  // pairs this wave added in the current chunk
    __syncthreads();
    const int home = (int)(blockIdx.x % (unsigned)A.nseg);
    for (;;) {
        if (tid == 0) {
            // the next chunk of the first segment, from this workgroup's own
            // on, that is not drained (counters only grow: a drained segment
            // stays drained)
            int ci = A.nchunks;
            for (int k = s_seg; k < A.nseg; ++k) {
                const int sg = home + k < A.nseg ? home + k : home + k - A.nseg;
                const unsigned v = atomicAdd(A.counter + sg, 1u);
                if (v < (unsigned)(A.seg[sg + 1] - A.seg[sg])) {
                    ci = A.seg[sg] + (int)v;
                    break;
                }
                s_seg = k + 1;
            }
            s_chunk = ci;
            s_next = 0;
        }
        __syncthreads();
        // LDS reads count as divergent to the compiler: make the chunk index
        // (and everything derived from it) wave-uniform explicitly
        const int ci = uniform(s_chunk);
        if (ci >= A.nchunks) break;
        const Chunk ch = A.chunks[ci];
        const TermDesc T = A.terms[ch.term];
        const bool full = counts_pairs(HM) && uniform((int)(s_pairs >= A.flush_pairs)) != 0;
        if (ch.term != cur || full) {
            if (cur >= 0) flush_hist<HM>(A, cur, cur_weighted, t.hist, &s_pairs);
            if (cur_partner >= 0) flush_hist<HM>(A, cur_partner, partner_weighted, hist2, &s_pairs);
            cur = ch.term;
            cur_weighted = T.weighted;
            cur_partner = T.partner;
            partner_weighted = T.partner >= 0 ? A.terms[T.partner].weighted : 0;
        }
        for (;;) {
            // every lane takes part (lane 0 adds 1, the others 0): a branch on
            // lane == 0 here gets threaded through the loop by the optimiser,
            // which leaves the whole primary loop under divergent control
            const int p = uniform(atomicAdd(&s_next, lane == 0 ? 1 : 0)) + ch.begin;
            if (p >= ch.end) break;
            const double px = uniform(T.x1[p]), py = uniform(T.y1[p]), pz = uniform(T.z1[p]),
                         pw = uniform(T.w1[p]);
            bool bad = false;
            const bool clamp = !A.survey;  // device.cu:78-80 vs 84-87
            const int cx = axis_cell(px, A.cs[0], A.nc[0], clamp, bad);
            const int cy = axis_cell(py, A.cs[1], A.nc[1], clamp, bad);
            const int cz = axis_cell(pz, A.cs[2], A.nc[2], clamp, bad);
            if (bad) {
                if (lane == 0) atomicOr(A.err, kErrDomain);
                continue;
            }
            // the six-fold path, unless the primary lies in a last cell (box)
            if (T.tri == kTriSix && (A.survey || (cx != A.nc[0] - 1 && cy != A.nc[1] - 1 && cz != A.nc[2] - 1))) {
                bool ok = true;
                if (T.weighted != (int)WT) {  // never planned so
                    if (lane == 0) atomicOr(A.err, kErrMode);
                } else {
                    ok = tri_primary<MIXED, HM, WT>(A, t, T, p, px, py, pz, pw, cx, cy, cz, l2, l3, wlds, lane, wpairs);
                }
                if (!ok && lane == 0) atomicOr(A.err, kErrOverflow);
                wave_sync();
                continue;
            }
            // kTriSwap (surveys: no last cells): list 2 holds the entries of
            // index above the primary's
            constexpr int LM = MIXED ? kListMixed : kListPlain;
            int n2;
            unsigned lastmask = 0u;
            if (T.tri == kTriSwap) {
                int nb_none = 0;
                n2 = build_list<LM, kSelSplit, lut32_units(HM)>(A, T.x2, T.y2, T.z2, T.w2, T.st2, T.prune2, px, py, pz, cx, cy, cz, l2, lane, t.sq,
                                               t.lut, p, l3, &nb_none);
            } else {
                n2 = build_list<LM, kSelAll, lut32_units(HM)>(A, T.x2, T.y2, T.z2, T.w2, T.st2, T.prune2, px, py, pz, cx,
                                                              cy, cz, l2, lane, t.sq, t.lut, 0, nullptr, nullptr, 0,
                                                              &lastmask);
            }
            const int n3 =
                T.sym ? n2
                      : build_list<LM, kSelAll, lut32_units(HM)>(A, T.x3, T.y3, T.z3, T.w3, T.st3, T.prune3, px, py, pz,
                                                                 cx, cy, cz, l3, lane, t.sq, t.lut);
            // wave-uniform lengths keep the pair loops' control in scalar registers
            const int n2u = uniform(n2), n3u = uniform(n3);
            if (n2u > A.nmax || n3u > A.nmax) {
                if (lane == 0) atomicOr(A.err, kErrOverflow);
                continue;
            }
            if (n3u == 0 || (n2u == 0 && T.partner < 0)) continue;
            wave_sync();
            const char *lb = T.sym ? l2 : l3;
            const CatRef c2 = {T.x2, T.y2, T.z2, T.w2, T.wunit2};
            const CatRef c3 = T.sym ? c2 : CatRef{T.x3, T.y3, T.z3, T.w3, T.wunit3};
            const PrimRef pr = {cx, cy, cz, px, py, pz, (unsigned)uniform((int)lastmask)};
            // a last-cell primary of a six-fold term (its other triples are counted six-fold)
            const bool lastp = T.tri == kTriSix && A.tri_last6;
            if constexpr (!WT) {
                if (T.weighted) {  // never planned so
                    if (lane == 0) atomicOr(A.err, kErrMode);
                    continue;
                }
            }
            if constexpr (counts_pairs(HM)) {
                const unsigned long long a2 = (unsigned long long)n2u, a3 = (unsigned long long)n3u;
                wpairs += T.tri == kTriSwap ? 2ull * a2 * a3 : T.sym ? a2 * (a2 - 1ull) : a2 * a3;
            }
            if (n2u == 0) {
                // only the partner term has pairs (below)
            } else if (lastp) {
                if (T.weighted) {
                    if constexpr (WT)
                        pair_loop<MIXED, HM, true, true, kOrdSym, false, true>(A, t, pw, l2, n2u, lb, n3u, wlds, lane,
                                                                               c2, c3, pr);
                } else {
                    pair_loop<MIXED, HM, true, false, kOrdSym, false, true>(A, t, pw, l2, n2u, lb, n3u, wlds, lane, c2,
                                                                            c3, pr);
                }
            } else if (T.sym) {
                if (T.weighted) {
                    if constexpr (WT && primw_ok(HM)) {
                        if (T.wprim)
                            pair_loop<MIXED, HM, true, true, kOrdSym, true>(A, t, pw, l2, n2u, lb, n3u, wlds, lane, c2,
                                                                            c3, pr);
                        else
                            pair_loop<MIXED, HM, true, true>(A, t, pw, l2, n2u, lb, n3u, wlds, lane, c2, c3, pr);
                    } else if constexpr (WT) {
                        pair_loop<MIXED, HM, true, true>(A, t, pw, l2, n2u, lb, n3u, wlds, lane, c2, c3, pr);
                    }
                } else {
                    pair_loop<MIXED, HM, true, false>(A, t, pw, l2, n2u, lb, n3u, wlds, lane, c2, c3, pr);
                }
            } else if (T.tri == kTriSwap) {
                // (the planner takes this only for terms of the variant's weighting)
                if (T.weighted == (int)WT)
                    pair_loop<MIXED, HM, false, WT, kOrdSwap>(A, t, pw, l2, n2u, lb, n3u, wlds, lane, c2, c3, pr);
                else if (lane == 0)
                    atomicOr(A.err, kErrMode);
            } else {
                if (T.weighted) {
                    if constexpr (WT && primw_ok(HM)) {
                        if (T.wprim)
                            pair_loop<MIXED, HM, false, true, kOrdOne, true>(A, t, pw, l2, n2u, lb, n3u, wlds, lane,
                                                                             c2, c3, pr);
                        else
                            pair_loop<MIXED, HM, false, true>(A, t, pw, l2, n2u, lb, n3u, wlds, lane, c2, c3, pr);
                    } else if constexpr (WT) {
                        pair_loop<MIXED, HM, false, true>(A, t, pw, l2, n2u, lb, n3u, wlds, lane, c2, c3, pr);
                    }
                } else {
                    pair_loop<MIXED, HM, false, false>(A, t, pw, l2, n2u, lb, n3u, wlds, lane, c2, c3, pr);
                }
            }
            if constexpr (primw_ok(HM)) {
                // the partner term (c1, c3, c3) over the list of c3 just built,
                // into the second histogram
                if (T.partner >= 0 && n3u > 1) {
                    wave_sync();
                    const TermDesc &TB = A.terms[T.partner];
                    LdsTabs t2 = t;
                    t2.hist = hist2;
                    if (uniform(TB.weighted)) {
                        if constexpr (WT) {
                            if (uniform(TB.wprim))
                                pair_loop<MIXED, HM, true, true, kOrdSym, true>(A, t2, pw, l3, n3u, l3, n3u, wlds,
                                                                                lane, c3, c3, pr);
                            else
                                pair_loop<MIXED, HM, true, true>(A, t2, pw, l3, n3u, l3, n3u, wlds, lane, c3, c3, pr);
                        }
                    } else {
                        pair_loop<MIXED, HM, true, false>(A, t2, pw, l3, n3u, l3, n3u, wlds, lane, c3, c3, pr);
                    }
                }
            }
            wave_sync();
        }
        if constexpr (counts_pairs(HM)) {
            atomicAdd(&s_pairs, lane == 0 ? wpairs : 0ull);  // every lane: no branch on lane == 0
            wpairs = 0ull;
        }
        __syncthreads();
    }
    if (cur >= 0) flush_hist<HM>(A, cur, cur_weighted, t.hist, &s_pairs);
    if (cur_partner >= 0) flush_hist<HM>(A, cur_partner, partner_weighted, hist2, &s_pairs);
}

// ---------------------------------------------------------------------------
// rppi binning (projected separation rp, line-of-sight separation pi).

// getTriangleIndexRppi (device.cu:203-255): sort the sides by rp^2 carrying
// their pi^2, FP32 triangle test on rp, rp bins, pi bins floor(sqrtf(pi^2))
// of the two smallest-rp sides.  The reference's flat index is
// T*npi^2 + rem with the dense rp triple T = k3 + k2*nb + k1*nb^2 and
// rem = p2 + p1*npi; a pi index equal to npi makes rem >= npi^2, which
// carries into the next triple exactly as the flat index does, so rem is
// returned normalised below npi^2 and T is the triple the flat index lands
// in (T >= nb^3: past the histogram, where the reference writes out of
// bounds).  false: rejected by the triangle test.
__device__ __forceinline__ bool rppi_cell(const KArgs &A, const LdsTabs &t, double r1, double r2, double r3,
                                          double q1, double q2, double q3, int &T, int &rem) {
    double x;
    if (r1 > r2) {
        x = r1; r1 = r2; r2 = x;
        x = q1; q1 = q2; q2 = x;
    }
    if (r1 > r3) {
        x = r1; r1 = r3; r3 = x;
        x = q1; q1 = q3; q3 = x;
    }
    if (r2 > r3) {
        x = r2; r2 = r3; r3 = x;
        x = q2; q2 = q3; q3 = x;
    }
    if (__builtin_sqrtf((float)r1) + __builtin_sqrtf((float)r2) <= __builtin_sqrtf((float)r3)) return false;
    const int k1 = bin_any<false>(A, t, r1), k2 = bin_any<false>(A, t, r2), k3 = bin_any<false>(A, t, r3);
    const int p1 = (int)floorf(__builtin_sqrtf((float)q1));
    const int p2 = (int)floorf(__builtin_sqrtf((float)q2));
    const int np = A.npi, np2 = np * np, nb = A.nb;
    T = k3 + k2 * nb + k1 * nb * nb;
    rem = p2 + p1 * np;
    if (rem >= np2) {  // rare: a pi index equal to npi
        const int c = rem / np2;
        T += c;
        rem -= c * np2;
    }
    return true;
}

// rppi_cell for box sides whose root, rp bin and pi bin are known: the same
// sort (the same swaps on rp^2, so ties keep their order), triangle test and
// flat index.
__device__ __forceinline__ bool rppi_cell_sides(const KArgs &A, double r1, float f1, int k1, int p1, double r2,
                                                float f2, int k2, int p2, double r3, float f3, int k3, int p3,
                                                int &T, int &rem) {
    if (r1 > r2) {
        const double x = r1; r1 = r2; r2 = x;
        const float y = f1; f1 = f2; f2 = y;
        int z = k1; k1 = k2; k2 = z;
        z = p1; p1 = p2; p2 = z;
    }
    if (r1 > r3) {
        const double x = r1; r1 = r3; r3 = x;
        const float y = f1; f1 = f3; f3 = y;
        int z = k1; k1 = k3; k3 = z;
        z = p1; p1 = p3; p3 = z;
    }
    if (r2 > r3) {
        const float y = f2; f2 = f3; f3 = y;
        int z = k2; k2 = k3; k3 = z;
        z = p2; p2 = p3; p3 = z;
    }
    if (f1 + f2 <= f3) return false;
    const int np = A.npi, np2 = np * np, nb = A.nb;
    T = k3 + k2 * nb + k1 * nb * nb;
    rem = p2 + p1 * np;
    if (rem >= np2) {
        const int c = rem / np2;
        T += c;
        rem -= c * np2;
    }
    return true;
}

// Views of a wave's stream state, after its staged keys: the flush
// permutation, then per bucket of the current term (nbm slots, the launch's
// largest bucket count rounded up to 8) counts and run positions.
struct StreamLds {
    uint16_t *perm;  // the flush's keys in bucket order
    unsigned *wcnt, *pre, *len1;
    unsigned long long *pos1, *pos2, *cur, *end, *base, *cap;  // base, cap: the term's regions

    __device__ void bind(char *p, int nbm) {
        perm = reinterpret_cast<uint16_t *>(p);
        p += (size_t)kStage * 2;
        wcnt = reinterpret_cast<unsigned *>(p);
        pre = wcnt + nbm;
        len1 = pre + nbm;
        p += (size_t)nbm * 12;
        unsigned long long *q = reinterpret_cast<unsigned long long *>(p);
        pos1 = q;
        pos2 = q + nbm;
        cur = q + 2 * nbm;
        end = q + 3 * nbm;
        base = q + 4 * nbm;
        cap = q + 5 * nbm;
    }
};

// Exclusive prefix sum over the 64 lanes of a wave; *total: the sum.
__device__ __forceinline__ unsigned wave_excl_scan(unsigned v, int lane, unsigned *total) {
    unsigned x = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const unsigned y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    *total = __shfl(x, kWave - 1);
    return x - v;
}

// Per-wave key stream state of the current term.
struct KeyStream {
    StreamLds L;
    int tbase = 0, nbt = 0, shift = 0, term = -1;

    // Ends the term: unused reserved slots get kRsPad; dry runs add their
    // key counts to the bucket cursors.
    __device__ void close(const KArgs &A, int lane) {
        if (term < 0) return;
        wave_sync();
        for (int b = 0; b < nbt; ++b) {
            const unsigned long long c = L.cur[b], e = L.end[b];
            if (A.rs_dry) {
                if (lane == 0 && c) atomicAdd(&A.rs_cursor[tbase + b], c);
                continue;
            }
            const unsigned long long b0 = L.base[b], cap = L.cap[b];
            const unsigned long long stop = e < cap ? e : cap;
            for (unsigned long long i = c + lane; i < stop; i += kWave) A.rs_keys[b0 + i] = kRsPad;
        }
        wave_sync();
        term = -1;
    }

    __device__ void open(const KArgs &A, int lane, int t, int4 tr) {
        if (t == term) return;
        close(A, lane);
        term = t;
        tbase = tr.x;
        nbt = tr.y;
        shift = tr.z;
        for (int b = lane; b < nbt; b += kWave) {
            L.cur[b] = L.end[b] = 0ull;
            if (!A.rs_dry) {
                const unsigned long long b0 = A.rs_base[tbase + b];
                L.base[b] = b0;
                L.cap[b] = A.rs_base[tbase + b + 1] - b0;
            }
        }
        wave_sync();
    }
};

// Pair evaluation of the rppi kernels on the plain FP64 path: rows from the
// wave's LDS chunk (broadcast), lane entries in registers.  Box (SURVEY
// false): rp^2 and pi^2 of the third side from the image positions; with
// p2 and p3 the same catalogue each unordered pair is evaluated once and
// counted for both orders (evaluated twice when two rp^2 tie, since the sort
// then keeps the input order of the tied sides).  Survey: the line of sight
// (p1 + p2 + p3 - 3 ori) depends on the order of the FP64 sum, so with p2 and
// p3 the same catalogue each unordered pair evaluates both orders and leaves
// one key of count 2 when they fall in the same cell, else one key each.
//
// Binned triplets leave either as atomics on the term's dense global
// histogram (STREAM false) or as keys staged in LDS and flushed, grouped by
// bucket, into the bucket regions that k_rppi_bin adds up (STREAM true).
template <bool SURVEY, bool SYM, bool WEIGHTED, bool STREAM>
struct RppiPol {
    static constexpr bool kSym = SYM;
    static constexpr int U = 2;
    const KArgs &A;
    const LdsTabs &t;
    const RElem *L2, *L3;
    RElem *la;
    unsigned long long *out;
    double px, py, pz, ppx, w1;
    int C, lane, row_first;
    RElem eb;
    double bsx, bsy, bsz;  // p1 - lane entry (survey)
    bool vb;
    int nb3, np2;
    // key stream
    const uint16_t *rank;
    unsigned *sk;
    double *sv;
    uint16_t *srk;
    StreamLds S;
    int cnt, tbase, nbt, shift;
    unsigned long long emitted;  // keys flushed by this primary

    __device__ RppiPol(const KArgs &A_, const LdsTabs &t_, const char *l2, const char *l3, char *wlds, int lane_,
                       unsigned long long *out_, double px_, double py_, double pz_, double w1_, int4 term)
        : A(A_), t(t_), out(out_), px(px_), py(py_), pz(pz_), w1(w1_), lane(lane_) {
        C = A.chunk_elems;
        L2 = reinterpret_cast<const RElem *>(l2);
        L3 = reinterpret_cast<const RElem *>(l3);
        la = reinterpret_cast<RElem *>(wlds);
        ppx = px - A.ori_x;
        row_first = 0;
        nb3 = A.nb * A.nb * A.nb;
        np2 = A.npi * A.npi;
        cnt = 0;
        emitted = 0ull;
        tbase = term.x;
        nbt = term.y;
        shift = term.z;
        if (STREAM) {
            rank = t.rk;
            char *b = wlds + A.off_stage;
            sk = reinterpret_cast<unsigned *>(b);
            b += (size_t)kStage * 4;
            if (WEIGHTED) {
                sv = reinterpret_cast<double *>(b);
                b += (size_t)kStage * 8;
            }
            srk = reinterpret_cast<uint16_t *>(b);
            S.bind(wlds + A.off_stage + A.rs_lds_off, A.rs_nbm);
        }
    }

    __device__ void stage(bool from3, int first, int last) {
        const RElem *L = from3 ? L3 : L2;
        wave_sync();
        for (int i = lane; i < last - first + 4; i += kWave) {
            RElem e{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            if (i < last - first) e = L[first + i];
            la[i] = e;
        }
        row_first = first;
        wave_sync();
    }

    __device__ void load_lanes(bool from3, int b0, int end) {
        const int b = b0 + lane;
        vb = b < end;
        eb = RElem{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (vb) eb = (from3 ? L3 : L2)[b];
        bsx = px - eb.x;
        bsy = py - eb.y;
        bsz = pz - eb.z;
    }

    // dense flat index of (T, rem) on the term's histogram
    __device__ __forceinline__ void add_direct(int T, int rem, double v, unsigned long long n) {
        const long long idx = (long long)T * np2 + rem;
        if (WEIGHTED)
            unsafeAtomicAdd(reinterpret_cast<double *>(&out[idx]), v);
        else
            atomicAdd(&out[idx], n);
    }

    // Called by every lane of the wave (has: this lane holds a key).
    __device__ __forceinline__ void emit(bool has, int T, int rem, int n, double v) {
        const bool keep = has & (T < nb3);  // past the histogram: dropped
        if constexpr (!STREAM) {
            if (keep) add_direct(T, rem, v, (unsigned long long)n);
        } else {
            const unsigned long long m = __ballot(keep);
            if (m != 0) {
                if (keep) {
                    const int pos = cnt + lane_rank(m);
                    sk[pos] = ((unsigned)rank[T] * (unsigned)np2 + (unsigned)rem) | (n == 2 ? 0x80000000u : 0u);
                    if (WEIGHTED) sv[pos] = v;
                }
                cnt += __popcll(m);
            }
        }
    }

    // The staged keys, grouped by bucket (a counting sort in LDS), go to the
    // bucket regions with one cursor atomic per bucket; keys past a region's
    // capacity are added to the dense histogram directly.
    __device__ void flush() {
        if constexpr (STREAM) {
            const int n = cnt;
            emitted += (unsigned long long)n;
            wave_sync();
            for (int b = lane; b < nbt; b += kWave) S.wcnt[b] = 0u;
            wave_sync();
            for (int i = lane; i < n; i += kWave)
                srk[i] = (uint16_t)atomicAdd(&S.wcnt[(sk[i] & 0x7fffffffu) >> shift], 1u);
            wave_sync();
            // per bucket: the keys of this flush fill the rest of the wave's
            // run (pos1, len1), then a new run of whole rs_blk blocks (pos2)
            for (int b = lane; b < nbt; b += kWave) {
                const unsigned c = S.wcnt[b];
                if (c == 0u) continue;
                if (A.rs_dry) {
                    S.cur[b] += c;
                    continue;
                }
                const unsigned long long cur = S.cur[b], end = S.end[b];
                const unsigned long long avail = end - cur;
                S.pos1[b] = cur;
                if (avail >= c) {
                    S.len1[b] = c;
                    S.cur[b] = cur + c;
                } else {
                    const unsigned long long more = c - avail, blk = (unsigned long long)A.rs_blk;
                    const unsigned long long take = (more + blk - 1) / blk * blk;
                    const unsigned long long p2 = atomicAdd(&A.rs_cursor[tbase + b], take);
                    S.len1[b] = (unsigned)avail;
                    S.pos2[b] = p2;
                    S.cur[b] = p2 + more;
                    S.end[b] = p2 + take;
                }
            }
            if (!A.rs_dry) {
                // bucket order, so that neighbouring lanes store to neighbouring slots
                unsigned tot = 0u;
                const unsigned v0 = lane < nbt ? S.wcnt[lane] : 0u;
                const unsigned p0 = wave_excl_scan(v0, lane, &tot);
                if (lane < nbt) S.pre[lane] = p0;
                if (nbt > kWave) {
                    unsigned tot1 = 0u;
                    const unsigned v1 = lane + kWave < nbt ? S.wcnt[lane + kWave] : 0u;
                    const unsigned p1 = wave_excl_scan(v1, lane, &tot1);
                    if (lane + kWave < nbt) S.pre[lane + kWave] = tot + p1;
                }
                wave_sync();
                for (int i = lane; i < n; i += kWave)
                    S.perm[S.pre[(sk[i] & 0x7fffffffu) >> shift] + srk[i]] = (uint16_t)i;
            }
            wave_sync();
            if (!A.rs_dry) {
                for (int j = lane; j < n; j += kWave) {
                    const int i = S.perm[j];
                    const unsigned key = sk[i];
                    const int b = (int)((key & 0x7fffffffu) >> shift);
                    const unsigned long long b0 = S.base[b];
                    const unsigned r = (unsigned)j - S.pre[b], l1 = S.len1[b];
                    const unsigned long long dst = r < l1 ? S.pos1[b] + r : S.pos2[b] + (r - l1);
                    if (dst < S.cap[b]) {
                        A.rs_keys[b0 + dst] = key;
                        if (WEIGHTED) A.rs_vals[b0 + dst] = sv[i];
                    } else {  // region full: the reference's own atomic
                        const unsigned c = key & 0x7fffffffu;
                        const unsigned rt = c / (unsigned)np2;
                        add_direct(A.rs_inv[rt], (int)(c - rt * (unsigned)np2), WEIGHTED ? sv[i] : 0.0,
                                   1ull + (key >> 31));
                    }
                }
            }
            wave_sync();
            cnt = 0;
        }
    }

    // one (p2, p3) pair; s2a: p1 - p2.  All lanes call it (ok: the lane's pair is real).
    __device__ __forceinline__ void pair(const RElem &e2, const RElem &e3, double s2x, double s2y, double s2z,
                                         double s3x, double s3y, double s3z, bool ok) {
        const double dx = e2.x - e3.x, dy = e2.y - e3.y, dz = e2.z - e3.z;  // p2 - p3
        int T0 = 0, rem0 = 0, T1 = 0, rem1 = 0, n0 = 1;
        bool h0 = false, h1 = false;
        double v0 = 0.0, v1 = 0.0;
        if constexpr (!SURVEY) {
            const double q3 = dz * dz, r3 = rp2(dx, dy);  // device.cu:357-361
            if (ok && q3 < A.pimax2 && r3 < A.r2max && r3 > A.r2min) {
                float f2, f3;
                int k2, p2, k3, p3;
                rp_side_unpack(e2.q, f2, k2, p2);
                rp_side_unpack(e3.q, f3, k3, p3);
                const float f = __builtin_sqrtf((float)r3);
                const int k = bin_any<false>(A, t, r3), p = (int)floorf(__builtin_sqrtf((float)q3));
                h0 = rppi_cell_sides(A, e2.r, f2, k2, p2, e3.r, f3, k3, p3, r3, f, k, p, T0, rem0);
                if (SYM) {
                    const bool tie = e2.r == e3.r || e2.r == r3 || e3.r == r3;
                    if (tie) {
                        h1 = rppi_cell_sides(A, e3.r, f3, k3, p3, e2.r, f2, k2, p2, r3, f, k, p, T1, rem1);
                        v0 = (w1 * e2.w) * e3.w;
                        v1 = (w1 * e3.w) * e2.w;
                    } else {
                        n0 = 2;
                        v0 = (w1 * e2.w) * e3.w + (w1 * e3.w) * e2.w;
                    }
                } else {
                    v0 = (w1 * e2.w) * e3.w;
                }
            }
        } else {
            const double s23 = sqrsep(dx, dy, dz);  // device.cu:465-466
            if (ok && s23 < A.smax2 && s23 > 0.0) {
                // getLOS (device.cu:51-60): pos = p - ori with ori = (ori.x, 0, 0)
                const double lx = (ppx + e2.q) + e3.q, ly = (py + e2.y) + e3.y, lz = (pz + e2.z) + e3.z;
                const double sl = dot3(lx, ly, lz, lx, ly, lz);
                const double lim = A.pimax2 * sl;
                double t1 = dot3(lx, ly, lz, s2x, s2y, s2z);
                t1 = t1 * t1;
                double t2 = dot3(lx, ly, lz, s3x, s3y, s3z);
                t2 = t2 * t2;
                double t3 = dot3(lx, ly, lz, dx, dy, dz);
                t3 = t3 * t3;
                if (!(lim < t1 || lim < t2 || lim < t3)) {  // device.cu:470-475
                    const double q1 = t1 / sl, q2 = t2 / sl, q3 = t3 / sl;  // getSqrRpSqrPi, device.cu:68-71
                    const double r1 = e2.r - q1, r2 = e3.r - q2, r3 = s23 - q3;
                    if (r1 < A.r2max && r2 < A.r2max && r3 < A.r2max && r1 > A.r2min && r2 > A.r2min &&
                        r3 > A.r2min && q1 < A.pimax2 && q2 < A.pimax2) {
                        h0 = rppi_cell(A, t, r1, r2, r3, q1, q2, q3, T0, rem0);
                        v0 = (w1 * e2.w) * e3.w;
                    }
                }
                if (SYM) {
                    // the other order (p2, p3) = (e3, e2), with its own line-of-sight
                    // sum; one key of count 2 when both orders land in the same cell
                    const double mx = (ppx + e3.q) + e2.q, my = (py + e3.y) + e2.y, mz = (pz + e3.z) + e2.z;
                    const double sm = dot3(mx, my, mz, mx, my, mz);
                    const double limm = A.pimax2 * sm;
                    double u1 = dot3(mx, my, mz, s3x, s3y, s3z);
                    u1 = u1 * u1;
                    double u2 = dot3(mx, my, mz, s2x, s2y, s2z);
                    u2 = u2 * u2;
                    double u3 = dot3(mx, my, mz, -dx, -dy, -dz);
                    u3 = u3 * u3;
                    if (!(limm < u1 || limm < u2 || limm < u3)) {
                        const double q1 = u1 / sm, q2 = u2 / sm, q3 = u3 / sm;
                        const double r1 = e3.r - q1, r2 = e2.r - q2, r3 = s23 - q3;
                        if (r1 < A.r2max && r2 < A.r2max && r3 < A.r2max && r1 > A.r2min && r2 > A.r2min &&
                            r3 > A.r2min && q1 < A.pimax2 && q2 < A.pimax2) {
                            h1 = rppi_cell(A, t, r1, r2, r3, q1, q2, q3, T1, rem1);
                            v1 = (w1 * e3.w) * e2.w;
                        }
                    }
                    if (h0 && h1 && T0 == T1 && rem0 == rem1) {
                        n0 = 2;
                        v0 += v1;
                        h1 = false;
                    }
                }
            }
        }
        emit(h0, T0, rem0, n0, v0);
        if (SYM) emit(h1, T1, rem1, 1, v1);
        if (STREAM && cnt > kStage - 2 * kWave) flush();
    }

    template <bool SWAP>
    __device__ __forceinline__ void row(const RElem &ea, bool ok) {
        if (SWAP)  // row = p3, lane = p2
            pair(eb, ea, bsx, bsy, bsz, px - ea.x, py - ea.y, pz - ea.z, ok);
        else
            pair(ea, eb, px - ea.x, py - ea.y, pz - ea.z, bsx, bsy, bsz, ok);
    }

    template <bool SWAP>
    __device__ void rect(int r0, int r1) {
        for (int a = r0; a < r1; ++a) row<SWAP>(la[a - row_first], vb);
    }

    __device__ void circ(int b0, int m) {
        const int smax = m / 2;
        for (int st = 1; st <= smax; ++st) {
            int off = lane + st;
            off -= off >= m ? m : 0;
            const bool ok = vb && (2 * st != m || lane < st);
            row<false>(la[b0 - row_first + (ok ? off : 0)], ok);
        }
    }
};

// Returns the keys the primary emitted (key stream).
template <bool SURVEY, bool SYM, bool WEIGHTED, bool STREAM>
__device__ __forceinline__ unsigned long long rppi_primary(const KArgs &A, const LdsTabs &t, const char *l2,
                                                           const char *lb, char *wlds, int lane,
                                                           unsigned long long *out, double px, double py, double pz,
                                                           double pw, int4 term, int n2u, int n3u) {
    RppiPol<SURVEY, SYM, WEIGHTED, STREAM> P(A, t, l2, lb, wlds, lane, out, px, py, pz, pw, term);
    traverse(P, n2u, n3u);
    if (STREAM && P.cnt > 0) P.flush();
    return P.emitted;
}

// Persistent rppi kernel (countTrianglesRppiBox / countTrianglesRppiSurvey,
// device.cu:317-379, 434-505): the work queue, list build and traversal of
// k_triplets with the rppi filters and pair arithmetic.
// The survey key-stream variant is built for workgroups of at most 8 waves
// and 168 VGPRs (three waves per SIMD): its per-wave key staging already
// holds a CU to about three waves per SIMD, and 128 VGPRs spilled 39.
template <bool SURVEY, bool STREAM>
__global__ __launch_bounds__(SURVEY && STREAM ? 512 : 1024)
    __attribute__((amdgpu_waves_per_eu(SURVEY && STREAM ? 3 : 4))) void k_rppi(KArgs A) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    __shared__ int s_chunk;
    __shared__ int s_next;
    __shared__ unsigned long long s_ckeys;  // dry run: keys of the current work item
    LdsTabs t;
    t.hist = nullptr;
    double *sq = reinterpret_cast<double *>(lds + A.off_sq);
    uint16_t *lut = reinterpret_cast<uint16_t *>(lds + A.off_lut);
    t.sq = sq;
    t.lut = lut;
    t.ctab = nullptr;
    t.lut32 = nullptr;
    t.t3 = nullptr;
    t.rk = nullptr;
    int prev_ci = -1;  // thread 0: the work item whose key count is pending
    const int tid = threadIdx.x, nthr = blockDim.x;
    for (int i = tid; i <= A.nb; i += nthr) sq[i] = A.sq[i];
    for (int i = tid; i <= A.lut_m; i += nthr) lut[i] = A.lut[i];
    if (STREAM) {
        uint16_t *rk = reinterpret_cast<uint16_t *>(lds + A.off_rank);
        t.rk = rk;
        const int nb3 = A.nb * A.nb * A.nb;
        for (int i = tid; i < nb3; i += nthr) rk[i] = A.rs_rank[i];
    }
    constexpr int LM = SURVEY ? kListRppiSurvey : kListRppiBox;
    const int lane = tid & (kWave - 1), wave = tid / kWave, nw = nthr / kWave;
    char *wlds = lds + A.off_wave + wave * A.wave_bytes;
    char *slot = A.scratch + (long long)(blockIdx.x * nw + wave) * A.slot_bytes;
    char *l2 = slot;
    char *l3 = slot + (size_t)A.cap * list_entry_bytes(LM);
    KeyStream ks;
    if (STREAM) ks.L.bind(wlds + A.off_stage + A.rs_lds_off, A.rs_nbm);
    __syncthreads();
    for (;;) {
        if (tid == 0) {
            if (STREAM && A.rs_dry && prev_ci >= 0) A.rs_chunk_keys[prev_ci] = s_ckeys;
            s_chunk = (int)atomicAdd(A.counter, 1u);
            s_next = 0;
            s_ckeys = 0ull;
            prev_ci = s_chunk;
        }
        __syncthreads();
        const int ci = uniform(s_chunk);
        if (ci >= A.nchunks) break;
        const Chunk ch = A.chunks[ci];
        const TermDesc T = A.terms[ch.term];
        const int4 term = STREAM ? A.rs_term[ch.term] : make_int4(0, 0, 0, 0);
        if (STREAM) ks.open(A, lane, ch.term, term);
        unsigned long long *out = A.out + (long long)ch.term * A.nbc;
        for (;;) {
            const int p = uniform(atomicAdd(&s_next, lane == 0 ? 1 : 0)) + ch.begin;
            if (p >= ch.end) break;
            const double px = uniform(T.x1[p]), py = uniform(T.y1[p]), pz = uniform(T.z1[p]),
                         pw = uniform(T.w1[p]);
            bool bad = false;
            const int cx = axis_cell(px, A.cs[0], A.nc[0], !SURVEY, bad);
            const int cy = axis_cell(py, A.cs[1], A.nc[1], !SURVEY, bad);
            const int cz = axis_cell(pz, A.cs[2], A.nc[2], !SURVEY, bad);
            if (bad) {
                if (lane == 0) atomicOr(A.err, kErrDomain);
                continue;
            }
            // p2 and p3 from one catalogue: each unordered pair once (survey: both
            // orders evaluated in it, the line-of-sight sum being order dependent)
            const bool sym = T.sym;
            const int n2 = build_list<LM>(A, T.x2, T.y2, T.z2, T.w2, T.st2, T.prune2, px, py, pz, cx, cy, cz, l2, lane, t.sq,
                                          t.lut);
            const int n3 = T.sym ? n2
                                 : build_list<LM>(A, T.x3, T.y3, T.z3, T.w3, T.st3, T.prune3, px, py, pz, cx, cy, cz, l3, lane,
                                                  t.sq, t.lut);
            const int n2u = uniform(n2), n3u = uniform(n3);
            if (n2u > A.cap || n3u > A.cap) {
                if (lane == 0) atomicOr(A.err, kErrOverflow);
                continue;
            }
            if (n2u == 0 || n3u == 0) continue;
            wave_sync();
            const char *lb = T.sym ? l2 : l3;
            unsigned long long nk;
            if (sym) {
                if (T.weighted)
                    nk = rppi_primary<SURVEY, true, true, STREAM>(A, t, l2, lb, wlds, lane, out, px, py, pz, pw, term,
                                                                  n2u, n3u);
                else
                    nk = rppi_primary<SURVEY, true, false, STREAM>(A, t, l2, lb, wlds, lane, out, px, py, pz, pw,
                                                                   term, n2u, n3u);
            } else {
                if (T.weighted)
                    nk = rppi_primary<SURVEY, false, true, STREAM>(A, t, l2, lb, wlds, lane, out, px, py, pz, pw,
                                                                   term, n2u, n3u);
                else
                    nk = rppi_primary<SURVEY, false, false, STREAM>(A, t, l2, lb, wlds, lane, out, px, py, pz, pw,
                                                                    term, n2u, n3u);
            }
            if (STREAM && A.rs_dry) atomicAdd(&s_ckeys, lane == 0 ? nk : 0ull);  // every lane: no lane-0 branch
            wave_sync();
        }
        __syncthreads();
    }
    if (STREAM) ks.close(A, lane);
}

template __global__ void k_rppi<false, false>(KArgs);
template __global__ void k_rppi<true, false>(KArgs);
template __global__ void k_rppi<false, true>(KArgs);
template __global__ void k_rppi<true, true>(KArgs);

// Binning of the rppi key stream: work item blockIdx.x adds its slice of one
// bucket's keys in an LDS histogram of 2^shift bins (u32 counts or f64 sums),
// then adds the non-zero bins to the term's dense histogram.  A slice holds
// fewer than 2^30 keys (each counting at most 2), so u32 counts cannot wrap.
__global__ __launch_bounds__(1024) void k_rppi_bin(KArgs A) {
    extern __shared__ __attribute__((aligned(16))) char lds[];
    const RBItem it = A.rs_items[blockIdx.x];
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int nbin = 1 << it.shift;
    unsigned *h32 = reinterpret_cast<unsigned *>(lds);
    double *h64 = reinterpret_cast<double *>(lds);
    for (int i = tid; i < (it.weighted ? 2 * nbin : nbin); i += nthr) h32[i] = 0u;
    __syncthreads();
    const unsigned long long b0 = A.rs_base[it.bucket];
    const unsigned long long capb = A.rs_base[it.bucket + 1] - b0;
    const unsigned long long cur = A.rs_cursor[it.bucket];
    const unsigned long long n = cur < capb ? cur : capb;
    const unsigned long long lo = n * (unsigned long long)it.j / (unsigned long long)it.s;
    const unsigned long long hi = n * (unsigned long long)(it.j + 1) / (unsigned long long)it.s;
    const unsigned mask = (unsigned)nbin - 1u;
    if (it.weighted) {
        for (unsigned long long i = lo + tid; i < hi; i += nthr) {
            const unsigned key = A.rs_keys[b0 + i];
            if (key != kRsPad) unsafeAtomicAdd(&h64[key & mask], A.rs_vals[b0 + i]);
        }
    } else {
        for (unsigned long long i = lo + tid; i < hi; i += nthr) {
            const unsigned key = A.rs_keys[b0 + i];
            if (key != kRsPad) atomicAdd(&h32[key & mask], 1u + (key >> 31));
        }
    }
    __syncthreads();
    unsigned long long *out = A.out + (long long)it.term * A.nbc;
    const unsigned np2 = (unsigned)(A.npi * A.npi);
    for (int i = tid; i < nbin; i += nthr) {
        const bool nz = it.weighted ? h64[i] != 0.0 : h32[i] != 0u;
        if (!nz) continue;
        const unsigned c = (unsigned)it.cbase + (unsigned)i;
        const unsigned r = c / np2;
        const long long idx = (long long)A.rs_inv[r] * np2 + (c - r * np2);
        if (it.weighted)
            unsafeAtomicAdd(reinterpret_cast<double *>(&out[idx]), h64[i]);
        else
            atomicAdd(&out[idx], (unsigned long long)h32[i]);
    }
}

template __global__ void k_triplets<false, kHistCtab, false>(KArgs);
template __global__ void k_triplets<false, kHistT3, false>(KArgs);
template __global__ void k_triplets<false, kHistDense, false>(KArgs);
template __global__ void k_triplets<true, kHistCtab, false>(KArgs);
template __global__ void k_triplets<true, kHistT3, false>(KArgs);
template __global__ void k_triplets<true, kHistDense, false>(KArgs);
template __global__ void k_triplets<false, kHistCtab, true>(KArgs);
template __global__ void k_triplets<false, kHistT3, true>(KArgs);
template __global__ void k_triplets<true, kHistCtab, true>(KArgs);
template __global__ void k_triplets<true, kHistT3, true>(KArgs);
template __global__ void k_triplets<false, kHistExact, true>(KArgs);
template __global__ void k_triplets<true, kHistExact, true>(KArgs);
template __global__ void k_triplets<false, kHistF32, true>(KArgs);
template __global__ void k_triplets<true, kHistF32, true>(KArgs);
template __global__ void k_triplets<false, kHistF32C, true>(KArgs);
template __global__ void k_triplets<true, kHistF32C, true>(KArgs);

// ---------------------------------------------------------------------------
// Scratch of nc_count_noalloc: device memory of a fixed size that the library
// holds on every GPU it runs on until the process ends, so that entry point
// allocates nothing.  In host code `noalloc_scratch` names a host-side
// stand-in, not device memory: the device address is looked up per device
// with hipGetSymbolAddress (noalloc_base) and only that address is used.
static __device__ __align__(256) char noalloc_scratch[kNoAllocBytes];

// Device address of noalloc_scratch on the current device (the host-side
// name is a stand-in, registered with this translation unit's code object).
hipError_t noalloc_symbol_address(void **p) { return hipGetSymbolAddress(p, HIP_SYMBOL(noalloc_scratch)); }

