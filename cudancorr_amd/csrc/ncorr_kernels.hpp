// Device-side definitions shared by the HIP kernels of libncorr_hip.so.
//
// Data layout in HBM (per catalogue, per count geometry):
//   raw    : n x double4 {x,y,z,w}, as handed over by the caller
//   key    : n x u32 slab key (z + nz*(y + ny*x)) * kSlabs + slab, the cell of
//            cudaNcorr.cu:50-58 and the point's z slab within it
//   rank   : n x u32 arrival slot of the point inside its slab (atomic)
//   perm   : n x u32 point indices in slab order, then placed stably (by index)
//   start  : (ncells * kSlabs rounded up to kScanTile) + 1 x u32, exclusive scan
//            of the per-slab counts; cell c is [start[c*kSlabs], start[(c+1)*kSlabs])
//   x,y,z,w: n x double each, points sorted by slab key (SoA, coalesced)
// Per persistent wave, a scratch slot holds the compacted neighbour lists of
// the primary it is working on (Elem + weight per entry).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nck {

constexpr int kWave = 64;
constexpr unsigned kErrDomain = 1u;
constexpr unsigned kErrOverflow = 2u;
constexpr unsigned kErrMode = 4u;  // a term the kernel variant was not planned for (weighting)
// a catalogue planned as exactly gridded (every cell index the exact floor of
// coordinate / cell size) no longer is: a replayed plan must be rebuilt
constexpr unsigned kErrReplan = 8u;
constexpr int kMaxWavesPerGroup = 8;

// Every cell is cut into kSlabs equal slabs along z, and its points are
// sorted by slab, so that a neighbour list can skip the slabs of a neighbour
// cell that lie beyond rmax of the primary.
constexpr int kSlabs = 8;
// Entries per workgroup of the multi-block scan of the slab counts.
constexpr int kScanTile = 4096;

// One compacted neighbour of a primary: the image position used by the
// reference (point + periodic shift, FP64), the exact bin of its squared
// separation from the primary and sqrtf(float(r^2)) of the FP32 triangle test.
struct __attribute__((aligned(16))) Elem {
    double x, y, z;
    int k;
    float f;
};

constexpr int kTriNone = 0, kTriSix = 1, kTriSwap = 2;  // TermDesc::tri

struct TermDesc {
    const double *x1, *y1, *z1, *w1;
    const double *x2, *y2, *z2, *w2;
    const uint32_t *st2;
    const double *x3, *y3, *z3, *w3;
    const uint32_t *st3;
    int n1;
    int sym;       // c2 == c3: count unordered (p2,p3) pairs once, doubled
    int weighted;  // accumulate (w1*w2)*w3 in FP64, else integer counts
    // kTriSix: c1 == c2 == c3, exactly gridded: every unordered triple of
    // points outside the last cells is evaluated once, from its lowest-index
    // point, and counted six times (see k_triplets).
    // kTriSwap: survey, c1 == c2 != c3, exactly gridded: the triplets
    // (p1, p2, p3) and (p2, p1, p3) are evaluated once, from the lower index
    // of p1 and p2, and counted twice.
    int tri;
    // prune2 / prune3: the catalogue is exactly gridded (threeD), so its lists
    // may skip the slabs of a neighbour cell beyond rmax of the primary
    int prune2, prune3;
};

struct Chunk {
    int term;
    int begin;
    int end;
    int pad;
};

// rppi key stream: one work item of the binning kernel, a slice j of s of
// one bucket's keys.
struct RBItem {
    int bucket;    // global bucket id
    int term;
    int shift;     // log2 of the bucket's bins
    int weighted;  // f64 sums (else u32 counts)
    int cbase;     // first compact bin of the bucket within its term
    int j, s;
    int pad;
};

// Work-queue segments: one per XCD (the chip's eight L2 domains).
constexpr int kMaxSeg = 8;

struct KArgs {
    const TermDesc *terms;
    const Chunk *chunks;
    int nchunks;
    // chunks [seg[s], seg[s+1]) form segment s, pulled through counter[s];
    // workgroup b starts on segment b % nseg (its XCD) and moves on to the
    // next segment once its own is drained (seg in device memory: a runtime
    // index into an array of the kernel arguments would copy them to scratch)
    int nseg;
    const int *seg;
    unsigned *counter;
    unsigned *err;
    unsigned long long *out;  // nterms x nbc, u64 counts or f64 sums (bit pattern)
    char *scratch;
    long long slot_bytes;
    int cap;
    int nmax;         // longest list the launch was planned for (<= cap); longer ones flag an overflow
    int chunk_elems;  // a-entries staged in LDS per wave
    double box[3];
    double cs[3];
    int nc[3];
    int survey;
    double r2min, r2max;
    double prune_r2;  // slab pruning: squared radius, rmax with a margin above any rounding of the bounds
    int nb, nbc;
    int nslots;                   // u64 words of the LDS histogram (nbc compact slots, or nb^3 u32 dense counts)
    int out_stride;               // u64 words of out per term: nbc, or 2*nbc (exact layout: low words, high words)
    double xscale;                // exact layout: 2^K, the fixed-point scale of the weight products
    // rppi binning (cudaNcorr.cu:159-216): npi unit-width pi bins; nbc is then
    // the dense nb^3 * npi^2 output length per term
    int npi;
    double pimax2;  // npi^2
    double smax2;   // rpMax^2 + piMax^2 (survey list filter)
    double ori_x;   // survey origin, x only (cudaNcorr.cu:254)
    unsigned long long flush_pairs;  // dense counts: merge once this many pairs were added (<= 2^31)
    double lut_scale;
    int lut_m;
    // mixed path
    // Quantities in units scaled by sigma^2, a power of 4 that brings rmax^2
    // into [1, 4): list coordinates are stored as sigma*(q - p1), so r23^2
    // arrives pre-scaled.  Lookup cell of a scaled r23^2 d: the float bits of
    // med3(d, lut32_lo, 4.0) shifted right by lut32_shift, minus lut32_c0
    // (2^(23-lut32_shift) log-spaced cells per octave).
    double sigma32;  // power of 2
    float lut32_lo;
    int lut32_shift;
    int lut32_c0;
    int lut32_n;  // entries
    float g32;     // guard band: |fp32 r23^2 - fp64 r23^2| bound (incl. edge rounding), scaled
    float tri_t;   // collinearity threshold: s_a*s_b - dot^2 <= tri_t => recheck, scaled
    const double *sq;             // nb+1 squared edges
    const uint16_t *lut;          // lut_m+1 lower-bound bins (FP64 path)
    const int *ctab;              // nb*nb compact base index
    const uint2 *lut32;           // lut32_n entries (mixed path): {bin below the edge, scaled edge as float bits}
    const uint16_t *t3;           // nb*nb*(nb+2): (ka*nb+kb)*(nb+2)+k+1 -> byte offset of the compact slot
    // dynamic LDS layout (bytes)
    int off_sq, off_lut, off_ctab, off_lut32, off_t3, off_wave, wave_bytes, lds_bytes;
    // rppi key stream (k_rppi with STREAM): binned triplets leave as compact
    // keys c = rank(T)*npi^2 + rem, grouped by bucket c >> shift into
    // per-bucket regions, then k_rppi_bin adds each bucket in LDS.
    int rs_dry;                            // count keys per bucket only (region sizing)
    int rs_blk;                            // keys a wave reserves in a bucket region at a time
    int rs_lds_off;                        // per-wave run state, from the start of the staging
    int rs_nbm;                            // per-wave bucket slots (>= buckets of any term)
    int rs_nitems;                         // k_rppi_bin work items
    int off_rank, off_stage;               // LDS: rank table, per-wave staging (inside the wave slice)
    const int4 *rs_term;                   // per term: first bucket, buckets, shift, 0
    const uint16_t *rs_rank;               // nb^3: rank of a dense rp triple among the reachable ones
    const int *rs_inv;                     // rank -> dense rp triple
    unsigned long long *rs_cursor;         // slots reserved per bucket (may pass the capacity); dry: keys
    const unsigned long long *rs_base;     // region of bucket b: [base[b], base[b+1]) in rs_keys / rs_vals
    unsigned *rs_keys;                     // c | (count 2 ? 1 << 31 : 0); kRsPad in unused reserved slots
    double *rs_vals;                       // weighted terms: the summed weight products
    const RBItem *rs_items;
    unsigned long long *rs_chunk_keys;     // dry: keys of each work item (region sizing per round)
};

}  // namespace nck
