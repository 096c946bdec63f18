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
    // weighted, with unit weights on c2 and c3 (FP64 accumulation): every
    // ordered triplet adds w1, so the pair loops stage no weights
    int wprim;
    int wunit2, wunit3;  // c2 / c3 has unit weights (its weights are not read)
    // partner: a term (c1, c3, c3) counted with this one (c1, c2, c3 != c2)
    // from the same primaries and the same list of c3, into the workgroup's
    // second histogram (-1: none); a partner term has no work items of its own
    int partner;
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
    int off_hist2;  // byte offset of the partner terms' histogram (0: none)
    // one-catalogue terms count their last-cell triples six-fold too (box,
    // at least three cells per axis); last-cell primaries skip those triples
    int tri_last6;
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


// ---------------------------------------------------------------------------
// Definitions shared by the kernels and the host-side planner.

struct GridGeo {
    double cs[3];
    int nc[3];
    int survey;
};

constexpr int kQueueCap = 320;  // deferred exact rechecks per wave (mixed path), >= 4*64 + 64

// Neighbour-list flavours: threeD plain (Elem + weight), threeD mixed (MList),
// rppi box and rppi survey (RElem).
constexpr int kListPlain = 0, kListMixed = 1, kListRppiBox = 2, kListRppiSurvey = 3;

// One rppi list entry: image position, to the primary rp^2 (box) or the
// squared 3D separation (survey) r, then q, and the weight.  Box: q holds the
// side's FP32 root sqrtf(rp^2), rp bin and pi bin floor(sqrtf(pi^2)) packed
// (rp_side_pack), so a pair only evaluates its third side; survey: x - ori.x.
struct RElem {
    double x, y, z, r, q, w;
};

// Bytes of one list entry in a wave's scratch slot.
__host__ __device__ constexpr size_t list_entry_bytes(int lm) {
    return lm == kListMixed ? sizeof(float4) + sizeof(int)
                            : lm == kListPlain ? sizeof(Elem) + sizeof(double) : sizeof(RElem);
}

// LDS histogram layouts of the triplet kernel (HM):
//   kHistCtab   compact C(nb+2,3) sorted slots (u64 counts / f64 sums), slot
//               index through ctab; the plain path's FP64 bin lookup may take
//               several steps;
//   kHistT3     compact sorted slots, slot byte offset from the t3 table;
//   kHistDense  dense nb^3 u32 counts indexed (ka, kb, k) unsorted, folded
//               into the sorted slots when flushed (unit-weight terms only).
// kHistT3 and kHistDense also take the single-step FP64 bin lookup.
constexpr int kHistCtab = 0, kHistT3 = 1, kHistDense = 2;
constexpr int kHistGlobal = 3;  // rppi: the dense histogram in global memory, no LDS copy
constexpr int kHistStream = 4;  // rppi: key stream binned per bucket in LDS (reported layout)
// kHistExact: the sorted slots of kHistT3 as 128-bit fixed-point integers (a
// low-word array, then a high-word array): weight products are truncated to
// multiples of 2^-K and summed exactly, so the sums do not depend on the order
// of the additions (nc_set_accumulation NC_ACCUM_EXACT).
constexpr int kHistExact = 5;
// kHistF32: the sorted slots of kHistT3 as FP32 running sums with FP64
// compensation (nc_set_accumulation NC_ACCUM_FP32): an array of nbc FP64
// compensations at the slot offsets of kHistT3, then (at byte xhi = 8 nbc) an
// array of nbc FP32 sums.  Each accepted pair's weight product is formed in
// FP32 as a float-float pair (hi, lo) and hi is added to the FP32 sum with a
// returning LDS atomic; the rounding error of that addition (TwoSum of the old
// sum and hi, exact) and lo go to the compensation.  Sum + compensation is the
// exact sum of the float-float products up to the FP64 rounding of the
// compensation's own additions; the two are folded into the FP64 totals at
// every merge.
constexpr int kHistF32 = 6;
// kHistF32C: the layout of kHistF32 with the slot offsets of kHistCtab (the
// reference's index arithmetic), for bin counts whose t3 table and FP32 sums
// do not fit in LDS together.
constexpr int kHistF32C = 7;
__host__ __device__ constexpr bool is_f32(int hm) { return hm == kHistF32 || hm == kHistF32C; }
// layouts addressed through the t3 slot table, or through ctab
// layouts whose weighted terms may take the unit-secondary (wprim) pair loops
__host__ __device__ constexpr bool primw_ok(int hm) { return hm == kHistT3 || hm == kHistCtab; }
__host__ __device__ constexpr bool t3_like(int hm) { return hm == kHistT3 || hm == kHistExact || hm == kHistF32; }
__host__ __device__ constexpr bool ctab_like(int hm) { return hm == kHistCtab || hm == kHistF32C; }
// layouts whose merges are also forced by a count of pairs added (flush_pairs)
__host__ __device__ constexpr bool counts_pairs(int hm) { return hm == kHistDense || is_f32(hm); }
// The mixed path's FP32 lookup (lut32, copied into LDS) holds in its entries'
// .x the bin count in the layout's units: bytes of a dense count (4), of a t3
// entry (2), or bins (ctab); the shift of .x.
__host__ __device__ constexpr int lut32_units(int hm) { return hm == kHistDense ? 2 : t3_like(hm) ? 1 : 0; }

// Per-wave staging of the rppi key stream: up to kStage keys (and values);
// a row adds at most 2*64, so the wave flushes once it holds more than
// kStage - 128.  kStreamBuckets bounds the buckets of one term.  Per bucket
// of its current term a wave holds a reserved run [cur, end) of its region
// (taken rs_blk slots at a time, so the bucket cursors see one atomic per
// run, not one per flush); slots it leaves unused get kRsPad.
constexpr int kStage = 512;
constexpr int kStreamBuckets = 128;
constexpr unsigned kRsPad = 0xffffffffu;

__host__ __device__ constexpr size_t stream_lds_offset(bool weighted) {
    return ((size_t)kStage * 4 + (weighted ? (size_t)kStage * 8 : 0) + (size_t)kStage * 2 + 15) & ~(size_t)15;
}
__host__ __device__ constexpr size_t stream_wave_bytes(bool weighted, int nbm) {
    return stream_lds_offset(weighted) + (size_t)kStage * 2 + (size_t)nbm * 60;
}

// An FP32 weight for the FP32 accumulation (NC_ACCUM_FP32): a value FP32
// holds exactly, 0 or of magnitude in [2^-40, 2^40], so that products of
// three such weights (and six times them) stay normal FP32 numbers.
__host__ __device__ inline bool fp32_weight(double w) {
    const double a = w < 0.0 ? -w : w;
    return (double)(float)w == w && (w == 0.0 || (a >= 0x1p-40 && a <= 0x1p40));
}

constexpr size_t kNoAllocBytes = (size_t)256 << 20;

}  // namespace nck
