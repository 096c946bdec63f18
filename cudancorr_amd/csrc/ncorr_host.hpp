// Host-side state of libncorr_hip.so shared by its host translation units:
// device buffers, the geometry of a count, the catalogue and context handles,
// error helpers, and the functions of the planner and launcher
// (ncorr_geometry.hip, ncorr_count.hip) that the C-ABI (ncorr_api.hip) calls.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ncorr_hip.h"
#include "ncorr_kernels.hpp"
#include "ncorr_launch.hpp"

namespace nch {

using namespace nck;

struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

constexpr int kNoExact = -100000;  // no exact layout (see convert_out)

// Grow b to at least `need` bytes with hipMalloc (contents not kept).
hipError_t dgrow(DBuf &b, size_t need);

void dfree(DBuf &b);

// Where a count takes its device buffers: grown with hipMalloc and kept
// (the default), or carved from a fixed arena that is never freed
// (nc_count_noalloc: the static per-device scratch array).  An arena request
// past its end fails with `exhausted` set and allocates nothing.
// Test knob (NC_FAIL_DEVICE_ALLOC=n at nc_create): the context's n-th device
// allocation from then on (counting from 0) fails as out of memory, so the
// error paths can be driven.  *fail < 0: off.
hipError_t dgrow_checked(long long *fail, DBuf &b, size_t need);

struct Mem {
    char *base = nullptr;
    size_t size = 0, used = 0;
    bool exhausted = false;
    long long *fail = nullptr;  // the context's allocation-failure countdown
    bool arena() const { return base != nullptr; }
    size_t left() const { return size - std::min(size, (used + 255) & ~(size_t)255); }
    hipError_t grow(DBuf &b, size_t need) {
        if (!base) return dgrow_checked(fail, b, need);
        if (need == 0) need = 16;
        const size_t at = (used + 255) & ~(size_t)255;
        if (at > size || need > size - at) {
            exhausted = true;
            return hipErrorOutOfMemory;
        }
        b.p = base + at;
        b.bytes = need;
        used = at + need;
        return hipSuccess;
    }
};

// Device buffers of one count (per context, reused across counts).
struct CountBufs {
    DBuf tables, terms, chunks, out, counter, errw, scratch, bound, wflags;
    DBuf cflags;  // per catalogue of a count: some cell index is not an exact floor
};

struct Geometry {
    double box[3];
    double cs[3];
    int nc[3];
    long long ncells;
    int survey;
    int nb, nbc;
    int rppi = 0, npi = 0;      // rppi binning: npi unit-width pi bins
    double ori_x = 0.0;         // survey origin (x only reaches the device in the reference)
    long long out_words = 0;    // output words per term: nbc compact bins, or nb^3 * npi^2 (rppi)
    std::vector<double> edges, sq;
    std::vector<uint16_t> lut;
    double lut_scale;
    int lut_m;
    std::vector<int> ctab;
    bool lut_single_step = false;  // one step over the edges always suffices after the lookup
    // mixed path: guard bands and the FP32 lookup (valid when mixed_ok)
    bool mixed_ok = false;
    std::vector<uint2> lut32;
    double sigma32 = 0.0;
    float lut32_lo = 0.f;
    int lut32_shift = 0, lut32_c0 = 0;
    float g32 = 0.f, tri_t = 0.f;
    std::vector<uint16_t> t3;  // (ka*nb+kb)*(nb+2)+k+1 -> byte offset of the compact slot
};

}  // namespace nch

using nch::CountBufs;
using nch::DBuf;
using nch::kNoExact;

// The autotuner times contiguous runs of kSampleRun primaries (whole work
// items from the run's first primary on), so that the workgroups running at
// one time share neighbourhoods (in L2) as in a full count, and so that every
// shape, whatever its work-item size, times the same primaries.
constexpr long long kSampleRun = 1024;
// Shapes whose sample time is within this fraction of the fastest count as
// tied; among them the autotuner takes the largest list chunk.
constexpr float kTuneTie = 0.005f;

struct nc_catalog {
    nc_context *ctx = nullptr;  // null once the context is destroyed
    int device = 0;
    uint64_t uid = 0;  // unique per handle (graph plans key on it)
    int64_t n = 0;
    const double *d_raw = nullptr;
    bool owns = false;
    bool unit_w = true;
    bool exact = false;  // last build: every cell index the exact floor of coordinate / cell size
    double wmax = 0.0;     // largest finite |w|
    bool wfinite = true;   // every weight finite
    bool wf32 = true;      // every weight an FP32 weight (fp32_weight)
    DBuf raw, key, rank, perm, cnt, start, scan, x, y, z, w;
};

// Captured launch sequence of nc_count_batched: replayed while the same
// terms, params, path and tuning are passed again.
struct BatchPlan {
    bool valid = false;
    std::vector<uint64_t> uids;
    std::vector<int> wkey;  // per term: unit-weight flags of c1, c2, c3 (bits 0-2)
    std::vector<double> pkey;
    std::vector<int> tkey;
    hipGraphExec_t build_exec = nullptr, count_exec = nullptr;
    unsigned long long *h_out = nullptr;
    unsigned *h_err = nullptr;
    std::vector<int> weighted;
    int nb = 0, nkernel = 0, hist_layout = 0, rppi = 0, lds_bytes = 0, path = 0, tri_terms = 0, swap_terms = 0,
        partner_terms = 0;
    int xk = kNoExact;  // exact layout: fixed-point scale exponent
    long long nbc = 0;  // output words per term (slots; twice as many u64 words in the exact layout)
};

struct nc_context {
    int device = 0;
    BatchPlan plan;
    uint64_t next_uid = 1;
    std::vector<nc_catalog *> cats;  // live catalogues, detached by nc_destroy
    hipStream_t stream = nullptr;
    std::string err;
    int ncu = 0;
    size_t lds_max = 0;  // LDS budget of a workgroup with two per CU
    size_t lds_hw = 0;   // the per-workgroup maximum
    CountBufs buf;  // device buffers of a count, kept at their largest size
    char *noalloc = nullptr;  // this device's noalloc_scratch (device address), set by nc_create
    std::string noalloc_err;  // why it is missing, if it is
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    nc_timing timing{};
    int waves_per_group = 8;
    int chunk = 64;
    int wgs_per_cu = 0;
    int chunk_elems = 128;
    // dense counts are merged once this many pairs were added (< 2^31 keeps
    // u32 counts from wrapping; NC_DENSE_FLUSH_PAIRS lowers it for tests)
    unsigned long long dense_flush = 1ull << 31;
    int sample_stride = 1;    // autotuner: queue only the work items of every n-th run of kSampleRun primaries
    int slice = 0, nslices = 1;  // nc_set_primary_slice: primaries of slice `slice` of `nslices` only
    bool autotuning = false;  // in nc_autotune's timed runs
    // the tuning of the unit-weight launch of a split batch (count_terms), when
    // nc_autotune timed the two launches apart; nc_set_tuning clears it
    struct {
        bool on = false;
        int waves_per_group = 0, chunk = 0, wgs_per_cu = 0, list_chunk = 0;
    } tune_unit;
    // longest list of each launch's terms (keyed by their catalogues), found
    // once per nc_autotune call: its runs share the inputs
    std::map<std::vector<uint64_t>, int> tune_list_bound;
    std::string dev_name;
    // rppi key stream: buffers, and the per-bucket capacities last sized for
    // a (catalogues, params) key; NC_RPPI_STREAM=0 turns the stream off
    int rppi_stream = -1;  // rppi key stream: 1 on, 0 off (direct atomics), -1 box on, survey off
    bool batch_split = true;  // unit-weight terms of a mixed-weight batch as their own launch
    bool fuse_terms = true;   // (c1, c3, c3) counted beside (c1, c2, c3) from the same lists (TermDesc.partner)
    bool tri = true;          // NC_TRI=0: terms over one catalogue never use the six-fold symmetry
    bool tri_last = true;     // NC_TRI_LAST=0: their last-cell triples stay two-fold
    int accum = NC_ACCUM_FP64;  // nc_set_accumulation
    long long fail_alloc = -1;  // NC_FAIL_DEVICE_ALLOC (test knob; see dgrow_checked)
    double rs_scale = 1.0;       // NC_RPPI_REGION_SCALE: region capacities x this (tests of the overflow path)
    double rs_round_keys = 0.0;  // NC_RPPI_ROUND_KEYS: keys per round (tests of several rounds)
    DBuf rs_tab, rs_cursor, rs_base, rs_items, rs_sample, rs_chunk_keys, rs_keys, rs_vals;
    std::vector<double> rs_key;
    std::vector<unsigned long long> rs_caps;
    std::vector<double> rs_cest;  // estimated keys per work item
};

namespace nch {

inline int fail(nc_context *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define NC_HIP(ctx, call)                                                                               \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail(ctx, e_ == hipErrorOutOfMemory ? NC_ERR_NOMEM : NC_ERR_HIP,                     \
                        std::string(#call) + ": " + hipGetErrorString(e_));                             \
    } while (0)

// A buffer request: on the arena, running past its end is NC_ERR_LIMIT.
#define NC_MEM(ctx, m, b, bytes)                                                                       \
    do {                                                                                                \
        const hipError_t e2_ = (m).grow((b), (bytes));                                                  \
        if (e2_ != hipSuccess && (m).exhausted)                                                         \
            return fail(ctx, NC_ERR_LIMIT,                                                              \
                        "the count needs more device memory than the no-alloc scratch holds (" +         \
                            std::to_string((m).size >> 20) + " MiB)");                                  \
        NC_HIP(ctx, e2_);                                                                               \
    } while (0)


inline int64_t choose_compact(int nb) { return (int64_t)nb * (nb + 1) * (nb + 2) / 6; }

struct Term {
    nc_catalog *c1, *c2, *c3;
};

// ncorr_geometry.hip: the reference's constructor geometry and the kernels' lookup tables.
int make_geometry(nc_context *ctx, const nc_params *P, Geometry &G);
void make_mixed_tables(Geometry &G);

// ncorr_count.hip: cell lists, planning, launches and graph replay of a count.
TermDesc term_desc(const Term &t);
void plan_drop(nc_context *ctx);
int run_count(nc_context *ctx, const nc_params *P, const std::vector<Term> &terms, int path, double *hist,
              uint64_t *counts, bool use_graph, Mem &m);
int refresh_weights(nc_context *ctx, Mem &m, const std::vector<Term> &terms);
// Whether count_terms runs the unit-weight terms of t (iu) and the weighted
// ones (iw) as two launches.
bool split_batch(const nc_context *ctx, const nc_params *params, const std::vector<Term> &t, std::vector<int> &iu,
                 std::vector<int> &iw);
int count_terms(nc_context *ctx, const nc_params *params, const std::vector<Term> &t, int path, double *hist,
                uint64_t *counts, bool use_graph);
hipError_t noalloc_base(int device, char **out);

}  // namespace nch
