/*
 * ncorr_hip.h -- C-ABI of the MI355X three-point triangle-count library
 * (libncorr_hip.so).  Plain C types only; no HIP or C++ types cross it.
 *
 * The reference has no C-ABI: its boundary is the C++ class
 * `ncorr` (include/cudaNcorr.hpp:10-68), whose calculateD1D2D3
 * (source/cudaNcorr.cu:220-430) uploads three double4 catalogues, runs one of
 * countTrianglesThreeDBox / countTrianglesThreeDSurvey (source/device.cu:259-315,
 * 381-431) and copies back a dense nb^3 double histogram.  This library is the
 * layer under the drop-in `ncorr` class of this project (include/cudaNcorr.hpp):
 * each entry point below names the reference code it replaces.
 *
 * Conventions
 *   - Every function returns NC_OK (0) or an NC_ERR_* code; nc_last_error()
 *     gives a message for the last failure on a context.
 *   - Host buffers are owned by the caller and only read (inputs) or written
 *     (outputs) during the call.  Device buffers are owned by the context or
 *     by a catalogue handle and are reused across calls.
 *   - One context per host thread; calls on one context are serialised on its
 *     own HIP stream.  There is no process-global device state, so several
 *     contexts may run side by side (unlike the reference's __constant__
 *     parameter block, device.cu:9-18).
 *   - Points are double4 {x, y, z, w} (w = weight), 32 bytes each, as the
 *     reference's std::vector<double4>.
 *   - Histograms are dense nb*nb*nb arrays indexed
 *     [i_smallest*nb*nb + i_middle*nb + i_largest] (device.cu:198), filled only
 *     on i_smallest <= i_middle <= i_largest.
 */
#ifndef NCORR_HIP_H
#define NCORR_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum nc_status {
    NC_OK = 0,
    NC_ERR_ARG = 1,     /* bad argument (null pointer, bad size, bad box) */
    NC_ERR_BINS = 2,    /* edges not >= 0 and strictly increasing, or fewer than 2 */
    NC_ERR_DOMAIN = 3,  /* a point's cell key falls outside the cell grid */
    NC_ERR_NOMEM = 4,   /* device or host allocation failed */
    NC_ERR_HIP = 5,     /* a HIP runtime call failed */
    NC_ERR_LIMIT = 6    /* a size exceeds what the kernels support */
};

/* Arithmetic of the triplet kernel. */
enum nc_path {
    NC_PATH_PLAIN = 0, /* every separation, bin and triangle test in FP64 as the reference */
    NC_PATH_MIXED = 1  /* FP32 separations on re-centred coordinates, FP64 recheck in a guard band */
};

typedef struct nc_context nc_context;
typedef struct nc_catalog nc_catalog;

/* Binning of a count (the reference's ncorr::binType, cudaNcorr.hpp:12). */
enum nc_bin_type {
    NC_BIN_THREED = 0, /* three sides binned in r */
    NC_BIN_RPPI = 1    /* three sides binned in projected rp, two in line-of-sight pi */
};

/* Geometry and binning of one count: the arguments of the reference's
 * constructors ncorr(box, 0, bins, isSurvey, ori) (threeD, cudaNcorr.hpp:14,
 * cudaNcorr.cu:118-157) and ncorr(box, 1, bins, numMinorBins, isSurvey, ori)
 * (rppi, cudaNcorr.hpp:15, cudaNcorr.cu:159-216).  A zeroed tail (bin_type 0)
 * is the threeD count. */
typedef struct {
    double box[3];        /* periodic box (box mode) or padded survey bounding box */
    int survey;           /* 0: periodic box, 1: survey (open boundaries) */
    int nedges;           /* number of bin edges = nb + 1 */
    const double *edges;  /* host pointer, bin edges in r (rppi: in rp), not squared */
    int bin_type;         /* nc_bin_type */
    int num_pi_bins;      /* rppi: unit-width pi bins 0..num_pi_bins (numMinorBins) */
    double ori[3];        /* survey origin of the rppi line of sight; as in the
                             reference only ori[0] is used (cudaNcorr.cu:254) */
} nc_params;

/* Device time of the last count call on a context, from HIP events on the
 * context's stream. */
typedef struct {
    float ms_build;      /* cell-list builds (key, scan, scatter) */
    float ms_count;      /* triplet kernel(s) */
    float ms_total;      /* first build kernel to end of the count */
    int64_t n_kernel;    /* triplet kernel launches in that span */
    int32_t hist_layout; /* workgroup histogram of the triplet kernel: 0 sorted slots via an index table,
                            1 sorted slots via a per-bin-triple slot table, 2 dense nb^3 32-bit counts
                            (unit weights only; merged into the sorted slots before they can wrap),
                            3 rppi: atomics on the dense global histogram, 4 rppi: binned triplets
                            streamed as keys into per-bucket regions and added up in LDS */
    int32_t lds_bytes;   /* LDS of one workgroup of the triplet kernel (above half the CU's LDS: one
                            workgroup per CU) */
    int32_t path;        /* nc_path that ran: NC_PATH_MIXED was asked for but NC_PATH_PLAIN ran when the
                            bins are too narrow for its guard bands; rppi counts always run NC_PATH_PLAIN */
    int32_t tri_terms;   /* terms over one catalogue (c1 == c2 == c3) whose unordered triples away from the
                            last cells were each evaluated once and counted six times (every cell index the
                            exact floor of coordinate / cell size; threeD; NC_TRI=0 turns it off) */
    int32_t swap_terms;  /* survey terms whose first two catalogues are one (c1 == c2 != c3, DDR): the
                            triplets (p1, p2, p3) and (p2, p1, p3) were evaluated once, from the lower
                            index, and counted twice (the same conditions as tri_terms) */
    int32_t partner_terms; /* terms (c1, c3, c3) counted beside a term (c1, c2, c3), c2 != c3, of the same
                            call from its primaries and its lists of c3 (DRR beside DDR), into a second
                            workgroup histogram; sorted-slot layouts, NC_FUSE_TERMS=0 turns it off */
} nc_timing;

/* Create a context on HIP device `device` (replaces the implicit CUDA context
 * and the __constant__ parameter block of the reference, device.cu:9-18).
 * Environment read here (test and tuning knobs; unset means the default):
 *   NC_RPPI_STREAM=0|1        rppi through direct global atomics (0) or the key stream (1);
 *                             unset: the key stream for boxes, direct atomics for surveys
 *   NC_RPPI_ROUND_KEYS=<n>    key-stream keys per round (default: 60% of free device memory)
 *   NC_RPPI_REGION_SCALE=<f>  key-stream regions at f (0 < f <= 1) of the sampled need
 *   NC_DENSE_FLUSH_PAIRS=<n>  dense counts merged once n pairs were added (default and max 2^31)
 *   NC_BATCH_SPLIT=0          nc_count_batched keeps mixed-weight batches in one launch
 *   NC_TRI=0                  terms over one catalogue count each ordered triple's (p2, p3) pair
 *                             once per primary (no six-fold symmetry)
 *   NC_FAIL_DEVICE_ALLOC=<n>  test knob: the context's n-th device allocation from creation on
 *                             (counting from 0) fails as out of memory (NC_ERR_NOMEM) */
int nc_create(nc_context **ctx, int device);
/* Catalogues of a destroyed context may still be passed to nc_catalog_destroy
 * (in any order); nothing else. */
int nc_destroy(nc_context *ctx);
const char *nc_last_error(const nc_context *ctx);

/* Upload a host catalogue of n double4 points into a device-resident handle.
 * Replaces the per-cell cudaMalloc/cudaMemcpy of cudaNcorr.cu:279-313: the
 * points are copied once, as one buffer, and gridded on the device. */
int nc_catalog_create(nc_context *ctx, const double *xyzw, int64_t n, nc_catalog **cat);
/* Wrap a device buffer of n double4 points (not copied; must outlive the handle). */
int nc_catalog_wrap_device(nc_context *ctx, const double *d_xyzw, int64_t n, nc_catalog **cat);
int nc_catalog_destroy(nc_catalog *cat);
int64_t nc_catalog_size(const nc_catalog *cat);

/* Dense histogram length nb^3 and compact (sorted-bin) length C(nb+2,3) for a
 * number of edges; 0 if nedges < 2. */
int64_t nc_hist_size(int nedges);
int64_t nc_compact_size(int nedges);
/* Dense histogram length per term for params: nb^3 (threeD) or nb^3 * npi^2
 * (rppi, cudaNcorr.cu:189); 0 if invalid. */
int64_t nc_hist_len(const nc_params *params);

/* The count (replaces ncorr::calculateD1D2D3, cudaNcorr.cu:220-430, and the
 * kernels device.cu:259-315 / 381-431 (threeD), 317-379 / 434-505 (rppi)).
 * Ordered triplets (p1 in c1, p2 in c2, p3 in c3); each binned triplet adds
 * (w1*w2)*w3.
 *   hist   [nc_hist_len] double, required: the reference's triCountsAll_
 *          (threeD [i_smallest*nb^2 + i_middle*nb + i_largest]; rppi
 *          [((i1*nb + i2)*nb + i3)*npi^2 + pi1*npi + pi2], device.cu:250-252).
 *   counts [nc_hist_len] uint64, optional (NULL): exact triplet counts (unit
 *          weights); filled only when every weight of c1, c2, c3 is 1.
 * When every weight is 1 the kernel counts in integers and hist holds those
 * counts exactly (below 2^53), as the reference's double sums of 1.0.  rppi
 * runs the FP64 path whatever `path` asks for. */
int nc_count(nc_context *ctx, const nc_params *params, nc_catalog *c1, nc_catalog *c2, nc_catalog *c3,
             int path, double *hist, uint64_t *counts);

/* Host-buffer convenience form of nc_count: uploads three double4 arrays,
 * counts, and frees them (the reference's calculateD1D2D3 signature). */
int nc_count_host(nc_context *ctx, const nc_params *params, const double *p1, int64_t n1, const double *p2,
                  int64_t n2, const double *p3, int64_t n3, int path, double *hist, uint64_t *counts);

/* Allocation-free form of nc_count.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The caller
 * passes three device buffers of n double4 points (d_p1 == d_p2 with n1 == n2
 * is one catalogue, as in nc_count_host) and host outputs as nc_count; threeD
 * bins only.  Every device buffer of the count (cell lists, tables, work
 * queue, histogram, neighbour lists) is carved from a fixed scratch array of
 * nc_noalloc_capacity() bytes (256 MiB) that the library holds on each GPU it
 * runs on until the process ends; the call makes no hipMalloc/hipFree and no
 * other allocating device call.  A count that needs more than the scratch
 * fails with NC_ERR_LIMIT (about 100 B per point plus 64 B per cell, then
 * neighbour lists for as many persistent workgroups as the rest holds).
 * All calls on one GPU share its scratch: calls that overlap on one GPU (from
 * several contexts or threads) race; serialise them.  Never captured in a
 * graph by the library. */
int nc_count_noalloc(nc_context *ctx, const nc_params *params, const double *d_p1, int64_t n1, const double *d_p2,
                     int64_t n2, const double *d_p3, int64_t n3, int path, double *hist, uint64_t *counts);
/* Bytes of the scratch behind nc_count_noalloc. */
int64_t nc_noalloc_capacity(void);

/* Batched count (no counterpart in the reference, which makes one
 * calculateD1D2D3 call per term and per catalogue, main.cu:78-95): nterms
 * terms (c1[t], c2[t], c3[t]) with shared params run in one persistent launch
 * that pulls (term, primary-tile) items from a device work queue.  The launch
 * sequence is captured in a HIP graph and replayed while the same terms are
 * passed again.  On NC_PATH_MIXED (threeD), unit-weight terms beside weighted
 * ones run as a second launch, in the dense 32-bit layout that weighted terms
 * cannot use, whenever a dense nb^3 histogram could fit in a CU's LDS; those
 * two launches are not captured, and nc_last_timing then sums their times and
 * reports the unit-weight launch's layout.  hist: nterms*nb^3 doubles; counts:
 * nterms*nb^3 uint64 or NULL (a term's counts are filled when all its weights
 * are 1). */
int nc_count_batched(nc_context *ctx, const nc_params *params, int nterms, nc_catalog *const *c1,
                     nc_catalog *const *c2, nc_catalog *const *c3, int path, double *hist, uint64_t *counts);

/* Device timing of the last nc_count / nc_count_batched on ctx. */
int nc_last_timing(const nc_context *ctx, nc_timing *out);

/* Tuning knobs of the triplet kernel.  Values <= 0 keep the current setting
 * (wgs_per_cu: 0 = as many as fit, < 0 keeps).
 *   waves_per_group  in {1, 2, 4, 8, 16}; a count whose kernel variant is built for smaller
 *                    workgroups (the rppi survey key stream: at most 8) fails with NC_ERR_LIMIT
 *   chunk            primaries per work item of the device queue
 *   wgs_per_cu       persistent workgroups per compute unit
 *   list_chunk       neighbour-list entries staged in LDS per wave (multiple of 64) */
int nc_set_tuning(nc_context *ctx, int waves_per_group, int chunk, int wgs_per_cu, int list_chunk);
int nc_get_tuning(const nc_context *ctx, int *waves_per_group, int *chunk, int *wgs_per_cu, int *list_chunk);

/* Accumulation of weighted sums (no counterpart in the reference, which adds
 * every FP64 product (w1*w2)*w3 with a global FP64 atomic, device.cu:305-306,
 * in a run-dependent order).
 *   NC_ACCUM_FP64   (default) per-workgroup FP64 sums in LDS, merged into FP64
 *                   totals: the reference's arithmetic, order not reproducible.
 *   NC_ACCUM_EXACT  every product, computed in FP64 in the reference's order, is
 *                   truncated to a multiple of 2^-K and summed exactly as a
 *                   128-bit integer, then rounded once to double.  K is the
 *                   largest integer with every product of a weighted term below
 *                   2^(60-K) (from the catalogues' largest |w|), so each product
 *                   is off by less than 2^-K: relative 2^(e-60)/|p| for a
 *                   product p when the largest possible one is below 2^e
 *                   (about 5e-14 for lognormal(0, 0.3) weights).  Results are
 *                   identical on every run and on both arithmetic paths.  Needs
 *                   finite weights and at most 35 bins; threeD only (rppi keeps
 *                   FP64).  Unit-weight terms count exactly either way.
 *   NC_ACCUM_FP32   weight products in FP32 and per-workgroup FP32 running sums
 *                   with FP64 compensation, folded into FP64 totals at every
 *                   merge.  Each pair's product (c w1 w2) w3 is formed as an
 *                   FP32 pair hi + lo (relative error below 2^-46); hi goes to
 *                   the slot's FP32 sum with a returning LDS atomic, and the
 *                   exact rounding error of that addition (TwoSum) plus lo to
 *                   the slot's FP64 compensation.  Merges come at least every
 *                   2^22 pairs a workgroup adds (at work-item boundaries), so
 *                   the compensation's own FP64 rounding stays below
 *                   m^2 2^-77 of a slot's sum for m additions to the slot
 *                   between merges.  Needs FP32 weights: values FP32 holds
 *                   exactly, 0 or of magnitude in [2^-40, 2^40]
 *                   (NC_ERR_ARG otherwise); at most 35 bins; threeD only.
 *                   On MI355X an FP64 product issues at the rate of an FP32
 *                   one, so this mode costs more instructions per pair than
 *                   NC_ACCUM_FP64 and is not the default. */
enum nc_accum { NC_ACCUM_FP64 = 0, NC_ACCUM_EXACT = 1, NC_ACCUM_FP32 = 2 };
int nc_set_accumulation(nc_context *ctx, int mode);

/* Primary slices (no counterpart in the reference, whose kernel takes every
 * primary of p1 in one launch, cudaNcorr.cu:327-345): later counts on ctx
 * take, for every term, only the primaries of slice `slice` of `nslices`
 * equal slices of c1 in cell-sorted order; the nslices results add up to the
 * full count (exactly for counts, to rounding for weighted sums).  The
 * catalogues stay the same handles, so a term over one catalogue keeps its
 * six-fold symmetry.  For long counts split into steps (progress,
 * checkpoints) and for dividing one count between devices.  (0, 1) = all
 * primaries, the default. */
int nc_set_primary_slice(nc_context *ctx, int slice, int nslices);

/* Autotuner (no counterpart in the reference): times the triplet kernel (the
 * better of two runs) on a strided sample of the work items of these terms,
 * sized by a first run to about 80 ms, for every candidate of
 * waves_per_group {4, 8} (and 16 when 8 waves need more than half the CU's LDS) x chunk {16, 64} x list_chunk {64, 128, 256} (wgs_per_cu = as many as fit;
 * shapes whose LDS does not fit are skipped), keeps the fastest in ctx and returns it in *best (ms =
 * device time of the sample).  With the environment variable NC_TUNE_CACHE
 * naming a file, results are looked up in and appended to that file, keyed by
 * device, path, survey flag, bin count and catalogue sizes.  Where the sample
 * is a small part of the count, the best shape of each chunk is timed again
 * on one contiguous primary slice eight times the sample's size.  A batch that
 * nc_count_batched runs as two launches (unit-weight terms beside weighted
 * ones) is tuned launch by launch: the context keeps the unit-weight launch's
 * shape for that launch, *best is the weighted launch's (ms: the sum of the
 * two), and nc_set_tuning returns to one shape for every launch. */
typedef struct {
    int waves_per_group, chunk, wgs_per_cu, list_chunk;
    float ms;
    int from_cache;
} nc_tuning;
int nc_autotune(nc_context *ctx, const nc_params *params, int nterms, nc_catalog *const *c1, nc_catalog *const *c2,
                nc_catalog *const *c3, int path, nc_tuning *best);

/* Host evaluation of the mixed path's FP32 bin lookup (no counterpart in the
 * reference; a check of the lookup tables, no GPU used): for each FP32 squared
 * separation d[i] (unscaled), out[i] = the bin the lookup assigns, -1 outside
 * (edges[0]^2, edges[nb]^2) or NaN, or -2 when d[i] lies in an edge's guard
 * band and the pair would be re-evaluated in FP64.  *entries (optional) receives the
 * table length.  NC_ERR_LIMIT when the mixed path is disabled for these edges
 * (bins narrower than the guard bands). */
int nc_mixed_classify(const double *edges, int nedges, const float *d, int64_t n, int32_t *out, int32_t *entries);

#ifdef __cplusplus
}
#endif

#endif /* NCORR_HIP_H */
