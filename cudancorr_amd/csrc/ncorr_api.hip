// The C-ABI of libncorr_hip.so (include/ncorr_hip.h): contexts, catalogues,
// counts, timing, tuning and the host evaluation of the FP32 lookup.
#include <cstdio>
#include <cstdlib>
#include <new>

#include "ncorr_host.hpp"

using namespace nch;

extern "C" {

int nc_create(nc_context **out, int device) {
    if (!out) return NC_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return NC_ERR_HIP;
    nc_context *ctx = new (std::nothrow) nc_context();
    if (!ctx) return NC_ERR_NOMEM;
    ctx->device = device;
    hipDeviceProp_t prop;
    if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return NC_ERR_HIP;
    }
    {
        const hipError_t e = noalloc_base(device, &ctx->noalloc);
        if (e != hipSuccess) ctx->noalloc_err = std::string("hipGetSymbolAddress(noalloc_scratch): ") + hipGetErrorString(e);
    }
    ctx->ncu = prop.multiProcessorCount;
    ctx->dev_name = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount);
    ctx->lds_max = std::min<size_t>(prop.sharedMemPerBlock, 81920) - 256;
    ctx->lds_hw = prop.sharedMemPerBlock - 256;
    if (const char *e = std::getenv("NC_RPPI_STREAM")) ctx->rppi_stream = std::strcmp(e, "0") != 0 ? 1 : 0;
    if (const char *e = std::getenv("NC_RPPI_ROUND_KEYS")) {
        const double v = std::strtod(e, nullptr);
        if (v >= 1.0) ctx->rs_round_keys = v;
    }
    if (const char *e = std::getenv("NC_RPPI_REGION_SCALE")) {
        const double v = std::strtod(e, nullptr);
        if (v > 0.0 && v <= 1.0) ctx->rs_scale = v;
    }
    if (const char *e = std::getenv("NC_BATCH_SPLIT")) ctx->batch_split = std::strcmp(e, "0") != 0;
    if (const char *e = std::getenv("NC_FUSE_TERMS")) ctx->fuse_terms = std::strcmp(e, "0") != 0;
    if (const char *e = std::getenv("NC_TRI")) ctx->tri = std::strcmp(e, "0") != 0;
    if (const char *e = std::getenv("NC_TRI_LAST")) ctx->tri_last = std::strcmp(e, "0") != 0;
    if (const char *e = std::getenv("NC_FAIL_DEVICE_ALLOC")) ctx->fail_alloc = std::strtoll(e, nullptr, 10);
    if (const char *e = std::getenv("NC_DENSE_FLUSH_PAIRS")) {
        const unsigned long long v = std::strtoull(e, nullptr, 10);
        if (v > 0 && v <= (1ull << 31)) ctx->dense_flush = v;
    }
    for (auto &e : ctx->ev)
        if (hipEventCreate(&e) != hipSuccess) {
            nc_destroy(ctx);
            return NC_ERR_HIP;
        }
    *out = ctx;
    return NC_OK;
}

int nc_destroy(nc_context *ctx) {
    if (!ctx) return NC_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    plan_drop(ctx);
    for (DBuf *b : {&ctx->buf.tables, &ctx->buf.terms, &ctx->buf.chunks, &ctx->buf.out, &ctx->buf.counter,
                    &ctx->buf.errw, &ctx->buf.scratch, &ctx->buf.bound, &ctx->buf.wflags, &ctx->buf.cflags, &ctx->rs_tab, &ctx->rs_cursor, &ctx->rs_base, &ctx->rs_items, &ctx->rs_sample,
                    &ctx->rs_chunk_keys, &ctx->rs_keys, &ctx->rs_vals})
        dfree(*b);
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    for (nc_catalog *c : ctx->cats) c->ctx = nullptr;
    delete ctx;
    return NC_OK;
}

const char *nc_last_error(const nc_context *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int64_t nc_hist_size(int nedges) { return nedges < 2 ? 0 : (int64_t)(nedges - 1) * (nedges - 1) * (nedges - 1); }

int64_t nc_compact_size(int nedges) { return nedges < 2 ? 0 : choose_compact(nedges - 1); }

int64_t nc_hist_len(const nc_params *p) {
    if (!p || p->nedges < 2) return 0;
    const int64_t nb = p->nedges - 1;
    if (p->bin_type == NC_BIN_RPPI) return p->num_pi_bins < 1 ? 0 : nb * nb * nb * p->num_pi_bins * p->num_pi_bins;
    return p->bin_type == NC_BIN_THREED ? nb * nb * nb : 0;
}

int nc_catalog_create(nc_context *ctx, const double *xyzw, int64_t n, nc_catalog **cat) {
    if (!ctx || !cat || n < 0 || (n > 0 && !xyzw)) return ctx ? fail(ctx, NC_ERR_ARG, "bad catalogue") : NC_ERR_ARG;
    if (n > 2147483647LL) return fail(ctx, NC_ERR_LIMIT, "catalogues are limited to 2^31-1 points");
    *cat = nullptr;
    NC_HIP(ctx, hipSetDevice(ctx->device));
    nc_catalog *c = new (std::nothrow) nc_catalog();
    if (!c) return fail(ctx, NC_ERR_NOMEM, "host allocation");
    c->ctx = ctx;
    c->uid = ctx->next_uid++;
    c->n = n;
    c->owns = true;
    for (int64_t i = 0; i < n; ++i) {
        const double w = xyzw[4 * i + 3];
        c->unit_w = c->unit_w && w == 1.0;
        c->wf32 = c->wf32 && fp32_weight(w);
        if (std::isfinite(w))
            c->wmax = std::max(c->wmax, std::fabs(w));
        else
            c->wfinite = false;
    }
    hipError_t e = dgrow_checked(&ctx->fail_alloc, c->raw, sizeof(double) * 4 * (size_t)n);
    if (e == hipSuccess && n > 0)
        e = hipMemcpyAsync(c->raw.p, xyzw, sizeof(double) * 4 * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        dfree(c->raw);
        delete c;
        return fail(ctx, e == hipErrorOutOfMemory ? NC_ERR_NOMEM : NC_ERR_HIP, hipGetErrorString(e));
    }
    c->d_raw = (const double *)c->raw.p;
    c->device = ctx->device;
    ctx->cats.push_back(c);
    *cat = c;
    return NC_OK;
}

int nc_catalog_wrap_device(nc_context *ctx, const double *d_xyzw, int64_t n, nc_catalog **cat) {
    if (!ctx || !cat || n < 0 || (n > 0 && !d_xyzw)) return ctx ? fail(ctx, NC_ERR_ARG, "bad catalogue") : NC_ERR_ARG;
    if (n > 2147483647LL) return fail(ctx, NC_ERR_LIMIT, "catalogues are limited to 2^31-1 points");
    *cat = nullptr;
    NC_HIP(ctx, hipSetDevice(ctx->device));
    nc_catalog *c = new (std::nothrow) nc_catalog();
    if (!c) return fail(ctx, NC_ERR_NOMEM, "host allocation");
    c->ctx = ctx;
    c->uid = ctx->next_uid++;
    c->n = n;
    c->owns = false;
    c->d_raw = d_xyzw;
    // weight statistics: derived from the current contents before every count
    c->device = ctx->device;
    ctx->cats.push_back(c);
    *cat = c;
    return NC_OK;
}

int nc_catalog_destroy(nc_catalog *c) {
    if (!c) return NC_ERR_ARG;
    if (c->ctx) {
        (void)hipSetDevice(c->ctx->device);
        (void)hipStreamSynchronize(c->ctx->stream);
        const std::vector<uint64_t> &u = c->ctx->plan.uids;
        if (std::find(u.begin(), u.end(), c->uid) != u.end()) plan_drop(c->ctx);
        auto &v = c->ctx->cats;
        v.erase(std::remove(v.begin(), v.end(), c), v.end());
    } else {
        (void)hipSetDevice(c->device);
    }
    for (DBuf *b : {&c->raw, &c->key, &c->rank, &c->perm, &c->cnt, &c->start, &c->scan, &c->x, &c->y, &c->z, &c->w})
        dfree(*b);
    delete c;
    return NC_OK;
}

int64_t nc_catalog_size(const nc_catalog *c) { return c ? c->n : -1; }

int nc_count(nc_context *ctx, const nc_params *params, nc_catalog *c1, nc_catalog *c2, nc_catalog *c3, int path,
             double *hist, uint64_t *counts) {
    if (!ctx) return NC_ERR_ARG;
    std::vector<Term> t{{c1, c2, c3}};
    Mem m;
    m.fail = &ctx->fail_alloc;
    if (const int rc = refresh_weights(ctx, m, t)) return rc;
    return run_count(ctx, params, t, path, hist, counts, false, m);
}

int nc_count_host(nc_context *ctx, const nc_params *params, const double *p1, int64_t n1, const double *p2, int64_t n2,
                  const double *p3, int64_t n3, int path, double *hist, uint64_t *counts) {
    if (!ctx) return NC_ERR_ARG;
    nc_catalog *c[3] = {nullptr, nullptr, nullptr};
    // identical host buffers share one device catalogue (and with it the
    // (p2,p3) exchange symmetry)
    int rc = nc_catalog_create(ctx, p1, n1, &c[0]);
    if (!rc) {
        if (p2 == p1 && n2 == n1)
            c[1] = c[0];
        else
            rc = nc_catalog_create(ctx, p2, n2, &c[1]);
    }
    if (!rc) {
        if (p3 == p2 && n3 == n2)
            c[2] = c[1];
        else if (p3 == p1 && n3 == n1)
            c[2] = c[0];
        else
            rc = nc_catalog_create(ctx, p3, n3, &c[2]);
    }
    if (!rc) rc = nc_count(ctx, params, c[0], c[1], c[2], path, hist, counts);
    if (c[2] && c[2] != c[1] && c[2] != c[0]) nc_catalog_destroy(c[2]);
    if (c[1] && c[1] != c[0]) nc_catalog_destroy(c[1]);
    if (c[0]) nc_catalog_destroy(c[0]);
    return rc;
}

int nc_count_noalloc(nc_context *ctx, const nc_params *params, const double *d_p1, int64_t n1, const double *d_p2,
                     int64_t n2, const double *d_p3, int64_t n3, int path, double *hist, uint64_t *counts) {
    if (!ctx) return NC_ERR_ARG;
    if (!ctx->noalloc) return fail(ctx, NC_ERR_HIP, "no-alloc scratch unavailable: " + ctx->noalloc_err);
    if (!params || params->bin_type != NC_BIN_THREED)
        return fail(ctx, NC_ERR_ARG, "nc_count_noalloc counts threeD bins only");
    const double *d[3] = {d_p1, d_p2, d_p3};
    const int64_t n[3] = {n1, n2, n3};
    for (int i = 0; i < 3; ++i) {
        if (n[i] < 0 || (n[i] > 0 && !d[i])) return fail(ctx, NC_ERR_ARG, "bad catalogue");
        if (n[i] > 2147483647LL) return fail(ctx, NC_ERR_LIMIT, "catalogues are limited to 2^31-1 points");
    }
    // catalogues wrapping the caller's points, on the stack (never registered
    // with the context, never freed: their buffers live in the arena); the
    // same pointer and size passed twice is one catalogue, as in nc_count_host
    nc_catalog cat[3];
    nc_catalog *c[3];
    for (int i = 0; i < 3; ++i) {
        c[i] = &cat[i];
        for (int j = 0; j < i; ++j)
            if (d[j] == d[i] && n[j] == n[i]) {
                c[i] = c[j];
                break;
            }
        if (c[i] != &cat[i]) continue;
        cat[i].ctx = ctx;
        cat[i].device = ctx->device;
        cat[i].n = n[i];
        cat[i].d_raw = d[i];
        cat[i].owns = false;
    }
    Mem m;
    m.base = ctx->noalloc;
    m.size = kNoAllocBytes;
    NC_HIP(ctx, hipSetDevice(ctx->device));
    std::vector<Term> t{{c[0], c[1], c[2]}};
    if (const int rc = refresh_weights(ctx, m, t)) return rc;
    return run_count(ctx, params, t, path, hist, counts, false, m);
}

int64_t nc_noalloc_capacity(void) { return (int64_t)kNoAllocBytes; }

int nc_count_batched(nc_context *ctx, const nc_params *params, int nterms, nc_catalog *const *c1,
                     nc_catalog *const *c2, nc_catalog *const *c3, int path, double *hist, uint64_t *counts) {
    if (!ctx) return NC_ERR_ARG;
    if (nterms < 1 || !c1 || !c2 || !c3) return fail(ctx, NC_ERR_ARG, "bad term list");
    std::vector<Term> t(nterms);
    for (int i = 0; i < nterms; ++i) t[i] = {c1[i], c2[i], c3[i]};
    return count_terms(ctx, params, t, path, hist, counts, true);
}

int nc_last_timing(const nc_context *ctx, nc_timing *out) {
    if (!ctx || !out) return NC_ERR_ARG;
    *out = ctx->timing;
    return NC_OK;
}

// The lookup of MixedPol::group on the host: same clamp, cell, entry and
// comparisons, in FP32 on scaled values (the scaling is exact).
int nc_mixed_classify(const double *edges, int nedges, const float *d, int64_t n, int32_t *out, int32_t *entries) {
    if (!edges || nedges < 2 || n < 0 || (n > 0 && (!d || !out))) return NC_ERR_ARG;
    if (nedges > 41) return NC_ERR_LIMIT;
    Geometry G;
    G.edges.assign(edges, edges + nedges);
    if (!(G.edges[0] >= 0.0)) return NC_ERR_BINS;
    for (int i = 1; i < nedges; ++i)
        if (!(G.edges[i] > G.edges[i - 1])) return NC_ERR_BINS;
    G.nb = nedges - 1;
    G.sq.resize(nedges);
    for (int i = 0; i < nedges; ++i) G.sq[i] = G.edges[i] * G.edges[i];
    make_mixed_tables(G);
    if (!G.mixed_ok) return NC_ERR_LIMIT;
    if (entries) *entries = (int32_t)G.lut32.size();
    const float s = (float)(G.sigma32 * G.sigma32);
    for (int64_t i = 0; i < n; ++i) {
        const float x = d[i] * s;
        if (std::isnan(x)) {  // no separation: outside every bin
            out[i] = -1;
            continue;
        }
        const float c = std::min(std::max(x, G.lut32_lo), 4.0f);
        uint32_t bits;
        std::memcpy(&bits, &c, 4);
        const uint2 e = G.lut32[(bits >> G.lut32_shift) - (uint32_t)G.lut32_c0];
        float E;
        std::memcpy(&E, &e.y, 4);
        const float dE = x - E;
        const int k = (int)e.x - (std::signbit(dE) ? 1 : 0);
        out[i] = std::fabs(dE) <= G.g32 ? -2 : (unsigned)k < (unsigned)G.nb ? k : -1;
    }
    return NC_OK;
}

int nc_set_tuning(nc_context *ctx, int waves_per_group, int chunk, int wgs_per_cu, int list_chunk) {
    if (!ctx) return NC_ERR_ARG;
    if (list_chunk > 0) {
        if (list_chunk % 64 != 0 || list_chunk > 4096)
            return fail(ctx, NC_ERR_ARG, "list_chunk must be a multiple of 64, at most 4096");
        ctx->chunk_elems = list_chunk;
    }
    if (waves_per_group > 0) {
        if (waves_per_group != 1 && waves_per_group != 2 && waves_per_group != 4 && waves_per_group != 8 &&
            waves_per_group != 16)
            return fail(ctx, NC_ERR_ARG, "waves_per_group must be 1, 2, 4, 8 or 16");
        ctx->waves_per_group = waves_per_group;
    }
    if (chunk > 0) ctx->chunk = chunk;
    if (wgs_per_cu >= 0) ctx->wgs_per_cu = wgs_per_cu;
    ctx->tune_unit.on = false;  // one tuning for every launch again
    return NC_OK;
}

int nc_set_accumulation(nc_context *ctx, int mode) {
    if (!ctx) return NC_ERR_ARG;
    if (mode != NC_ACCUM_FP64 && mode != NC_ACCUM_EXACT && mode != NC_ACCUM_FP32)
        return fail(ctx, NC_ERR_ARG, "unknown accumulation mode");
    ctx->accum = mode;
    return NC_OK;
}

int nc_set_primary_slice(nc_context *ctx, int slice, int nslices) {
    if (!ctx) return NC_ERR_ARG;
    if (nslices < 1 || slice < 0 || slice >= nslices) return fail(ctx, NC_ERR_ARG, "need 0 <= slice < nslices");
    ctx->slice = slice;
    ctx->nslices = nslices;
    return NC_OK;
}

int nc_get_tuning(const nc_context *ctx, int *waves_per_group, int *chunk, int *wgs_per_cu, int *list_chunk) {
    if (!ctx) return NC_ERR_ARG;
    if (list_chunk) *list_chunk = ctx->chunk_elems;
    if (waves_per_group) *waves_per_group = ctx->waves_per_group;
    if (chunk) *chunk = ctx->chunk;
    if (wgs_per_cu) *wgs_per_cu = ctx->wgs_per_cu;
    return NC_OK;
}

// The autotuner over one set of terms (nc_autotune), leaving the best shape
// in ctx.
static int tune_terms(nc_context *ctx, const nc_params *params, const std::vector<Term> &t, int path,
                      nc_tuning *best) {
    const int nterms = (int)t.size();
    int64_t npts = 0, nprim = 0;
    for (const Term &x : t) {
        npts += x.c1->n + x.c2->n + x.c3->n;
        nprim += x.c1->n;
    }
    // cache key: device, path, survey, bins, sizes
    char key[256];
    std::snprintf(key, sizeof(key), "%s path=%d survey=%d nb=%d terms=%d points=%lld", ctx->dev_name.c_str(), path,
                  params->survey ? 1 : 0, params->nedges - 1, nterms, (long long)npts);
    const char *cache = std::getenv("NC_TUNE_CACHE");
    if (cache && *cache) {
        if (FILE *f = std::fopen(cache, "r")) {
            char line[512];
            while (std::fgets(line, sizeof(line), f)) {
                const size_t kl = std::strlen(key);
                int w, c, pc, lc;
                float ms;
                if (std::strncmp(line, key, kl) == 0 && line[kl] == '|' &&
                    std::sscanf(line + kl + 1, "%d %d %d %d %f", &w, &c, &pc, &lc, &ms) == 5) {
                    std::fclose(f);
                    ctx->waves_per_group = w;
                    ctx->chunk = c;
                    ctx->wgs_per_cu = pc;
                    ctx->chunk_elems = lc;
                    *best = {w, c, pc, lc, ms, 1};
                    return NC_OK;
                }
            }
            std::fclose(f);
        }
    }
    const int saved[4] = {ctx->waves_per_group, ctx->chunk, ctx->wgs_per_cu, ctx->chunk_elems};
    std::vector<double> hist((size_t)nterms * nc_hist_len(params));
    // sample: to start with, about two runs of kSampleRun primaries per CU
    const long long items = std::max<long long>(1, nprim / kSampleRun);  // runs of kSampleRun primaries
    ctx->sample_stride = (int)std::max<long long>(1, items / ((long long)ctx->ncu * 2));
    ctx->autotuning = true;
    ctx->tune_list_bound.clear();
    const int saved_slice[2] = {ctx->slice, ctx->nslices};  // the sample is drawn from all primaries
    ctx->slice = 0;
    ctx->nslices = 1;
    nc_tuning b = {saved[0], saved[1], saved[2], saved[3], 1e30f, 0};
    // the best shape for each work-item size, for the second stage
    nc_tuning best_c[2] = {{0, 0, 0, 0, 1e30f, 0}, {0, 0, 0, 0, 1e30f, 0}};
    std::vector<nc_tuning> runs;  // every shape that ran, with its time
    int rc = NC_OK;
    // 16-wave workgroups only where one workgroup per CU is forced anyway
    // (more than half the CU's LDS); elsewhere two 8-wave ones do better
    bool big_lds = false;
    {
        // a first run (8 waves per workgroup, the current chunks) sizes the
        // sample to about 80 ms, enough to separate shapes a few percent apart
        ctx->waves_per_group = 8;
        ctx->wgs_per_cu = 0;
        rc = count_terms(ctx, params, t, path, hist.data(), nullptr, false);
        big_lds = rc == NC_OK && (size_t)ctx->timing.lds_bytes > ctx->lds_max;
        const double ms = ctx->timing.ms_count;
        if (rc == NC_OK && ms > 0.0)
            ctx->sample_stride = (int)std::max(
                1.0, std::min<double>((double)items, std::floor(ctx->sample_stride * (ms / 80.0))));
        if (rc != NC_OK && rc != NC_ERR_LIMIT) goto done;
        rc = NC_OK;
    }
    for (int w : {4, 8, 16})
        for (int c : {16, 64})
            for (int lc : {64, 128, 256}) {
                if (w == 16 && !big_lds) continue;
                ctx->waves_per_group = w;
                ctx->chunk = c;
                ctx->wgs_per_cu = 0;
                ctx->chunk_elems = lc;
                float ms = 1e30f;
                int r = NC_OK;
                for (int rep = 0; rep < 2 && r == NC_OK; ++rep) {  // the better of two runs
                    r = count_terms(ctx, params, t, path, hist.data(), nullptr, false);
                    if (r == NC_OK) ms = std::min(ms, ctx->timing.ms_count);
                }
                if (r == NC_ERR_LIMIT) {  // this shape does not fit: try the next
                    if (rc == NC_OK) rc = r;
                    continue;
                }
                if (r != NC_OK) {
                    rc = r;
                    goto done;
                }
                runs.push_back({w, c, 0, lc, ms, 0});
            }
    // The fastest shape, except that a larger list chunk within kTuneTie of
    // it is preferred: it re-reads the HBM lists fewer times (list chunks 64
    // and 128 often time within a percent of each other).
    for (int g = -1; g < 2; ++g) {  // all shapes, then each work-item size
        nc_tuning &dst = g < 0 ? b : best_c[g];
        float fastest = 1e30f;
        for (const nc_tuning &x : runs)
            if (g < 0 || (x.chunk == 16) == (g == 0)) fastest = std::min(fastest, x.ms);
        for (const nc_tuning &x : runs)
            if ((g < 0 || (x.chunk == 16) == (g == 0)) && x.ms <= fastest * (1.0f + kTuneTie) &&
                (dst.ms >= 1e30f || x.list_chunk > dst.list_chunk ||
                 (x.list_chunk == dst.list_chunk && x.ms < dst.ms)))
                dst = x;
    }
    if (b.ms < 1e30f) rc = NC_OK;  // at least one shape ran
    // Second stage: where the sample is a small part of the count, runs of
    // primaries scattered over the catalogues need not rank the work-item
    // sizes as a count over consecutive cells does, so the best shape of each
    // item size is timed again on one contiguous primary slice of eight times
    // the sample's size (the whole count where that is less).
    if (rc == NC_OK && ctx->sample_stride > 1 && best_c[0].ms < 1e30f && best_c[1].ms < 1e30f) {
        const int ns = std::max(1, ctx->sample_stride / 8);
        ctx->sample_stride = 1;
        ctx->slice = ns / 2;
        ctx->nslices = ns;
        nc_tuning b2 = {0, 0, 0, 0, 1e30f, 0};
        for (const nc_tuning &cand : best_c) {
            ctx->waves_per_group = cand.waves_per_group;
            ctx->chunk = cand.chunk;
            ctx->wgs_per_cu = 0;
            ctx->chunk_elems = cand.list_chunk;
            float ms = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                const int r = count_terms(ctx, params, t, path, hist.data(), nullptr, false);
                if (r != NC_OK) {
                    rc = r;
                    goto done;
                }
                ms = std::min(ms, ctx->timing.ms_count);
            }
            if (ms < b2.ms) b2 = {cand.waves_per_group, cand.chunk, 0, cand.list_chunk, ms, 0};
        }
        b = b2;
    }
done:
    ctx->sample_stride = 1;
    ctx->autotuning = false;
    ctx->tune_list_bound.clear();
    ctx->slice = saved_slice[0];
    ctx->nslices = saved_slice[1];
    plan_drop(ctx);
    if (rc) {
        ctx->waves_per_group = saved[0];
        ctx->chunk = saved[1];
        ctx->wgs_per_cu = saved[2];
        ctx->chunk_elems = saved[3];
        return rc;
    }
    ctx->waves_per_group = b.waves_per_group;
    ctx->chunk = b.chunk;
    ctx->wgs_per_cu = b.wgs_per_cu;
    ctx->chunk_elems = b.list_chunk;
    *best = b;
    if (cache && *cache)
        if (FILE *f = std::fopen(cache, "a")) {
            std::fprintf(f, "%s|%d %d %d %d %.4f\n", key, b.waves_per_group, b.chunk, b.wgs_per_cu, b.list_chunk,
                         b.ms);
            std::fclose(f);
        }
    return NC_OK;
}

int nc_autotune(nc_context *ctx, const nc_params *params, int nterms, nc_catalog *const *c1, nc_catalog *const *c2,
                nc_catalog *const *c3, int path, nc_tuning *best) {
    if (!ctx) return NC_ERR_ARG;
    if (nterms < 1 || !c1 || !c2 || !c3 || !params || !best) return fail(ctx, NC_ERR_ARG, "bad autotune arguments");
    std::vector<Term> t(nterms);
    for (int i = 0; i < nterms; ++i) {
        t[i] = {c1[i], c2[i], c3[i]};
        if (!c1[i] || !c2[i] || !c3[i]) return fail(ctx, NC_ERR_ARG, "null catalogue");
    }
    Mem m;
    m.fail = &ctx->fail_alloc;
    if (const int rc = refresh_weights(ctx, m, t)) return rc;
    ctx->tune_unit.on = false;
    std::vector<int> iu, iw;
    if (!split_batch(ctx, params, t, iu, iw)) return tune_terms(ctx, params, t, path, best);
    // the unit-weight and the weighted launch of a split batch, each tuned on its own
    std::vector<Term> tu, tw;
    for (int i : iu) tu.push_back(t[i]);
    for (int i : iw) tw.push_back(t[i]);
    nc_tuning bu, bw;
    if (const int rc = tune_terms(ctx, params, tu, path, &bu)) return rc;
    if (const int rc = tune_terms(ctx, params, tw, path, &bw)) return rc;
    ctx->tune_unit.on = true;
    ctx->tune_unit.waves_per_group = bu.waves_per_group;
    ctx->tune_unit.chunk = bu.chunk;
    ctx->tune_unit.wgs_per_cu = bu.wgs_per_cu;
    ctx->tune_unit.list_chunk = bu.list_chunk;
    *best = bw;
    best->ms = bu.ms + bw.ms;
    best->from_cache = bu.from_cache && bw.from_cache;
    return NC_OK;
}

}  // extern "C"

