// A count on the device: cell-list builds, the plan (histogram layout, launch
// shape, work queue), the launches, graph capture and replay, and the
// conversion of the results into the reference's dense layout.
#include <mutex>

#include "ncorr_host.hpp"

namespace nch {

hipError_t dgrow(DBuf &b, size_t need) {
    if (need == 0) need = 16;
    if (need <= b.bytes) return hipSuccess;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    hipError_t e = hipMalloc(&b.p, need);
    if (e == hipSuccess) b.bytes = need;
    return e;
}

void dfree(DBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

hipError_t dgrow_checked(long long *fail, DBuf &b, size_t need) {
    if (fail && *fail >= 0 && (need == 0 ? 16 : need) > b.bytes && (*fail)-- == 0) return hipErrorOutOfMemory;
    return dgrow(b, need);
}

struct CatGrid {
    nc_catalog *cat;
    bool is_p1;
    bool expect_exact;  // a captured build: the plan counted this catalogue as exactly gridded
};

int build_grid(nc_context *ctx, Mem &m, CountBufs &B, const Geometry &G, nc_catalog *c, bool check_unclamped,
               unsigned *inexact, bool expect_exact) {
    const int n = (int)c->n;
    // slab counts, padded to whole scan tiles (the padding stays zero)
    const size_t ntiles = ((size_t)G.ncells * kSlabs + kScanTile - 1) / kScanTile;
    const size_t ns = ntiles * kScanTile;
    NC_MEM(ctx, m, c->key, sizeof(uint32_t) * (size_t)n);
    NC_MEM(ctx, m, c->rank, sizeof(uint32_t) * (size_t)n);
    NC_MEM(ctx, m, c->perm, sizeof(uint32_t) * (size_t)n);
    NC_MEM(ctx, m, c->cnt, sizeof(uint32_t) * ns);
    NC_MEM(ctx, m, c->start, sizeof(uint32_t) * (ns + 1));
    NC_MEM(ctx, m, c->scan, sizeof(uint32_t) * (2 * ntiles + 1));  // tile sums, then their scan
    NC_MEM(ctx, m, c->x, sizeof(double) * (size_t)n);
    NC_MEM(ctx, m, c->y, sizeof(double) * (size_t)n);
    NC_MEM(ctx, m, c->z, sizeof(double) * (size_t)n);
    NC_MEM(ctx, m, c->w, sizeof(double) * (size_t)n);
    GridGeo g;
    for (int a = 0; a < 3; ++a) {
        g.cs[a] = G.cs[a];
        g.nc[a] = G.nc[a];
    }
    g.survey = G.survey;
    const int fb = (int)std::min<long long>(((long long)ns + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(fb), dim3(256), 0, ctx->stream, (uint32_t *)c->cnt.p, 0u, (long long)ns);
    if (n > 0) {
        const int nb = (n + 255) / 256;
        hipLaunchKernelGGL(k_cell_keys, dim3(nb), dim3(256), 0, ctx->stream, (const double4 *)c->d_raw, n, g,
                           check_unclamped ? 1 : 0, (uint32_t *)c->key.p, (uint32_t *)c->rank.p, (uint32_t *)c->cnt.p,
                           (unsigned *)B.errw.p, inexact, expect_exact ? 1 : 0);
    }
    uint32_t *part = (uint32_t *)c->scan.p, *pstart = part + ntiles;
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(256), 0, ctx->stream, (const uint32_t *)c->cnt.p, part);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, ctx->stream, (const uint32_t *)part, pstart, (int)ntiles);
    hipLaunchKernelGGL(k_scan_apply, dim3((unsigned)ntiles), dim3(256), 0, ctx->stream, (const uint32_t *)c->cnt.p,
                       (const uint32_t *)pstart, (uint32_t *)c->start.p, (int)ntiles);
    if (n > 0) {
        const int nb = (n + 255) / 256;
        hipLaunchKernelGGL(k_place, dim3(nb), dim3(256), 0, ctx->stream, n, (const uint32_t *)c->key.p,
                           (const uint32_t *)c->rank.p, (const uint32_t *)c->start.p, (uint32_t *)c->perm.p);
        hipLaunchKernelGGL(k_stable_scatter, dim3(nb), dim3(256), 0, ctx->stream, (const double4 *)c->d_raw, n,
                           (const uint32_t *)c->key.p, (const uint32_t *)c->start.p, (const uint32_t *)c->perm.p,
                           (double *)c->x.p, (double *)c->y.p, (double *)c->z.p, (double *)c->w.p);
    }
    NC_HIP(ctx, hipGetLastError());
    return NC_OK;
}


// Device view of a term: the cell-sorted arrays of its three catalogues.
TermDesc term_desc(const Term &t) {
    TermDesc d;
    d.x1 = (const double *)t.c1->x.p;
    d.y1 = (const double *)t.c1->y.p;
    d.z1 = (const double *)t.c1->z.p;
    d.w1 = (const double *)t.c1->w.p;
    d.x2 = (const double *)t.c2->x.p;
    d.y2 = (const double *)t.c2->y.p;
    d.z2 = (const double *)t.c2->z.p;
    d.w2 = (const double *)t.c2->w.p;
    d.st2 = (const uint32_t *)t.c2->start.p;
    d.x3 = (const double *)t.c3->x.p;
    d.y3 = (const double *)t.c3->y.p;
    d.z3 = (const double *)t.c3->z.p;
    d.w3 = (const double *)t.c3->w.p;
    d.st3 = (const uint32_t *)t.c3->start.p;
    d.n1 = (int)t.c1->n;
    d.sym = t.c2 == t.c3;
    d.weighted = !(t.c1->unit_w && t.c2->unit_w && t.c3->unit_w);
    d.wprim = 0;  // set by run_count for FP64 accumulation
    d.wunit2 = t.c2->unit_w;
    d.wunit3 = t.c3->unit_w;
    d.partner = -1;  // set by run_count
    d.tri = 0;  // set by run_count once the builds have shown exact gridding
    d.prune2 = d.prune3 = 0;  // likewise
    return d;
}

void plan_drop(nc_context *ctx) {
    BatchPlan &p = ctx->plan;
    if (p.build_exec) (void)hipGraphExecDestroy(p.build_exec);
    if (p.count_exec) (void)hipGraphExecDestroy(p.count_exec);
    if (p.h_out) (void)hipHostFree(p.h_out);
    if (p.h_err) (void)hipHostFree(p.h_err);
    p = BatchPlan();
}

// A 128-bit fixed-point sum (high word, low word) in units of 2^-k, rounded
// once to double.
double fixed_to_double(unsigned long long lo, unsigned long long hi, int k) {
    const __int128 v = (__int128)(((unsigned __int128)hi << 64) | (unsigned __int128)lo);
    return std::ldexp((double)v, -k);
}

// Compact sorted-bin results -> dense nb^3 histogram (device.cu:198 layout).
// Exact layout (xk != kNoExact): per term nbc low words then nbc high words,
// weighted sums in units of 2^-xk, counts in the low words.
void convert_out(int nb, long long nbc, int rppi, int nterms, const std::vector<int> &weighted,
                 const unsigned long long *res, double *hist, uint64_t *counts, int xk = kNoExact) {
    const bool exact = xk != kNoExact;
    const long long stride = exact ? 2 * nbc : nbc;
    if (rppi) {  // already the dense layout
        for (int i = 0; i < nterms; ++i) {
            const unsigned long long *r = res + (size_t)i * nbc;
            double *h = hist + (size_t)i * nbc;
            uint64_t *cn = counts ? counts + (size_t)i * nbc : nullptr;
            for (long long k = 0; k < nbc; ++k) {
                if (weighted[i]) {
                    std::memcpy(&h[k], &r[k], 8);
                } else {
                    h[k] = (double)r[k];
                    if (cn) cn[k] = r[k];
                }
            }
            if (weighted[i] && cn) std::fill(cn, cn + nbc, (uint64_t)0);
        }
        return;
    }
    const size_t nh = (size_t)nb * nb * nb;
    for (int i = 0; i < nterms; ++i) {
        double *h = hist + (size_t)i * nh;
        uint64_t *cn = counts ? counts + (size_t)i * nh : nullptr;
        std::fill(h, h + nh, 0.0);
        if (cn) std::fill(cn, cn + nh, (uint64_t)0);
        const unsigned long long *r = res + (size_t)i * stride;
        int c = 0;
        for (int a = 0; a < nb; ++a)
            for (int b = a; b < nb; ++b)
                for (int k = b; k < nb; ++k, ++c) {
                    const size_t di = (size_t)k + (size_t)b * nb + (size_t)a * nb * nb;
                    if (weighted[i] && exact) {
                        h[di] = fixed_to_double(r[c], r[nbc + c], xk);
                    } else if (weighted[i]) {
                        double v;
                        std::memcpy(&v, &r[c], 8);
                        h[di] = v;
                    } else {
                        h[di] = (double)r[c];
                        if (cn) cn[di] = r[c];
                    }
                }
    }
}

// Replay a captured plan: cell-list builds graph, triplet graph, copy-back.
int plan_replay(nc_context *ctx, double *hist, uint64_t *counts, bool &retry) {
    BatchPlan &p = ctx->plan;
    retry = false;
    NC_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    NC_HIP(ctx, hipGraphLaunch(p.build_exec, ctx->stream));
    NC_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    NC_HIP(ctx, hipGraphLaunch(p.count_exec, ctx->stream));
    NC_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    const int nterms = (int)p.weighted.size();
    const long long ostride = p.xk != kNoExact ? 2 * p.nbc : p.nbc;
    NC_HIP(ctx, hipMemcpyAsync(p.h_out, ctx->buf.out.p, sizeof(unsigned long long) * (size_t)nterms * ostride,
                               hipMemcpyDeviceToHost, ctx->stream));
    NC_HIP(ctx, hipMemcpyAsync(p.h_err, ctx->buf.errw.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    NC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms_b = 0.f, ms_c = 0.f;
    NC_HIP(ctx, hipEventElapsedTime(&ms_b, ctx->ev[0], ctx->ev[1]));
    NC_HIP(ctx, hipEventElapsedTime(&ms_c, ctx->ev[1], ctx->ev[3]));
    ctx->timing.ms_build = ms_b;
    ctx->timing.ms_count = ms_c;
    ctx->timing.ms_total = ms_b + ms_c;
    ctx->timing.n_kernel = p.nkernel;
    ctx->timing.hist_layout = p.hist_layout;
    ctx->timing.lds_bytes = p.lds_bytes;
    ctx->timing.path = p.path;
    ctx->timing.tri_terms = p.tri_terms;
    ctx->timing.swap_terms = p.swap_terms;
    ctx->timing.partner_terms = p.partner_terms;
    if (*p.h_err & kErrDomain) return fail(ctx, NC_ERR_DOMAIN, "a point lies outside the cell grid of the box");
    if (*p.h_err & (kErrOverflow | kErrReplan)) {  // catalogue contents changed under a wrapped pointer: rebuild eagerly
        plan_drop(ctx);
        retry = true;
        return NC_OK;
    }
    if (*p.h_err & kErrMode) return fail(ctx, NC_ERR_HIP, "internal error: a term does not match the weighting of its kernel variant");
    convert_out(p.nb, p.nbc, p.rppi, nterms, p.weighted, p.h_out, hist, counts, p.xk);
    return NC_OK;
}

// ---------------------------------------------------------------------------
// A count in stages (run_count at the end): the plan key, the rppi key-stream
// tables, the LDS layout, the cell lists, the symmetric paths, the histogram
// layout, the work queue and launch arguments, the key-stream regions, the
// launches and the graph capture.

namespace {

// Key of a captured plan: the catalogues, their weight status, the params and
// accumulation, the path and the tuning; the plan is replayed while all of
// them come back.
struct PlanKey {
    std::vector<uint64_t> uids;
    std::vector<int> wkey;  // per term: unit-weight flags of c1, c2, c3 (bits 0-2), FP32-weight flags (bits 3-5)
    std::vector<double> pkey;
    std::vector<int> tkey;
};

// Accumulation of the weighted sums of a count (nc_set_accumulation).
struct Accum {
    bool all_unit = true;  // every term has unit weights
    bool exact = false;    // NC_ACCUM_EXACT over weighted terms
    bool f32 = false;      // NC_ACCUM_FP32 over weighted terms
    int xk = kNoExact;     // exact: the fixed-point scale exponent
};

bool unit_term(const Term &t) { return t.c1->unit_w && t.c2->unit_w && t.c3->unit_w; }

int plan_key(nc_context *ctx, const nc_params *P, const std::vector<Term> &terms, int path, const Geometry &G,
             PlanKey &key, Accum &acc) {
    for (const Term &t : terms) key.uids.insert(key.uids.end(), {t.c1->uid, t.c2->uid, t.c3->uid});
    // weight status of each catalogue (wrapped ones may change it between calls)
    for (const Term &t : terms)
        key.wkey.push_back((t.c1->unit_w ? 1 : 0) | (t.c2->unit_w ? 2 : 0) | (t.c3->unit_w ? 4 : 0) |
                           (t.c1->wf32 ? 8 : 0) | (t.c2->wf32 ? 16 : 0) | (t.c3->wf32 ? 32 : 0));
    key.pkey.assign(P->box, P->box + 3);
    key.pkey.push_back((double)G.survey);
    key.pkey.push_back((double)G.rppi);
    key.pkey.push_back((double)G.npi);
    key.pkey.push_back(G.ori_x);
    key.pkey.insert(key.pkey.end(), G.edges.begin(), G.edges.end());
    // exact accumulation: every weight product is truncated to a multiple of
    // 2^-xk with xk set so that the largest product any weighted term can form
    // stays below 2^60 units (below 2^63 for the six orders of a triple together)
    for (const Term &t : terms) acc.all_unit = acc.all_unit && unit_term(t);
    acc.exact = ctx->accum == NC_ACCUM_EXACT && !G.rppi && !acc.all_unit;
    if (acc.exact) {
        double pmax = 0.0;
        for (const Term &t : terms) {
            if (unit_term(t)) continue;
            if (!t.c1->wfinite || !t.c2->wfinite || !t.c3->wfinite)
                return fail(ctx, NC_ERR_ARG, "exact accumulation needs finite weights");
            pmax = std::max(pmax, t.c1->wmax * t.c2->wmax * t.c3->wmax);
        }
        int e = 0;
        if (pmax > 0.0) std::frexp(pmax, &e);  // pmax < 2^e
        acc.xk = std::min(std::max(60 - e, -900), 900);
    }
    key.pkey.push_back((double)acc.xk);
    // FP32 accumulation: FP32 weight products and running sums with FP64
    // compensation, for FP32 weights only
    acc.f32 = ctx->accum == NC_ACCUM_FP32 && !G.rppi && !acc.all_unit;
    if (acc.f32)
        for (const Term &t : terms)
            if (!unit_term(t) && !(t.c1->wf32 && t.c2->wf32 && t.c3->wf32))
                return fail(ctx, NC_ERR_ARG,
                            "FP32 accumulation needs FP32 weights (values FP32 holds exactly, 0 or of magnitude in "
                            "[2^-40, 2^40])");
    key.pkey.push_back((double)ctx->accum);
    key.tkey = {path, ctx->waves_per_group, ctx->chunk, ctx->wgs_per_cu, ctx->chunk_elems, ctx->slice, ctx->nslices};
    return NC_OK;
}

size_t al16(size_t v) { return (v + 15) & ~(size_t)15; }

// The rppi key stream's tables: the reachable dense rp triples (sorted ones
// and the next triple, where a pi index equal to npi carries the flat index),
// their ranks, and per term a run of buckets of 2^shift compact bins.
// rppi output: the key stream by default for boxes (6x the direct atomics on
// the 1M-point box); direct atomics for surveys, whose pair arithmetic bounds
// both paths (2.0 s direct against 2.4 s streamed on 500k data + 2.5M
// randoms), unless NC_RPPI_STREAM forces one.
struct RppiStream {
    bool on = false;
    std::vector<uint16_t> rank;
    std::vector<int> inv;
    std::vector<int4> term;  // per term: first bucket, buckets, shift, 0
    int nbk = 0;             // buckets of all terms
    bool anyw = false;       // some term weighted
    int maxshift = 0;        // log2 of the binning LDS bytes
    int nbm = 8;             // per-wave bucket slots: the largest bucket count of a term, rounded up to 8
};

void rppi_stream_tables(const nc_context *ctx, const Geometry &G, const std::vector<Term> &terms, RppiStream &rs) {
    const int nterms = (int)terms.size();
    rs.on = G.rppi && (ctx->rppi_stream > 0 || (ctx->rppi_stream < 0 && !G.survey));
    rs.term.assign(nterms, make_int4(0, 0, 0, 0));
    if (!rs.on) return;
    const int nb = G.nb, dmax = G.npi == 1 ? 2 : 1;
    const long long n3 = (long long)nb * nb * nb;
    std::vector<char> reach((size_t)n3, 0);
    for (int k1 = 0; k1 < nb; ++k1)
        for (int k2 = k1; k2 < nb; ++k2)
            for (int k3 = k2; k3 < nb; ++k3) {
                const long long T = k3 + (long long)k2 * nb + (long long)k1 * nb * nb;
                for (int d = 0; d <= dmax; ++d)
                    if (T + d < n3) reach[(size_t)(T + d)] = 1;
            }
    rs.rank.assign((size_t)n3, 0xffff);
    for (long long T = 0; T < n3; ++T)
        if (reach[(size_t)T]) {
            rs.rank[(size_t)T] = (uint16_t)rs.inv.size();
            rs.inv.push_back((int)T);
        }
    const long long cbins = (long long)rs.inv.size() * G.npi * G.npi;
    int sh_u = 0, sh_w = 0;
    while (sh_u < 15 && ((size_t)4 << (sh_u + 1)) <= ctx->lds_hw) ++sh_u;
    while (sh_w < 14 && ((size_t)8 << (sh_w + 1)) <= ctx->lds_hw) ++sh_w;
    if (rs.inv.size() >= 65535 || cbins >= (1LL << 31)) rs.on = false;
    for (int i = 0; i < nterms && rs.on; ++i) {
        const bool w = !unit_term(terms[i]);
        const int sh = w ? sh_w : sh_u;
        const long long nbt = (cbins + (1LL << sh) - 1) >> sh;
        if (nbt > kStreamBuckets) rs.on = false;
        rs.term[i] = make_int4(rs.nbk, (int)nbt, sh, 0);
        rs.nbk += (int)nbt;
        rs.anyw = rs.anyw || w;
        rs.maxshift = std::max(rs.maxshift, sh + (w ? 3 : 2));
    }
    for (int i = 0; i < nterms && rs.on; ++i) rs.nbm = std::max(rs.nbm, (rs.term[i].y + 7) & ~7);
}

// LDS of a launch of k_triplets / k_rppi: the histogram, the tables the
// variant reads in its pair loop (plain sq + lut, mixed lut32, and the slot
// table t3 or ctab; the rppi rank table ends them), then per wave the staged
// list chunk (mixed: float4 + bin, + weight unless the layout is the
// unit-weight kHistDense), the recheck queue (4- or 8-byte entries, as
// QueueEntry) and the rppi key staging.
struct LdsOffsets {
    size_t sq = 0, lut = 0, ctab = 0, lut32 = 0, t3 = 0;
};
struct LdsLayout {
    const Geometry &G;
    bool mixed;
    size_t stage_bytes = 0, rank_bytes = 0;
    int nhist = 1;  // 2: a second histogram for partner terms

    size_t wave(int c, int hm) const {
        const size_t cp = (size_t)c + 4;
        if (G.rppi) return al16(cp * sizeof(RElem)) + stage_bytes;
        return mixed ? al16(cp * (sizeof(float4) + sizeof(int) + (hm == kHistDense ? 0 : sizeof(double))) +
                            kQueueCap * (hm == kHistDense ? sizeof(uint32_t) : sizeof(uint2)))
                     : al16(cp * (sizeof(Elem) + sizeof(double)));
    }
    size_t hist(int hm) const {
        const size_t nb3 = (size_t)G.nb * G.nb * G.nb;
        return hm == kHistGlobal    ? (size_t)0
               : hm == kHistDense   ? al16(nb3 * 4)
               : hm == kHistExact   ? al16((size_t)G.nbc * 16)
               : is_f32(hm)         ? al16((size_t)G.nbc * 12)
                                    : al16((size_t)G.nbc * 8);
    }
    size_t tables(int hm, LdsOffsets &o) const {
        size_t x = 0;
        o = LdsOffsets();
        if (!mixed) {
            o.sq = x;
            x = al16(x + sizeof(double) * G.sq.size());
            o.lut = x;
            x = al16(x + sizeof(uint16_t) * G.lut.size());
        } else {
            o.lut32 = x;
            x = al16(x + sizeof(uint2) * G.lut32.size());
        }
        if (t3_like(hm)) {
            o.t3 = x;
            x = al16(x + sizeof(uint16_t) * G.t3.size());
        } else if (ctab_like(hm)) {
            o.ctab = x;
            x = al16(x + sizeof(int) * G.ctab.size());
        }
        return x + (hm == kHistGlobal ? rank_bytes : 0);
    }
    bool fits(int hm, int W, size_t budget) const {
        LdsOffsets o;
        return nhist * hist(hm) + tables(hm, o) + W * wave(64, hm) <= budget;
    }
};

// Global copies of every table, one upload: offsets into it.
struct TableImage {
    size_t sq, lut, ctab, lut32, t3, bytes;
    std::vector<char> data;
};
TableImage table_image(const Geometry &G) {
    TableImage T;
    T.sq = 0;
    T.lut = al16(T.sq + sizeof(double) * G.sq.size());
    T.ctab = al16(T.lut + sizeof(uint16_t) * G.lut.size());
    T.lut32 = al16(T.ctab + sizeof(int) * G.ctab.size());
    T.t3 = al16(T.lut32 + sizeof(uint2) * G.lut32.size());
    T.bytes = al16(T.t3 + sizeof(uint16_t) * G.t3.size());
    T.data.assign(T.bytes, 0);
    std::memcpy(T.data.data() + T.sq, G.sq.data(), sizeof(double) * G.sq.size());
    std::memcpy(T.data.data() + T.lut, G.lut.data(), sizeof(uint16_t) * G.lut.size());
    std::memcpy(T.data.data() + T.ctab, G.ctab.data(), sizeof(int) * G.ctab.size());
    if (!G.lut32.empty()) std::memcpy(T.data.data() + T.lut32, G.lut32.data(), sizeof(uint2) * G.lut32.size());
    if (!G.t3.empty()) std::memcpy(T.data.data() + T.t3, G.t3.data(), sizeof(uint16_t) * G.t3.size());
    return T;
}

// One cell list per distinct catalogue of the terms.
std::vector<CatGrid> collect_grids(const std::vector<Term> &terms) {
    std::vector<CatGrid> grids;
    for (const Term &t : terms) {
        nc_catalog *cs[3] = {t.c1, t.c2, t.c3};
        for (int r = 0; r < 3; ++r) {
            bool found = false;
            for (CatGrid &g : grids)
                if (g.cat == cs[r]) {
                    found = true;
                    g.is_p1 = g.is_p1 || r == 0;
                }
            if (!found) grids.push_back({cs[r], r == 0, false});
        }
    }
    return grids;
}

int enqueue_builds(nc_context *ctx, Mem &m, CountBufs &bufs, const Geometry &G, const std::vector<CatGrid> &grids) {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t *)bufs.errw.p, 0u, 4LL);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, ctx->stream, (uint32_t *)bufs.cflags.p, 0u,
                       (long long)grids.size());
    for (size_t gi = 0; gi < grids.size(); ++gi) {
        const CatGrid &g = grids[gi];
        const int r = build_grid(ctx, m, bufs, G, g.cat, G.survey && g.is_p1, (unsigned *)bufs.cflags.p + gi,
                                 g.expect_exact);
        if (r) return r;
    }
    return NC_OK;
}

// Lists of an exactly gridded catalogue skip the slabs beyond rmax (threeD).
bool prunes(const Geometry &G, const nc_catalog *c) { return !G.rppi && c->exact; }

// Terms over one catalogue (c1 == c2 == c3) count every unordered triple
// once when its gridding is exact (threeD only); survey terms with c1 == c2 !=
// c3 every unordered (p1, p2) pair.  A launch with weighted terms runs its
// unit-weight terms on the plain two-fold path (its kernel variant carries
// these paths for weighted terms only).  The catalogues these paths and the
// slab cut rely on are checked again by a replayed build.
struct Symmetry {
    std::vector<char> tri;
    int ntri = 0, nswap = 0;
};
Symmetry plan_symmetry(const nc_context *ctx, const Geometry &G, const std::vector<Term> &terms, bool all_unit,
                       std::vector<CatGrid> &grids) {
    const int nterms = (int)terms.size();
    Symmetry S;
    S.tri.assign(nterms, kTriNone);
    for (int i = 0; i < nterms; ++i) {
        const Term &t = terms[i];
        const bool own = all_unit || !unit_term(t);
        const bool one = ctx->tri && !G.rppi && own && t.c1 == t.c2 && t.c1->exact && t.c1->n > 0;
        S.tri[i] = !one ? kTriNone : t.c2 == t.c3 ? kTriSix : G.survey ? kTriSwap : kTriNone;
        S.ntri += S.tri[i] == kTriSix;
        S.nswap += S.tri[i] == kTriSwap;
    }
    for (CatGrid &g : grids) {
        g.expect_exact = false;
        for (int i = 0; i < nterms; ++i)
            g.expect_exact = g.expect_exact || (S.tri[i] && terms[i].c1 == g.cat) ||
                             (prunes(G, g.cat) && (terms[i].c2 == g.cat || terms[i].c3 == g.cat));
    }
    return S;
}

// Slab pruning radius: rmax plus a margin far above the rounding of the cell
// and slab bounds (and of a point's slab) at these coordinates.
double prune_radius2(const Geometry &G) {
    double bmax = 0.0;
    for (int a = 0; a < 3; ++a) bmax = std::max(bmax, std::fabs(G.box[a]));
    const double r = G.edges[G.nb] * (1.0 + 1e-9) + 1e-9 * bmax;
    return r * r;
}

// The primaries [p0, p1) of the slice counted of a catalogue of n points.
void slice_range(const nc_context *ctx, int n, int &p0, int &p1) {
    p0 = (int)((long long)n * ctx->slice / ctx->nslices);
    p1 = (int)((long long)n * (ctx->slice + 1) / ctx->nslices);
}

// The exact longest list over the counted primaries (k_list_bound: build_list
// with no storage), when only it can admit the dense layout; kept per
// catalogues for the runs of one nc_autotune call.
int longest_list(nc_context *ctx, CountBufs &bufs, const Geometry &G, const std::vector<Term> &terms,
                 const std::vector<uint64_t> &uids, int cap, int &nmax) {
    if (ctx->autotuning && ctx->tune_list_bound.count(uids)) {
        nmax = ctx->tune_list_bound[uids];
        return NC_OK;
    }
    KArgs B;
    std::memset(&B, 0, sizeof(B));
    for (int a = 0; a < 3; ++a) {
        B.box[a] = G.box[a];
        B.cs[a] = G.cs[a];
        B.nc[a] = G.nc[a];
    }
    B.survey = G.survey;
    B.r2min = G.sq[0];
    B.r2max = G.sq[G.nb];
    B.prune_r2 = prune_radius2(G);
    B.cap = 0;  // count only
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t *)bufs.bound.p, 0u, 1LL);
    for (const Term &t : terms) {
        TermDesc d = term_desc(t);
        d.prune2 = prunes(G, t.c2);
        d.prune3 = prunes(G, t.c3);
        if (d.n1 == 0 || t.c2->n == 0 || t.c3->n == 0) continue;
        int p0, p1;  // the same split as the work queue
        slice_range(ctx, d.n1, p0, p1);
        if (p1 <= p0) continue;
        const int grid = (int)std::min<long long>((p1 - p0 + 3) / 4, (long long)ctx->ncu * 32);
        hipLaunchKernelGGL(k_list_bound, dim3(grid), dim3(256), 0, ctx->stream, B, d, p0, p1,
                           (unsigned *)bufs.bound.p);
    }
    NC_HIP(ctx, hipGetLastError());
    unsigned lb = 0;
    NC_HIP(ctx, hipMemcpyAsync(&lb, bufs.bound.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    NC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    nmax = std::min(cap, (int)std::max(1u, lb));
    if (ctx->autotuning) ctx->tune_list_bound[uids] = nmax;
    return NC_OK;
}

// The kernel variant of a launch.
const void *launch_kernel(const Geometry &G, bool stream, int hm, bool mixed, bool all_unit) {
    if (G.rppi)
        return G.survey ? (stream ? (const void *)k_rppi<true, true> : (const void *)k_rppi<true, false>)
                        : (stream ? (const void *)k_rppi<false, true> : (const void *)k_rppi<false, false>);
    if (hm == kHistExact)
        return mixed ? (const void *)k_triplets<true, kHistExact, true> : (const void *)k_triplets<false, kHistExact, true>;
    if (hm == kHistF32)
        return mixed ? (const void *)k_triplets<true, kHistF32, true> : (const void *)k_triplets<false, kHistF32, true>;
    if (hm == kHistF32C)
        return mixed ? (const void *)k_triplets<true, kHistF32C, true> : (const void *)k_triplets<false, kHistF32C, true>;
    // [weighted terms present][mixed][layout]
    const void *kfns[2][2][3] = {
        {{(const void *)k_triplets<false, kHistCtab, false>, (const void *)k_triplets<false, kHistT3, false>,
          (const void *)k_triplets<false, kHistDense, false>},
         {(const void *)k_triplets<true, kHistCtab, false>, (const void *)k_triplets<true, kHistT3, false>,
          (const void *)k_triplets<true, kHistDense, false>}},
        {{(const void *)k_triplets<false, kHistCtab, true>, (const void *)k_triplets<false, kHistT3, true>, nullptr},
         {(const void *)k_triplets<true, kHistCtab, true>, (const void *)k_triplets<true, kHistT3, true>, nullptr}}};
    return hm >= 0 && hm < 3 ? kfns[all_unit ? 0 : 1][mixed ? 1 : 0][hm] : nullptr;
}

// Queue segments, one per XCD (workgroup b runs on XCD b % 8): each XCD works
// through its own run of cell-ordered work items, so the neighbourhoods its
// primaries scan stay in its L2.  The item list is cut at equal estimated work
// (primaries x |c2| x |c3|, plus primaries x |c3|^2 / 2 for a partner term);
// drained segments are helped with by the other XCDs.  seg: nseg + 1 bounds.
int queue_segments(const std::vector<Chunk> &chunks, const std::vector<Term> &terms, const std::vector<int> &partner,
                   int grid, std::vector<int> &seg) {
    const int nseg = std::max(1, std::min(kMaxSeg, grid));
    seg.assign(kMaxSeg + 1, 0);
    std::vector<double> cw(chunks.size());
    double wtot = 0.0;
    for (size_t c = 0; c < chunks.size(); ++c) {
        const Term &t = terms[chunks[c].term];
        const double n3 = (double)t.c3->n;
        cw[c] = (double)(chunks[c].end - chunks[c].begin) * ((double)t.c2->n * n3 + (partner[chunks[c].term] >= 0 ? 0.5 * n3 * n3 : 0.0));
        wtot += cw[c];
    }
    int sgi = 1;
    double acc = 0.0;
    for (size_t c = 0; c < chunks.size() && sgi < nseg; ++c) {
        acc += cw[c];
        while (sgi < nseg && acc >= wtot * sgi / nseg) seg[sgi++] = (int)c + 1;
    }
    while (sgi <= nseg) seg[sgi++] = (int)chunks.size();
    return nseg;
}

// A launch (or the rounds of one, rppi key stream): enough to enqueue it
// again under graph capture.
struct Round {
    int first, n;       // work items [first, first + n)
    int item0, nitems;  // binning work items
};
struct Launch {
    KArgs A;
    const void *kfn = nullptr;
    int grid = 0, block = 0;
    size_t lds_bytes = 0;
    bool stream = false;
    std::vector<Round> rounds;
    int rs_nbk = 0;
    size_t bin_lds = 0;
    int nterms = 0;
    long long ostride = 0;  // u64 words of out per term
};

// The launch arguments of round r of the key stream.
KArgs round_args(const nc_context *ctx, const CountBufs &bufs, const Launch &L, int r) {
    KArgs Ar = L.A;
    const Round &rd = L.rounds[r];
    Ar.chunks = (const Chunk *)bufs.chunks.p + rd.first;
    Ar.nchunks = rd.n;
    Ar.rs_base = (const unsigned long long *)ctx->rs_base.p + (size_t)r * (L.rs_nbk + 1);
    Ar.rs_cursor = (unsigned long long *)ctx->rs_cursor.p + (size_t)r * L.rs_nbk;
    Ar.rs_items = (const RBItem *)ctx->rs_items.p + rd.item0;
    Ar.rs_nitems = rd.nitems;
    return Ar;
}

int enqueue_count(nc_context *ctx, const CountBufs &bufs, const Launch &L) {
    hipLaunchKernelGGL(k_fill_u64, dim3(64), dim3(256), 0, ctx->stream, (unsigned long long *)bufs.out.p, 0ull,
                       (long long)L.nterms * L.ostride);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t *)bufs.counter.p, 0u, 16LL);
    if (L.A.nchunks > 0 && !L.stream) {
        KArgs A = L.A;
        void *args[] = {&A};
        if (hipLaunchKernel(L.kfn, dim3(L.grid), dim3(L.block), args, L.lds_bytes, ctx->stream) != hipSuccess)
            return NC_ERR_HIP;
    }
    if (L.A.nchunks > 0 && L.stream) {
        hipLaunchKernelGGL(k_fill_u64, dim3(8), dim3(256), 0, ctx->stream, (unsigned long long *)ctx->rs_cursor.p,
                           0ull, (long long)L.rounds.size() * L.rs_nbk);
        for (int r = 0; r < (int)L.rounds.size(); ++r) {
            KArgs Ar = round_args(ctx, bufs, L, r);
            void *args[] = {&Ar};
            if (r > 0)
                hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t *)bufs.counter.p, 0u, 16LL);
            if (hipLaunchKernel(L.kfn, dim3(L.grid), dim3(L.block), args, L.lds_bytes, ctx->stream) != hipSuccess)
                return NC_ERR_HIP;
            if (Ar.rs_nitems > 0 && hipLaunchKernel((const void *)k_rppi_bin, dim3(Ar.rs_nitems), dim3(1024), args,
                                                    L.bin_lds, ctx->stream) != hipSuccess)
                return NC_ERR_HIP;
        }
    }
    return hipGetLastError() == hipSuccess ? NC_OK : NC_ERR_HIP;
}

// rppi key stream: tables, rounds, bucket regions, binning work items.  A dry
// run over every S-th work item counts the keys per bucket and per work item
// (kept for the same catalogues and params).  The work queue is cut into
// rounds of consecutive work items whose keys fit 60% of the free device
// memory; a round's regions get each bucket's keys in proportion to its
// term's share of the round, plus the runs' slack.  Keys beyond a region's
// capacity still count, through direct atomics.
struct RegionPlan {
    std::vector<unsigned long long> caps, rcaps;  // wanted keys per bucket; allocated per (round, bucket)
    std::vector<double> cest;                     // estimated keys per work item
    std::vector<double> ckey;                     // key of the kept sizes
};
int rppi_regions(nc_context *ctx, CountBufs &bufs, const RppiStream &rs, const std::vector<Term> &terms,
                 const PlanKey &key, const std::vector<Chunk> &chunks, const std::vector<int> &weighted,
                 long long nwaves, Launch &L, RegionPlan &R) {
    KArgs &A = L.A;
    const int nterms = (int)terms.size();
    const int rs_nbk = rs.nbk;
    const size_t tb_rank = al16(sizeof(int4) * (size_t)nterms);
    const size_t tb_inv = al16(tb_rank + sizeof(uint16_t) * rs.rank.size());
    const size_t tb_bytes = al16(tb_inv + sizeof(int) * rs.inv.size());
    std::vector<char> tb(tb_bytes);
    std::memcpy(tb.data(), rs.term.data(), sizeof(int4) * (size_t)nterms);
    std::memcpy(tb.data() + tb_rank, rs.rank.data(), sizeof(uint16_t) * rs.rank.size());
    std::memcpy(tb.data() + tb_inv, rs.inv.data(), sizeof(int) * rs.inv.size());
    NC_HIP(ctx, dgrow(ctx->rs_tab, tb_bytes));
    NC_HIP(ctx, hipMemcpyAsync(ctx->rs_tab.p, tb.data(), tb_bytes, hipMemcpyHostToDevice, ctx->stream));
    A.rs_term = (const int4 *)ctx->rs_tab.p;
    A.rs_rank = (const uint16_t *)((char *)ctx->rs_tab.p + tb_rank);
    A.rs_inv = (const int *)((char *)ctx->rs_tab.p + tb_inv);
    const size_t nch = chunks.size();
    R.ckey = key.pkey;
    for (uint64_t u : key.uids) R.ckey.push_back((double)u);
    for (int u : key.wkey) R.ckey.push_back((double)u);
    R.ckey.push_back((double)ctx->chunk);  // the work items the estimates refer to
    if (ctx->sample_stride == 1 && ctx->rs_key == R.ckey && ctx->rs_caps.size() == (size_t)rs_nbk &&
        ctx->rs_cest.size() == nch) {
        R.caps = ctx->rs_caps;
        R.cest = ctx->rs_cest;
    } else {
        const size_t S = std::max<size_t>(1, std::min<size_t>(32, nch / 256));
        std::vector<Chunk> sample;
        for (size_t i = 0; i < nch; i += S) sample.push_back(chunks[i]);
        NC_HIP(ctx, dgrow(ctx->rs_sample, sizeof(Chunk) * sample.size()));
        NC_HIP(ctx, dgrow(ctx->rs_cursor, sizeof(unsigned long long) * (size_t)rs_nbk));
        NC_HIP(ctx, dgrow(ctx->rs_chunk_keys, sizeof(unsigned long long) * sample.size()));
        NC_HIP(ctx, hipMemcpyAsync(ctx->rs_sample.p, sample.data(), sizeof(Chunk) * sample.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
        KArgs As = A;
        As.chunks = (const Chunk *)ctx->rs_sample.p;
        As.nchunks = (int)sample.size();
        As.rs_dry = 1;
        As.rs_cursor = (unsigned long long *)ctx->rs_cursor.p;
        As.rs_chunk_keys = (unsigned long long *)ctx->rs_chunk_keys.p;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t *)bufs.counter.p, 0u, 16LL);
        hipLaunchKernelGGL(k_fill_u64, dim3(8), dim3(256), 0, ctx->stream, As.rs_cursor, 0ull, (long long)rs_nbk);
        hipLaunchKernelGGL(k_fill_u64, dim3(64), dim3(256), 0, ctx->stream, As.rs_chunk_keys, 0ull,
                           (long long)sample.size());
        void *args[] = {&As};
        NC_HIP(ctx, hipLaunchKernel(L.kfn, dim3(L.grid), dim3(L.block), args, L.lds_bytes, ctx->stream));
        std::vector<unsigned long long> n(rs_nbk), ck(sample.size());
        NC_HIP(ctx, hipMemcpyAsync(n.data(), As.rs_cursor, sizeof(unsigned long long) * n.size(), hipMemcpyDeviceToHost,
                                   ctx->stream));
        NC_HIP(ctx, hipMemcpyAsync(ck.data(), As.rs_chunk_keys, sizeof(unsigned long long) * ck.size(),
                                   hipMemcpyDeviceToHost, ctx->stream));
        NC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        R.caps.resize(rs_nbk);
        for (int b = 0; b < rs_nbk; ++b)
            R.caps[b] = S == 1 ? n[b] + 64 : (unsigned long long)((double)n[b] * (double)S * 1.3) + 4096;
        R.cest.resize(nch);
        for (size_t i = 0; i < nch; ++i) R.cest[i] = (double)ck[i / S] * (S == 1 ? 1.0 : 1.3);
    }
    // memory: 60% of the free device memory, counting the context's own
    // region buffers; the runs' slack (every bucket, every wave) first
    size_t freeb = 0, totalb = 0;
    NC_HIP(ctx, hipMemGetInfo(&freeb, &totalb));
    const double per = rs.anyw ? 12.0 : 4.0;
    const double room = 0.6 * ((double)freeb + (double)ctx->rs_keys.bytes + (double)ctx->rs_vals.bytes) / per;
    double need = 0.0;
    for (unsigned long long c : R.caps) need += (double)c;
    int blk = 64;  // a wave reserves blk slots of a region at a time
    while (blk < 4096 && 2.0 * blk <= need / std::max(1, rs_nbk) / (4.0 * (double)nwaves)) blk *= 2;
    while (blk > 64 && (double)rs_nbk * (double)nwaves * blk > 0.25 * room) blk /= 2;
    A.rs_blk = blk;
    const unsigned long long slack = (unsigned long long)nwaves * (unsigned long long)blk;
    double K = std::max(0.0, room - (double)rs_nbk * (double)slack);  // keys per round
    double total_est = 0.0;
    for (double e : R.cest) total_est += e;
    K = std::max({K, 1e7, total_est / 256.0});  // at most ~256 rounds (past the room: allocation fails)
    if (ctx->rs_round_keys > 0.0) K = std::max(ctx->rs_round_keys, total_est / 256.0);
    // per term: estimated keys, and work items (for terms the sample saw no key of)
    std::vector<double> test(nterms, 0.0), tcnt(nterms, 0.0);
    for (size_t i = 0; i < nch; ++i) {
        test[chunks[i].term] += R.cest[i];
        tcnt[chunks[i].term] += 1.0;
    }
    std::vector<Round> &rounds = L.rounds;
    for (size_t i = 0; i < nch;) {
        double acc = 0.0;
        size_t j = i;
        while (j < nch && (j == i || acc + R.cest[j] <= K)) acc += R.cest[j++];
        rounds.push_back({(int)i, (int)(j - i), 0, 0});
        i = j;
    }
    const int NR = (int)rounds.size();
    R.rcaps.assign((size_t)NR * rs_nbk, 0ull);
    std::vector<unsigned long long> base((size_t)NR * (rs_nbk + 1), 0ull);
    unsigned long long kmax = 0;
    std::vector<RBItem> items;
    for (int r = 0; r < NR; ++r) {
        std::vector<double> fe(nterms, 0.0), fc(nterms, 0.0);
        for (int i = rounds[r].first; i < rounds[r].first + rounds[r].n; ++i) {
            fe[chunks[i].term] += R.cest[i];
            fc[chunks[i].term] += 1.0;
        }
        double want = 0.0;
        std::vector<double> w(rs_nbk, 0.0);
        std::vector<char> live(rs_nbk, 0);  // buckets of the terms with work in this round
        for (int t = 0; t < nterms; ++t) {
            if (fc[t] == 0.0) continue;
            const double f = test[t] > 0.0 ? fe[t] / test[t] : fc[t] / tcnt[t];
            for (int bb = 0; bb < rs.term[t].y; ++bb) {
                w[rs.term[t].x + bb] = (double)R.caps[rs.term[t].x + bb] * f * ctx->rs_scale;
                live[rs.term[t].x + bb] = 1;
                want += w[rs.term[t].x + bb];
            }
        }
        const double sc = want > K ? K / want : 1.0;  // one work item past the room: overflow
        unsigned long long *bs = base.data() + (size_t)r * (rs_nbk + 1);
        for (int b = 0; b < rs_nbk; ++b) {
            const unsigned long long c = live[b] ? (((unsigned long long)(w[b] * sc) + slack + 63) & ~63ull) : 0ull;
            R.rcaps[(size_t)r * rs_nbk + b] = c;
            bs[b + 1] = bs[b] + c;
        }
        kmax = std::max(kmax, bs[rs_nbk]);
        // binning work items: slices in proportion to the region sizes
        // (about four per CU in all), each under 2^30 keys
        rounds[r].item0 = (int)items.size();
        const double total = std::max(1.0, (double)bs[rs_nbk]);
        for (int t = 0; t < nterms; ++t) {
            if (fc[t] == 0.0) continue;
            const int4 tr = rs.term[t];
            for (int bb = 0; bb < tr.y; ++bb) {
                const int b = tr.x + bb;
                const unsigned long long cb = bs[b + 1] - bs[b];
                if (cb == 0) continue;
                long long sl = std::llround(4.0 * ctx->ncu * (double)cb / total);
                sl = std::max<long long>({1LL, sl, (long long)((cb + (1ull << 30) - 1) >> 30)});
                for (long long j = 0; j < sl; ++j)
                    items.push_back({b, t, tr.z, weighted[t] ? 1 : 0, bb << tr.z, (int)j, (int)sl, 0});
            }
        }
        rounds[r].nitems = (int)items.size() - rounds[r].item0;
    }
    NC_HIP(ctx, dgrow(ctx->rs_base, sizeof(unsigned long long) * base.size()));
    NC_HIP(ctx, dgrow(ctx->rs_cursor, sizeof(unsigned long long) * (size_t)NR * rs_nbk));
    NC_HIP(ctx, hipMemcpyAsync(ctx->rs_base.p, base.data(), sizeof(unsigned long long) * base.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    const size_t kbytes = sizeof(unsigned) * (size_t)kmax;
    const size_t vbytes = rs.anyw ? sizeof(double) * (size_t)kmax : 0;
    if (kbytes > ctx->rs_keys.bytes || vbytes > ctx->rs_vals.bytes || vbytes == 0) {
        // release both before growing either: the room above counted them
        NC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (kbytes > ctx->rs_keys.bytes) dfree(ctx->rs_keys);
        if (vbytes > ctx->rs_vals.bytes || vbytes == 0) dfree(ctx->rs_vals);
    }
    NC_HIP(ctx, dgrow(ctx->rs_keys, kbytes));
    if (rs.anyw) NC_HIP(ctx, dgrow(ctx->rs_vals, vbytes));
    A.rs_keys = (unsigned *)ctx->rs_keys.p;
    A.rs_vals = rs.anyw ? (double *)ctx->rs_vals.p : nullptr;
    NC_HIP(ctx, dgrow(ctx->rs_items, sizeof(RBItem) * std::max<size_t>(1, items.size())));
    if (!items.empty())
        NC_HIP(ctx, hipMemcpyAsync(ctx->rs_items.p, items.data(), sizeof(RBItem) * items.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
    return NC_OK;
}

// After a key-stream count: keep the region sizes; after an overflow
// (counted, but through direct atomics) scale the buckets that overflowed by
// what this run reserved.  Returns whether some bucket overflowed.
bool keep_region_sizes(nc_context *ctx, const Launch &L, RegionPlan &R, int &rc) {
    const int nrounds = (int)L.rounds.size(), rs_nbk = L.rs_nbk;
    std::vector<unsigned long long> n((size_t)nrounds * rs_nbk);
    rc = NC_OK;
    const hipError_t e =
        hipMemcpy(n.data(), ctx->rs_cursor.p, sizeof(unsigned long long) * n.size(), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        rc = fail(ctx, NC_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
        return false;
    }
    bool overflow = false;
    for (int b = 0; b < rs_nbk; ++b) {
        double grow = 1.0;
        for (int r = 0; r < nrounds; ++r) {  // n: slots reserved, at least the keys
            const unsigned long long got = n[(size_t)r * rs_nbk + b], had = R.rcaps[(size_t)r * rs_nbk + b];
            if (got > had) grow = std::max(grow, (double)got / (double)std::max(had, 1ull));
        }
        if (grow > 1.0) {
            overflow = true;
            R.caps[b] = (unsigned long long)((double)R.caps[b] * grow * 1.125) + 64;
        }
    }
    if (ctx->sample_stride == 1) {  // not from an autotuning sample
        ctx->rs_key = R.ckey;
        ctx->rs_caps = R.caps;
        ctx->rs_cest = R.cest;
    }
    return overflow;
}

// Capture the launch sequence for replay (builds, then count); any failure
// leaves the eager result valid and no plan cached.
void capture_plan(nc_context *ctx, Mem &m, CountBufs &bufs, const Geometry &G, const std::vector<CatGrid> &grids,
                  const Launch &L, BatchPlan p) {
    bool ok = hipHostMalloc((void **)&p.h_out,
                            sizeof(unsigned long long) * std::max<size_t>(1, (size_t)L.nterms * L.ostride)) ==
                  hipSuccess &&
              hipHostMalloc((void **)&p.h_err, 16) == hipSuccess;
    for (int part = 0; part < 2 && ok; ++part) {
        hipGraph_t g = nullptr;
        ok = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        if (!ok) break;
        const int erc = part == 0 ? enqueue_builds(ctx, m, bufs, G, grids) : enqueue_count(ctx, bufs, L);
        ok = hipStreamEndCapture(ctx->stream, &g) == hipSuccess && erc == NC_OK && g != nullptr;
        if (ok)
            ok = hipGraphInstantiate(part == 0 ? &p.build_exec : &p.count_exec, g, nullptr, nullptr, 0) == hipSuccess;
        if (g) (void)hipGraphDestroy(g);
    }
    (void)hipGetLastError();
    p.valid = ok;
    ctx->plan = p;
    if (!ok) plan_drop(ctx);
}

}  // namespace

int run_count(nc_context *ctx, const nc_params *P, const std::vector<Term> &terms, int path, double *hist,
              uint64_t *counts, bool use_graph, Mem &m) {
    if (!ctx) return NC_ERR_ARG;
    // buffers: the context's own, or (arena) fresh ones carved from it; an
    // arena count is never captured
    CountBufs arena_bufs;
    CountBufs &bufs = m.arena() ? arena_bufs : ctx->buf;
    if (m.arena()) use_graph = false;
    if (!P || !hist) return fail(ctx, NC_ERR_ARG, "null params or output");
    if (path != NC_PATH_PLAIN && path != NC_PATH_MIXED) return fail(ctx, NC_ERR_ARG, "unknown path");
    for (const Term &t : terms)
        if (!t.c1 || !t.c2 || !t.c3 || t.c1->ctx != ctx || t.c2->ctx != ctx || t.c3->ctx != ctx)
            return fail(ctx, NC_ERR_ARG, "null catalogue or catalogue of another context");
    Geometry G;
    int rc = make_geometry(ctx, P, G);
    if (rc) return rc;
    NC_HIP(ctx, hipSetDevice(ctx->device));

    // --- graph plan: replay while the key comes back
    PlanKey key;
    Accum acc;
    rc = plan_key(ctx, P, terms, path, G, key, acc);
    if (rc) return rc;
    if (use_graph && ctx->plan.valid && ctx->plan.uids == key.uids && ctx->plan.wkey == key.wkey &&
        ctx->plan.pkey == key.pkey && ctx->plan.tkey == key.tkey) {
        bool retry = false;
        rc = plan_replay(ctx, hist, counts, retry);
        if (rc || !retry) return rc;
    }
    plan_drop(ctx);
    const int nterms = (int)terms.size();
    const int W = ctx->waves_per_group;
    const int block = W * kWave;
    // the mixed lookup table must leave LDS room for the smallest layout
    // (ctab) and 64-entry chunks, else the count runs the plain path
    const size_t mixed_min_lds =
        al16(sizeof(uint2) * G.lut32.size()) + al16((size_t)G.nbc * 8) + al16(sizeof(int) * G.ctab.size()) +
        (size_t)W * al16(68 * (sizeof(float4) + sizeof(int) + sizeof(double)) + kQueueCap * sizeof(uint2));
    const bool mixed = path == NC_PATH_MIXED && G.mixed_ok && mixed_min_lds <= ctx->lds_hw;
    const int lmode = G.rppi ? (G.survey ? kListRppiSurvey : kListRppiBox) : mixed ? kListMixed : kListPlain;
    RppiStream rs;
    rppi_stream_tables(ctx, G, terms, rs);
    LdsLayout lay{G, mixed};
    lay.rank_bytes = rs.on ? al16(sizeof(uint16_t) * rs.rank.size()) : 0;
    lay.stage_bytes = rs.on ? al16(stream_wave_bytes(rs.anyw, rs.nbm)) : 0;

    // --- tables and small buffers
    const TableImage tabs = table_image(G);
    NC_MEM(ctx, m, bufs.tables, tabs.bytes);
    NC_HIP(ctx, hipMemcpyAsync(bufs.tables.p, tabs.data.data(), tabs.bytes, hipMemcpyHostToDevice, ctx->stream));
    NC_MEM(ctx, m, bufs.errw, 16);
    // 16 queue counters (zeroed before every launch), then the queue's
    // kMaxSeg + 1 segment bounds
    NC_MEM(ctx, m, bufs.counter, sizeof(unsigned) * (16 + kMaxSeg + 1));
    NC_MEM(ctx, m, bufs.bound, 16);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, ctx->stream, (uint32_t *)bufs.bound.p, 0u, 4LL);
    for (const Term &t : terms)
        for (nc_catalog *c : {t.c1, t.c2, t.c3})
            if (c->n > 2147483647LL) return fail(ctx, NC_ERR_LIMIT, "catalogues are limited to 2^31-1 points");

    // --- cell lists, one per distinct catalogue, and the neighbourhood bound
    NC_HIP(ctx, hipEventRecord(ctx->ev[0], ctx->stream));
    std::vector<CatGrid> grids = collect_grids(terms);
    NC_MEM(ctx, m, bufs.cflags, sizeof(unsigned) * grids.size());
    rc = enqueue_builds(ctx, m, bufs, G, grids);
    if (rc) return rc;
    NC_HIP(ctx, hipEventRecord(ctx->ev[1], ctx->stream));
    GridGeo gg;
    for (int a = 0; a < 3; ++a) {
        gg.cs[a] = G.cs[a];
        gg.nc[a] = G.nc[a];
    }
    gg.survey = G.survey;
    const int cb = (int)((G.ncells + 255) / 256);
    for (const Term &t : terms) {
        hipLaunchKernelGGL(k_nbr_bound, dim3(cb), dim3(256), 0, ctx->stream, (const uint32_t *)t.c1->start.p,
                           (const uint32_t *)t.c2->start.p, gg, (unsigned *)bufs.bound.p);
        if (t.c3 != t.c2)
            hipLaunchKernelGGL(k_nbr_bound, dim3(cb), dim3(256), 0, ctx->stream, (const uint32_t *)t.c1->start.p,
                               (const uint32_t *)t.c3->start.p, gg, (unsigned *)bufs.bound.p);
    }
    NC_HIP(ctx, hipGetLastError());
    unsigned host_flags[2] = {0, 0};
    std::vector<unsigned> cflags(grids.size());
    NC_HIP(ctx, hipMemcpyAsync(&host_flags[0], bufs.errw.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    NC_HIP(ctx, hipMemcpyAsync(&host_flags[1], bufs.bound.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    NC_HIP(ctx, hipMemcpyAsync(cflags.data(), bufs.cflags.p, sizeof(unsigned) * cflags.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    NC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (host_flags[0] & kErrDomain) return fail(ctx, NC_ERR_DOMAIN, "a point lies outside the cell grid of the box");
    const int cap = (int)((std::max(1u, host_flags[1]) + 15u) & ~15u);
    for (size_t gi = 0; gi < grids.size(); ++gi) grids[gi].cat->exact = cflags[gi] == 0u;
    const Symmetry sym = plan_symmetry(ctx, G, terms, acc.all_unit, grids);

    // --- histogram layout: dense u32 counts when every term has unit weights,
    // they fit and no count can wrap between merges (one chunk adds at most
    // chunk * cap^2 pairs); else the slot table; else ctab.  The plain path
    // needs the single-step bin lookup for either of the first two.
    // The budget is first half the CU's LDS (two workgroups per CU); a layout
    // that only fits the whole of it runs one workgroup per CU.
    auto fits = [&](int hm, size_t budget) { return lay.fits(hm, W, budget); };
    // Longest list the launch may meet: the 27-cell candidate bound, or, when
    // only the exact lengths can admit the dense layout, the longest list
    // k_list_bound finds (one extra pass over the primaries).
    // A primary adds at most n(n-1) to the counts (n: its list length), three
    // times that on the six-fold path, twice on the pair-swap path.
    const double kWrap = 2147483648.0;
    const double mult = sym.ntri > 0 ? 3.0 : sym.nswap > 0 ? 2.0 : 1.0;
    const bool dense_cand = !G.rppi && (mixed || G.lut_single_step) && acc.all_unit;
    int nmax = cap;
    // (the mixed path takes the dense layout over a whole CU, the plain path
    // only within half of one)
    if (dense_cand && (double)ctx->chunk * mult * (double)cap * (double)cap >= kWrap &&
        fits(kHistDense, mixed ? ctx->lds_hw : ctx->lds_max)) {
        rc = longest_list(ctx, bufs, G, terms, key.uids, cap, nmax);
        if (rc) return rc;
    }
    // primaries per work item: the tuned chunk, cut where the dense layout
    // needs it so that one work item cannot add 2^31 (merges follow work items)
    int qchunk = ctx->chunk;
    const double per_primary = mult * (double)nmax * (double)nmax;
    const bool dense_ok = dense_cand && per_primary < kWrap;
    if (dense_ok && (double)qchunk * per_primary >= kWrap) qchunk = std::max(1, (int)((kWrap - 1.0) / per_primary));
    auto choose = [&](size_t budget) {
        if (G.rppi) return kHistGlobal;
        if (dense_ok && fits(kHistDense, budget)) return kHistDense;
        if ((mixed || G.lut_single_step) && !G.t3.empty() && fits(kHistT3, budget)) return kHistT3;
        return kHistCtab;
    };
    size_t budget = ctx->lds_max;
    int hm = choose(budget);
    if (acc.exact) {
        if (G.t3.empty() || !(mixed || G.lut_single_step))
            return fail(ctx, NC_ERR_LIMIT, "exact accumulation supports at most 35 bins per side");
        hm = kHistExact;
        if (!fits(hm, budget)) budget = ctx->lds_hw;
    } else if (acc.f32) {
        if (G.t3.empty() || !(mixed || G.lut_single_step))
            return fail(ctx, NC_ERR_LIMIT, "FP32 accumulation supports at most 35 bins per side");
        // the t3 table where it fits beside the FP32 sums, else ctab
        hm = fits(kHistF32, budget) || (!fits(kHistF32C, budget) && fits(kHistF32, ctx->lds_hw)) ? kHistF32
                                                                                                 : kHistF32C;
        if (!fits(hm, budget)) budget = ctx->lds_hw;
    } else if (mixed && dense_ok && hm != kHistDense && fits(kHistDense, ctx->lds_hw)) {
        // mixed path: dense counts over a whole CU beat the sorted-slot
        // layouts at half of it (25 bins, survey randoms: 4405 against 5428 ms)
        budget = ctx->lds_hw;
        hm = kHistDense;
    } else if (!fits(hm, budget) && ctx->lds_hw > budget) {
        budget = ctx->lds_hw;
        hm = choose(budget);
    }
    if (rs.on && !fits(hm, budget)) {  // no room for the key stream: direct atomics
        // (the autotuner skips such shapes rather than time the other path)
        if (ctx->autotuning) return fail(ctx, NC_ERR_LIMIT, "no LDS room for the rppi key stream");
        rs.on = false;
        lay.rank_bytes = lay.stage_bytes = 0;
        budget = fits(hm, ctx->lds_max) ? ctx->lds_max : ctx->lds_hw;
    }
    // Partner terms: (c1, c3, c3) beside (c1, c2, c3), c2 != c3 (DRR beside
    // DDR): the primaries and the list of c3 are the same, so the work items
    // of the first count both, into two histograms; the partner gets no work
    // items.  Sorted-slot layouts only, where the second histogram fits.
    std::vector<int> partner(nterms, -1), absorbed(nterms, 0);
    if (ctx->fuse_terms && !G.rppi && primw_ok(hm)) {
        bool any = false;
        for (int i = 0; i < nterms; ++i) {
            const Term &a = terms[i];
            if (absorbed[i] || a.c2 == a.c3 || sym.tri[i] == kTriSix || a.c2->n == 0 || a.c3->n == 0) continue;
            for (int j = 0; j < nterms; ++j) {
                const Term &b = terms[j];
                if (j == i || absorbed[j] || partner[j] >= 0 || sym.tri[j] != kTriNone) continue;
                if (b.c1 == a.c1 && b.c2 == a.c3 && b.c3 == a.c3) {
                    partner[i] = j;
                    absorbed[j] = 1;
                    any = true;
                    break;
                }
            }
        }
        if (any) {
            lay.nhist = 2;
            if (!fits(hm, budget)) {
                lay.nhist = 1;
                partner.assign(nterms, -1);
                absorbed.assign(nterms, 0);
            }
        }
    }
    int nabsorbed = 0;
    for (int a : absorbed) nabsorbed += a;
    LdsOffsets off;
    const size_t hist_bytes = lay.nhist * lay.hist(hm);
    const size_t lds_tab_bytes = lay.tables(hm, off);

    // --- launch shape
    Launch L;
    L.stream = rs.on;
    L.rs_nbk = rs.nbk;
    L.bin_lds = (size_t)1 << rs.maxshift;
    L.nterms = nterms;
    L.ostride = acc.exact ? 2 * G.out_words : G.out_words;
    L.block = block;
    int C = ctx->chunk_elems;
    const size_t lds_tab = al16(hist_bytes + lds_tab_bytes);
    // chunks are whole 64-entry blocks (the circulant diagonal blocks need it)
    C = std::max(64, C - C % 64);
    while (C > 64 && lds_tab + W * lay.wave(C, hm) > budget) C -= 64;
    const size_t wave_bytes = lay.wave(C, hm);
    L.lds_bytes = lds_tab + W * wave_bytes;
    if (L.lds_bytes > budget) return fail(ctx, NC_ERR_LIMIT, "histogram does not fit in LDS");
    L.kfn = launch_kernel(G, L.stream, hm, mixed, acc.all_unit);
    if (L.kfn == nullptr) return fail(ctx, NC_ERR_LIMIT, "no kernel variant for the planned layout");
    {
        // some variants are compiled for smaller workgroups (more VGPRs per lane)
        hipFuncAttributes fa;
        NC_HIP(ctx, hipFuncGetAttributes(&fa, L.kfn));
        if (block > fa.maxThreadsPerBlock)
            return fail(ctx, NC_ERR_LIMIT, "waves_per_group " + std::to_string(W) + " exceeds the " +
                                               std::to_string(fa.maxThreadsPerBlock / kWave) +
                                               " waves this kernel variant is built for");
    }

    // --- terms and work queue
    std::vector<TermDesc> td(nterms);
    std::vector<Chunk> chunks;
    std::vector<int> weighted(nterms);
    for (int i = 0; i < nterms; ++i) {
        const Term &t = terms[i];
        TermDesc &d = td[i];
        d = term_desc(t);
        d.wprim = d.weighted && t.c2->unit_w && t.c3->unit_w && !acc.exact && !acc.f32;
        d.partner = partner[i];
        weighted[i] = d.weighted;
        if (t.c2->n == 0 || t.c3->n == 0 || absorbed[i]) continue;
        d.tri = sym.tri[i];
        d.prune2 = prunes(G, t.c2);
        d.prune3 = prunes(G, t.c3);
        int p0, p1;  // the primary slice: positions [p0, p1) of the cell-sorted catalogue
        slice_range(ctx, d.n1, p0, p1);
        for (int b = p0; b < p1; b += qchunk)
            if (((b - p0) / kSampleRun) % ctx->sample_stride == 0) chunks.push_back({i, b, std::min(p1, b + qchunk), 0});
    }
    NC_MEM(ctx, m, bufs.terms, sizeof(TermDesc) * td.size());
    NC_MEM(ctx, m, bufs.chunks, sizeof(Chunk) * std::max<size_t>(1, chunks.size()));
    NC_MEM(ctx, m, bufs.out, sizeof(unsigned long long) * (size_t)nterms * L.ostride);
    if (!td.empty())
        NC_HIP(ctx, hipMemcpyAsync(bufs.terms.p, td.data(), sizeof(TermDesc) * td.size(), hipMemcpyHostToDevice,
                                   ctx->stream));
    if (!chunks.empty())
        NC_HIP(ctx, hipMemcpyAsync(bufs.chunks.p, chunks.data(), sizeof(Chunk) * chunks.size(), hipMemcpyHostToDevice,
                                   ctx->stream));

    // --- persistent grid and the per-wave list scratch (on an arena: as many
    // workgroups as the rest of it holds lists for)
    int per_cu = ctx->wgs_per_cu;
    if (per_cu <= 0) {
        int occ = 0;
        NC_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, L.kfn, block, L.lds_bytes));
        per_cu = std::max(1, occ);
    }
    L.grid = ctx->ncu * per_cu;
    const long long slot_bytes = (long long)((2 * (size_t)cap * list_entry_bytes(lmode) + 255) & ~(size_t)255);
    if (m.arena()) {
        const long long fit = (long long)(m.left() / (size_t)slot_bytes) / W;
        if (fit < 1)
            return fail(ctx, NC_ERR_LIMIT, "the neighbour lists of one workgroup do not fit in the no-alloc scratch (" +
                                               std::to_string(m.size >> 20) + " MiB)");
        L.grid = (int)std::min<long long>(L.grid, fit);
    }
    const long long nwaves = (long long)L.grid * W;
    if ((double)nwaves * (double)slot_bytes > 64.0 * (1ULL << 30))
        return fail(ctx, NC_ERR_LIMIT, "neighbour lists need more than 64 GiB of scratch");
    NC_MEM(ctx, m, bufs.scratch, (size_t)(nwaves * slot_bytes));

    // --- launch arguments
    KArgs &A = L.A;
    std::memset(&A, 0, sizeof(A));
    A.terms = (const TermDesc *)bufs.terms.p;
    A.chunks = (const Chunk *)bufs.chunks.p;
    A.nchunks = (int)chunks.size();
    A.counter = (unsigned *)bufs.counter.p;
    std::vector<int> seg;  // host copy, kept (like chunks) until the copy has run
    A.nseg = queue_segments(chunks, terms, partner, L.grid, seg);
    A.seg = (const int *)(A.counter + 16);
    NC_HIP(ctx, hipMemcpyAsync((void *)A.seg, seg.data(), sizeof(int) * (size_t)(A.nseg + 1), hipMemcpyHostToDevice,
                               ctx->stream));
    A.err = (unsigned *)bufs.errw.p;
    A.out = (unsigned long long *)bufs.out.p;
    A.scratch = (char *)bufs.scratch.p;
    A.slot_bytes = slot_bytes;
    A.cap = cap;
    A.nmax = hm == kHistDense ? nmax : cap;
    A.chunk_elems = C;
    for (int a = 0; a < 3; ++a) {
        A.box[a] = G.box[a];
        A.cs[a] = G.cs[a];
        A.nc[a] = G.nc[a];
    }
    A.survey = G.survey;
    A.r2min = G.sq[0];
    A.r2max = G.sq[G.nb];
    A.prune_r2 = prune_radius2(G);
    A.nb = G.nb;
    A.nbc = (int)G.out_words;
    A.npi = G.npi;
    A.pimax2 = (double)G.npi * (double)G.npi;
    A.smax2 = G.sq[G.nb] + A.pimax2;  // rpMax*rpMax + piMax*piMax, cudaNcorr.cu:186
    A.ori_x = G.ori_x;
    A.lut_scale = G.lut_scale;
    A.lut_m = G.lut_m;
    A.sq = (const double *)((char *)bufs.tables.p + tabs.sq);
    A.lut = (const uint16_t *)((char *)bufs.tables.p + tabs.lut);
    A.ctab = (const int *)((char *)bufs.tables.p + tabs.ctab);
    A.lut32 = (const uint2 *)((char *)bufs.tables.p + tabs.lut32);
    A.t3 = (const uint16_t *)((char *)bufs.tables.p + tabs.t3);
    A.sigma32 = G.sigma32;
    A.lut32_lo = G.lut32_lo;
    A.lut32_shift = G.lut32_shift;
    A.lut32_c0 = G.lut32_c0;
    A.lut32_n = (int)G.lut32.size();
    A.g32 = G.g32;
    A.tri_t = G.tri_t;
    const size_t nb3 = (size_t)G.nb * G.nb * G.nb;
    A.nslots = hm == kHistDense    ? (int)((nb3 + 1) / 2)
               : hm == kHistGlobal ? 0
               : hm == kHistExact  ? 2 * G.nbc
               : is_f32(hm)        ? (int)(lay.hist(hm) / 8)
                                   : G.nbc;
    A.out_stride = (int)L.ostride;
    A.xscale = acc.exact ? std::ldexp(1.0, acc.xk) : 0.0;
    // FP32 accumulation: merged at the first work-item boundary after 2^22
    // pairs, which bounds the FP64 rounding of the compensations (see the
    // header) far below the FP32 products' own 2^-46
    A.flush_pairs = is_f32(hm) ? std::min<unsigned long long>(ctx->dense_flush, 1ull << 22) : ctx->dense_flush;
    A.off_sq = (int)hist_bytes;
    A.off_lut = (int)(hist_bytes + off.lut);
    A.off_ctab = (int)(hist_bytes + off.ctab);
    A.off_lut32 = (int)(hist_bytes + off.lut32);
    A.off_t3 = (int)(hist_bytes + off.t3);
    A.off_wave = (int)lds_tab;
    A.off_hist2 = lay.nhist == 2 ? (int)lay.hist(hm) : 0;
    A.tri_last6 = ctx->tri_last && !G.survey && G.nc[0] >= 3 && G.nc[1] >= 3 && G.nc[2] >= 3;
    A.wave_bytes = (int)wave_bytes;
    A.lds_bytes = (int)L.lds_bytes;
    NC_HIP(ctx, hipEventRecord(ctx->ev[2], ctx->stream));

    // --- rppi key stream: rounds and regions
    RegionPlan R;
    if (L.stream && A.nchunks > 0) {
        A.off_rank = (int)(hist_bytes + lds_tab_bytes - lay.rank_bytes);
        A.off_stage = (int)(wave_bytes - lay.stage_bytes);
        A.rs_lds_off = (int)stream_lds_offset(rs.anyw);
        A.rs_nbm = rs.nbm;
        rc = rppi_regions(ctx, bufs, rs, terms, key, chunks, weighted, nwaves, L, R);
        if (rc) return rc;
    }

    // --- the count
    if (enqueue_count(ctx, bufs, L) != NC_OK) return fail(ctx, NC_ERR_HIP, "triplet kernel launch failed");
    NC_HIP(ctx, hipEventRecord(ctx->ev[3], ctx->stream));
    std::vector<unsigned long long> res((size_t)nterms * L.ostride);
    NC_HIP(ctx, hipMemcpyAsync(res.data(), bufs.out.p, sizeof(unsigned long long) * res.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    NC_HIP(ctx, hipMemcpyAsync(&host_flags[0], bufs.errw.p, 4, hipMemcpyDeviceToHost, ctx->stream));
    NC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms_b = 0.f, ms_c = 0.f;
    NC_HIP(ctx, hipEventElapsedTime(&ms_b, ctx->ev[0], ctx->ev[1]));
    NC_HIP(ctx, hipEventElapsedTime(&ms_c, ctx->ev[2], ctx->ev[3]));
    const int nrounds = L.stream && A.nchunks > 0 ? (int)L.rounds.size() : 0;
    const int nkernel = A.nchunks > 0 ? (L.stream ? 2 * nrounds : 1) : 0;
    ctx->timing.ms_build = ms_b;
    ctx->timing.ms_count = ms_c;
    ctx->timing.ms_total = ms_b + ms_c;
    ctx->timing.n_kernel = nkernel;
    ctx->timing.hist_layout = L.stream ? kHistStream : hm;
    ctx->timing.lds_bytes = (int32_t)L.lds_bytes;
    ctx->timing.path = mixed ? NC_PATH_MIXED : NC_PATH_PLAIN;
    ctx->timing.tri_terms = sym.ntri;
    ctx->timing.swap_terms = sym.nswap;
    ctx->timing.partner_terms = nabsorbed;
    // key stream: keep the region sizes, and capture no graph after an
    // overflow so that the next call uses the grown ones
    bool rs_overflow = false;
    if (nrounds > 0) {
        rs_overflow = keep_region_sizes(ctx, L, R, rc);
        if (rc) return rc;
    }
    if (host_flags[0] & kErrDomain) return fail(ctx, NC_ERR_DOMAIN, "a point lies outside the cell grid of the box");
    if (host_flags[0] & kErrOverflow) return fail(ctx, NC_ERR_HIP, "internal error: neighbour list overflow");
    if (host_flags[0] & kErrMode)
        return fail(ctx, NC_ERR_HIP, "internal error: a term does not match the weighting of its kernel variant");

    convert_out(G.nb, G.out_words, G.rppi, nterms, weighted, res.data(), hist, counts, acc.xk);

    if (use_graph && !rs_overflow) {
        BatchPlan p;
        p.uids = key.uids;
        p.wkey = key.wkey;
        p.pkey = key.pkey;
        p.tkey = key.tkey;
        p.weighted = weighted;
        p.nb = G.nb;
        p.nbc = G.out_words;
        p.xk = acc.xk;
        p.rppi = G.rppi;
        p.nkernel = nkernel;
        p.hist_layout = L.stream ? kHistStream : hm;
        p.lds_bytes = (int)L.lds_bytes;
        p.path = mixed ? NC_PATH_MIXED : NC_PATH_PLAIN;
        p.tri_terms = sym.ntri;
        p.swap_terms = sym.nswap;
        p.partner_terms = nabsorbed;
        capture_plan(ctx, m, bufs, G, grids, L, p);
    }
    return NC_OK;
}

// Device address of noalloc_scratch on `device` (current): looked up once
// per device with hipGetSymbolAddress and kept; each GPU holds its own copy.
static std::mutex g_noalloc_mu;
static std::vector<char *> g_noalloc_ptr;
hipError_t noalloc_base(int device, char **out) {
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if ((size_t)device >= g_noalloc_ptr.size()) g_noalloc_ptr.resize((size_t)device + 1, nullptr);
    if (!g_noalloc_ptr[(size_t)device]) {
        void *p = nullptr;
        const hipError_t e = noalloc_symbol_address(&p);
        if (e != hipSuccess) return e;
        g_noalloc_ptr[(size_t)device] = static_cast<char *>(p);
    }
    *out = g_noalloc_ptr[(size_t)device];
    return hipSuccess;
}

// The points of a catalogue that wraps caller-owned device memory
// (nc_catalog_wrap_device) may be overwritten between counts, weights
// included.  Whether all weights are 1 decides the histogram layout, the
// batch split and whether integer counts are returned, so it is derived again
// from the current contents before every count over such catalogues (one
// read of their weights and one synchronisation).  A captured plan keys on
// these flags, so a changed weight status is never replayed.
int refresh_weights(nc_context *ctx, Mem &m, const std::vector<Term> &terms) {
    std::vector<nc_catalog *> w;
    for (const Term &t : terms)
        for (nc_catalog *c : {t.c1, t.c2, t.c3})
            if (c && !c->owns && c->n > 0 && std::find(w.begin(), w.end(), c) == w.end()) w.push_back(c);
    if (w.empty()) return NC_OK;
    NC_HIP(ctx, hipSetDevice(ctx->device));
    DBuf arena_flags;
    DBuf &fl = m.arena() ? arena_flags : ctx->buf.wflags;
    NC_MEM(ctx, m, fl, 4 * sizeof(unsigned) * w.size());
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, ctx->stream, (uint32_t *)fl.p, 0u,
                       4LL * (long long)w.size());
    for (size_t i = 0; i < w.size(); ++i) {
        const int grid = (int)std::min<long long>((w[i]->n + 255) / 256, 1024);
        hipLaunchKernelGGL(k_weight_stats, dim3(grid), dim3(256), 0, ctx->stream, (const double4 *)w[i]->d_raw,
                           (long long)w[i]->n, (unsigned *)fl.p + 4 * i);
    }
    NC_HIP(ctx, hipGetLastError());
    std::vector<unsigned> f(4 * w.size());
    NC_HIP(ctx, hipMemcpyAsync(f.data(), fl.p, sizeof(unsigned) * f.size(), hipMemcpyDeviceToHost,
                               ctx->stream));
    NC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < w.size(); ++i) {
        w[i]->unit_w = f[4 * i] == 0u;
        w[i]->wfinite = (f[4 * i + 1] & 1u) == 0u;
        w[i]->wf32 = (f[4 * i + 1] & 2u) == 0u;
        const unsigned long long b = (unsigned long long)f[4 * i + 2] | ((unsigned long long)f[4 * i + 3] << 32);
        std::memcpy(&w[i]->wmax, &b, 8);
    }
    return NC_OK;
}

// A count over terms, as nc_count_batched runs it.  Unit-weight terms beside
// weighted ones (the survey's RRR beside DDD, DDR, DRR) may take the dense u32
// layout, which the weighted ones cannot, and the six-fold path, which a
// launch with weighted terms keeps for those: they run as a launch of their
// own whenever a dense histogram could fit in a CU's LDS, on either path (so
// both paths give every primary the same share of the triplets, and the
// primary slices of the two paths agree).  Those two launches are not
// captured in a graph.
bool split_batch(const nc_context *ctx, const nc_params *params, const std::vector<Term> &t, std::vector<int> &iu,
                 std::vector<int> &iw) {
    iu.clear();
    iw.clear();
    if (!ctx->batch_split || !params || params->bin_type != NC_BIN_THREED || params->nedges < 2 ||
        params->nedges > 41)
        return false;
    for (size_t i = 0; i < t.size(); ++i) {
        if (!t[i].c1 || !t[i].c2 || !t[i].c3) return false;
        (t[i].c1->unit_w && t[i].c2->unit_w && t[i].c3->unit_w ? iu : iw).push_back((int)i);
    }
    const size_t nb = (size_t)params->nedges - 1;
    return !iu.empty() && !iw.empty() && nb * nb * nb * 4 + 48 * 1024 <= ctx->lds_hw;
}

int count_terms(nc_context *ctx, const nc_params *params, const std::vector<Term> &t, int path, double *hist,
                uint64_t *counts, bool use_graph) {
    Mem m;  // the context's own buffers
    m.fail = &ctx->fail_alloc;
    if (const int rc = refresh_weights(ctx, m, t)) return rc;
    for (const Term &x : t)
        if (!x.c1 || !x.c2 || !x.c3) return fail(ctx, NC_ERR_ARG, "null catalogue");
    {
        std::vector<int> iu, iw;
        if (split_batch(ctx, params, t, iu, iw)) {
            if (!hist) return fail(ctx, NC_ERR_ARG, "null params or output");
            const size_t len = (size_t)nc_hist_len(params);
            std::vector<double> h[2];
            std::vector<uint64_t> c[2];
            nc_timing tm[2];
            for (int g = 0; g < 2; ++g) {
                const std::vector<int> &idx = g ? iw : iu;
                std::vector<Term> sub;
                for (int i : idx) sub.push_back(t[i]);
                h[g].resize(idx.size() * len);
                if (counts) c[g].resize(idx.size() * len);
                // the unit-weight launch with its own tuning, where nc_autotune found one
                const int saved[4] = {ctx->waves_per_group, ctx->chunk, ctx->wgs_per_cu, ctx->chunk_elems};
                const bool own = g == 0 && ctx->tune_unit.on;
                if (own) {
                    ctx->waves_per_group = ctx->tune_unit.waves_per_group;
                    ctx->chunk = ctx->tune_unit.chunk;
                    ctx->wgs_per_cu = ctx->tune_unit.wgs_per_cu;
                    ctx->chunk_elems = ctx->tune_unit.list_chunk;
                }
                const int rc =
                    run_count(ctx, params, sub, path, h[g].data(), counts ? c[g].data() : nullptr, false, m);
                if (own) {
                    ctx->waves_per_group = saved[0];
                    ctx->chunk = saved[1];
                    ctx->wgs_per_cu = saved[2];
                    ctx->chunk_elems = saved[3];
                }
                if (rc) return rc;
                tm[g] = ctx->timing;
                for (size_t k = 0; k < idx.size(); ++k) {
                    std::memcpy(hist + (size_t)idx[k] * len, h[g].data() + k * len, sizeof(double) * len);
                    if (counts)
                        std::memcpy(counts + (size_t)idx[k] * len, c[g].data() + k * len, sizeof(uint64_t) * len);
                }
            }
            // the sum of both launches; layout and LDS of the unit-weight one
            ctx->timing.ms_build = tm[0].ms_build + tm[1].ms_build;
            ctx->timing.ms_count = tm[0].ms_count + tm[1].ms_count;
            ctx->timing.ms_total = tm[0].ms_total + tm[1].ms_total;
            ctx->timing.n_kernel = tm[0].n_kernel + tm[1].n_kernel;
            ctx->timing.hist_layout = tm[0].hist_layout;
            ctx->timing.lds_bytes = tm[0].lds_bytes;
            ctx->timing.path = tm[0].path;
            ctx->timing.tri_terms = tm[0].tri_terms + tm[1].tri_terms;
            ctx->timing.swap_terms = tm[0].swap_terms + tm[1].swap_terms;
            ctx->timing.partner_terms = tm[0].partner_terms + tm[1].partner_terms;
            return NC_OK;
        }
    }
    return run_count(ctx, params, t, path, hist, counts, use_graph, m);
}

}  // namespace nch
