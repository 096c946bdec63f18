// Geometry of a count (the reference's ncorr constructor, cudaNcorr.cu:118-216)
// and the lookup tables of the kernels: the FP64 bin lookup, the compact slot
// index, the mixed path's FP32 lookup with its guard bands, the slot table.
#include "ncorr_host.hpp"

namespace nch {


// Guard band, collinearity bound and FP32 lookup of the mixed path.
//
// Let u = 2^-24, R = rmax, A, B the FP64 image positions of two list entries,
// P the primary.  Stored: r = fl32(A - P) per component (|error| <= u*R, as
// |A - P| < R), s = fl32(|r|^2).  Per pair: dot = fl32(r_a . r_b) and
// d32 = fl32(s_a + s_b - 2*dot).  Then, with S = |A-P|^2 and D = (A-P).(B-P)
// exact,  |s - S| <= 6.5uR^2,  |dot - D| <= 6.5uR^2,  and
//     |d32 - |A-B|^2| <= 32uR^2,
// while the reference's FP64 |A-B|^2 is within ~1e-15 R^2 of the exact value.
// Every squared edge E (rounded to FP32, error <= u*R^2) gets the guard band
// g32 = 80uR^2 (2x the sum): d32 outside it is on the same side of E as the
// reference's value.
//
// Triangle test.  The reference rejects fl(f1 + f2) <= f3 on sorted
// f = sqrtf(float(r^2)); with exact lengths a = |A-P|, b = |B-P|, c = |A-B|,
// it accepts whenever min(a + b - c, c - |a - b|) > 20uR.  Both slacks are at
// least ab*sin^2(t)/(2(a+b)) = (S_a S_b - D^2)/(2ab(a+b)) and ab(a+b) <=
// R(S_a + S_b), so  S_a S_b - D^2 > 40uR^2 (S_a + S_b)  is sufficient.  The
// FP32 value s_a*s_b - dot^2 is within 28uR^4 of the exact one, so
// s_a*s_b - dot^2 > 80uR^2 (s_a + s_b) + 64uR^4  (both terms doubled)
// suffices; with s_a, s_b <= R^2 (1 + 8u) (list entries lie within R) the
// one constant  tri_t = 80uR^2 * 2R^2 (1 + 1e-5) + 64uR^4  bounds the right
// side for every pair.  A pair with s_a*s_b - dot^2 > tri_t is accepted,
// any other re-evaluated exactly.
//
// Scaling: s32 = sigma^2 is a power of 4 and the lists hold sigma*(A - P)
// (rounded once, from the exact FP64 product), so every FP32 quantity above
// is the unscaled one times a power of 2, with the same rounding; the
// thresholds g32, tri_t and the edges are stored scaled (by s32, s32^2,
// s32).
//
// Lookup.  For a positive float x, bits(x) >> sh is monotone in x: cell c
// holds the floats [val(c << sh), val((c+1) << sh)), 2^(23-sh) log-spaced
// cells per octave, fine near zero where squared edges crowd and coarse where
// they are sparse.  The kernel clamps d32' = s32*d32 to [val(c0 << sh), 4]
// (cell c0 thereby takes every smaller value, negative ones included, and
// the cell c1 of 4.0, above the scaled rmax^2 < 4, every larger one) and
// reads entry c - c0 = {number of edges whose bands lie wholly below the
// cell, the edge s32*E of the one band that meets the cell as float bits,
// +inf if none}: the bin is that number, minus one when d32' < E.  c0 is at most 2^w - 1 cells below c1 and at least the cell of the
// smallest normal float.  The smallest table (w = 8..13) for which some
// cell density (8..4096 per octave) leaves no cell meeting two bands is kept;
// bins narrower than the bands, or no such table, disable the path.
void make_mixed_tables(Geometry &G) {
    G.mixed_ok = false;
    const double u = std::ldexp(1.0, -24);
    const double R = G.edges.back();
    const int ne = G.nb + 1;
    const double g = 80.0 * u * R * R;
    std::vector<float> E(ne);
    for (int k = 0; k < ne; ++k) E[k] = (float)G.sq[k];
    for (int k = 0; k + 1 < ne; ++k)
        if (!((double)E[k] + g < (double)E[k + 1] - g)) return;  // bins narrower than the guard band
    // s32 = 4^kk with s32 * rmax^2 in [1, 4)
    const int kk = -(int)std::floor(std::log2(G.sq[G.nb]) / 2.0);
    const double s = std::ldexp(1.0, 2 * kk);
    std::vector<double> Es(ne);
    for (int k = 0; k < ne; ++k) Es[k] = (double)E[k] * s;  // exact: s is a power of 4
    const double gs = g * s;
    auto fbits = [](float x) {
        uint32_t b;
        std::memcpy(&b, &x, 4);
        return b;
    };
    auto fval = [](uint32_t b) {
        float x;
        std::memcpy(&x, &b, 4);
        return (double)x;
    };
    const uint32_t top = fbits(4.0f);  // the kernel's upper clamp, above s32 * rmax^2
    for (int w = 8; w <= 13; ++w)
        for (int p = 3; p <= 12; ++p) {
            const int sh = 23 - p;
            const int64_t c1 = top >> sh;
            const int64_t c0 = std::max<int64_t>(c1 - ((int64_t)1 << w) + 1, (int64_t)(0x00800000u >> sh) + 1);
            const int n = (int)(c1 - c0 + 1);
            std::vector<uint2> lut(n);
            bool ok = n >= 2;
            // cells and bands both ascend: edges [0, below) lie under the
            // cell, [below, over) meet it
            int below = 0, over = 0;
            for (int j = 0; j < n && ok; ++j) {
                const uint32_t c = (uint32_t)(c0 + j);
                const double a = j == 0 ? -INFINITY : fval(c << sh) * (1.0 - 1e-6);
                const double b = j == n - 1 ? INFINITY : fval((c + 1) << sh) * (1.0 + 1e-6);
                while (below < ne && Es[below] + gs < a) ++below;
                over = std::max(over, below);
                while (over < ne && Es[over] - gs <= b) ++over;
                if (over - below > 1) ok = false;
                const float edge = over > below ? (float)Es[below] : INFINITY;
                lut[j] = make_uint2((uint32_t)below, fbits(edge));
            }
            if (!ok) continue;
            G.lut32 = std::move(lut);
            G.lut32_shift = sh;
            G.lut32_c0 = (int)c0;
            G.lut32_lo = (float)fval((uint32_t)c0 << sh);
            G.sigma32 = std::ldexp(1.0, kk);
            G.g32 = (float)gs;
            const double R2s = R * R * s;
            G.tri_t = std::nextafter((float)(80.0 * u * R2s * 2.0 * R2s * (1.0 + 1e-5) + 64.0 * u * R2s * R2s),
                                     INFINITY);
            G.mixed_ok = true;
            return;
        }
}

static void make_slot_table(Geometry &G) {
    const int nb = G.nb;
    G.t3.assign((size_t)nb * nb * (nb + 2), 0);
    for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b)
            for (int k = 0; k < nb; ++k) {
                const int lo = std::min(std::min(a, b), k), hi = std::max(std::max(a, b), k);
                const int mid = a + b + k - lo - hi;
                G.t3[((size_t)a * nb + b) * (nb + 2) + k + 1] =
                    (uint16_t)((G.ctab[lo * nb + mid] + (hi - mid)) * sizeof(unsigned long long));
            }
}

// Geometry of the reference's threeD constructor (cudaNcorr.cu:118-157) plus
// the lookup tables of the kernels.
int make_geometry(nc_context *ctx, const nc_params *P, Geometry &G) {
    if (!P || !P->edges || P->nedges < 2) return fail(ctx, NC_ERR_BINS, "need at least two bin edges");
    const int ne = P->nedges;
    G.edges.assign(P->edges, P->edges + ne);
    if (!(G.edges[0] >= 0.0)) return fail(ctx, NC_ERR_BINS, "first bin edge must be >= 0");
    for (int i = 1; i < ne; ++i)
        if (!(G.edges[i] > G.edges[i - 1])) return fail(ctx, NC_ERR_BINS, "bin edges must be strictly increasing");
    G.nb = ne - 1;
    if (G.nb > 40) return fail(ctx, NC_ERR_LIMIT, "at most 40 bins per side are supported");
    G.nbc = (int)choose_compact(G.nb);
    G.out_words = G.nbc;
    G.sq.resize(ne);
    for (int i = 0; i < ne; ++i) G.sq[i] = G.edges[i] * G.edges[i];  // cudaNcorr.cu:124-127
    G.survey = P->survey ? 1 : 0;
    if (P->bin_type != NC_BIN_THREED && P->bin_type != NC_BIN_RPPI) return fail(ctx, NC_ERR_ARG, "unknown bin type");
    double cell[3];
    for (int a = 0; a < 3; ++a) cell[a] = G.edges[ne - 1];  // cellSize = rmax, cudaNcorr.cu:137
    if (P->bin_type == NC_BIN_RPPI) {
        // rppi constructor, cudaNcorr.cu:159-216: pi bins 0..npi, cells
        // (rpMax, rpMax, piMax) in a box, sqrtf(rpMax^2 + piMax^2) (FP32) in a survey
        if (P->num_pi_bins < 1 || P->num_pi_bins > 4096) return fail(ctx, NC_ERR_BINS, "need 1..4096 pi bins");
        G.rppi = 1;
        G.npi = P->num_pi_bins;
        G.ori_x = P->ori[0];
        const double rpmax = G.edges[ne - 1], pimax = (double)G.npi;
        const double rmax = (double)sqrtf((float)(rpmax * rpmax + pimax * pimax));
        for (int a = 0; a < 3; ++a) cell[a] = G.survey ? rmax : (a == 2 ? pimax : rpmax);
        G.out_words = (long long)G.nb * G.nb * G.nb * G.npi * G.npi;
        if (G.out_words > (1LL << 27)) return fail(ctx, NC_ERR_LIMIT, "rppi histogram larger than 2^27 bins");
    }
    G.ncells = 1;
    for (int a = 0; a < 3; ++a) {
        G.box[a] = P->box[a];
        G.cs[a] = cell[a];
        const double q = std::ceil(G.box[a] / G.cs[a]);
        if (!(q >= 1.0 && q <= 1048576.0)) return fail(ctx, NC_ERR_ARG, "box must be positive and finite");
        G.nc[a] = (int)q;
        G.ncells *= G.nc[a];
    }
    if (G.ncells > (1LL << 24)) return fail(ctx, NC_ERR_LIMIT, "cell grid larger than 2^24 cells");
    // lower-bound bin lookup over d*scale (see bin_of / bin_any): the table
    // is refined until one step over the squared edges always suffices.
    const double r2max = G.sq[G.nb];
    for (int M = 2048;; M *= 2) {
        G.lut_m = M;
        G.lut_scale = (double)M / r2max;
        G.lut.assign(M + 1, 0);
        int worst = 0;
        for (int j = 0; j <= M; ++j) {
            int k = 0;
            for (int e = 1; e < G.nb; ++e)
                if (G.sq[e] * G.lut_scale * (1.0 + 1e-12) <= (double)(j - 1)) k = e;
            G.lut[j] = (uint16_t)k;
            int steps = 0;
            for (int e = k + 1; e < G.nb; ++e)
                if (G.sq[e] <= (double)(j + 1) / G.lut_scale * (1.0 + 1e-9)) ++steps;
            worst = std::max(worst, steps);
        }
        G.lut_single_step = worst <= 1;
        if (G.lut_single_step || M >= 8192) break;
    }
    // compact index of sorted (i <= j <= k) bins in lexicographic order
    G.ctab.assign((size_t)G.nb * G.nb, 0);
    int c = 0;
    for (int i = 0; i < G.nb; ++i)
        for (int j = i; j < G.nb; ++j) {
            G.ctab[i * G.nb + j] = c;
            c += G.nb - j;
        }
    if (G.rppi) return NC_OK;  // plain FP64 path only, histogram in global memory
    make_mixed_tables(G);
    if ((size_t)G.nbc * sizeof(unsigned long long) <= 65535) make_slot_table(G);  // byte offsets fit 16 bits
    return NC_OK;
}

}  // namespace nch
