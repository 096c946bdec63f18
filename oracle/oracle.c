/*
 * oracle.c -- CPU restatement of the reference three-point (threeD) triangle
 * count.  TEST INFRASTRUCTURE ONLY: this file is the checker that the HIP
 * library is compared against, and the CPU baseline that bench.py times.  No
 * product code path links or calls it.
 *
 * It restates, in plain C99 + OpenMP, the semantics of:
 *   source/device.cu:30-36    sqrSep: dx*dx+dy*dy+dz*dz in FP64
 *   source/device.cu:75-87    initCellIndexBox (clamp ==numCells) / Survey (no clamp)
 *   source/device.cu:89-153   shiftCellIndexBox (one wrap, +/-L image shift) /
 *                             shiftCellIndexSurvey (out-of-grid => skip)
 *   source/device.cu:154-201  getTriangleIndexThreeD: sort the three squared
 *                             sides, FP32 sqrtf triangle test, half-open bins
 *                             on squared edges, flat index idx3+idx2*nb+idx1*nb^2
 *   source/device.cu:259-315  countTrianglesThreeDBox loop structure: primary p1,
 *                             27 neighbour entries for p2 filtered on r12, 27
 *                             for p3 filtered on r13 and r23, strict
 *                             rmin^2 < r^2 < rmax^2, weight (w1*w2)*w3
 *   source/device.cu:381-431  countTrianglesThreeDSurvey (no shift, skip cells)
 *   source/cudaNcorr.cu:37-58 shift order (x-major) and host cell key with clamp
 *   source/cudaNcorr.cu:118-157  geometry from the edges: sqrBins = e*e,
 *                             cellSize = rmax, numCells = ceil(box/cellSize)
 *
 * Evaluation order of sqrSep.  The reference leaves FMA contraction to nvcc
 * (-fmad=true by default, makefile:2), which cannot be observed here.  The
 * oracle supports three orders and the HIP library implements order 0:
 *   0: fma(dz,dz, fma(dy,dy, dx*dx))   left-to-right contraction (canonical)
 *   1: (dx*dx + dy*dy) + dz*dz         no contraction
 *   2: fma(dz,dz, fma(dx,dx, dy*dy))   clang's contraction of the same source
 * For coordinates whose differences and squares are exact in FP64 (e.g.
 * float32-valued coordinates) all three agree; nc_oracle_sensitivity() counts
 * the pairs for which they do not.
 *
 * Weighted sums: every ordered binned triplet adds (w1*w2)*w3 (FP64 product,
 * as device.cu:305).  Sums are accumulated with Neumaier compensation per bin
 * and per thread and folded in thread order, so the result is the FP64 value
 * of the exact sum to within a few ulp: this is the "FP64 oracle" that
 * weighted GPU results are measured against.  Unit counts are accumulated
 * separately as uint64.
 *
 * Validation that the reference does not do (its behaviour there is
 * undefined): edges must be >= 0 and strictly increasing with at least two
 * edges; every cell key must land inside the grid (box and survey secondaries
 * after the ==numCells clamp of cudaNcorr.cu:50-58, survey primaries without
 * it, device.cu:84-87).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OC_OK 0
#define OC_ERR_ARG 1
#define OC_ERR_BINS 2
#define OC_ERR_DOMAIN 3
#define OC_ERR_NOMEM 4

typedef struct {
    double x, y, z, w;
} oc_point;

typedef struct {
    double box[3];
    double cell[3];
    int ncell[3];
    int survey;
    int nb;
    const double *sq; /* nb+1 squared edges */
    int contract;
} oc_geom;

static inline double oc_sqrsep(double dx, double dy, double dz, int mode) {
    /* device.cu:30-36 in the chosen evaluation order (see header). */
    if (mode == 1) {
        double a = dx * dx;
        double b = dy * dy;
        double c = dz * dz;
        double s = a + b;
        return s + c;
    }
    if (mode == 2) return fma(dz, dz, fma(dx, dx, dy * dy));
    return fma(dz, dz, fma(dy, dy, dx * dx));
}

/* Host cell key per axis, cudaNcorr.cu:50-58 (clamp ==numCells to numCells-1).
 * Returns -1 if the point falls outside the grid (undefined in the reference). */
static inline int oc_axis_cell(double v, double cs, int n, int clamp) {
    double q = v / cs;
    if (!(q > -1.0 && q < (double)n + 1.0)) return -1;
    int c = (int)q; /* truncation toward zero, as static_cast<int> */
    if (clamp && c == n) c = n - 1;
    if (c < 0 || c >= n) return -1;
    return c;
}

/* Triangle index, device.cu:154-201.  Returns -1 for a rejected triangle. */
static inline int oc_tri_index(double r1, double r2, double r3, const double *sq, int nb) {
    double t;
    if (r1 > r2) { t = r1; r1 = r2; r2 = t; }
    if (r1 > r3) { t = r1; r1 = r3; r3 = t; }
    if (r2 > r3) { t = r2; r2 = r3; r3 = t; }
    /* FP32 triangle test: each squared side is rounded to float, sqrtf is the
     * IEEE correctly rounded square root, the sum is a float add. */
    float s1 = sqrtf((float)r1);
    float s2 = sqrtf((float)r2);
    float s3 = sqrtf((float)r3);
    float sum = s1 + s2;
    if (sum <= s3) return -1;
    int i1 = -1, i2 = -1, i3 = -1;
    for (int i = 0; i < nb; ++i) {
        if (r1 >= sq[i] && r1 < sq[i + 1]) i1 = i;
        if (r2 >= sq[i] && r2 < sq[i + 1]) i2 = i;
        if (r3 >= sq[i] && r3 < sq[i + 1]) i3 = i;
    }
    if (i1 < 0 || i2 < 0 || i3 < 0) return -1;
    return i3 + i2 * nb + i1 * nb * nb;
}

typedef struct {
    int *start;   /* ncells+1 */
    oc_point *pt; /* points grouped by cell, input order within a cell */
} oc_grid;

static int oc_grid_build(const oc_geom *g, const oc_point *p, long n, oc_grid *out) {
    long nc = (long)g->ncell[0] * g->ncell[1] * g->ncell[2];
    int *key = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    out->start = (int *)calloc((size_t)nc + 1, sizeof(int));
    out->pt = (oc_point *)malloc(sizeof(oc_point) * (size_t)(n > 0 ? n : 1));
    if (!key || !out->start || !out->pt) {
        free(key);
        return OC_ERR_NOMEM;
    }
    for (long i = 0; i < n; ++i) {
        int cx = oc_axis_cell(p[i].x, g->cell[0], g->ncell[0], 1);
        int cy = oc_axis_cell(p[i].y, g->cell[1], g->ncell[1], 1);
        int cz = oc_axis_cell(p[i].z, g->cell[2], g->ncell[2], 1);
        if (cx < 0 || cy < 0 || cz < 0) {
            free(key);
            return OC_ERR_DOMAIN;
        }
        key[i] = cz + g->ncell[2] * (cy + g->ncell[1] * cx);
        out->start[key[i] + 1]++;
    }
    for (long c = 0; c < nc; ++c) out->start[c + 1] += out->start[c];
    int *cur = (int *)malloc(sizeof(int) * (size_t)nc);
    if (!cur) {
        free(key);
        return OC_ERR_NOMEM;
    }
    memcpy(cur, out->start, sizeof(int) * (size_t)nc);
    for (long i = 0; i < n; ++i) out->pt[cur[key[i]]++] = p[i];
    free(cur);
    free(key);
    return OC_OK;
}

static void oc_grid_free(oc_grid *gr) {
    free(gr->start);
    free(gr->pt);
}

/* One of the 27 neighbour entries of cell (cx,cy,cz): shifts in x-major order
 * (cudaNcorr.cu:37-48).  Box: one wrap with +/-L image shift (device.cu:89-122).
 * Survey: out-of-grid => -1 (device.cu:123-153). */
static inline int oc_neighbour(const oc_geom *g, int cx, int cy, int cz, int s, double sh[3]) {
    int d[3] = {s / 9 - 1, (s / 3) % 3 - 1, s % 3 - 1};
    int c[3] = {cx + d[0], cy + d[1], cz + d[2]};
    sh[0] = sh[1] = sh[2] = 0.0;
    for (int a = 0; a < 3; ++a) {
        if (g->survey) {
            if (c[a] == g->ncell[a] || c[a] == -1) return -1;
        } else {
            if (c[a] == g->ncell[a]) {
                c[a] = 0;
                sh[a] = g->box[a];
            }
            if (c[a] == -1) {
                c[a] = g->ncell[a] - 1;
                sh[a] = -g->box[a];
            }
        }
    }
    return c[2] + g->ncell[2] * (c[1] + g->ncell[1] * c[0]);
}

static inline void oc_neumaier(double *s, double *c, double v) {
    double t = *s + v;
    if (fabs(*s) >= fabs(v))
        *c += (*s - t) + v;
    else
        *c += (v - t) + *s;
    *s = t;
}

static int oc_setup(const double *edges, int nedges, const double box[3], int survey, int contract,
                    oc_geom *g, double **sq_out) {
    if (!edges || nedges < 2 || !box) return OC_ERR_ARG;
    if (!(edges[0] >= 0.0)) return OC_ERR_BINS;
    for (int i = 1; i < nedges; ++i)
        if (!(edges[i] > edges[i - 1])) return OC_ERR_BINS;
    double *sq = (double *)malloc(sizeof(double) * (size_t)nedges);
    if (!sq) return OC_ERR_NOMEM;
    for (int i = 0; i < nedges; ++i) sq[i] = edges[i] * edges[i]; /* cudaNcorr.cu:124-127 */
    g->nb = nedges - 1;
    g->sq = sq;
    g->survey = survey;
    g->contract = contract;
    long total = 1;
    for (int a = 0; a < 3; ++a) {
        g->box[a] = box[a];
        g->cell[a] = edges[nedges - 1]; /* cellSize = rmax, cudaNcorr.cu:137 */
        double q = ceil(box[a] / g->cell[a]);
        if (!(q >= 1.0 && q < 2147483647.0)) {
            free(sq);
            return OC_ERR_ARG;
        }
        g->ncell[a] = (int)q;
        total *= g->ncell[a];
        if (total > (1L << 28)) {
            free(sq);
            return OC_ERR_ARG;
        }
    }
    *sq_out = sq;
    return OC_OK;
}

/*
 * Count ordered triplets (p1,p2,p3) into the dense nb^3 histogram.
 *   counts[nb^3]  uint64 number of binned triplets (may be NULL)
 *   wsum[nb^3]    compensated FP64 sum of (w1*w2)*w3 (may be NULL)
 * Returns OC_OK or an OC_ERR_* code.
 */
int nc_oracle_count(const oc_point *p1, long n1, const oc_point *p2, long n2, const oc_point *p3, long n3,
                    const double *edges, int nedges, const double *box, int survey, int contract,
                    int nthreads, uint64_t *counts, double *wsum) {
    oc_geom g;
    double *sq = NULL;
    int rc = oc_setup(edges, nedges, box, survey, contract, &g, &sq);
    if (rc) return rc;
    if (n1 < 0 || n2 < 0 || n3 < 0 || (n1 && !p1) || (n2 && !p2) || (n3 && !p3)) {
        free(sq);
        return OC_ERR_ARG;
    }
    const int nb = g.nb;
    const long nh = (long)nb * nb * nb;
    const double r2min = sq[0], r2max = sq[nb]; /* rMin*rMin, rMax*rMax: same values */

    oc_grid g2 = {0}, g3 = {0};
    rc = oc_grid_build(&g, p2, n2, &g2);
    if (!rc) rc = oc_grid_build(&g, p3, n3, &g3);
    /* primary cells: box clamps (device.cu:78-80), survey does not (device.cu:84-87) */
    int *pc = (int *)malloc(sizeof(int) * 3 * (size_t)(n1 > 0 ? n1 : 1));
    if (!rc && !pc) rc = OC_ERR_NOMEM;
    for (long i = 0; !rc && i < n1; ++i) {
        int clamp = !survey;
        int cx = oc_axis_cell(p1[i].x, g.cell[0], g.ncell[0], clamp);
        int cy = oc_axis_cell(p1[i].y, g.cell[1], g.ncell[1], clamp);
        int cz = oc_axis_cell(p1[i].z, g.cell[2], g.ncell[2], clamp);
        if (cx < 0 || cy < 0 || cz < 0) rc = OC_ERR_DOMAIN;
        pc[3 * i] = cx;
        pc[3 * i + 1] = cy;
        pc[3 * i + 2] = cz;
    }
    if (rc) {
        free(pc);
        oc_grid_free(&g2);
        oc_grid_free(&g3);
        free(sq);
        return rc;
    }
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    int nt = 1;
    (void)nthreads;
#endif
    uint64_t *tc = (uint64_t *)calloc((size_t)nt * nh, sizeof(uint64_t));
    double *ts = (double *)calloc((size_t)nt * nh, sizeof(double));
    double *tcmp = (double *)calloc((size_t)nt * nh, sizeof(double));
    if (!tc || !ts || !tcmp) {
        free(tc);
        free(ts);
        free(tcmp);
        free(pc);
        oc_grid_free(&g2);
        oc_grid_free(&g3);
        free(sq);
        return OC_ERR_NOMEM;
    }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        uint64_t *c = tc + (size_t)tid * nh;
        double *s = ts + (size_t)tid * nh;
        double *e = tcmp + (size_t)tid * nh;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (long i = 0; i < n1; ++i) {
            const oc_point a = p1[i];
            const int cx = pc[3 * i], cy = pc[3 * i + 1], cz = pc[3 * i + 2];
            for (int s2 = 0; s2 < 27; ++s2) {
                double sh2[3];
                int k2 = oc_neighbour(&g, cx, cy, cz, s2, sh2);
                if (k2 < 0) continue;
                for (int j = g2.start[k2]; j < g2.start[k2 + 1]; ++j) {
                    oc_point b = g2.pt[j];
                    b.x += sh2[0];
                    b.y += sh2[1];
                    b.z += sh2[2];
                    double r12 = oc_sqrsep(a.x - b.x, a.y - b.y, a.z - b.z, contract);
                    if (!(r12 < r2max && r12 > r2min)) continue;
                    const double w12 = a.w * b.w;
                    for (int s3 = 0; s3 < 27; ++s3) {
                        double sh3[3];
                        int k3 = oc_neighbour(&g, cx, cy, cz, s3, sh3);
                        if (k3 < 0) continue;
                        for (int l = g3.start[k3]; l < g3.start[k3 + 1]; ++l) {
                            oc_point d = g3.pt[l];
                            d.x += sh3[0];
                            d.y += sh3[1];
                            d.z += sh3[2];
                            double r13 = oc_sqrsep(a.x - d.x, a.y - d.y, a.z - d.z, contract);
                            double r23 = oc_sqrsep(b.x - d.x, b.y - d.y, b.z - d.z, contract);
                            if (!(r13 < r2max && r23 < r2max && r13 > r2min && r23 > r2min)) continue;
                            int idx = oc_tri_index(r12, r13, r23, sq, nb);
                            if (idx < 0) continue;
                            c[idx] += 1;
                            oc_neumaier(&s[idx], &e[idx], w12 * d.w);
                        }
                    }
                }
            }
        }
    }
    for (long h = 0; h < nh; ++h) {
        uint64_t cc = 0;
        double ss = 0.0, ee = 0.0;
        for (int t = 0; t < nt; ++t) {
            cc += tc[(size_t)t * nh + h];
            oc_neumaier(&ss, &ee, ts[(size_t)t * nh + h]);
            ee += tcmp[(size_t)t * nh + h];
        }
        if (counts) counts[h] = cc;
        if (wsum) wsum[h] = ss + ee;
    }
    free(tc);
    free(ts);
    free(tcmp);
    free(pc);
    oc_grid_free(&g2);
    oc_grid_free(&g3);
    free(sq);
    return OC_OK;
}

/*
 * Contraction sensitivity: over the same (p1,p2) and (p2,p3) pairs the count
 * visits, count pairs whose squared separation differs between evaluation
 * orders 0, 1 and 2 AND lands on different sides of some squared edge (i.e. a
 * pair where the unobservable nvcc contraction choice could change a count).
 * Only pairs with both points within the range of p1 are considered, as in
 * the count.  Returns the number of such pairs, or a negative error code.
 */
long nc_oracle_sensitivity(const oc_point *p1, long n1, const oc_point *p2, long n2, const double *edges,
                           int nedges, const double *box, int survey) {
    oc_geom g;
    double *sq = NULL;
    int rc = oc_setup(edges, nedges, box, survey, 0, &g, &sq);
    if (rc) return -rc;
    oc_grid g2 = {0};
    rc = oc_grid_build(&g, p2, n2, &g2);
    if (rc) {
        oc_grid_free(&g2);
        free(sq);
        return -rc;
    }
    const int nb = g.nb;
    long bad = 0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : bad) schedule(dynamic, 64)
#endif
    for (long i = 0; i < n1; ++i) {
        const oc_point a = p1[i];
        int cx = oc_axis_cell(a.x, g.cell[0], g.ncell[0], !survey);
        int cy = oc_axis_cell(a.y, g.cell[1], g.ncell[1], !survey);
        int cz = oc_axis_cell(a.z, g.cell[2], g.ncell[2], !survey);
        if (cx < 0 || cy < 0 || cz < 0) continue;
        for (int s2 = 0; s2 < 27; ++s2) {
            double sh[3];
            int k = oc_neighbour(&g, cx, cy, cz, s2, sh);
            if (k < 0) continue;
            for (int j = g2.start[k]; j < g2.start[k + 1]; ++j) {
                oc_point b = g2.pt[j];
                b.x += sh[0];
                b.y += sh[1];
                b.z += sh[2];
                double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
                double v0 = oc_sqrsep(dx, dy, dz, 0);
                double v1 = oc_sqrsep(dx, dy, dz, 1);
                double v2 = oc_sqrsep(dx, dy, dz, 2);
                if (v0 == v1 && v0 == v2) continue;
                double lo = v0 < v1 ? v0 : v1;
                lo = lo < v2 ? lo : v2;
                double hi = v0 > v1 ? v0 : v1;
                hi = hi > v2 ? hi : v2;
                for (int e = 0; e <= nb; ++e)
                    if (lo < sq[e] && sq[e] <= hi) {
                        ++bad;
                        break;
                    }
            }
        }
    }
    oc_grid_free(&g2);
    free(sq);
    return bad;
}

/* Output compaction, cudaNcorr.cu:72-86 and 145-154: keep sorted (i<=j<=k)
 * bins with mid_k <= mid_i + mid_j, in lexicographic order.  Writes at most
 * cap values and returns the number of output bins. */
int nc_oracle_compact(const double *edges, int nedges, const double *dense, double *out, int cap) {
    if (!edges || nedges < 2) return -OC_ERR_ARG;
    int nb = nedges - 1, n = 0;
    for (int i = 0; i < nb; ++i)
        for (int j = i; j < nb; ++j)
            for (int k = j; k < nb; ++k) {
                double mi = (edges[i + 1] + edges[i]) / 2.0;
                double mj = (edges[j + 1] + edges[j]) / 2.0;
                double mk = (edges[k + 1] + edges[k]) / 2.0;
                if (mk <= mi + mj) {
                    if (out && dense && n < cap) out[n] = dense[k + j * nb + i * nb * nb];
                    ++n;
                }
            }
    return n;
}

/* ------------------------------------------------------------------------
 * rppi binning: projected separation rp and line-of-sight separation pi.
 *
 * Restates:
 *   source/cudaNcorr.cu:159-216  rppi constructor: sqrBins = e*e on the rp
 *                                edges; pi bins 0,1,...,numMinorBins (unit
 *                                width), piMax = numMinorBins;
 *                                sqrRMax = rpMax^2 + piMax^2, rMax =
 *                                sqrtf(sqrRMax) (FP32); cells (rpMax, rpMax,
 *                                piMax) in box mode, (rMax, rMax, rMax) in
 *                                survey mode; histogram nb^3 * npi^2
 *   source/device.cu:38-49     sqrRpBox = dx*dx + dy*dy, sqrPiBox = dz*dz
 *   source/device.cu:51-71     getLOS (los = pos1 + pos2 + pos3, pos = p - ori,
 *                                sqrL = los.los), getSqrSdotL ((s.los)^2 with
 *                                s = pa - pb), getSqrRpSqrPi (pi^2 =
 *                                (s.L)^2 / L^2, rp^2 = s^2 - pi^2)
 *   source/device.cu:203-255   getTriangleIndexRppi: sort the three sides by
 *                                rp^2 carrying their pi^2, FP32 sqrtf triangle
 *                                test on rp, half-open rp bins, pi bins
 *                                floor(sqrtf(pi^2)) of the two smallest-rp
 *                                sides, flat index
 *                                idxPi2 + idxPi1*npi + (idx3 + idx2*nb +
 *                                idx1*nb^2)*npi^2
 *   source/device.cu:317-379   countTrianglesRppiBox: p2 kept when
 *                                pi12^2 < piMax^2 and rpMin^2 < rp12^2 <
 *                                rpMax^2, p3 when the same holds for (1,3)
 *                                and (2,3)
 *   source/device.cu:434-505   countTrianglesRppiSurvey: p2, p3 kept on the
 *                                3D separations in (sqrRMin, sqrRMax); pi of
 *                                each side along the triplet's LOS; skip when
 *                                piMax^2 * L^2 < (s.L)^2 for any side; keep
 *                                when the three rp^2 are in range and
 *                                pi12^2, pi13^2 < piMax^2
 * Quirks of the reference and the choices made here:
 *   - sqrRMin is never set for rppi (cudaNcorr.hpp:38, cudaNcorr.cu:229
 *     uploads an indeterminate value): taken as 0 here;
 *   - only ori.x reaches the device (cudaNcorr.cu:254 copies
 *     sizeof(double) into the double3 d_ori): the LOS uses (ori.x, 0, 0);
 *   - a pi index equal to npi (pi^2 just below piMax^2 with sqrtf rounding
 *     up) aliases into the next flat bin exactly as the flat index does;
 *     an index past the end of the histogram (the reference writes out of
 *     bounds) is dropped and counted in *dropped.
 * The dot products and rp^2 follow the same contraction modes as sqrSep.
 * ------------------------------------------------------------------------ */

static inline double oc_dot3(double ax, double ay, double az, double bx, double by, double bz, int mode) {
    if (mode == 1) {
        double s = ax * bx + ay * by;
        return s + az * bz;
    }
    if (mode == 2) return fma(az, bz, fma(ax, bx, ay * by));
    return fma(az, bz, fma(ay, by, ax * bx));
}

static inline double oc_rp2(double dx, double dy, int mode) {
    if (mode == 1) {
        double a = dx * dx;
        double b = dy * dy;
        return a + b;
    }
    if (mode == 2) return fma(dx, dx, dy * dy);
    return fma(dy, dy, dx * dx);
}

/* device.cu:203-255; -1 for a rejected triangle. */
static inline long oc_rppi_index(double r1, double r2, double r3, double q1, double q2, double q3,
                                 const double *sq, int nb, int npi) {
    double t;
    if (r1 > r2) {
        t = r1; r1 = r2; r2 = t;
        t = q1; q1 = q2; q2 = t;
    }
    if (r1 > r3) {
        t = r1; r1 = r3; r3 = t;
        t = q1; q1 = q3; q3 = t;
    }
    if (r2 > r3) {
        t = r2; r2 = r3; r3 = t;
        t = q2; q2 = q3; q3 = t;
    }
    if (sqrtf((float)r1) + sqrtf((float)r2) <= sqrtf((float)r3)) return -1;
    int i1 = -1, i2 = -1, i3 = -1;
    for (int i = 0; i < nb; ++i) {
        if (r1 >= sq[i] && r1 < sq[i + 1]) i1 = i;
        if (r2 >= sq[i] && r2 < sq[i + 1]) i2 = i;
        if (r3 >= sq[i] && r3 < sq[i + 1]) i3 = i;
    }
    if (i1 < 0 || i2 < 0 || i3 < 0) return -1; /* cannot happen after the range filters */
    const long p1i = (long)floor(sqrtf((float)q1)), p2i = (long)floor(sqrtf((float)q2));
    const long np = npi, n = nb;
    return p2i + p1i * np + (long)i3 * np * np + (long)i2 * n * np * np + (long)i1 * n * n * np * np;
}

/*
 * Ordered-triplet rppi histogram (dense nb^3 * npi^2, the reference's layout).
 *   npi   number of unit-width pi bins (numMinorBins)
 *   ori   survey origin (only ori[0] is used, as in the reference)
 *   counts, wsum  as nc_oracle_count; dropped: triplets whose flat index
 *   falls past the histogram (may be NULL)
 */
int nc_oracle_count_rppi(const oc_point *p1, long n1, const oc_point *p2, long n2, const oc_point *p3, long n3,
                         const double *edges, int nedges, int npi, const double *box, int survey,
                         const double *ori, int contract, int nthreads, uint64_t *counts, double *wsum,
                         long *dropped) {
    if (!edges || nedges < 2 || !box || npi < 1 || npi > 4096) return OC_ERR_ARG;
    if (!(edges[0] >= 0.0)) return OC_ERR_BINS;
    for (int i = 1; i < nedges; ++i)
        if (!(edges[i] > edges[i - 1])) return OC_ERR_BINS;
    if (n1 < 0 || n2 < 0 || n3 < 0 || (n1 && !p1) || (n2 && !p2) || (n3 && !p3)) return OC_ERR_ARG;
    const int nb = nedges - 1;
    const long nh = (long)nb * nb * nb * npi * npi;
    if (nh > (1L << 28)) return OC_ERR_ARG;
    double *sq = (double *)malloc(sizeof(double) * (size_t)nedges);
    if (!sq) return OC_ERR_NOMEM;
    for (int i = 0; i < nedges; ++i) sq[i] = edges[i] * edges[i]; /* cudaNcorr.cu:163-165 */
    const double rpmax = edges[nb], rpmin = edges[0];
    const double sqr_rpmax = rpmax * rpmax, sqr_rpmin = rpmin * rpmin;
    const double pimax = (double)npi; /* piBins 0,1,...,npi: cudaNcorr.cu:175-182 */
    const double sqr_pimax = pimax * pimax;
    const double sqr_rmax = rpmax * rpmax + pimax * pimax;   /* cudaNcorr.cu:186 */
    const double rmax = (double)sqrtf((float)sqr_rmax);      /* cudaNcorr.cu:187 (FP32) */
    const double sqr_rmin = 0.0;                              /* never set in the reference */
    const double ox = ori ? ori[0] : 0.0, oy = 0.0, oz = 0.0; /* only ori.x reaches the device */
    oc_geom g;
    g.survey = survey;
    g.contract = contract;
    g.nb = nb;
    g.sq = sq;
    const double cell[3] = {survey ? rmax : rpmax, survey ? rmax : rpmax, survey ? rmax : pimax};
    long total = 1;
    for (int a = 0; a < 3; ++a) {
        g.box[a] = box[a];
        g.cell[a] = cell[a];
        double q = ceil(box[a] / cell[a]);
        if (!(q >= 1.0 && q < 2147483647.0)) {
            free(sq);
            return OC_ERR_ARG;
        }
        g.ncell[a] = (int)q;
        total *= g.ncell[a];
        if (total > (1L << 28)) {
            free(sq);
            return OC_ERR_ARG;
        }
    }
    oc_grid g2 = {0}, g3 = {0};
    int rc = oc_grid_build(&g, p2, n2, &g2);
    if (!rc) rc = oc_grid_build(&g, p3, n3, &g3);
    int *pc = (int *)malloc(sizeof(int) * 3 * (size_t)(n1 > 0 ? n1 : 1));
    if (!rc && !pc) rc = OC_ERR_NOMEM;
    for (long i = 0; !rc && i < n1; ++i) {
        int clamp = !survey;
        int cx = oc_axis_cell(p1[i].x, g.cell[0], g.ncell[0], clamp);
        int cy = oc_axis_cell(p1[i].y, g.cell[1], g.ncell[1], clamp);
        int cz = oc_axis_cell(p1[i].z, g.cell[2], g.ncell[2], clamp);
        if (cx < 0 || cy < 0 || cz < 0) rc = OC_ERR_DOMAIN;
        pc[3 * i] = cx;
        pc[3 * i + 1] = cy;
        pc[3 * i + 2] = cz;
    }
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    int nt = 1;
    (void)nthreads;
#endif
    uint64_t *tc = NULL;
    double *ts = NULL, *tcmp = NULL;
    if (!rc) {
        tc = (uint64_t *)calloc((size_t)nt * nh, sizeof(uint64_t));
        ts = (double *)calloc((size_t)nt * nh, sizeof(double));
        tcmp = (double *)calloc((size_t)nt * nh, sizeof(double));
        if (!tc || !ts || !tcmp) rc = OC_ERR_NOMEM;
    }
    if (rc) {
        free(tc);
        free(ts);
        free(tcmp);
        free(pc);
        oc_grid_free(&g2);
        oc_grid_free(&g3);
        free(sq);
        return rc;
    }
    long lost = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt) reduction(+ : lost)
#endif
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        uint64_t *c = tc + (size_t)tid * nh;
        double *s = ts + (size_t)tid * nh;
        double *e = tcmp + (size_t)tid * nh;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 16)
#endif
        for (long i = 0; i < n1; ++i) {
            const oc_point a = p1[i];
            const int cx = pc[3 * i], cy = pc[3 * i + 1], cz = pc[3 * i + 2];
            for (int s2 = 0; s2 < 27; ++s2) {
                double sh2[3];
                int k2 = oc_neighbour(&g, cx, cy, cz, s2, sh2);
                if (k2 < 0) continue;
                for (int j = g2.start[k2]; j < g2.start[k2 + 1]; ++j) {
                    oc_point b = g2.pt[j];
                    b.x += sh2[0];
                    b.y += sh2[1];
                    b.z += sh2[2];
                    double q12 = 0.0, r12 = 0.0, s12 = 0.0;
                    if (survey) {
                        s12 = oc_sqrsep(a.x - b.x, a.y - b.y, a.z - b.z, contract);
                        if (!(s12 < sqr_rmax && s12 > sqr_rmin)) continue;
                    } else {
                        const double dz = a.z - b.z;
                        q12 = dz * dz;
                        if (!(q12 < sqr_pimax)) continue;
                        r12 = oc_rp2(a.x - b.x, a.y - b.y, contract);
                        if (!(r12 < sqr_rpmax && r12 > sqr_rpmin)) continue;
                    }
                    const double w12 = a.w * b.w;
                    for (int s3 = 0; s3 < 27; ++s3) {
                        double sh3[3];
                        int k3 = oc_neighbour(&g, cx, cy, cz, s3, sh3);
                        if (k3 < 0) continue;
                        for (int l = g3.start[k3]; l < g3.start[k3 + 1]; ++l) {
                            oc_point d = g3.pt[l];
                            d.x += sh3[0];
                            d.y += sh3[1];
                            d.z += sh3[2];
                            double r13, r23, q13, q23, r12t = r12, q12t = q12;
                            if (survey) {
                                const double s13 = oc_sqrsep(a.x - d.x, a.y - d.y, a.z - d.z, contract);
                                const double s23 = oc_sqrsep(b.x - d.x, b.y - d.y, b.z - d.z, contract);
                                if (!(s13 < sqr_rmax && s23 < sqr_rmax && s13 > sqr_rmin && s23 > sqr_rmin)) continue;
                                /* getLOS: pos = p - ori, los = pos1 + pos2 + pos3 */
                                const double lx = ((a.x - ox) + (b.x - ox)) + (d.x - ox);
                                const double ly = ((a.y - oy) + (b.y - oy)) + (d.y - oy);
                                const double lz = ((a.z - oz) + (b.z - oz)) + (d.z - oz);
                                const double sqr_l = oc_dot3(lx, ly, lz, lx, ly, lz, contract);
                                const double lim = sqr_pimax * sqr_l;
                                double t1 = oc_dot3(lx, ly, lz, a.x - b.x, a.y - b.y, a.z - b.z, contract);
                                t1 = t1 * t1;
                                if (lim < t1) continue;
                                double t2 = oc_dot3(lx, ly, lz, a.x - d.x, a.y - d.y, a.z - d.z, contract);
                                t2 = t2 * t2;
                                if (lim < t2) continue;
                                double t3 = oc_dot3(lx, ly, lz, b.x - d.x, b.y - d.y, b.z - d.z, contract);
                                t3 = t3 * t3;
                                if (lim < t3) continue;
                                q12t = t1 / sqr_l;
                                r12t = s12 - q12t;
                                q13 = t2 / sqr_l;
                                r13 = s13 - q13;
                                q23 = t3 / sqr_l;
                                r23 = s23 - q23;
                                if (!(r12t < sqr_rpmax && r13 < sqr_rpmax && r23 < sqr_rpmax && r12t > sqr_rpmin &&
                                      r13 > sqr_rpmin && r23 > sqr_rpmin && q12t < sqr_pimax && q13 < sqr_pimax))
                                    continue;
                            } else {
                                const double dz13 = a.z - d.z, dz23 = b.z - d.z;
                                q13 = dz13 * dz13;
                                q23 = dz23 * dz23;
                                if (!(q13 < sqr_pimax && q23 < sqr_pimax)) continue;
                                r13 = oc_rp2(a.x - d.x, a.y - d.y, contract);
                                r23 = oc_rp2(b.x - d.x, b.y - d.y, contract);
                                if (!(r13 < sqr_rpmax && r23 < sqr_rpmax && r13 > sqr_rpmin && r23 > sqr_rpmin)) continue;
                            }
                            const long idx = oc_rppi_index(r12t, r13, r23, q12t, q13, q23, sq, nb, npi);
                            if (idx < 0) continue;
                            if (idx >= nh) {
                                ++lost;
                                continue;
                            }
                            c[idx] += 1;
                            oc_neumaier(&s[idx], &e[idx], w12 * d.w);
                        }
                    }
                }
            }
        }
    }
    for (long h = 0; h < nh; ++h) {
        uint64_t cc = 0;
        double ss = 0.0, ee = 0.0;
        for (int t = 0; t < nt; ++t) {
            cc += tc[(size_t)t * nh + h];
            oc_neumaier(&ss, &ee, ts[(size_t)t * nh + h]);
            ee += tcmp[(size_t)t * nh + h];
        }
        if (counts) counts[h] = cc;
        if (wsum) wsum[h] = ss + ee;
    }
    if (dropped) *dropped = lost;
    free(tc);
    free(ts);
    free(tcmp);
    free(pc);
    oc_grid_free(&g2);
    oc_grid_free(&g3);
    free(sq);
    return OC_OK;
}

/* Output compaction for rppi, cudaNcorr.cu:88-105: the threeD selection of
 * sorted rp bins, each followed by its npi^2 pi bins (l outer, m inner). */
int nc_oracle_compact_rppi(const double *edges, int nedges, int npi, const double *dense, double *out, int cap) {
    if (!edges || nedges < 2 || npi < 1) return -OC_ERR_ARG;
    int nb = nedges - 1, n = 0;
    for (int i = 0; i < nb; ++i)
        for (int j = i; j < nb; ++j)
            for (int k = j; k < nb; ++k) {
                double mi = (edges[i + 1] + edges[i]) / 2.0;
                double mj = (edges[j + 1] + edges[j]) / 2.0;
                double mk = (edges[k + 1] + edges[k]) / 2.0;
                if (!(mk <= mi + mj)) continue;
                for (int l = 0; l < npi; ++l)
                    for (int m = 0; m < npi; ++m) {
                        long idx = m + (long)l * npi + (long)k * npi * npi + (long)j * npi * npi * nb +
                                   (long)i * npi * npi * nb * nb;
                        if (out && dense && n < cap) out[n] = dense[idx];
                        ++n;
                    }
            }
    return n;
}
