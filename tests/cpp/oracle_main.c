/*
 * Driver of the CPU oracle (oracle/oracle.c) for the sanitizer build of the
 * CPU suite (AddressSanitizer + UndefinedBehaviorSanitizer): every entry point
 * on seeded inputs, including the edge cases (empty catalogues, one and two
 * cells per side, points on cell boundaries, survey and box mode, rppi,
 * invalid arguments).  Prints one line per case, "<case> <total count>
 * <weighted sum>", which the test compares with the regular build.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

typedef struct {
    double x, y, z, w;
} pt;

int nc_oracle_count(const pt *p1, long n1, const pt *p2, long n2, const pt *p3, long n3, const double *edges,
                    int nedges, const double *box, int survey, int contract, int nthreads, uint64_t *counts,
                    double *wsum);
long nc_oracle_sensitivity(const pt *p1, long n1, const pt *p2, long n2, const double *edges, int nedges,
                           const double *box, int survey);
int nc_oracle_compact(const double *edges, int nedges, const double *dense, double *out, int cap);
int nc_oracle_count_rppi(const pt *p1, long n1, const pt *p2, long n2, const pt *p3, long n3, const double *edges,
                         int nedges, int npi, const double *box, int survey, const double *ori, int contract,
                         int nthreads, uint64_t *counts, double *wsum, long *dropped);
int nc_oracle_compact_rppi(const double *edges, int nedges, int npi, const double *dense, double *out, int cap);

static uint64_t state = 88172645463325252ull;
static double uni(void) { /* xorshift64*, [0, 1) on a 1/1024 grid */
    state ^= state >> 12;
    state ^= state << 25;
    state ^= state >> 27;
    return (double)((state * 2685821657736338717ull) >> 54) / 1024.0;
}

static pt *points(long n, double L, int weighted) {
    pt *p = malloc(sizeof(pt) * (size_t)(n > 0 ? n : 1));
    for (long i = 0; i < n; ++i) {
        p[i].x = uni() * L;
        p[i].y = uni() * L;
        p[i].z = uni() * L;
        p[i].w = weighted ? 0.5 + uni() : 1.0;
    }
    return p;
}

static void threed(const char *name, const pt *a, long na, const pt *b, long nb, const pt *c, long nc,
                   const double *edges, int ne, double L, int survey) {
    const long nh = (long)(ne - 1) * (ne - 1) * (ne - 1);
    uint64_t *cnt = calloc((size_t)nh, sizeof(uint64_t));
    double *ws = calloc((size_t)nh, sizeof(double));
    const double box[3] = {L, L, L};
    const int rc = nc_oracle_count(a, na, b, nb, c, nc, edges, ne, box, survey, 0, 2, cnt, ws);
    uint64_t t = 0;
    double s = 0.0;
    for (long i = 0; i < nh; ++i) {
        t += cnt[i];
        s += ws[i];
    }
    printf("%s %d %llu %.17g\n", name, rc, (unsigned long long)t, s);
    free(cnt);
    free(ws);
}

int main(void) {
    const double e5[] = {0.0, 0.5, 1.0, 1.5, 2.0, 2.5};
    const double e3[] = {0.25, 1.0, 2.0, 3.0};
    pt *p = points(400, 12.0, 1), *q = points(300, 12.0, 0);
    threed("box_ppp", p, 400, p, 400, p, 400, e5, 6, 12.0, 0);
    threed("box_pqq", p, 400, q, 300, q, 300, e5, 6, 12.0, 0);
    threed("survey_pqp", p, 400, q, 300, p, 400, e5, 6, 12.5, 1);
    threed("box_one_cell", p, 60, p, 60, p, 60, e3, 4, 2.5, 0);   /* L < rmax: one cell */
    threed("box_two_cells", p, 80, q, 80, q, 80, e3, 4, 5.0, 0);  /* two cells, repeated images */
    threed("empty", p, 0, q, 300, q, 300, e5, 6, 12.0, 0);
    pt edge[4] = {{0.0, 0.0, 0.0, 1.0}, {2.5, 0.0, 0.0, 1.0}, {2.5, 2.5, 0.0, 1.0}, {12.0, 2.5, 2.5, 1.0}};
    threed("on_boundaries", edge, 4, edge, 4, edge, 4, e5, 6, 12.0, 0);
    const double bad_edges[] = {1.0, 0.5};
    threed("bad_edges", p, 10, p, 10, p, 10, bad_edges, 2, 12.0, 0);
    const double box[3] = {12.0, 12.0, 12.0};
    printf("sensitivity %ld\n", nc_oracle_sensitivity(p, 400, q, 300, e5, 6, box, 0));
    double dense[125], out[125];
    for (int i = 0; i < 125; ++i) dense[i] = i;
    const int nout = nc_oracle_compact(e5, 6, dense, out, 125);
    double cs = 0.0;
    for (int i = 0; i < nout; ++i) cs += out[i];
    printf("compact %d %.17g\n", nout, cs);
    const double ori[3] = {-30.0, 0.0, 0.0};
    const int npi = 3;
    const long nr = 5L * 5 * 5 * npi * npi;
    uint64_t *rc = calloc((size_t)nr, sizeof(uint64_t));
    double *rw = calloc((size_t)nr, sizeof(double));
    long dropped = 0;
    for (int survey = 0; survey <= 1; ++survey) {
        const int r = nc_oracle_count_rppi(p, 400, p, 400, q, 300, e5, 6, npi, box, survey, ori, 0, 2, rc, rw,
                                           &dropped);
        uint64_t t = 0;
        for (long i = 0; i < nr; ++i) t += rc[i];
        printf("rppi_%d %d %llu %ld\n", survey, r, (unsigned long long)t, dropped);
    }
    double *rout = malloc(sizeof(double) * (size_t)nr);
    double *rdense = malloc(sizeof(double) * (size_t)nr);
    for (long i = 0; i < nr; ++i) rdense[i] = (double)i;
    printf("compact_rppi %d\n", nc_oracle_compact_rppi(e5, 6, npi, rdense, rout, (int)nr));
    free(rc);
    free(rw);
    free(rout);
    free(rdense);
    free(p);
    free(q);
    return 0;
}
