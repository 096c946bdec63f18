// survey.cpp -- the drop-in survey helper (include/survey.hpp).
//
// prep_cat_survey / find_min_max_elements / shift_catalog restate
// source/survey.cu:60-128 (joint min/max, box = max - min + 1.0, ori = -min,
// in-place shift by -min).  convert_coordinates restates survey.cu:39-57;
// the comoving distance (survey.cu:26-37, GSL QAGS with epsrel 1e-7) is
// computed here with an adaptive Gauss-Legendre rule of our own, converged to
// ~1e-12 relative, so it agrees with any 1e-7-accurate quadrature.
#include <algorithm>
#include <cmath>
#include <limits>

#include "survey.hpp"

namespace {

constexpr double kSpeedOfLight = 299792.458;  // km/s
constexpr double kDegToRad = M_PI / 180.0;

// Nodes/weights of 10-point Gauss-Legendre on [-1, 1], from Newton iteration
// on the Legendre polynomial P_10 (standard construction).
struct GaussLegendre {
    static constexpr int n = 10;
    double x[n], w[n];
    GaussLegendre() {
        for (int i = 0; i < n; ++i) {
            double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
            for (int it = 0; it < 100; ++it) {
                double p0 = 1.0, p1 = z;
                for (int k = 2; k <= n; ++k) {
                    double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                double dp = n * (z * p1 - p0) / (z * z - 1.0);
                double dz = p1 / dp;
                z -= dz;
                if (std::fabs(dz) < 1e-16) {
                    x[i] = z;
                    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
                    break;
                }
                x[i] = z;
                w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
            }
        }
    }
};

const GaussLegendre &gl() {
    static const GaussLegendre g;
    return g;
}

template <class F>
double gl_panel(F &&f, double a, double b) {
    const GaussLegendre &g = gl();
    const double h = 0.5 * (b - a), c = 0.5 * (a + b);
    double s = 0.0;
    for (int i = 0; i < GaussLegendre::n; ++i) s += g.w[i] * f(c + h * g.x[i]);
    return s * h;
}

template <class F>
double adaptive(F &&f, double a, double b, double whole, double tol, int depth) {
    const double m = 0.5 * (a + b);
    const double l = gl_panel(f, a, m), r = gl_panel(f, m, b);
    if (depth <= 0 || std::fabs(l + r - whole) <= tol) return l + r;
    return adaptive(f, a, m, l, 0.5 * tol, depth - 1) + adaptive(f, m, b, r, 0.5 * tol, depth - 1);
}

}  // namespace

survey::survey(const cosmology &cosmo) : cosmo_(cosmo) {}

double survey::comoving_distance(double z) {
    const double Om = cosmo_.OmegaM, OL = cosmo_.OmegaLambda;
    auto f = [Om, OL](double zz) { return 1.0 / std::sqrt(Om * std::pow(1.0 + zz, 3) + OL); };
    if (z == 0.0) return 0.0;
    const double whole = gl_panel(f, 0.0, z);
    const double I = adaptive(f, 0.0, z, whole, 1e-13 * std::fabs(whole), 40);
    return kSpeedOfLight * I / 100.0;  // Mpc/h
}

std::vector<double4> survey::convert_coordinates(const std::vector<double4> &catalog) {
    std::vector<double4> out(catalog.size());
#pragma omp parallel for
    for (long i = 0; i < (long)catalog.size(); ++i) {
        const double ra = catalog[i].x * kDegToRad;
        const double dec = catalog[i].y * kDegToRad;
        const double d = comoving_distance(catalog[i].z);
        out[i] = make_double4(d * std::cos(dec) * std::cos(ra), d * std::cos(dec) * std::sin(ra), d * std::sin(dec),
                              catalog[i].w);
    }
    return out;
}

void survey::find_min_max_elements(const std::vector<double4> &data, double3 &catMin, double3 &catMax) {
    double mn[3] = {std::numeric_limits<double>::max(), std::numeric_limits<double>::max(),
                    std::numeric_limits<double>::max()};
    double mx[3] = {std::numeric_limits<double>::lowest(), std::numeric_limits<double>::lowest(),
                    std::numeric_limits<double>::lowest()};
    for (const double4 &e : data) {
        mn[0] = std::min(mn[0], e.x);
        mn[1] = std::min(mn[1], e.y);
        mn[2] = std::min(mn[2], e.z);
        mx[0] = std::max(mx[0], e.x);
        mx[1] = std::max(mx[1], e.y);
        mx[2] = std::max(mx[2], e.z);
    }
    catMin = make_double3(mn[0], mn[1], mn[2]);
    catMax = make_double3(mx[0], mx[1], mx[2]);
}

void survey::shift_catalog(std::vector<double4> &catalog, double3 catMin) {
    for (double4 &e : catalog) {
        e.x -= catMin.x;
        e.y -= catMin.y;
        e.z -= catMin.z;
    }
}

void survey::prep_cat_survey(std::vector<double4> &catalog, double3 &box, double3 &ori) {
    double3 mn, mx;
    find_min_max_elements(catalog, mn, mx);
    box = make_double3(mx.x - mn.x + 1.0, mx.y - mn.y + 1.0, mx.z - mn.z + 1.0);
    ori = make_double3(-mn.x, -mn.y, -mn.z);
    shift_catalog(catalog, mn);
}

void survey::prep_cat_survey(std::vector<double4> &catalog1, std::vector<double4> &catalog2, double3 &box,
                             double3 &ori) {
    double3 mn1, mx1, mn2, mx2;
    find_min_max_elements(catalog1, mn1, mx1);
    find_min_max_elements(catalog2, mn2, mx2);
    const double3 mn = make_double3(std::min(mn1.x, mn2.x), std::min(mn1.y, mn2.y), std::min(mn1.z, mn2.z));
    const double3 mx = make_double3(std::max(mx1.x, mx2.x), std::max(mx1.y, mx2.y), std::max(mx1.z, mx2.z));
    box = make_double3(mx.x - mn.x + 1.0, mx.y - mn.y + 1.0, mx.z - mn.z + 1.0);
    ori = make_double3(-mn.x, -mn.y, -mn.z);
    shift_catalog(catalog1, mn);
    shift_catalog(catalog2, mn);
}
