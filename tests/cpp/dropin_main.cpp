// Test driver for the drop-in C++ classes (include/cudaNcorr.hpp,
// include/survey.hpp), written the way reference callers use them
// (source/main.cu:20-95).  Built by the tests with g++ against libncorr.so.
//
//   dropin_main ctor                     constructor / bin-list checks (no GPU)
//   dropin_main survey                   prep_cat_survey and comoving distance (no GPU)
//   dropin_main count <cat1> <cat2> <cat3> <survey> <precision> <bx> <by> <bz> <edge0> <edge1> ...
//       catN: binary files of n x double4; the same path passed twice means the
//       same std::vector object.  Prints getTriCountsOut(), one value per line.
//   dropin_main countbins ...            as count, each line "count mid1 mid2 mid3" from
//       getTriCountsOut() and getTriBinsOut() at the same index.
//   dropin_main rppi <cat1> <cat2> <cat3> <survey> <npi> <ori_x> <bx> <by> <bz> <edge0> <edge1> ...
//       the same with rppi binning (rp edges, npi pi bins).
#include <cstdio>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "cudaNcorr.hpp"
#include "survey.hpp"

static int fails = 0;
#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);     \
            ++fails;                                                    \
        }                                                               \
    } while (0)

template <class E, class F>
static bool throws(F f) {
    try {
        f();
    } catch (const E &) {
        return true;
    } catch (...) {
        return false;
    }
    return false;
}

static std::vector<double> lin(double lo, double hi, int nb) {
    std::vector<double> e;
    for (int i = 0; i <= nb; ++i) e.push_back(lo + i * ((hi - lo) / nb));
    return e;
}

static int mode_ctor() {
    double3 box = {250.0, 250.0, 250.0};
    struct Case {
        double lo, hi;
        int nb;
        size_t nout;
    } cases[] = {{2, 20, 10, 150}, {0, 30, 15, 372}, {0, 40, 20, 825}, {0, 50, 25, 1547}};
    for (auto &c : cases) {
        std::vector<double> e = lin(c.lo, c.hi, c.nb);
        ncorr corr(box, 0, e, false, {0.0, 0.0, 0.0});
        CHECK(corr.getTriBinsOut().size() == c.nout);
        CHECK(corr.getTriCountsOut().size() == c.nout);
        CHECK(corr.getPiBinsOut().empty());
        auto b = corr.getTriBinsOut();
        for (auto &t : b) CHECK(t.x <= t.y && t.y <= t.z && t.z <= t.x + t.y);
    }
    std::vector<double> e = lin(0, 30, 15);
    CHECK(throws<std::runtime_error>([&] { ncorr c(box, 1, e, false, {0, 0, 0}); }));
    CHECK(throws<std::runtime_error>([&] { ncorr c(box, 2, e, false, {0, 0, 0}); }));
    CHECK(throws<std::runtime_error>([&] { ncorr c(box, 0, e, 8, false, {0, 0, 0}); }));
    // rppi: the threeD selection of rp bins, each with npi^2 pi bins (cudaNcorr.cu:198-212)
    for (auto &c : cases) {
        std::vector<double> er = lin(c.lo, c.hi, c.nb);
        const int npi = 7;
        ncorr corr(box, 1, er, npi, false, {0.0, 0.0, 0.0});
        CHECK(corr.getTriBinsOut().size() == c.nout * npi * npi);
        CHECK(corr.getTriCountsOut().size() == c.nout * npi * npi);
        auto pb = corr.getPiBinsOut();
        CHECK(pb.size() == c.nout * npi * npi);
        CHECK(pb[0].x == 0.0 && pb[0].y == 0.0 && pb[1].y == 1.0 && pb[npi].x == 1.0);
        CHECK(pb[npi * npi - 1].x == npi - 1.0 && pb[npi * npi - 1].y == npi - 1.0);
        auto tb = corr.getTriBinsOut();
        CHECK(tb[0].z == tb[npi * npi - 1].z && tb[npi * npi].z != tb[0].z);
    }
    CHECK(throws<std::invalid_argument>([&] { ncorr c(box, 1, e, 0, false, {0, 0, 0}); }));
    std::vector<double> bad = {0.0, 2.0, 1.0};
    CHECK(throws<std::invalid_argument>([&] { ncorr c(box, 0, bad, false, {0, 0, 0}); }));
    std::vector<double> one = {1.0};
    CHECK(throws<std::invalid_argument>([&] { ncorr c(box, 0, one, false, {0, 0, 0}); }));
    std::printf(fails ? "FAILED\n" : "OK\n");
    return fails ? 1 : 0;
}

static int mode_survey() {
    cosmology fid{67.36, 0.313772, 0.686228};
    survey s(fid);
    std::vector<double4> d = {{10.0, -5.0, 3.0, 1.0}, {12.0, 1.0, 7.5, 2.0}};
    std::vector<double4> r = {{9.0, 0.0, 4.0, 1.0}, {11.0, 2.0, 20.0, 1.0}};
    double3 box, ori;
    s.prep_cat_survey(d, r, box, ori);
    std::printf("box %.17g %.17g %.17g\n", box.x, box.y, box.z);
    std::printf("ori %.17g %.17g %.17g\n", ori.x, ori.y, ori.z);
    std::printf("d0 %.17g %.17g %.17g\n", d[0].x, d[0].y, d[0].z);
    std::printf("r1 %.17g %.17g %.17g\n", r[1].x, r[1].y, r[1].z);
    for (double z : {0.1, 0.5, 1.0, 2.0, 3.0}) std::printf("dc %.3f %.17g\n", z, s.comoving_distance(z));
    std::vector<double4> radec = {{30.0, 20.0, 0.5, 1.5}};
    auto xyz = s.convert_coordinates(radec);
    std::printf("xyz %.17g %.17g %.17g %.17g\n", xyz[0].x, xyz[0].y, xyz[0].z, xyz[0].w);
    return 0;
}

static std::vector<double4> load(const std::string &path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open " + path);
    const size_t bytes = (size_t)f.tellg();
    std::vector<double4> v(bytes / sizeof(double4));
    f.seekg(0);
    f.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(double4)));
    return v;
}

static int mode_count(int argc, char **argv, bool with_bins) {
    std::map<std::string, std::vector<double4>> cats;
    std::string p[3] = {argv[2], argv[3], argv[4]};
    for (auto &s : p)
        if (!cats.count(s)) cats[s] = load(s);
    const bool surveyMode = std::stoi(argv[5]) != 0;
    const int prec = std::stoi(argv[6]);
    double3 box = {std::stod(argv[7]), std::stod(argv[8]), std::stod(argv[9])};
    std::vector<double> edges;
    for (int i = 10; i < argc; ++i) edges.push_back(std::stod(argv[i]));
    ncorr corr(box, 0, edges, surveyMode, {0.0, 0.0, 0.0});
    corr.setPrecision(prec ? ncorr::mixedFP32 : ncorr::plainFP64);
    corr.calculateD1D2D3(cats[p[0]], cats[p[1]], cats[p[2]]);
    const std::vector<double> c = corr.getTriCountsOut();
    const std::vector<double3> b = corr.getTriBinsOut();
    if (b.size() != c.size()) return 4;
    for (size_t i = 0; i < c.size(); ++i) {
        if (with_bins)
            std::printf("%.17g %.17g %.17g %.17g\n", c[i], b[i].x, b[i].y, b[i].z);
        else
            std::printf("%.17g\n", c[i]);
    }
    return 0;
}

static int mode_rppi(int argc, char **argv) {
    std::map<std::string, std::vector<double4>> cats;
    std::string p[3] = {argv[2], argv[3], argv[4]};
    for (auto &s : p)
        if (!cats.count(s)) cats[s] = load(s);
    const bool surveyMode = std::stoi(argv[5]) != 0;
    const int npi = std::stoi(argv[6]);
    const double3 ori = {std::stod(argv[7]), 0.0, 0.0};
    double3 box = {std::stod(argv[8]), std::stod(argv[9]), std::stod(argv[10])};
    std::vector<double> edges;
    for (int i = 11; i < argc; ++i) edges.push_back(std::stod(argv[i]));
    ncorr corr(box, 1, edges, npi, surveyMode, ori);
    corr.calculateD1D2D3(cats[p[0]], cats[p[1]], cats[p[2]]);
    for (double v : corr.getTriCountsOut()) std::printf("%.17g\n", v);
    return 0;
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    std::string m = argv[1];
    try {
        if (m == "ctor") return mode_ctor();
        if (m == "survey") return mode_survey();
        if (m == "count" && argc >= 12) return mode_count(argc, argv, false);
        if (m == "countbins" && argc >= 12) return mode_count(argc, argv, true);
        if (m == "rppi" && argc >= 13) return mode_rppi(argc, argv);
    } catch (const std::invalid_argument &e) {
        std::printf("EXCEPTION invalid_argument %s\n", e.what());
        return 3;
    } catch (const std::runtime_error &e) {
        std::printf("EXCEPTION runtime_error %s\n", e.what());
        return 3;
    } catch (const std::exception &e) {
        std::printf("EXCEPTION %s\n", e.what());
        return 3;
    }
    return 2;
}
