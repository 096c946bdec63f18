// Test driver for the include/utils.hpp helpers (reference
// source/utils.cu:16-208), called the way source/main.cu:20-95 calls them.
// Built by the tests with g++ against libncorr.so; needs no GPU.
//
//   utils_main logspace <start> <end> <n>                 edges, one per line
//   utils_main randpos <seed> <n> <bx> <by> <bz>          n generateRandomPosition draws from mt19937(seed)
//   utils_main randcat <out.bin> <np> <bx> <by> <bz> [seed]   catalogue as n x double4
//   utils_main write3d <out.txt> <in.bin> [extra]         in.bin: n x (r1 r2 r3 count); extra: one more count
//   utils_main writerppi <out.txt> <in.bin> [extra]       in.bin: n x (r1 r2 r3 pi1 pi2 count)
//   utils_main readtxt <file> <bx> <by> <bz> <c0> <c1> <c2> [c3]
//   utils_main readfits <file> <hdu> <zmin> <zmax> <name>...
// Points are printed "x y z w" with 17 significant digits; an exception
// prints "EXCEPTION <what>" and exits with 3.
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "utils.hpp"

static std::vector<double> load_doubles(const std::string &path) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::vector<double> v(static_cast<size_t>(f.tellg()) / sizeof(double));
    f.seekg(0);
    f.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(double)));
    return v;
}

static void print_points(const std::vector<double4> &p) {
    for (const double4 &q : p) std::printf("%.17g %.17g %.17g %.17g\n", q.x, q.y, q.z, q.w);
}

static double3 box_arg(char **a) { return {std::stod(a[0]), std::stod(a[1]), std::stod(a[2])}; }

static int run(int argc, char **argv) {
    const std::string m = argv[1];
    if (m == "logspace" && argc == 5) {
        for (double e : logspace(std::stod(argv[2]), std::stod(argv[3]), std::stoi(argv[4]))) std::printf("%.17g\n", e);
        return 0;
    }
    if (m == "randpos" && argc == 7) {
        std::mt19937 gen(static_cast<uint32_t>(std::stoul(argv[2])));
        const int n = std::stoi(argv[3]);
        const double3 box = box_arg(argv + 4);
        std::vector<double4> p;
        for (int i = 0; i < n; ++i) p.push_back(generateRandomPosition(gen, box));
        print_points(p);
        return 0;
    }
    if (m == "randcat" && (argc == 7 || argc == 8)) {
        const double3 box = box_arg(argv + 4);
        const int np = std::stoi(argv[3]);
        std::vector<double4> p = argc == 8 ? generateRandomCatalog(np, box, std::stoull(argv[7]))
                                           : generateRandomCatalog(np, box);
        std::ofstream f(argv[2], std::ios::binary);
        f.write(reinterpret_cast<const char *>(p.data()), static_cast<std::streamsize>(p.size() * sizeof(double4)));
        return f ? 0 : 1;
    }
    if ((m == "write3d" || m == "writerppi") && (argc == 4 || argc == 5)) {
        const bool rppi = m == "writerppi";
        const size_t k = rppi ? 6 : 4;
        const std::vector<double> v = load_doubles(argv[3]);
        std::vector<double3> r;
        std::vector<double2> pi;
        std::vector<double> c;
        for (size_t i = 0; i + k <= v.size(); i += k) {
            r.push_back({v[i], v[i + 1], v[i + 2]});
            if (rppi) pi.push_back({v[i + 3], v[i + 4]});
            c.push_back(v[i + k - 1]);
        }
        if (argc == 5) c.push_back(0.0);
        if (rppi)
            writeToFileRppi(argv[2], r, pi, c);
        else
            writeToFileThreeD(argv[2], r, c);
        return 0;
    }
    if (m == "readtxt" && (argc == 9 || argc == 10)) {
        std::vector<int> cols;
        for (int i = 6; i < argc; ++i) cols.push_back(std::stoi(argv[i]));
        print_points(readCatalogTxt(argv[2], box_arg(argv + 3), cols));
        return 0;
    }
    if (m == "readfits" && argc >= 6) {
        std::vector<std::string> names(argv + 6, argv + argc);
        print_points(readCatalogFits(argv[2], names, std::stoi(argv[3]), std::stod(argv[4]), std::stod(argv[5])));
        return 0;
    }
    return 2;
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    try {
        return run(argc, argv);
    } catch (const std::exception &e) {
        std::printf("EXCEPTION %s\n", e.what());
        return 3;
    }
}
