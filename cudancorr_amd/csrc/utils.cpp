// utils.cpp -- log-spaced edges, random catalogues, catalogue readers and
// result writers for the drop-in include/utils.hpp (reference
// source/utils.cu:16-208).  Host code only.
//
// The FITS reader is written from the FITS standard (version 4.0): 2880-byte
// blocks of 80-character header cards, big-endian binary-table rows.  It
// reads what the reference's CFITSIO calls read (utils.cu:66-107): the first
// element of each named scalar column, scaled by TSCALn/TZEROn, as a double.
#include "utils.hpp"

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <thread>

// ---------------------------------------------------------------- helpers --

std::vector<double> logspace(double start, double end, int Nedge) {
    // utils.cu:16-27: 10^(log10(start) + i*step), step over Nedge-1 intervals
    if (Nedge < 0) throw std::length_error("logspace: Nedge is negative");
    const double lo = std::log10(start);
    const double step = (std::log10(end) - lo) / (Nedge - 1);
    std::vector<double> e(static_cast<size_t>(Nedge));
    for (int i = 0; i < Nedge; ++i) e[i] = std::pow(10.0, lo + i * step);
    return e;
}

double4 generateRandomPosition(std::mt19937 &gen, const double3 &box) {
    // utils.cu:29-39: x, y, z drawn in this order, each uniform in [0, box)
    std::uniform_real_distribution<double> ux(0.0, box.x), uy(0.0, box.y), uz(0.0, box.z);
    double4 p;
    p.x = ux(gen);
    p.y = uy(gen);
    p.z = uz(gen);
    p.w = 1.0;
    return p;
}

namespace {

// Points [b*kRandBlock, (b+1)*kRandBlock) come from one generator seeded by
// (seed, b), so the catalogue depends on the seed only: not on the thread
// count, and the first n points of a longer catalogue are the catalogue of n.
constexpr int kRandBlock = 1 << 16;

void fill_random(std::vector<double4> &p, const double3 &box, uint64_t seed) {
    const int n = static_cast<int>(p.size());
    const int nblocks = (n + kRandBlock - 1) / kRandBlock;
    auto work = [&](int first, int stride) {
        for (int b = first; b < nblocks; b += stride) {
            std::seed_seq seq{static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32), static_cast<uint32_t>(b)};
            std::mt19937 gen(seq);
            const int end = std::min(n, (b + 1) * kRandBlock);
            for (int i = b * kRandBlock; i < end; ++i) p[i] = generateRandomPosition(gen, box);
        }
    };
    const int nthreads = std::max(1, std::min<int>(nblocks, static_cast<int>(std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(work, t, nthreads);
    work(0, nthreads);
    for (auto &th : pool) th.join();
}

}  // namespace

std::vector<double4> generateRandomCatalog(int Np, const double3 &box, uint64_t seed) {
    if (Np < 0) throw std::length_error("generateRandomCatalog: Np is negative");
    std::vector<double4> p(static_cast<size_t>(Np));
    fill_random(p, box, seed);
    return p;
}

std::vector<double4> generateRandomCatalog(int Np, const double3 &box) {
    // utils.cu:41-52 seeds from the clock: a fresh catalogue on every call
    const uint64_t seed = static_cast<uint64_t>(std::chrono::system_clock::now().time_since_epoch().count()) ^
                          (static_cast<uint64_t>(std::random_device{}()) << 32);
    return generateRandomCatalog(Np, box, seed);
}

// ---------------------------------------------------------------- writers --

namespace {

void finish(std::ofstream &out, const std::string &filename) {
    out.close();
    if (out.fail()) throw std::runtime_error("Error writing file: " + filename);
}

}  // namespace

void writeToFileThreeD(const std::string &filename, const std::vector<double3> &double3_vec,
                       const std::vector<double> &double_vec) {
    // utils.cu:161-183: "r1 r2 r3 count", fixed with 6 decimals
    if (double3_vec.size() != double_vec.size()) throw std::runtime_error("Vectors have different sizes.");
    std::ofstream out(filename);
    if (!out.is_open()) throw std::runtime_error("Unable to open file: " + filename);
    out << std::fixed << std::setprecision(6);
    for (size_t i = 0; i < double_vec.size(); ++i)
        out << double3_vec[i].x << ' ' << double3_vec[i].y << ' ' << double3_vec[i].z << ' ' << double_vec[i] << '\n';
    finish(out, filename);
}

void writeToFileRppi(const std::string &filename, const std::vector<double3> &double3_vec,
                     const std::vector<double2> &double2_vec, const std::vector<double> &double_vec) {
    // utils.cu:185-208: "r1 r2 r3 pi1 pi2 count", fixed with 6 decimals
    if (double3_vec.size() != double_vec.size() || double2_vec.size() != double_vec.size())
        throw std::runtime_error("Vectors have different sizes.");
    std::ofstream out(filename);
    if (!out.is_open()) throw std::runtime_error("Unable to open file: " + filename);
    out << std::fixed << std::setprecision(6);
    for (size_t i = 0; i < double_vec.size(); ++i)
        out << double3_vec[i].x << ' ' << double3_vec[i].y << ' ' << double3_vec[i].z << ' ' << double2_vec[i].x << ' '
            << double2_vec[i].y << ' ' << double_vec[i] << '\n';
    finish(out, filename);
}

// ------------------------------------------------------------ text reader --

std::vector<double4> readCatalogTxt(const std::string &filename, const double3 &box,
                                    const std::vector<int> &pos_columns) {
    // utils.cu:119-159.  Column c of a line is the c-th number read from it
    // (reading stops at the first token that is not a number); the first of
    // x, y, z, w whose column number is c takes it.
    if (pos_columns.size() < 3)
        throw std::invalid_argument("readCatalogTxt: pos_columns needs the x, y and z column numbers");
    std::ifstream in(filename);
    if (!in.is_open()) throw std::runtime_error("Unable to open file: " + filename);
    const int cx = pos_columns[0], cy = pos_columns[1], cz = pos_columns[2];
    const int cw = pos_columns.size() > 3 ? pos_columns[3] : -1;

    std::vector<double4> out;
    std::string line;
    std::istringstream ls;
    size_t line_number = 0;
    while (std::getline(in, line)) {
        ++line_number;
        ls.clear();
        ls.str(line);
        double4 p;
        p.x = p.y = p.z = 0.0;
        p.w = 1.0;
        bool hx = false, hy = false, hz = false;
        int column = 0;
        double v;
        while (ls >> v) {
            if (column == cx) p.x = v, hx = true;
            else if (column == cy) p.y = v, hy = true;
            else if (column == cz) p.z = v, hz = true;
            else if (column == cw) p.w = v;
            ++column;
        }
        if (column == 0) continue;  // blank line, or a comment with no leading number
        if (!hx || !hy || !hz)
            throw std::runtime_error("Error at line " + std::to_string(line_number) +
                                     ": fewer columns than the x, y, z column numbers.");
        if (p.x < 0 || p.y < 0 || p.z < 0 || p.x > box.x || p.y > box.y || p.z > box.z)
            throw std::runtime_error("Error at line " + std::to_string(line_number) +
                                     ": double3 value is out of bounds.");
        out.push_back(p);
    }
    return out;
}

// ------------------------------------------------------------ FITS reader --

namespace {

constexpr size_t kFitsBlock = 2880;
constexpr size_t kCard = 80;

struct FitsHeader {
    std::vector<std::pair<std::string, std::string>> cards;  // keyword, value field (strings unquoted)

    const std::string *find(const std::string &key) const {
        for (const auto &c : cards)
            if (c.first == key) return &c.second;
        return nullptr;
    }
    bool has(const std::string &key) const { return find(key) != nullptr; }
    std::string str(const std::string &key) const {
        const std::string *v = find(key);
        return v ? *v : std::string();
    }
    // Integer keyword; `bad` set when absent (and no default) or not an integer.
    long long integer(const std::string &key, bool &bad, long long dflt = 0, bool required = true) const {
        const std::string *v = find(key);
        if (!v) {
            if (required) bad = true;
            return dflt;
        }
        char *end = nullptr;
        const long long x = std::strtoll(v->c_str(), &end, 10);
        if (end == v->c_str() || *end != '\0') bad = true;
        return x;
    }
    // Real keyword; FITS allows a 'D' exponent.
    double real(const std::string &key, double dflt, bool &bad) const {
        const std::string *v = find(key);
        if (!v) return dflt;
        std::string s = *v;
        for (char &ch : s)
            if (ch == 'D' || ch == 'd') ch = 'E';
        char *end = nullptr;
        const double x = std::strtod(s.c_str(), &end);
        if (end == s.c_str() || *end != '\0') bad = true;
        return x;
    }
};

std::string rtrim(std::string s) {
    while (!s.empty() && s.back() == ' ') s.pop_back();
    return s;
}

std::string trim(const std::string &s) {
    size_t a = 0;
    while (a < s.size() && s[a] == ' ') ++a;
    return rtrim(s.substr(a));
}

// Value field of a card (columns 11-80 after "= " in columns 9-10): a quoted
// string ('' is a quote; trailing blanks dropped), or the text before a '/'
// comment, trimmed.
std::string card_value(const char *card) {
    std::string f(card + 10, card + kCard);
    size_t i = 0;
    while (i < f.size() && f[i] == ' ') ++i;
    if (i < f.size() && f[i] == '\'') {
        std::string s;
        for (++i; i < f.size(); ++i) {
            if (f[i] == '\'') {
                if (i + 1 < f.size() && f[i + 1] == '\'') {
                    s.push_back('\'');
                    ++i;
                } else {
                    break;
                }
            } else {
                s.push_back(f[i]);
            }
        }
        return rtrim(s);
    }
    const size_t slash = f.find('/', i);
    return trim(f.substr(i, slash == std::string::npos ? std::string::npos : slash - i));
}

// Reads header blocks at the stream position up to the END card.  Returns
// false at end of file before any card, or on a truncated header.
bool read_header(std::ifstream &in, FitsHeader &h) {
    h.cards.clear();
    char block[kFitsBlock];
    bool first = true;
    for (;;) {
        in.read(block, kFitsBlock);
        if (in.gcount() != static_cast<std::streamsize>(kFitsBlock)) return false;
        for (size_t c = 0; c < kFitsBlock; c += kCard) {
            const char *card = block + c;
            std::string key = rtrim(std::string(card, card + 8));
            if (first && key.empty()) return false;
            first = false;
            if (key == "END") return true;
            if (card[8] == '=' && card[9] == ' ') h.cards.emplace_back(key, card_value(card));
        }
    }
}

// Data bytes after a header, padded to whole blocks (FITS 4.0 eqs. 1-2; a
// random-groups primary array has NAXIS1 = 0 and no axis-1 factor).
bool data_bytes(const FitsHeader &h, unsigned long long &bytes) {
    bool bad = false;
    const long long bitpix = h.integer("BITPIX", bad);
    const long long naxis = h.integer("NAXIS", bad);
    const long long pcount = h.integer("PCOUNT", bad, 0, false);
    const long long gcount = h.integer("GCOUNT", bad, 1, false);
    if (bad || naxis < 0 || naxis > 999 || pcount < 0 || gcount < 0) return false;
    const bool groups = h.str("GROUPS") == "T";
    unsigned long long prod = naxis > 0 ? 1 : 0;
    for (long long n = 1; n <= naxis; ++n) {
        const long long len = h.integer("NAXIS" + std::to_string(n), bad);
        if (bad || len < 0) return false;
        if (n == 1 && groups && len == 0) continue;
        prod *= static_cast<unsigned long long>(len);
    }
    const unsigned long long bits = static_cast<unsigned long long>(bitpix < 0 ? -bitpix : bitpix) *
                                    static_cast<unsigned long long>(gcount) *
                                    (static_cast<unsigned long long>(pcount) + prod);
    bytes = (bits / 8 + kFitsBlock - 1) / kFitsBlock * kFitsBlock;
    return true;
}

struct Column {
    std::string name;
    char type = 0;
    long long repeat = 0;
    size_t offset = 0;  // byte offset in the row
    double scale = 1.0, zero = 0.0;
};

// Binary-table column layout (FITS 4.0 sec. 7.3): TFORMn = rT..., widths
// from the type letter; the widths must add up to NAXIS1.
bool table_layout(const FitsHeader &h, long long &row_bytes, long long &nrows, std::vector<Column> &cols) {
    bool bad = false;
    row_bytes = h.integer("NAXIS1", bad);
    nrows = h.integer("NAXIS2", bad);
    const long long nfields = h.integer("TFIELDS", bad);
    if (bad || row_bytes < 0 || nrows < 0 || nfields < 0 || nfields > 999) return false;
    cols.assign(static_cast<size_t>(nfields), Column{});
    size_t off = 0;
    for (long long n = 1; n <= nfields; ++n) {
        Column &c = cols[n - 1];
        const std::string idx = std::to_string(n);
        const std::string form = trim(h.str("TFORM" + idx));
        size_t i = 0;
        while (i < form.size() && std::isdigit(static_cast<unsigned char>(form[i]))) ++i;
        if (i == form.size()) return false;
        c.repeat = i ? std::stoll(form.substr(0, i)) : 1;
        c.type = static_cast<char>(std::toupper(static_cast<unsigned char>(form[i])));
        long long width;
        switch (c.type) {
            case 'L': case 'B': case 'A': width = c.repeat; break;
            case 'X': width = (c.repeat + 7) / 8; break;
            case 'I': width = 2 * c.repeat; break;
            case 'J': case 'E': width = 4 * c.repeat; break;
            case 'K': case 'D': case 'C': case 'P': width = 8 * c.repeat; break;
            case 'M': case 'Q': width = 16 * c.repeat; break;
            default: return false;
        }
        c.name = rtrim(h.str("TTYPE" + idx));
        c.offset = off;
        c.scale = h.real("TSCAL" + idx, 1.0, bad);
        c.zero = h.real("TZERO" + idx, 0.0, bad);
        off += static_cast<size_t>(width);
    }
    return !bad && static_cast<long long>(off) == row_bytes;
}

bool ieq(char a, char b) {
    return std::toupper(static_cast<unsigned char>(a)) == std::toupper(static_cast<unsigned char>(b));
}

// Case-insensitive column-name template: '*' any run of characters, '?' one
// character, '#' a run of one or more decimal digits.
bool name_matches(const char *t, const char *s) {
    for (; *t; ++t, ++s) {
        if (*t == '*') {
            for (const char *k = s;; ++k) {
                if (name_matches(t + 1, k)) return true;
                if (!*k) return false;
            }
        }
        if (*t == '#') {
            if (!std::isdigit(static_cast<unsigned char>(*s))) return false;
            for (const char *k = s + 1;; ++k) {
                if (name_matches(t + 1, k)) return true;
                if (!std::isdigit(static_cast<unsigned char>(*k))) return false;
            }
        }
        if (!*s) return false;
        if (*t != '?' && !ieq(*t, *s)) return false;
    }
    return *s == '\0';
}

// First element of a scalar numeric column in one big-endian row.
double column_value(const Column &c, const unsigned char *row) {
    const unsigned char *p = row + c.offset;
    uint64_t u = 0;
    double raw;
    switch (c.type) {
        case 'B': raw = p[0]; break;
        case 'I': raw = static_cast<int16_t>(static_cast<uint16_t>(p[0] << 8 | p[1])); break;
        case 'J': {
            const uint32_t w = uint32_t(p[0]) << 24 | uint32_t(p[1]) << 16 | uint32_t(p[2]) << 8 | p[3];
            raw = static_cast<int32_t>(w);
            break;
        }
        case 'E': {
            const uint32_t w = uint32_t(p[0]) << 24 | uint32_t(p[1]) << 16 | uint32_t(p[2]) << 8 | p[3];
            float f;
            std::memcpy(&f, &w, 4);
            raw = f;
            break;
        }
        case 'K':
        case 'D':
            for (int b = 0; b < 8; ++b) u = u << 8 | p[b];
            if (c.type == 'K') {
                raw = static_cast<double>(static_cast<int64_t>(u));
            } else {
                std::memcpy(&raw, &u, 8);
            }
            break;
        default: raw = 0.0;  // excluded by readable()
    }
    if (c.scale == 1.0 && c.zero == 0.0) return raw;
    return raw * c.scale + c.zero;
}

bool readable(const Column &c) {
    return c.repeat > 0 && std::strchr("BIJKED", c.type) != nullptr;
}

}  // namespace

std::vector<double4> readCatalogFits(const std::string &filename, const std::vector<std::string> &column_names,
                                     int hdu_index, double zmin, double zmax) {
    // utils.cu:54-115
    std::ifstream in(filename, std::ios::binary);
    FitsHeader h;
    if (!in.is_open() || !read_header(in, h) || h.cards.empty() || h.cards.front().first != "SIMPLE" ||
        h.cards.front().second != "T")
        throw std::runtime_error("Unable to open FITS file: " + filename);

    // utils.cu:73-76: hdu_index is absolute, 1 = the primary array
    if (hdu_index < 1) throw std::runtime_error("Error moving to the specified HDU.");
    for (int k = 1; k < hdu_index; ++k) {
        unsigned long long skip = 0;
        if (!data_bytes(h, skip)) throw std::runtime_error("Error moving to the specified HDU.");
        in.seekg(static_cast<std::streamoff>(skip), std::ios::cur);
        if (!in || !read_header(in, h) || h.cards.empty() || h.cards.front().first != "XTENSION")
            throw std::runtime_error("Error moving to the specified HDU.");
    }

    // utils.cu:78-81
    const std::string xt = hdu_index == 1 ? std::string() : h.str("XTENSION");
    if (xt != "BINTABLE" && xt != "A3DTABLE") throw std::runtime_error("The specified HDU is not a Binary Table.");
    long long row_bytes = 0, nrows = 0;
    std::vector<Column> cols;
    if (!table_layout(h, row_bytes, nrows, cols)) throw std::runtime_error("Error moving to the specified HDU.");

    std::vector<double4> result;
    if (nrows == 0) return result;  // utils.cu:88: no row, so no column is looked up

    // utils.cu:91-102: each name is looked up, then read, in the order given;
    // a name matching no column or several is an error
    std::vector<const Column *> sel(column_names.size());
    for (size_t j = 0; j < column_names.size(); ++j) {
        const Column *hit = nullptr;
        int nhit = 0;
        for (const Column &c : cols)
            if (name_matches(column_names[j].c_str(), c.name.c_str()) && nhit++ == 0) hit = &c;
        if (nhit != 1) throw std::runtime_error("Error finding column named " + column_names[j]);
        if (!readable(*hit)) throw std::runtime_error("Error reading FITS file data.");
        sel[j] = hit;
    }

    // Rows in chunks of about 64 MiB.
    const long long rows_per_chunk = std::max<long long>(1, (64ll << 20) / std::max<long long>(row_bytes, 1));
    std::vector<unsigned char> buf;
    for (long long r0 = 0; r0 < nrows; r0 += rows_per_chunk) {
        const long long nr = std::min(rows_per_chunk, nrows - r0);
        buf.resize(static_cast<size_t>(nr * row_bytes));
        in.read(reinterpret_cast<char *>(buf.data()), static_cast<std::streamsize>(buf.size()));
        if (in.gcount() != static_cast<std::streamsize>(buf.size()))
            throw std::runtime_error("Error reading FITS file data.");
        for (long long r = 0; r < nr; ++r) {
            const unsigned char *row = buf.data() + r * row_bytes;
            double4 p;
            p.x = p.y = p.z = 0.0;
            p.w = 1.0;
            for (size_t j = 0; j < sel.size(); ++j) {
                const std::string &name = column_names[j];
                if (name == "RA") p.x = column_value(*sel[j], row);
                else if (name == "DEC") p.y = column_value(*sel[j], row);
                else if (name == "Z") p.z = column_value(*sel[j], row);
                else if (name == "WEIGHT") p.w = column_value(*sel[j], row);
            }
            if (p.z > zmin && p.z < zmax) result.push_back(p);  // utils.cu:108-110
        }
    }
    return result;
}
