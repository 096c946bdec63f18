// utils.hpp -- drop-in replacement of the reference's helper header
// (include/utils.hpp:9-44): log-spaced edges, random catalogues, the text and
// FITS catalogue readers and the result writers.  Host C++ only (libncorr.so).
#ifndef _UTILS_HPP_
#define _UTILS_HPP_

#include <cstdint>
#include <random>
#include <string>
#include <vector>

#include "nc_vector_types.hpp"

// Nedge edges 10^(log10(start) + i*step), step = (log10(end) - log10(start))/(Nedge-1)
// (utils.cu:16-27).
std::vector<double> logspace(double start, double end, int Nedge);

// Np points uniform in [0, box) with w = 1 (utils.cu:41-52).  The reference
// seeds each OpenMP thread from the clock; this two-argument form does the
// same (so it is not reproducible).  The three-argument form is reproducible:
// the same (Np, box, seed) gives the same catalogue on every machine and
// thread count.
std::vector<double4> generateRandomCatalog(int Np, const double3 &box);
std::vector<double4> generateRandomCatalog(int Np, const double3 &box, uint64_t seed);

// One uniform point of the box drawn from gen, w = 1 (utils.cu:29-39).
double4 generateRandomPosition(std::mt19937 &gen, const double3 &box);

// "x y z count" per line, fixed with 6 decimals (utils.cu:161-184).
void writeToFileThreeD(const std::string &filename, const std::vector<double3> &double3_vec,
                       const std::vector<double> &double_vec);

// "x y z pi1 pi2 count" per line, fixed with 6 decimals (utils.cu:186-208).
void writeToFileRppi(const std::string &filename, const std::vector<double3> &double3_vec,
                     const std::vector<double2> &double2_vec, const std::vector<double> &double_vec);

// Binary-table catalogue reader (utils.cu:54-115): the columns named
// "RA", "DEC", "Z", "WEIGHT" among column_names (matched case-insensitively in
// the file) go to x, y, z, w of the rows with zmin < Z < zmax.  hdu_index is
// the absolute HDU number (1 = primary).  Written from the FITS standard
// (no CFITSIO): scalar columns of type B, I, J, K, E, D with TSCALn/TZEROn.
// w is 1 when WEIGHT is not requested.
std::vector<double4> readCatalogFits(const std::string &filename, const std::vector<std::string> &column_names,
                                     int hdu_index, double zmin, double zmax);

// Whitespace-separated text catalogue (utils.cu:119-159): columns
// pos_columns[0..3] of each line are x, y, z, w; a point outside [0, box]
// throws std::runtime_error naming the line.  Blank lines are skipped; a
// line without the x, y or z column throws; w is 1 when its column is absent.
std::vector<double4> readCatalogTxt(const std::string &filename, const double3 &box,
                                    const std::vector<int> &pos_columns);

#endif
