// nc_vector_types.hpp -- plain-old-data vector types with the layout and
// member names of the CUDA <vector_types.h> types the reference's public
// headers use (include/cudaNcorr.hpp:5, include/survey.hpp:7): double2,
// double3, double4 (32 bytes, {x,y,z,w}) and int3, plus make_double3/4.
// Host-only: no HIP or CUDA header is included, so callers of the drop-in
// classes compile with any C++17 compiler.
#ifndef NC_VECTOR_TYPES_HPP
#define NC_VECTOR_TYPES_HPP

#if !defined(__HIPCC__) && !defined(__CUDACC__) && !defined(HIP_INCLUDE_HIP_HIP_VECTOR_TYPES_H)
struct double2 {
    double x, y;
};
struct double3 {
    double x, y, z;
};
struct alignas(16) double4 {
    double x, y, z, w;
};
struct int3 {
    int x, y, z;
};
inline double3 make_double3(double x, double y, double z) { return double3{x, y, z}; }
inline double4 make_double4(double x, double y, double z, double w) { return double4{x, y, z, w}; }
static_assert(sizeof(double4) == 32, "double4 must be 32 bytes");
#endif

#endif
