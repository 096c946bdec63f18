// Host-side declarations of the kernels of libncorr_hip.so, defined (with
// their launch bounds and explicit instantiations) in ncorr_device.hip.  Only
// the host translation units include this header.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ncorr_kernels.hpp"

using nck::GridGeo;
using nck::KArgs;
using nck::TermDesc;

__global__ void k_fill_u32(uint32_t *p, uint32_t v, long long n);
__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, long long n);
__global__ void k_cell_keys(const double4 *__restrict__ raw, int n, GridGeo g, int check_unclamped,
                            uint32_t *__restrict__ key, uint32_t *__restrict__ rank, uint32_t *__restrict__ cnt,
                            unsigned *__restrict__ err, unsigned *__restrict__ inexact, int expect_exact);
__global__ void k_scan(const uint32_t *__restrict__ cnt, uint32_t *__restrict__ start, int ncells);
__global__ void k_scan_tiles(const uint32_t *__restrict__ cnt, uint32_t *__restrict__ part);
__global__ void k_scan_apply(const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ pstart,
                             uint32_t *__restrict__ start, int ntiles);
__global__ void k_place(int n, const uint32_t *__restrict__ key, const uint32_t *__restrict__ rank,
                        const uint32_t *__restrict__ start, uint32_t *__restrict__ perm);
__global__ void k_stable_scatter(const double4 *__restrict__ raw, int n, const uint32_t *__restrict__ key,
                                 const uint32_t *__restrict__ start, const uint32_t *__restrict__ perm,
                                 double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
                                 double *__restrict__ w);
__global__ void k_nbr_bound(const uint32_t *__restrict__ st1, const uint32_t *__restrict__ st2, GridGeo g,
                            unsigned *__restrict__ out);
__global__ void k_weight_stats(const double4 *__restrict__ raw, long long n, unsigned *__restrict__ s);
__global__ void k_list_bound(KArgs A, TermDesc T, int p0, int p1, unsigned *__restrict__ out);
template <bool MIXED, int HM, bool WT>
__global__ void k_triplets(KArgs A);
template <bool SURVEY, bool STREAM>
__global__ void k_rppi(KArgs A);
__global__ void k_rppi_bin(KArgs A);

// Device address of the no-alloc scratch array on the current device.
hipError_t noalloc_symbol_address(void **p);
