// vec.hpp -- launchers of the Krylov vector kernels (vec.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace tailor {

constexpr int kDotBlocks = 256;  // partial-sum blocks of vec_dots

hipError_t vec_nat_to_elim(hipStream_t s, long long n, const int *perm, const double *src, double *dr,
                           double *di);
hipError_t vec_elim_to_nat(hipStream_t s, long long n, const int *perm, const double *sr,
                           const double *si, double *dst);
hipError_t vec_bmul(hipStream_t s, long long nodes, const int *order, const double *bvals,
                    const double *xr, const double *xi, double *yr, double *yi);
// y = (A - mu B) x in elimination order; xcol[k]: block k's column node in
// elimination order, pos: node -> elimination position
hipError_t vec_kmul(hipStream_t s, long long nodes, int per, const double *avals, const int *xcol,
                    const double *bvals, const int *pos, double mu_re, double mu_im,
                    const double *xr, const double *xi, double *yr, double *yi);
hipError_t vec_dots(hipStream_t s, long long n, int ncols, const double *Vr, const double *Vi,
                    long long ldv, const double *wr, const double *wi, double *partials,
                    int nblocks);
hipError_t vec_update(hipStream_t s, long long n, int ncols, const double *Vr, const double *Vi,
                      long long ldv, const double *h, double *wr, double *wi);
hipError_t vec_combine(hipStream_t s, long long n, int m, int k, const double *Vr, const double *Vi,
                       long long ldv, const double *C, double *Or, double *Oi, long long ldo);
hipError_t vec_axpby(hipStream_t s, long long n, double a_r, double a_i, const double *xr,
                     const double *xi, double b_r, double b_i, double *yr, double *yi);
// out[j] = V[:, j]^H w for j < ncols, summed on the device (partials: scratch
// of 2 nblocks ncols doubles; out: ncols interleaved complex).
hipError_t vec_dots_device(hipStream_t s, long long n, int ncols, const double *Vr,
                           const double *Vi, long long ldv, const double *wr, const double *wi,
                           double *partials, int nblocks, double *out);
// y = x / sqrt(Re s2[0]) with s2 in device memory.
hipError_t vec_scale_inv_norm(hipStream_t s, long long n, const double *s2, const double *xr,
                              const double *xi, double *yr, double *yi);

}  // namespace tailor
