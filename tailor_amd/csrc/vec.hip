// vec.hip -- vector kernels of the shift-invert Krylov-Schur iteration:
// permutations between natural (k_j_dof, interleaved complex) and
// elimination (split planes) order, the block-diagonal B product, the BSR
// product K x = A x - mu B x, and the tall-skinny dot / update / combine
// kernels of classical Gram-Schmidt with reorthogonalisation.
#include "vec.hpp"

namespace tailor {
namespace {

constexpr int VT = 256;
constexpr int kMaxCols = 130;  // Krylov columns a kernel may take (ncv <= 128, plus one)

__global__ void k_nat_to_elim(long long n, const int *__restrict__ perm, const double *__restrict__ src,
                              double *__restrict__ dr, double *__restrict__ di) {
  const long long k = static_cast<long long>(blockIdx.x) * VT + threadIdx.x;
  if (k >= n) return;
  const long long g = perm[k];
  dr[k] = src[2 * g];
  di[k] = src[2 * g + 1];
}

__global__ void k_elim_to_nat(long long n, const int *__restrict__ perm, const double *__restrict__ sr,
                              const double *__restrict__ si, double *__restrict__ dst) {
  const long long k = static_cast<long long>(blockIdx.x) * VT + threadIdx.x;
  if (k >= n) return;
  const long long g = perm[k];
  dst[2 * g] = sr[k];
  dst[2 * g + 1] = si[k];
}

// y = B x, elimination order; order[q] = node at position q.
__global__ void k_bmul(long long nodes, const int *__restrict__ order, const double *__restrict__ bvals,
                       const double *__restrict__ xr, const double *__restrict__ xi,
                       double *__restrict__ yr, double *__restrict__ yi) {
  const long long t = static_cast<long long>(blockIdx.x) * VT + threadIdx.x;
  if (t >= nodes * 5) return;
  const long long q = t / 5;
  const int r = static_cast<int>(t - 5 * q);
  const double *b = bvals + 50 * static_cast<long long>(order[q]) + 10 * r;
  double sr = 0.0, si = 0.0;
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    const double br = b[2 * c], bi = b[2 * c + 1];
    const double x_r = xr[5 * q + c], x_i = xi[5 * q + c];
    sr += br * x_r - bi * x_i;
    si += br * x_i + bi * x_r;
  }
  yr[t] = sr;
  yi[t] = si;
}

// y = A x - mu B x in elimination order.  One wave per block row: the row's
// 77 x 25 entries are contiguous and each lane takes every 64th of them,
// accumulating per block-row entry r; the wave then reduces the five row
// sums.  xcol[k] is block k's column node in elimination order
// (pos[colidx[k]], built once per structure), so an entry's x needs one
// dependent load, and each lane issues kKmulSteps steps' loads (entries
// idx, idx + 64, ...) before using them; the sums run over a lane's entries
// in the same order either way.
// Steps of loads in flight per lane (200x100, 633 MB per application:
// 2 steps 133 us, 4 steps 139 us, 8 steps 150 us; one step, with the column
// position looked up through pos[colidx[k]], 174 us)
constexpr int kKmulSteps = 2;
__global__ __launch_bounds__(VT) void k_kmul(long long nodes, int per, const double *__restrict__ avals,
                                             const int *__restrict__ xcol,
                                             const double *__restrict__ bvals,
                                             const int *__restrict__ pos, double mu_re, double mu_im,
                                             const double *__restrict__ xr,
                                             const double *__restrict__ xi, double *__restrict__ yr,
                                             double *__restrict__ yi) {
  const long long u = static_cast<long long>(blockIdx.x) * (VT / 64) + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  if (u >= nodes) return;
  double ar[5] = {0, 0, 0, 0, 0}, ai[5] = {0, 0, 0, 0, 0};
  const double2 *arow = reinterpret_cast<const double2 *>(avals) + u * per * 25;
  const int *xrow = xcol + u * per;
  const int total = per * 25;
  auto add = [&](int idx, double2 a, double x_r, double x_i) {
    const int e = idx - 25 * (idx / 25), r = e / 5;
    const double pr = a.x * x_r - a.y * x_i, pi = a.x * x_i + a.y * x_r;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
      if (q == r) {
        ar[q] += pr;
        ai[q] += pi;
      }
    }
  };
  constexpr int U = kKmulSteps;
  int idx = lane;
  for (; idx + 64 * (U - 1) < total; idx += 64 * U) {
    double2 a[U];
    long long xc[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      const int i = idx + 64 * q, k = i / 25;
      a[q] = arow[i];
      xc[q] = 5LL * xrow[k] + (i - 25 * k) % 5;
    }
    double x_r[U], x_i[U];
#pragma unroll
    for (int q = 0; q < U; ++q) {
      x_r[q] = xr[xc[q]];
      x_i[q] = xi[xc[q]];
    }
#pragma unroll
    for (int q = 0; q < U; ++q) add(idx + 64 * q, a[q], x_r[q], x_i[q]);
  }
  for (; idx < total; idx += 64) {
    const int k = idx / 25;
    const long long x0 = 5LL * xrow[k] + (idx - 25 * k) % 5;
    add(idx, arow[idx], xr[x0], xi[x0]);
  }
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    for (int off = 32; off > 0; off >>= 1) {
      ar[q] += __shfl_xor(ar[q], off);
      ai[q] += __shfl_xor(ai[q], off);
    }
  }
  if (lane < 5) {
    const int r = lane;
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int q = 0; q < 5; ++q)
      if (q == r) {
        sr = ar[q];
        si = ai[q];
      }
    const long long pu = 5LL * pos[u];
    // - mu B x on the diagonal block
    const double *b = bvals + 50 * u + 10 * r;
    double bxr = 0.0, bxi = 0.0;
    for (int c = 0; c < 5; ++c) {
      const double br = b[2 * c], bi = b[2 * c + 1];
      bxr += br * xr[pu + c] - bi * xi[pu + c];
      bxi += br * xi[pu + c] + bi * xr[pu + c];
    }
    sr -= mu_re * bxr - mu_im * bxi;
    si -= mu_re * bxi + mu_im * bxr;
    yr[pu + r] = sr;
    yi[pu + r] = si;
  }
}

// Sum over the 64 lanes of a wave (fixed butterfly: every lane gets the
// same, run-to-run reproducible value).
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// partial h_j = sum_i conj(V_ij) w_i over this block's rows, for the
// kDotCols columns j of chunk blockIdx.y: each thread accumulates its rows
// of all the chunk's columns (w read once), then wave sums and one LDS step
// across the waves.  partials[b][j] complex, b = blockIdx.x.
constexpr int kDotCols = 8;
__global__ __launch_bounds__(VT) void k_dots(long long n, int ncols, const double *__restrict__ Vr,
                                             const double *__restrict__ Vi, long long ldv,
                                             const double *__restrict__ wr,
                                             const double *__restrict__ wi,
                                             double *__restrict__ partials) {
  constexpr int NW = VT / 64;
  __shared__ double red[NW][2 * kDotCols];
  const int j0 = blockIdx.y * kDotCols;
  const int nc = min(kDotCols, ncols - j0);
  double sr[kDotCols], si[kDotCols];
#pragma unroll
  for (int c = 0; c < kDotCols; ++c) {
    sr[c] = 0.0;
    si[c] = 0.0;
  }
  for (long long i = static_cast<long long>(blockIdx.x) * VT + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * VT) {
    const double x_r = wr[i], x_i = wi[i];
#pragma unroll
    for (int c = 0; c < kDotCols; ++c) {
      if (c < nc) {
        const long long o = static_cast<long long>(j0 + c) * ldv + i;
        const double vr = Vr[o], vi = Vi[o];
        sr[c] += vr * x_r + vi * x_i;
        si[c] += vr * x_i - vi * x_r;
      }
    }
  }
  const int w = threadIdx.x / 64, l = threadIdx.x % 64;
#pragma unroll
  for (int c = 0; c < kDotCols; ++c) {
    const double r = wave_sum(sr[c]), q = wave_sum(si[c]);
    if (l == 0) {
      red[w][2 * c] = r;
      red[w][2 * c + 1] = q;
    }
  }
  __syncthreads();
  if (threadIdx.x < 2 * nc) {
    double t = 0.0;
#pragma unroll
    for (int v = 0; v < NW; ++v) t += red[v][threadIdx.x];
    const int c = threadIdx.x / 2;
    partials[2 * (static_cast<long long>(blockIdx.x) * ncols + j0 + c) + threadIdx.x % 2] = t;
  }
}

// w -= V h (h complex, ncols <= kMaxCols, interleaved in device memory)
__global__ void k_update(long long n, int ncols, const double *__restrict__ Vr,
                         const double *__restrict__ Vi, long long ldv, const double *__restrict__ h,
                         double *__restrict__ wr, double *__restrict__ wi) {
  __shared__ double hr[kMaxCols], hi[kMaxCols];
  for (int t = threadIdx.x; t < ncols; t += blockDim.x) {
    hr[t] = h[2 * t];
    hi[t] = h[2 * t + 1];
  }
  __syncthreads();
  const long long i = static_cast<long long>(blockIdx.x) * VT + threadIdx.x;
  if (i >= n) return;
  double sr = 0.0, si = 0.0;
  for (int j = 0; j < ncols; ++j) {
    const double vr = Vr[j * ldv + i], vi = Vi[j * ldv + i];
    sr += vr * hr[j] - vi * hi[j];
    si += vr * hi[j] + vi * hr[j];
  }
  wr[i] -= sr;
  wi[i] -= si;
}

// Out[:, 0:k] = V[:, 0:m] C, C m x k column-major complex (interleaved).
// A workgroup covers VT rows and kCombineCols output columns (blockIdx.y
// selects the column chunk); its columns of C are staged in LDS and each
// output is summed over a = 0..m-1 in order.
constexpr int kCombineCols = 16;
__global__ __launch_bounds__(VT) void k_combine(long long n, int m, int k,
                                                const double *__restrict__ Vr,
                                                const double *__restrict__ Vi, long long ldv,
                                                const double *__restrict__ C,
                                                double *__restrict__ Or, double *__restrict__ Oi,
                                                long long ldo) {
  __shared__ double cr[kMaxCols * kCombineCols], ci[kMaxCols * kCombineCols];
  const int b0 = blockIdx.y * kCombineCols;
  const int kb = min(kCombineCols, k - b0);
  for (int t = threadIdx.x; t < m * kb; t += VT) {
    const int a = t % m, b = t / m;
    cr[b * m + a] = C[2 * ((b0 + b) * m + a)];
    ci[b * m + a] = C[2 * ((b0 + b) * m + a) + 1];
  }
  __syncthreads();
  const long long i = static_cast<long long>(blockIdx.x) * VT + threadIdx.x;
  if (i >= n) return;
  double sr[kCombineCols], si[kCombineCols];
#pragma unroll
  for (int b = 0; b < kCombineCols; ++b) {
    sr[b] = 0.0;
    si[b] = 0.0;
  }
  for (int a = 0; a < m; ++a) {
    const double vr = Vr[a * ldv + i], vi = Vi[a * ldv + i];
#pragma unroll
    for (int b = 0; b < kCombineCols; ++b) {
      if (b < kb) {
        const double c_r = cr[b * m + a], c_i = ci[b * m + a];
        sr[b] += vr * c_r - vi * c_i;
        si[b] += vr * c_i + vi * c_r;
      }
    }
  }
#pragma unroll
  for (int b = 0; b < kCombineCols; ++b) {
    if (b < kb) {
      Or[(b0 + b) * ldo + i] = sr[b];
      Oi[(b0 + b) * ldo + i] = si[b];
    }
  }
}

// out[j] = sum over blocks b of partials[b][j], complex: one wave per
// column j = blockIdx.x; lane l sums blocks l, l + 64, ... in order, then
// the wave's fixed butterfly.
__global__ __launch_bounds__(64) void k_reduce_partials(const double *__restrict__ partials,
                                                        int nblocks, int ncols,
                                                        double *__restrict__ out) {
  const int j = blockIdx.x, l = threadIdx.x;
  double sr = 0.0, si = 0.0;
  for (int b = l; b < nblocks; b += 64) {
    sr += partials[2 * (static_cast<long long>(b) * ncols + j)];
    si += partials[2 * (static_cast<long long>(b) * ncols + j) + 1];
  }
  sr = wave_sum(sr);
  si = wave_sum(si);
  if (l == 0) {
    out[2 * j] = sr;
    out[2 * j + 1] = si;
  }
}

// y = x / sqrt(max(0, Re s2[0])): normalisation by a norm computed on the
// device (s2 = x^H x from k_dots + k_reduce_partials).
__global__ void k_scale_inv_norm(long long n, const double *__restrict__ s2,
                                 const double *__restrict__ xr, const double *__restrict__ xi,
                                 double *__restrict__ yr, double *__restrict__ yi) {
  const long long i = static_cast<long long>(blockIdx.x) * VT + threadIdx.x;
  if (i >= n) return;
  const double inv = 1.0 / sqrt(fmax(0.0, s2[0]));
  yr[i] = inv * xr[i];
  yi[i] = inv * xi[i];
}

// y = alpha x + beta y (complex scalars)
__global__ void k_axpby(long long n, double a_r, double a_i, const double *__restrict__ xr,
                        const double *__restrict__ xi, double b_r, double b_i, double *__restrict__ yr,
                        double *__restrict__ yi) {
  const long long i = static_cast<long long>(blockIdx.x) * VT + threadIdx.x;
  if (i >= n) return;
  const double x_r = xr[i], x_i = xi[i], y_r = yr[i], y_i = yi[i];
  yr[i] = a_r * x_r - a_i * x_i + b_r * y_r - b_i * y_i;
  yi[i] = a_r * x_i + a_i * x_r + b_r * y_i + b_i * y_r;
}

}  // namespace

unsigned vblocks(long long n) { return static_cast<unsigned>((n + VT - 1) / VT); }

hipError_t vec_nat_to_elim(hipStream_t s, long long n, const int *perm, const double *src, double *dr,
                           double *di) {
  hipLaunchKernelGGL(k_nat_to_elim, dim3(vblocks(n)), dim3(VT), 0, s, n, perm, src, dr, di);
  return hipGetLastError();
}

hipError_t vec_elim_to_nat(hipStream_t s, long long n, const int *perm, const double *sr,
                           const double *si, double *dst) {
  hipLaunchKernelGGL(k_elim_to_nat, dim3(vblocks(n)), dim3(VT), 0, s, n, perm, sr, si, dst);
  return hipGetLastError();
}

hipError_t vec_bmul(hipStream_t s, long long nodes, const int *order, const double *bvals,
                    const double *xr, const double *xi, double *yr, double *yi) {
  hipLaunchKernelGGL(k_bmul, dim3(vblocks(nodes * 5)), dim3(VT), 0, s, nodes, order, bvals, xr, xi,
                     yr, yi);
  return hipGetLastError();
}

hipError_t vec_kmul(hipStream_t s, long long nodes, int per, const double *avals, const int *xcol,
                    const double *bvals, const int *pos, double mu_re, double mu_im,
                    const double *xr, const double *xi, double *yr, double *yi) {
  const unsigned g = static_cast<unsigned>((nodes + VT / 64 - 1) / (VT / 64));
  hipLaunchKernelGGL(k_kmul, dim3(g), dim3(VT), 0, s, nodes, per, avals, xcol, bvals, pos, mu_re,
                     mu_im, xr, xi, yr, yi);
  return hipGetLastError();
}

hipError_t vec_dots(hipStream_t s, long long n, int ncols, const double *Vr, const double *Vi,
                    long long ldv, const double *wr, const double *wi, double *partials,
                    int nblocks) {
  if (ncols <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_dots, dim3(nblocks, (ncols + kDotCols - 1) / kDotCols), dim3(VT), 0, s, n,
                     ncols, Vr, Vi, ldv, wr, wi, partials);
  return hipGetLastError();
}

hipError_t vec_update(hipStream_t s, long long n, int ncols, const double *Vr, const double *Vi,
                      long long ldv, const double *h, double *wr, double *wi) {
  if (ncols > kMaxCols) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_update, dim3(vblocks(n)), dim3(VT), 0, s, n, ncols, Vr, Vi, ldv, h, wr, wi);
  return hipGetLastError();
}

hipError_t vec_combine(hipStream_t s, long long n, int m, int k, const double *Vr, const double *Vi,
                       long long ldv, const double *C, double *Or, double *Oi, long long ldo) {
  if (m > kMaxCols || k <= 0) return k <= 0 ? hipSuccess : hipErrorInvalidValue;
  hipLaunchKernelGGL(k_combine, dim3(vblocks(n), (k + kCombineCols - 1) / kCombineCols), dim3(VT),
                     0, s, n, m, k, Vr, Vi, ldv, C, Or, Oi, ldo);
  return hipGetLastError();
}

hipError_t vec_dots_device(hipStream_t s, long long n, int ncols, const double *Vr,
                           const double *Vi, long long ldv, const double *wr, const double *wi,
                           double *partials, int nblocks, double *out) {
  if (ncols <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_dots, dim3(nblocks, (ncols + kDotCols - 1) / kDotCols), dim3(VT), 0, s, n,
                     ncols, Vr, Vi, ldv, wr, wi, partials);
  hipLaunchKernelGGL(k_reduce_partials, dim3(ncols), dim3(64), 0, s, partials, nblocks, ncols, out);
  return hipGetLastError();
}

hipError_t vec_scale_inv_norm(hipStream_t s, long long n, const double *s2, const double *xr,
                              const double *xi, double *yr, double *yi) {
  hipLaunchKernelGGL(k_scale_inv_norm, dim3(vblocks(n)), dim3(VT), 0, s, n, s2, xr, xi, yr, yi);
  return hipGetLastError();
}

hipError_t vec_axpby(hipStream_t s, long long n, double a_r, double a_i, const double *xr,
                     const double *xi, double b_r, double b_i, double *yr, double *yi) {
  hipLaunchKernelGGL(k_axpby, dim3(vblocks(n)), dim3(VT), 0, s, n, a_r, a_i, xr, xi, b_r, b_i, yr,
                     yi);
  return hipGetLastError();
}

}  // namespace tailor
