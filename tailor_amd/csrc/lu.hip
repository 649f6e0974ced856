// lu.hip -- multifrontal LU of K = A_bc - sigma B_bc on gfx950 and the
// triangular solves that drive the shift-invert Krylov-Schur iteration.
//
// This replaces PCLU (PETSc native or MUMPS) behind STSINVERT
// (EigenSolver.cpp:61-67).  Design:
//
// * every front is a dense column-major complex matrix held in HBM for the
//   whole factorization (no update-matrix stack: 288 GB leaves room), stored
//   as separate real and imaginary planes so the update GEMMs load operands
//   with unit stride;
// * fronts of one elimination-tree level are factored together: each kernel
//   of the blocked right-looking LU (panel, row TRSM, trailing GEMM) takes the
//   level's front list and covers all of them in one launch;
// * pivoting is partial, restricted to the fully summed rows of the front
//   (threshold-free LAPACK-style row choice inside the pivot block);
// * besides L and U the factorization keeps E = L11^-1 P (the row operations
//   of the pivot block applied to the identity) and Z = U11^-1, so each
//   triangular solve is a sequence of per-level batched GEMVs with no serial
//   dependency inside a front.
//
// Arena layout of front f at offset off (doubles):
//   F_re[m*m] F_im[m*m] E_re[p*p] E_im[p*p] Z_re[p*p] Z_im[p*p]
#include <cfloat>
#include <cmath>

#include "lu.hpp"

namespace tailor {
namespace {

__device__ inline double cabs1(double re, double im) { return fabs(re) + fabs(im); }

// ---------------------------------------------------------------- set-up

__global__ void k_fill(double *__restrict__ p, long long n, double v) {
  for (long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += static_cast<long long>(gridDim.x) * blockDim.x)
    p[k] = v;
}

// E = I and Z = I for every front (the planes were zero-filled before).
__global__ void k_init_identity(const FrontDev *__restrict__ fr, int nf, double *__restrict__ arena) {
  const int f = blockIdx.y;
  if (f >= nf) return;
  const FrontDev d = fr[f];
  double *E = arena + d.off + 2LL * d.m * d.m;
  double *Z = E + 2LL * d.p * d.p;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < d.p; k += gridDim.x * blockDim.x) {
    E[static_cast<long long>(k) * d.p + k] = 1.0;
    Z[static_cast<long long>(k) * d.p + k] = 1.0;
  }
}

// Scatter K = A - sigma B into the fronts.  Thread per (block, entry).
// amap[3k..3k+2] = front, local row node, local col node of BSR block k.
__global__ void k_assemble_fronts(long long nnzb, int per, const double *__restrict__ avals,
                                  const int *__restrict__ colidx, const double *__restrict__ bvals,
                                  double sig_re, double sig_im, const int *__restrict__ amap,
                                  const FrontDev *__restrict__ fr, double *__restrict__ arena) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nnzb * 25) return;
  const long long k = t / 25;
  const int e = static_cast<int>(t - 25 * k);
  const int f = amap[3 * k], lu = amap[3 * k + 1], lv = amap[3 * k + 2];
  const FrontDev d = fr[f];
  const int r = e / 5, c = e - 5 * (e / 5);
  double vr = avals[2 * (k * 25 + e)], vi = avals[2 * (k * 25 + e) + 1];
  const long long node = k / per;
  if (colidx[k] == node) {
    const double br = bvals[2 * (node * 25 + e)], bi = bvals[2 * (node * 25 + e) + 1];
    vr -= sig_re * br - sig_im * bi;
    vi -= sig_re * bi + sig_im * br;
  }
  const long long idx = static_cast<long long>(5 * lv + c) * d.m + 5 * lu + r;
  arena[d.off + idx] = vr;
  arena[d.off + static_cast<long long>(d.m) * d.m + idx] = vi;
}

// F_parent[rel, rel] += F_child[p:, p:] for the children listed in `tasks`.
__global__ void k_extend_add(const int *__restrict__ tasks, int ntasks,
                             const FrontDev *__restrict__ fr, const int *__restrict__ par,
                             const int *__restrict__ rel, const long long *__restrict__ rel_off,
                             double *__restrict__ arena) {
  const int ti = blockIdx.y;
  if (ti >= ntasks) return;
  const int ch = tasks[ti];
  const FrontDev dc = fr[ch], dp = fr[par[ch]];
  const int *rl = rel + rel_off[ch];
  const long long cc = static_cast<long long>(dc.c) * dc.c;
  const double *Cr = arena + dc.off, *Ci = Cr + static_cast<long long>(dc.m) * dc.m;
  double *Pr = arena + dp.off, *Pi = Pr + static_cast<long long>(dp.m) * dp.m;
  for (long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; idx < cc;
       idx += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(idx % dc.c), j = static_cast<int>(idx / dc.c);
    const long long src = static_cast<long long>(dc.p + j) * dc.m + dc.p + i;
    const int pi = 5 * rl[i / 5] + i % 5, pj = 5 * rl[j / 5] + j % 5;
    const long long dst = static_cast<long long>(pj) * dp.m + pi;
    Pr[dst] += Cr[src];
    Pi[dst] += Ci[src];
  }
}

// ------------------------------------------------------------ factorization

constexpr int kPanelThreads = 1024;

// Unblocked LU of the panel columns [k0, k0+kb) over the fully summed rows
// [k0, p) of each front; pivot rows are chosen among those rows.  Row
// exchanges touch only the panel columns here (k_trsm_lower applies them to
// the columns right of the panel and to E); the update rows [p, m) of the
// panel are finished by k_l21.
__global__ __launch_bounds__(kPanelThreads) void k_panel(const int *__restrict__ list, int nlist,
                                                         const FrontDev *__restrict__ fr, int k0,
                                                         int nb, double tiny, double perturb,
                                                         double *__restrict__ arena,
                                                         int *__restrict__ ipiv,
                                                         int *__restrict__ nperturbed) {
  const int f = list[blockIdx.x];
  const FrontDev d = fr[f];
  if (k0 >= d.p) return;
  const int kb = min(nb, d.p - k0);
  const long long mm = static_cast<long long>(d.m) * d.m;
  double *Fr = arena + d.off, *Fi = Fr + mm;
  __shared__ double sv[kPanelThreads];
  __shared__ int si[kPanelThreads];
  __shared__ double prow_r[kLuNb], prow_i[kLuNb];
  __shared__ double inv_r, inv_i;
  const int tid = threadIdx.x;
  for (int j = k0; j < k0 + kb; ++j) {
    const long long cj = static_cast<long long>(j) * d.m;
    double best = -1.0;
    int bi = j;
    for (int r = j + tid; r < d.p; r += kPanelThreads) {
      const double a = cabs1(Fr[cj + r], Fi[cj + r]);
      if (a > best) {
        best = a;
        bi = r;
      }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int w = kPanelThreads / 2; w > 0; w >>= 1) {
      if (tid < w) {
        const double o = sv[tid + w];
        const int oi = si[tid + w];
        if (o > sv[tid] || (o == sv[tid] && oi < si[tid])) {
          sv[tid] = o;
          si[tid] = oi;
        }
      }
      __syncthreads();
    }
    const int piv = si[0];
    const bool small = !(sv[0] > tiny);
    __syncthreads();
    if (!small && piv != j && tid < kb) {
      const long long col = static_cast<long long>(k0 + tid) * d.m;
      double t = Fr[col + j]; Fr[col + j] = Fr[col + piv]; Fr[col + piv] = t;
      t = Fi[col + j]; Fi[col + j] = Fi[col + piv]; Fi[col + piv] = t;
    }
    __syncthreads();
    if (tid == 0) {
      ipiv[d.piv_off + j] = small ? j : piv;
      double pr = Fr[cj + j], pi = Fi[cj + j];
      if (small) {
        // static pivot perturbation: refinement recovers the accuracy
        pr = pr >= 0.0 ? perturb : -perturb;
        pi = 0.0;
        Fr[cj + j] = pr;
        Fi[cj + j] = pi;
        atomicAdd(nperturbed, 1);
      }
      const double den = pr * pr + pi * pi;
      inv_r = pr / den;
      inv_i = -pi / den;
    }
    if (tid < k0 + kb - j - 1) {
      const long long cc = static_cast<long long>(j + 1 + tid) * d.m + j;
      prow_r[tid] = Fr[cc];
      prow_i[tid] = Fi[cc];
    }
    __syncthreads();
    const double ir = inv_r, ii = inv_i;
    const int ncol = k0 + kb - j - 1;
    for (int r = j + 1 + tid; r < d.p; r += kPanelThreads) {
      const double ar = Fr[cj + r], ai = Fi[cj + r];
      const double lr = ar * ir - ai * ii, li = ar * ii + ai * ir;
      Fr[cj + r] = lr;
      Fi[cj + r] = li;
      for (int q = 0; q < ncol; ++q) {
        const long long cq = static_cast<long long>(j + 1 + q) * d.m + r;
        const double ur = prow_r[q], ui = prow_i[q];
        Fr[cq] -= lr * ur - li * ui;
        Fi[cq] -= lr * ui + li * ur;
      }
    }
    __syncthreads();
  }
}

// Update rows [p, m) of the panel columns: X <- X U_D^-1 with U_D the upper
// kb x kb pivot block (row-wise forward substitution, thread per row).
__global__ __launch_bounds__(256) void k_l21(const int *__restrict__ list, int nlist,
                                             const FrontDev *__restrict__ fr, int k0, int nb,
                                             double *__restrict__ arena) {
  const int f = list[blockIdx.y];
  const FrontDev d = fr[f];
  if (k0 >= d.p || d.c == 0) return;
  const int kb = min(nb, d.p - k0);
  const long long mm = static_cast<long long>(d.m) * d.m;
  double *Fr = arena + d.off, *Fi = Fr + mm;
  __shared__ double Ur[kLuNb * kLuNb], Ui[kLuNb * kLuNb], Dr[kLuNb], Di[kLuNb];
  for (int k = threadIdx.x; k < kb * kb; k += blockDim.x) {
    const int r = k % kb, s = k / kb;
    const long long idx = static_cast<long long>(k0 + s) * d.m + k0 + r;
    Ur[s * kLuNb + r] = Fr[idx];
    Ui[s * kLuNb + r] = Fi[idx];
  }
  __syncthreads();
  if (threadIdx.x < kb) {
    const double pr = Ur[threadIdx.x * kLuNb + threadIdx.x], pi = Ui[threadIdx.x * kLuNb + threadIdx.x];
    const double den = pr * pr + pi * pi;
    Dr[threadIdx.x] = pr / den;
    Di[threadIdx.x] = -pi / den;
  }
  __syncthreads();
  const int row = d.p + blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= d.m) return;
  double xr[kLuNb], xi[kLuNb];
#pragma unroll
  for (int j = 0; j < kLuNb; ++j) {
    if (j < kb) {
      const long long idx = static_cast<long long>(k0 + j) * d.m + row;
      double ar = Fr[idx], ai = Fi[idx];
#pragma unroll
      for (int i = 0; i < kLuNb; ++i) {
        if (i < j) {
          const double ur = Ur[j * kLuNb + i], ui = Ui[j * kLuNb + i];
          ar -= xr[i] * ur - xi[i] * ui;
          ai -= xr[i] * ui + xi[i] * ur;
        }
      }
      xr[j] = ar * Dr[j] - ai * Di[j];
      xi[j] = ar * Di[j] + ai * Dr[j];
      Fr[idx] = xr[j];
      Fi[idx] = xi[j];
    }
  }
}

// Rows [k0, k0+kb) of the F columns right of the panel and of all E columns:
// apply the panel's row exchanges to the column, then X <- L_D^-1 X with L_D
// the unit lower kb x kb pivot block.
__global__ __launch_bounds__(256) void k_trsm_lower(const int *__restrict__ list, int nlist,
                                                    const FrontDev *__restrict__ fr, int k0, int nb,
                                                    const int *__restrict__ ipiv,
                                                    double *__restrict__ arena) {
  const int f = list[blockIdx.y];
  const FrontDev d = fr[f];
  if (k0 >= d.p) return;
  const int kb = min(nb, d.p - k0);
  const long long mm = static_cast<long long>(d.m) * d.m;
  double *Fr = arena + d.off, *Fi = Fr + mm;
  double *Er = Fi + mm, *Ei = Er + static_cast<long long>(d.p) * d.p;
  __shared__ double Lr[kLuNb * kLuNb], Li[kLuNb * kLuNb];
  __shared__ int pv[kLuNb];
  for (int k = threadIdx.x; k < kb * kb; k += blockDim.x) {
    const int r = k % kb, s = k / kb;
    const long long idx = static_cast<long long>(k0 + s) * d.m + k0 + r;
    Lr[s * kLuNb + r] = Fr[idx];
    Li[s * kLuNb + r] = Fi[idx];
  }
  if (threadIdx.x < kb) pv[threadIdx.x] = ipiv[d.piv_off + k0 + threadIdx.x];
  __syncthreads();
  const int nF = d.m - k0 - kb;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= nF + d.p) return;
  double *Xr, *Xi;
  if (col < nF) {
    const long long base = static_cast<long long>(k0 + kb + col) * d.m;
    Xr = Fr + base;
    Xi = Fi + base;
  } else {
    const long long base = static_cast<long long>(col - nF) * d.p;
    Xr = Er + base;
    Xi = Ei + base;
  }
  for (int j = 0; j < kb; ++j) {
    const int pr = pv[j], r = k0 + j;
    if (pr != r) {
      double t = Xr[r]; Xr[r] = Xr[pr]; Xr[pr] = t;
      t = Xi[r]; Xi[r] = Xi[pr]; Xi[pr] = t;
    }
  }
  Xr += k0;
  Xi += k0;
  double xr[kLuNb], xi[kLuNb];
#pragma unroll
  for (int r = 0; r < kLuNb; ++r) {
    if (r < kb) {
      xr[r] = Xr[r];
      xi[r] = Xi[r];
    }
  }
#pragma unroll
  for (int r = 1; r < kLuNb; ++r) {
    if (r < kb) {
      double ar = xr[r], ai = xi[r];
#pragma unroll
      for (int s = 0; s < kLuNb; ++s) {
        if (s < r) {
          const double lr = Lr[s * kLuNb + r], li = Li[s * kLuNb + r];
          ar -= lr * xr[s] - li * xi[s];
          ai -= lr * xi[s] + li * xr[s];
        }
      }
      xr[r] = ar;
      xi[r] = ai;
    }
  }
#pragma unroll
  for (int r = 0; r < kLuNb; ++r) {
    if (r < kb) {
      Xr[r] = xr[r];
      Xi[r] = xi[r];
    }
  }
}

// Rows [k0, k0+kb) of Z columns [k0, p): X <- U_D^-1 X with U_D the upper
// (non-unit) kb x kb pivot block.
__global__ __launch_bounds__(256) void k_trsm_upper(const int *__restrict__ list, int nlist,
                                                    const FrontDev *__restrict__ fr, int k0, int nb,
                                                    double *__restrict__ arena) {
  const int f = list[blockIdx.y];
  const FrontDev d = fr[f];
  if (k0 >= d.p) return;
  const int kb = min(nb, d.p - k0);
  const long long mm = static_cast<long long>(d.m) * d.m;
  const double *Fr = arena + d.off, *Fi = Fr + mm;
  double *Zr = arena + d.off + 2 * mm + 2LL * d.p * d.p;
  double *Zi = Zr + static_cast<long long>(d.p) * d.p;
  __shared__ double Ur[kLuNb * kLuNb], Ui[kLuNb * kLuNb];
  for (int k = threadIdx.x; k < kb * kb; k += blockDim.x) {
    const int r = k % kb, s = k / kb;
    const long long idx = static_cast<long long>(k0 + s) * d.m + k0 + r;
    Ur[s * kLuNb + r] = Fr[idx];
    Ui[s * kLuNb + r] = Fi[idx];
  }
  __syncthreads();
  const int col = k0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= d.p) return;
  double *Xr = Zr + static_cast<long long>(col) * d.p + k0;
  double *Xi = Zi + static_cast<long long>(col) * d.p + k0;
  double xr[kLuNb], xi[kLuNb];
#pragma unroll
  for (int r = 0; r < kLuNb; ++r) {
    if (r < kb) {
      xr[r] = Xr[r];
      xi[r] = Xi[r];
    }
  }
#pragma unroll
  for (int r = kLuNb - 1; r >= 0; --r) {
    if (r < kb) {
      double ar = xr[r], ai = xi[r];
#pragma unroll
      for (int s = 0; s < kLuNb; ++s) {
        if (s > r && s < kb) {
          const double ur = Ur[s * kLuNb + r], ui = Ui[s * kLuNb + r];
          ar -= ur * xr[s] - ui * xi[s];
          ai -= ur * xi[s] + ui * xr[s];
        }
      }
      const double dr = Ur[r * kLuNb + r], di = Ui[r * kLuNb + r];
      const double den = dr * dr + di * di;
      xr[r] = (ar * dr + ai * di) / den;
      xi[r] = (ai * dr - ar * di) / den;
    }
  }
#pragma unroll
  for (int r = 0; r < kLuNb; ++r) {
    if (r < kb) {
      Xr[r] = xr[r];
      Xi[r] = xi[r];
    }
  }
}

// Batched complex GEMM  C -= A B  on the update regions of one panel step:
//   region 0: F[k0+kb:m, k0+kb:m] -= F[k0+kb:m, panel] F[panel, k0+kb:m]
//   region 1: E[k0+kb:p, :]       -= F[k0+kb:p, panel] E[panel, :]
//   region 2: Z[0:k0, k0:p]       -= F[0:k0, panel]    Z[panel, k0:p]
// 64x64 output tile per workgroup, 4x4 complex outputs per thread, K staged
// through LDS in chunks of 16.
constexpr int GT = 64, GK = 16;

__global__ __launch_bounds__(256) void k_gemm_update(const int *__restrict__ list, int nlist,
                                                     const FrontDev *__restrict__ fr, int k0, int nb,
                                                     int region, double *__restrict__ arena) {
  const int f = list[blockIdx.y];
  const FrontDev d = fr[f];
  if (k0 >= d.p) return;
  const int kb = min(nb, d.p - k0);
  const long long mm = static_cast<long long>(d.m) * d.m, pp = static_cast<long long>(d.p) * d.p;
  double *Fr = arena + d.off, *Fi = Fr + mm;
  double *Er = Fi + mm, *Ei = Er + pp;
  double *Zr = Ei + pp, *Zi = Zr + pp;
  int M, N;
  const double *Ar, *Ai, *Br, *Bi;
  double *Cr, *Ci;
  long long lda, ldb, ldc;
  if (region == 0) {
    M = N = d.m - k0 - kb;
    Ar = Fr + static_cast<long long>(k0) * d.m + k0 + kb;
    Ai = Fi + static_cast<long long>(k0) * d.m + k0 + kb;
    Br = Fr + static_cast<long long>(k0 + kb) * d.m + k0;
    Bi = Fi + static_cast<long long>(k0 + kb) * d.m + k0;
    Cr = Fr + static_cast<long long>(k0 + kb) * d.m + k0 + kb;
    Ci = Fi + static_cast<long long>(k0 + kb) * d.m + k0 + kb;
    lda = ldb = ldc = d.m;
  } else if (region == 1) {
    M = d.p - k0 - kb;
    N = d.p;
    Ar = Fr + static_cast<long long>(k0) * d.m + k0 + kb;
    Ai = Fi + static_cast<long long>(k0) * d.m + k0 + kb;
    Br = Er + k0;
    Bi = Ei + k0;
    Cr = Er + k0 + kb;
    Ci = Ei + k0 + kb;
    lda = d.m;
    ldb = ldc = d.p;
  } else {
    M = k0;
    N = d.p - k0;
    Ar = Fr + static_cast<long long>(k0) * d.m;
    Ai = Fi + static_cast<long long>(k0) * d.m;
    Br = Zr + static_cast<long long>(k0) * d.p + k0;
    Bi = Zi + static_cast<long long>(k0) * d.p + k0;
    Cr = Zr + static_cast<long long>(k0) * d.p;
    Ci = Zi + static_cast<long long>(k0) * d.p;
    lda = d.m;
    ldb = ldc = d.p;
  }
  if (M <= 0 || N <= 0) return;
  const int tiles_m = (M + GT - 1) / GT, tiles_n = (N + GT - 1) / GT;
  const int tile = blockIdx.x;
  if (tile >= tiles_m * tiles_n) return;
  const int tm = tile % tiles_m, tn = tile / tiles_m;
  const int row0 = tm * GT, col0 = tn * GT;
  __shared__ double sAr[GK][GT + 1], sAi[GK][GT + 1], sBr[GK][GT + 1], sBi[GK][GT + 1];
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  double cr[4][4], ci[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) cr[a][b] = ci[a][b] = 0.0;
  for (int kk = 0; kk < kb; kk += GK) {
    for (int e = threadIdx.x; e < GT * GK; e += 256) {
      const int r = e % GT, k = e / GT;
      const bool ok = row0 + r < M && kk + k < kb;
      const long long ia = static_cast<long long>(kk + k) * lda + row0 + r;
      sAr[k][r] = ok ? Ar[ia] : 0.0;
      sAi[k][r] = ok ? Ai[ia] : 0.0;
      const int cb = e / GK, kb2 = e % GK;
      const bool okb = col0 + cb < N && kk + kb2 < kb;
      const long long ib = static_cast<long long>(col0 + cb) * ldb + kk + kb2;
      sBr[kb2][cb] = okb ? Br[ib] : 0.0;
      sBi[kb2][cb] = okb ? Bi[ib] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < GK; ++k) {
      double ar[4], ai[4], br[4], bi[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        ar[a] = sAr[k][ty + 16 * a];
        ai[a] = sAi[k][ty + 16 * a];
        br[a] = sBr[k][tx + 16 * a];
        bi[a] = sBi[k][tx + 16 * a];
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          cr[a][b] += ar[a] * br[b] - ai[a] * bi[b];
          ci[a][b] += ar[a] * bi[b] + ai[a] * br[b];
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int r = row0 + ty + 16 * a;
    if (r >= M) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int cidx = col0 + tx + 16 * b;
      if (cidx >= N) continue;
      const long long ic = static_cast<long long>(cidx) * ldc + r;
      Cr[ic] -= cr[a][b];
      Ci[ic] -= ci[a][b];
    }
  }
}

// ------------------------------------------------------------------ solves
//
// Vectors are split complex planes in elimination (dof) order:
// position 5*pos[node] + dof.  Per-front workspace W (m entries per front).

// Forward set-up of a level: W_f = [b_f; 0] + sum over children of their
// update vectors (extend-add).
__global__ void k_fwd_init(const int *__restrict__ list, int nlist, const FrontDev *__restrict__ fr,
                           const int *__restrict__ child_ptr, const int *__restrict__ child_list,
                           const int *__restrict__ rel, const long long *__restrict__ rel_off,
                           const double *__restrict__ br, const double *__restrict__ bi,
                           double *__restrict__ Wr, double *__restrict__ Wi,
                           double *__restrict__ Yr, double *__restrict__ Yi) {
  const int f = list[blockIdx.x];
  const FrontDev d = fr[f];
  double *wr = Wr + d.woff, *wi = Wi + d.woff;
  for (int k = threadIdx.x; k < d.m; k += blockDim.x) {
    wr[k] = k < d.p ? br[d.dof0 + k] : 0.0;
    wi[k] = k < d.p ? bi[d.dof0 + k] : 0.0;
    if (k < d.p) {
      Yr[d.dof0 + k] = 0.0;  // output of the E product
      Yi[d.dof0 + k] = 0.0;
    }
  }
  __syncthreads();
  for (int q = child_ptr[f]; q < child_ptr[f + 1]; ++q) {
    const int ch = child_list[q];
    const FrontDev dc = fr[ch];
    const int *rl = rel + rel_off[ch];
    const double *ur = Wr + dc.woff + dc.p, *ui = Wi + dc.woff + dc.p;
    for (int k = threadIdx.x; k < dc.c; k += blockDim.x) {
      const int t = 5 * rl[k / 5] + k % 5;
      wr[t] += ur[k];
      wi[t] += ui[k];
    }
    __syncthreads();
  }
}

// Backward set-up of a level: W_f[p:] = x[update rows of f] (ancestors,
// already solved), T_f = y_f, and x_f = 0 (accumulated by the Z product).
__global__ void k_bwd_gather(const int *__restrict__ list, int nlist, const FrontDev *__restrict__ fr,
                             const int *__restrict__ upd_pos, const long long *__restrict__ upd_off,
                             double *__restrict__ xr, double *__restrict__ xi,
                             double *__restrict__ Wr, double *__restrict__ Wi,
                             const double *__restrict__ Yr, const double *__restrict__ Yi,
                             double *__restrict__ Tr, double *__restrict__ Ti) {
  const int f = list[blockIdx.x];
  const FrontDev d = fr[f];
  const int *up = upd_pos + upd_off[f];
  for (int k = threadIdx.x; k < d.c; k += blockDim.x) {
    const int g = 5 * up[k / 5] + k % 5;
    Wr[d.woff + d.p + k] = xr[g];
    Wi[d.woff + d.p + k] = xi[g];
  }
  for (int k = threadIdx.x; k < d.p; k += blockDim.x) {
    Tr[d.dof0 + k] = Yr[d.dof0 + k];
    Ti[d.dof0 + k] = Yi[d.dof0 + k];
    xr[d.dof0 + k] = 0.0;
    xi[d.dof0 + k] = 0.0;
  }
}

// Batched complex GEMV over the fronts of a level, y (+)= alpha * A x.
//   op 0: y = b0 + E W[0:p]          (E = L11^-1 P; y written to the
//                                     solution plane at the front's dofs)
//   op 1: W[p:m] -= L21 y            (y read from the solution plane)
//   op 2: T = y - U12 W[p:m]         (T in the scratch plane at dof0)
//   op 3: x = Z T                    (x written at the front's dofs)
// Output rows are split over blockIdx.x (GVR rows per workgroup) and the
// reduction dimension over blockIdx.z (GVC columns); partial sums are added
// with FP64 atomics into an output that ops 0, 2, 3 first initialise
// (k_fwd_init / k_bwd_gather initialise them).
constexpr int GVR = 64, GVC = 64, GVW = 4;  // rows per WG, columns per wave, waves per WG
struct GemvView {
  const double *Ar, *Ai;
  long long lda;
  int M, N;
  const double *xr, *xi;
  double *yr, *yi;
  double sign;
};

__device__ inline bool gemv_view(int op, const FrontDev &d, double *arena, double *Wr, double *Wi,
                                 double *Sr, double *Si, double *Xr, double *Xi, GemvView &v) {
  const long long mm = static_cast<long long>(d.m) * d.m, pp = static_cast<long long>(d.p) * d.p;
  double *Fr = arena + d.off, *Fi = Fr + mm;
  double *Er = Fi + mm, *Ei = Er + pp;
  double *Zr = Ei + pp, *Zi = Zr + pp;
  switch (op) {
    case 0:
      v = {Er, Ei, d.p, d.p, d.p, Wr + d.woff, Wi + d.woff, Sr + d.dof0, Si + d.dof0, 1.0};
      return d.p > 0;
    case 1:
      v = {Fr + d.p, Fi + d.p, d.m, d.c, d.p, Sr + d.dof0, Si + d.dof0, Wr + d.woff + d.p,
           Wi + d.woff + d.p, -1.0};
      return d.c > 0 && d.p > 0;
    case 2:
      v = {Fr + static_cast<long long>(d.p) * d.m, Fi + static_cast<long long>(d.p) * d.m,
           d.m, d.p, d.c, Wr + d.woff + d.p, Wi + d.woff + d.p, Sr + d.dof0, Si + d.dof0, -1.0};
      return d.c > 0 && d.p > 0;
    default:
      v = {Zr, Zi, d.p, d.p, d.p, Sr + d.dof0, Si + d.dof0, Xr + d.dof0, Xi + d.dof0, 1.0};
      return d.p > 0;
  }
}

// One workgroup: GVR output rows x (GVW * GVC) reduction columns; each wave
// takes GVC of the columns and the four partial sums are combined in LDS
// before one FP64 atomic per row.
__global__ __launch_bounds__(GVR * GVW) void k_gemv(const int *__restrict__ list, int nlist,
                                                    const FrontDev *__restrict__ fr, int op,
                                                    double *__restrict__ arena, double *Wr, double *Wi,
                                                    double *Sr, double *Si, double *Xr, double *Xi) {
  const int f = list[blockIdx.y];
  const FrontDev d = fr[f];
  GemvView v;
  if (!gemv_view(op, d, arena, Wr, Wi, Sr, Si, Xr, Xi, v)) return;
  const int lane = threadIdx.x % GVR, wave = threadIdx.x / GVR;
  const int row = blockIdx.x * GVR + lane;
  const int cw = blockIdx.z * GVW * GVC;  // first column of this workgroup
  if (blockIdx.x * GVR >= v.M || cw >= v.N) return;
  const int c0 = cw + wave * GVC;
  const int c1 = min(v.N, c0 + GVC);
  __shared__ double sxr[GVW * GVC], sxi[GVW * GVC];
  __shared__ double pr[GVW][GVR], pi[GVW][GVR];
  for (int k = threadIdx.x; k < GVW * GVC; k += GVR * GVW) {
    const int col = cw + k;
    sxr[k] = col < v.N ? v.xr[col] : 0.0;
    sxi[k] = col < v.N ? v.xi[col] : 0.0;
  }
  __syncthreads();
  double accr = 0.0, acci = 0.0;
  if (row < v.M) {
    const double *ar = v.Ar + row, *ai = v.Ai + row;
    int k = c0;
    // 16 columns per step so that 32 loads per lane are in flight
    for (; k + 16 <= c1; k += 16) {
      double a_r[16], a_i[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const long long idx = static_cast<long long>(k + u) * v.lda;
        a_r[u] = ar[idx];
        a_i[u] = ai[idx];
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const double x_r = sxr[k + u - cw], x_i = sxi[k + u - cw];
        accr += a_r[u] * x_r - a_i[u] * x_i;
        acci += a_r[u] * x_i + a_i[u] * x_r;
      }
    }
    for (; k < c1; ++k) {
      const long long idx = static_cast<long long>(k) * v.lda;
      const double a_r = ar[idx], a_i = ai[idx];
      const double x_r = sxr[k - cw], x_i = sxi[k - cw];
      accr += a_r * x_r - a_i * x_i;
      acci += a_r * x_i + a_i * x_r;
    }
  }
  pr[wave][lane] = accr;
  pi[wave][lane] = acci;
  __syncthreads();
  if (wave == 0 && row < v.M) {
    double sr = 0.0, si = 0.0;
    for (int w = 0; w < GVW; ++w) {
      sr += pr[w][lane];
      si += pi[w][lane];
    }
    atomicAdd(v.yr + row, v.sign * sr);
    atomicAdd(v.yi + row, v.sign * si);
  }
}

}  // namespace

// ============================================================== host side

inline unsigned cdivu(long long a, long long b) {
  return static_cast<unsigned>(a <= 0 ? 1 : (a + b - 1) / b);
}

hipError_t lu_fill(hipStream_t s, double *p, long long n, double v) {
  hipLaunchKernelGGL(k_fill, dim3(2048), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t lu_factor(const LuPlan &P, const LuInputs &in, double sig_re, double sig_im,
                     double tiny, double perturb, hipStream_t s) {
  hipError_t e;
  if ((e = lu_fill(s, P.arena, P.arena_doubles, 0.0)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_init_identity, dim3(8, P.nsn), dim3(256), 0, s, P.d_front, P.nsn, P.arena);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const long long threads = in.nnzb * 25;
  hipLaunchKernelGGL(k_assemble_fronts, dim3(static_cast<unsigned>((threads + 255) / 256)),
                     dim3(256), 0, s, in.nnzb, in.per, in.avals, in.colidx, in.bvals, sig_re,
                     sig_im, P.d_amap, P.d_front, P.arena);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(1), 0, s, reinterpret_cast<double *>(P.d_counters), 1LL, 0.0);
  for (size_t lev = 0; lev < P.levels.size(); ++lev) {
    const LevelPlan &L = P.levels[lev];
    for (int slot = 0; slot < 2; ++slot) {
      const int nt = L.child_slot_count[slot];
      if (nt == 0) continue;
      hipLaunchKernelGGL(k_extend_add, dim3(256, nt), dim3(256), 0, s,
                         P.d_child_slots + L.child_slot_off[slot], nt, P.d_front, P.d_parent,
                         P.d_rel, P.d_rel_off, P.arena);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const int *list = P.d_level_list + L.list_off;
    for (int k0 = 0; k0 < L.pmax; k0 += P.nb) {
      hipLaunchKernelGGL(k_panel, dim3(L.count), dim3(kPanelThreads), 0, s, list, L.count,
                         P.d_front, k0, P.nb, tiny, perturb, P.arena, P.d_ipiv, P.d_counters);
      if (L.cmax > 0)
        hipLaunchKernelGGL(k_l21, dim3(cdivu(L.cmax, 256), L.count), dim3(256), 0, s, list,
                           L.count, P.d_front, k0, P.nb, P.arena);
      const int cols = L.mmax + L.pmax;
      hipLaunchKernelGGL(k_trsm_lower, dim3(cdivu(cols, 256), L.count), dim3(256), 0, s, list,
                         L.count, P.d_front, k0, P.nb, P.d_ipiv, P.arena);
      const int tm = (L.mmax + GT - 1) / GT;
      hipLaunchKernelGGL(k_gemm_update, dim3(static_cast<unsigned>(tm * tm), L.count), dim3(256), 0, s, list, L.count,
                         P.d_front, k0, P.nb, 0, P.arena);
      const int tp = (L.pmax + GT - 1) / GT;
      hipLaunchKernelGGL(k_gemm_update, dim3(tp * tp, L.count), dim3(256), 0, s, list, L.count,
                         P.d_front, k0, P.nb, 1, P.arena);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    // Z = U11^-1 by backward Jordan steps on the pivot blocks
    for (int k0 = ((L.pmax - 1) / P.nb) * P.nb; k0 >= 0; k0 -= P.nb) {
      hipLaunchKernelGGL(k_trsm_upper, dim3(cdivu(L.pmax, 256), L.count), dim3(256), 0, s, list,
                         L.count, P.d_front, k0, P.nb, P.arena);
      if (k0 > 0) {
        const int tp = (L.pmax + GT - 1) / GT;
        hipLaunchKernelGGL(k_gemm_update, dim3(tp * tp, L.count), dim3(256), 0, s, list, L.count,
                           P.d_front, k0, P.nb, 2, P.arena);
      }
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

// x = K^-1 b for vectors in elimination order (split planes).  y and t are
// solver scratch planes of the same length.
hipError_t lu_solve(const LuPlan &P, const double *br, const double *bi, double *xr, double *xi,
                    hipStream_t s) {
  double *Wr = P.W, *Wi = P.W + P.w_len;
  double *Yr = P.Y, *Yi = P.Y + P.n;
  double *Tr = P.T, *Ti = P.T + P.n;
  hipError_t e;
  for (size_t lev = 0; lev < P.levels.size(); ++lev) {
    const LevelPlan &L = P.levels[lev];
    const int *list = P.d_level_list + L.list_off;
    hipLaunchKernelGGL(k_fwd_init, dim3(L.count), dim3(256), 0, s, list, L.count, P.d_front,
                       P.d_child_ptr, P.d_child_list, P.d_rel, P.d_rel_off, br, bi, Wr, Wi, Yr, Yi);
    hipLaunchKernelGGL(k_gemv, dim3(cdivu(L.pmax, GVR), L.count, cdivu(L.pmax, GVC * GVW)),
                       dim3(GVR * GVW), 0, s, list, L.count, P.d_front, 0, P.arena, Wr, Wi, Yr, Yi, xr, xi);
    hipLaunchKernelGGL(k_gemv, dim3(cdivu(L.cmax, GVR), L.count, cdivu(L.pmax, GVC * GVW)),
                       dim3(GVR * GVW), 0, s, list, L.count, P.d_front, 1, P.arena, Wr, Wi, Yr, Yi, xr, xi);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  for (int lev = static_cast<int>(P.levels.size()) - 1; lev >= 0; --lev) {
    const LevelPlan &L = P.levels[lev];
    const int *list = P.d_level_list + L.list_off;
    hipLaunchKernelGGL(k_bwd_gather, dim3(L.count), dim3(256), 0, s, list, L.count, P.d_front,
                       P.d_upd_pos, P.d_upd_off, xr, xi, Wr, Wi, Yr, Yi, Tr, Ti);
    hipLaunchKernelGGL(k_gemv, dim3(cdivu(L.pmax, GVR), L.count, cdivu(L.cmax, GVC * GVW)),
                       dim3(GVR * GVW), 0, s, list, L.count, P.d_front, 2, P.arena, Wr, Wi, Tr, Ti, xr, xi);
    hipLaunchKernelGGL(k_gemv, dim3(cdivu(L.pmax, GVR), L.count, cdivu(L.pmax, GVC * GVW)),
                       dim3(GVR * GVW), 0, s, list, L.count, P.d_front, 3, P.arena, Wr, Wi, Tr, Ti, xr, xi);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace tailor
