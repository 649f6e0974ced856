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
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "lu.hpp"

namespace tailor {
namespace {

__device__ inline double cabs1(double re, double im) { return fabs(re) + fabs(im); }

// ---------------------------------------------------------------- set-up

// The shift of the next factorization (written on the stream before it, so
// that a captured factorization replays with any shift).
__global__ void k_set_sigma(double *__restrict__ fp, double sig_re, double sig_im) {
  fp[kFpSigRe] = sig_re;
  fp[kFpSigIm] = sig_im;
}

// Partial sums of squares of n doubles: one per workgroup (a tree over its
// 256 threads after a grid-strided loop).
constexpr int kParamBlocks = 512;
__global__ __launch_bounds__(256) void k_sumsq_flat(const double *__restrict__ data, long long n,
                                                    double *__restrict__ partials) {
  __shared__ double red[256];
  double acc = 0.0;
  for (long long r = static_cast<long long>(blockIdx.x) * 256 + threadIdx.x; r < n;
       r += static_cast<long long>(gridDim.x) * 256) {
    const double x = data[r];
    acc += x * x;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// Pivot threshold and static perturbation of the factorization, scaled by
// the RMS row norm of K = A - sigma B: scale = sqrt(|A|_F^2 / rows) +
// |sigma| sqrt(|B|_F^2 / rows); pivots below tiny = c_t * scale are replaced
// by +-perturb = c_p * scale (c_t = 1e-14, c_p = sqrt(eps) for FP64 fronts;
// 1e-7 and sqrt(eps_32) for FP32 fronts), and iterative refinement recovers
// the accuracy.  One thread sums the partials in order.
__global__ void k_factor_params(const double *__restrict__ partials, double rows, int fp32,
                                double *__restrict__ fp) {
  double ssa = 0.0, ssb = 0.0;
  for (int b = 0; b < kParamBlocks; ++b) ssa += partials[b];
  for (int b = 0; b < kParamBlocks; ++b) ssb += partials[kParamBlocks + b];
  const double sig = hypot(fp[kFpSigRe], fp[kFpSigIm]);
  const double scale = fmax(1e-300, sqrt(ssa / rows) + sig * sqrt(ssb / rows));
  fp[kFpTiny] = (fp32 ? 1e-7 : 1e-14) * scale;
  fp[kFpPerturb] = (fp32 ? 3.4526698300124393e-04 : 1.4901161193847656e-08) * scale;
}

__global__ void k_fill(double *__restrict__ p, long long n, double v) {
  for (long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += static_cast<long long>(gridDim.x) * blockDim.x)
    p[k] = v;
}

// E = I and Z = I for every front: the four pe x p planes (Er, Ei, Zr, Zi)
// written whole, two elements per store (pe is even, so a pair never
// straddles a column).
template <typename T>
__global__ void k_init_identity(const FrontDev *__restrict__ fr, int nf, T *__restrict__ arena) {
  typedef T t2 __attribute__((ext_vector_type(2)));
  const int f = blockIdx.y;
  if (f >= nf) return;
  const FrontDev d = fr[f];
  t2 *E = reinterpret_cast<t2 *>(arena + d.off + 2LL * d.m * d.m);
  const long long plane = static_cast<long long>(d.pe) * d.p / 2;  // pairs per plane
  for (long long k = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; k < 4 * plane;
       k += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int q = static_cast<int>(k / plane);                 // Er, Ei, Zr, Zi
    const long long e = 2 * (k - q * plane);                   // element in the plane
    const int col = static_cast<int>(e / d.pe), row = static_cast<int>(e - static_cast<long long>(col) * d.pe);
    const bool re = q == 0 || q == 2;
    E[k] = t2{re && row == col ? T(1) : T(0), re && row + 1 == col ? T(1) : T(0)};
  }
}

// Add K = A - sigma B into the built fronts of one level: the nblk BSR
// blocks listed in ablk, a thread per (block, entry).  amap[3k..3k+2] =
// front, local row node, local col node of block k; the front is
// fbase + amap[3k] of fr (fbase: the part's first front in a batch plan,
// else 0) and fp the part's parameter block.  Distinct blocks land on
// distinct entries, so the additions do not race.  K is formed in FP64 and
// rounded once to the front precision T.
template <typename T>
__global__ void k_assemble_fronts(long long nblk, const int *__restrict__ ablk, int per,
                                  const double *__restrict__ avals,
                                  const int *__restrict__ colidx, const double *__restrict__ bvals,
                                  const double *__restrict__ fp, const int *__restrict__ amap,
                                  const FrontDev *__restrict__ fr, int fbase,
                                  T *__restrict__ arena) {
  const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= nblk * 25) return;
  const double sig_re = fp[kFpSigRe], sig_im = fp[kFpSigIm];
  const long long k = ablk[t / 25];
  const int e = static_cast<int>(t - 25 * (t / 25));
  const int f = fbase + amap[3 * k], lu = amap[3 * k + 1], lv = amap[3 * k + 2];
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
  T *fr_ = arena + d.off + idx, *fi_ = fr_ + static_cast<long long>(d.m) * d.m;
  *fr_ += static_cast<T>(vr);
  *fi_ += static_cast<T>(vi);
}

// F of every front of a level, written once: the sum of its children's
// update blocks F_child[p:, p:] at the rows and columns of the front that
// the children's inverse maps give (front node -> child update node, -1
// where the child has none; d_build / d_inv of the plan), zero elsewhere,
// padding included.  K is added afterwards (k_assemble_fronts).  Workgroup:
// kBuildRows rows of kBuildCols columns, each wave a quarter of the columns
// with four rows per lane (64 apart); the lane's row maps are read once, a
// column's maps are uniform over the wave.  Each child value is read once
// and each front entry written once, where a zero fill and a
// read-modify-write per child would move F twice more.
constexpr int kBuildRows = 256, kBuildCols = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_build_fronts(const int *__restrict__ list, int nlist,
                                                      const FrontDev *__restrict__ fr,
                                                      const int *__restrict__ build,
                                                      const int *__restrict__ inv, int nrc,
                                                      T *__restrict__ arena) {
  constexpr int U = kBuildRows / 64;
  const int f = list[blockIdx.y];
  const FrontDev d = fr[f];
  const int r0 = (blockIdx.x % nrc) * kBuildRows, c0 = (blockIdx.x / nrc) * kBuildCols;
  if (r0 >= d.m || c0 >= d.m) return;
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const long long m = d.m;
  T *Fr = arena + d.off, *Fi = Fr + m * m;
  const int dofs = d.p + d.c;  // rows and columns with a node; the rest is padding
  int ch[2];
  const int *map[2];
  const T *Cr[2], *Ci[2];
  long long cm[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    ch[k] = build[kBuildInts * f + k];
    map[k] = inv + max(build[kBuildInts * f + 2 + k], 0);
    // the child's F22 (rows and columns [p, p + c)); the front's own F
    // stands in for an absent child (read at index 0 and discarded)
    const FrontDev dc = fr[max(ch[k], 0)];
    cm[k] = ch[k] >= 0 ? dc.m : m;
    Cr[k] = ch[k] >= 0 ? arena + dc.off + dc.p * (dc.m + 1LL) : Fr;
    Ci[k] = ch[k] >= 0 ? Cr[k] + static_cast<long long>(dc.m) * dc.m : Fr;
  }
  // child rows of the lane's rows (-1: none)
  int src[2][U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const int r = r0 + lane + 64 * u;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int n = ch[k] >= 0 && r < dofs ? map[k][r / 5] : -1;
      src[k][u] = n >= 0 ? 5 * n + r % 5 : -1;
    }
  }
  const int c1 = min(c0 + kBuildCols, d.m);
  for (int j = c0 + wave; j < c1; j += 4) {
    int jc[2];
    const T *pr[2], *pi[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
      const int n = ch[k] >= 0 && j < dofs ? map[k][j / 5] : -1;
      jc[k] = n >= 0 ? 5 * n + j % 5 : -1;
      pr[k] = Cr[k] + max(jc[k], 0) * cm[k];
      pi[k] = Ci[k] + max(jc[k], 0) * cm[k];
    }
    T vr[2][U], vi[2][U];
#pragma unroll
    for (int k = 0; k < 2; ++k)
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int q = jc[k] >= 0 ? max(src[k][u], 0) : 0;
        vr[k][u] = pr[k][q];
        vi[k][u] = pi[k][q];
      }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int r = r0 + lane + 64 * u;
      if (r >= d.m) continue;
      T ar = T(0), ai = T(0);
#pragma unroll
      for (int k = 0; k < 2; ++k)
        if (jc[k] >= 0 && src[k][u] >= 0) {
          ar += vr[k][u];
          ai += vi[k][u];
        }
      Fr[j * m + r] = ar;
      Fi[j * m + r] = ai;
    }
  }
}

// ------------------------------------------------------------ factorization
//
// Blocked right-looking LU per front.  The fully summed columns are taken in
// outer blocks of kLuNbOuter columns, each factored as kLuNb-column panels:
//
//   for each block [K0, K1):
//     for each panel [k0, k0+kb) of the block:
//       k_panel[_regs] LU of the panel over the fully summed rows [k0, p),
//                      row exchanges applied across the block's columns
//       k_diag_inv     L_D^-1 and U_D^-1 of the panel's pivot block
//       k_right_apply  rows [p, m) of the panel:  X <- X U_D^-1
//       k_left_apply   (mode 0) panel rows of the block columns right of it:
//                      X <- L_D^-1 X
//       k_gemm INNER   F[k0+kb:m, k0+kb:K1] -= L_panel U_panel
//     k_block_perm,    the block's row exchanges on the trailing F columns
//     k_apply_perm     and on E, as one permutation
//     k_block_rows     the block's rows of the trailing F columns and E:
//                      X <- L_blk^-1 X, all panels in one sweep
//     k_gemm TRAIL     F[K1:m, K1:m] -= L21_blk U12_blk and
//                      E[K1:p, :]    -= L21_blk E[K0:K1, :]   (K = K1 - K0)
//   then Z = U11^-1 block by block from the last: k_block_z (the block's
//   rows: per panel from the last, back substitution on its rows and their
//   update of the block's rows above), then k_gemm ZBLK (all rows above the
//   block).
//
// Row exchanges of a block are applied to the block's own columns at once
// but never to earlier blocks: L11 is stored block by block in the row order
// of its own block.  The solves never read L11 (they use E = L11^-1 P, which
// records every exchange), so this ordering is invisible outside the
// factorization.

constexpr int kPanelThreads = 1024;
constexpr int kPanelChunk = 8;  // panel columns updated per batch of loads

// Unblocked LU of the panel columns [k0, k0+kb) over the fully summed rows
// [k0, p) of each front; pivot rows are chosen among those rows and the
// exchange is applied to the columns [K0, K1) of the current block.  The
// update rows [p, m) of the panel are finished by k_right_apply.  The pivot
// search reduces within waves by shuffles; each row's update loads
// kPanelChunk columns before storing any, so the loads overlap.
template <typename T>
__global__ __launch_bounds__(kPanelThreads) void k_panel(const int *__restrict__ list, int nlist,
                                                         const FrontDev *__restrict__ fr, int k0,
                                                         int nb, int K0, int NB,
                                                         const double *__restrict__ fp,
                                                         T *__restrict__ arena,
                                                         int *__restrict__ ipiv,
                                                         int *__restrict__ nperturbed) {
  const int f = list[blockIdx.x];
  const FrontDev d = fr[f];
  if (k0 >= d.p) return;
  const double tiny = fp[d.inst * kFpCount + kFpTiny], perturb = fp[d.inst * kFpCount + kFpPerturb];
  const int kb = min(nb, d.p - k0);
  const int K1 = min(K0 + NB, d.p);
  const long long mm = static_cast<long long>(d.m) * d.m;
  T *__restrict__ Fr = arena + d.off;
  T *__restrict__ Fi = Fr + mm;
  constexpr int kWaves = kPanelThreads / 64;
  __shared__ double sv[kWaves];
  __shared__ int si[kWaves];
  __shared__ int spiv;
  __shared__ bool ssmall;
  __shared__ T prow_r[kLuNb], prow_i[kLuNb];
  __shared__ T inv_r, inv_i;
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  for (int j = k0; j < k0 + kb; ++j) {
    const long long cj = static_cast<long long>(j) * d.m;
    double best = -1.0;
    int bi = j;
    for (int r = j + tid; r < d.p; r += kPanelThreads) {
      const double a = cabs1(static_cast<double>(Fr[cj + r]), static_cast<double>(Fi[cj + r]));
      if (a > best) {
        best = a;
        bi = r;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ob = __shfl_xor(best, off);
      const int oi = __shfl_xor(bi, off);
      if (ob > best || (ob == best && oi < bi)) {
        best = ob;
        bi = oi;
      }
    }
    if (lane == 0) {
      sv[wave] = best;
      si[wave] = bi;
    }
    __syncthreads();
    if (wave == 0) {
      best = lane < kWaves ? sv[lane] : -1.0;
      bi = lane < kWaves ? si[lane] : 0x7fffffff;
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bi, off);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (lane == 0) {
        spiv = bi;
        ssmall = !(best > tiny);
      }
    }
    __syncthreads();
    const int piv = spiv;
    const bool small = ssmall;
    if (!small && piv != j && tid < K1 - K0) {
      const long long col = static_cast<long long>(K0 + tid) * d.m;
      T t = Fr[col + j]; Fr[col + j] = Fr[col + piv]; Fr[col + piv] = t;
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
        Fr[cj + j] = static_cast<T>(pr);
        Fi[cj + j] = static_cast<T>(pi);
        atomicAdd(nperturbed + d.inst, 1);
      }
      const double den = pr * pr + pi * pi;
      inv_r = static_cast<T>(pr / den);
      inv_i = static_cast<T>(-pi / den);
    }
    if (tid < k0 + kb - j - 1) {
      const long long cc = static_cast<long long>(j + 1 + tid) * d.m + j;
      prow_r[tid] = Fr[cc];
      prow_i[tid] = Fi[cc];
    }
    __syncthreads();
    const T ir = inv_r, ii = inv_i;
    const int ncol = k0 + kb - j - 1;
    for (int r = j + 1 + tid; r < d.p; r += kPanelThreads) {
      const T ar = Fr[cj + r], ai = Fi[cj + r];
      const T lr = ar * ir - ai * ii, li = ar * ii + ai * ir;
      Fr[cj + r] = lr;
      Fi[cj + r] = li;
      for (int q0 = 0; q0 < ncol; q0 += kPanelChunk) {
        T xr[kPanelChunk], xi[kPanelChunk];
#pragma unroll
        for (int u = 0; u < kPanelChunk; ++u) {
          const int q = q0 + u < ncol ? q0 + u : q0;  // clamped: loads stay in the panel
          const long long cq = static_cast<long long>(j + 1 + q) * d.m + r;
          xr[u] = Fr[cq];
          xi[u] = Fi[cq];
        }
#pragma unroll
        for (int u = 0; u < kPanelChunk; ++u) {
          if (q0 + u < ncol) {
            const long long cq = static_cast<long long>(j + 1 + q0 + u) * d.m + r;
            const T ur = prow_r[q0 + u], ui = prow_i[q0 + u];
            Fr[cq] = xr[u] - (lr * ur - li * ui);
            Fi[cq] = xi[u] - (lr * ui + li * ur);
          }
        }
      }
    }
    __syncthreads();
  }
}

// Register-resident panel LU for fronts whose fully summed rows [k0, p) fit
// the R rows of each of NT threads.  Same result as k_panel (partial pivoting
// over rows [j, p), exchanges across the block's columns [K0, K1)), organised
// so that one workgroup does not stream the panel through its CU once per
// column:
//
//   * the panel is taken in sub-panels of kSubW columns, held in registers
//     (thread t owns the logical rows k0 + t + s NT, s < R);
//   * a sub-panel is brought up to date left-looking: its rows [k0, c0) by
//     forward substitution with the unit lower L[k0:c0, k0:c0] (one wave per
//     column, broadcasts by readlane), the rows below by
//     X[r, :] -= L[r, k0:c0] U[k0:c0, :], one pass over the finished columns;
//   * its columns are factored in registers: pivot search by shuffles and
//     LDS, the pivot row broadcast through LDS, rank-1 updates in place;
//   * row exchanges are recorded as a logical -> physical row map in LDS, so
//     global memory is read and written through the map and never swapped
//     per pivot; at the end of the panel one gather moves the exchanged rows
//     of every block column [K0, K1) into their final (logical) positions.
constexpr int kSubW = 4;          // sub-panel width (columns held in registers)
constexpr int kPanelMapRows = 4096;  // rows of the LDS row map

template <typename T> __device__ inline T readlane_t(T v, int l);
template <> __device__ inline double readlane_t<double>(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), l);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), l);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
template <> __device__ inline float readlane_t<float>(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Wave-wide max / min by DPP: exchanges within quads (xor 1, xor 2), half-row
// and row mirrors, then row broadcasts 15 and 31, after which lane 63 holds
// the result; it is returned to every lane.  Lanes a step does not write keep
// their own value, which max / min leave unchanged.
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(v, v, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK = 0xf>
__device__ inline double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i<CTRL, ROW_MASK>(static_cast<int>(b));
  const int hi = dpp_i<CTRL, ROW_MASK>(static_cast<int>(b >> 32));
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned>(lo));
}
__device__ inline double wave_max(double v) {
  v = fmax(v, dpp_d<0xB1>(v));        // quad_perm [1,0,3,2]
  v = fmax(v, dpp_d<0x4E>(v));        // quad_perm [2,3,0,1]
  v = fmax(v, dpp_d<0x141>(v));       // row_half_mirror
  v = fmax(v, dpp_d<0x140>(v));       // row_mirror
  v = fmax(v, dpp_d<0x142, 0xA>(v));  // row_bcast:15
  v = fmax(v, dpp_d<0x143, 0xC>(v));  // row_bcast:31
  return readlane_t(v, 63);
}
__device__ inline int wave_min(int v) {
  v = min(v, dpp_i<0xB1>(v));
  v = min(v, dpp_i<0x4E>(v));
  v = min(v, dpp_i<0x141>(v));
  v = min(v, dpp_i<0x140>(v));
  v = min(v, dpp_i<0x142, 0xA>(v));
  v = min(v, dpp_i<0x143, 0xC>(v));
  return __builtin_amdgcn_readlane(v, 63);
}

template <typename T, int R, int NT>
__global__ __launch_bounds__(NT) void k_panel_regs(
    const int *__restrict__ list, int nlist, const FrontDev *__restrict__ fr, int k0, int nb,
    int K0, int NB, const double *__restrict__ fp, T *__restrict__ arena, int *__restrict__ ipiv,
    int *__restrict__ nperturbed) {
  const int f = list[blockIdx.x];
  const FrontDev d = fr[f];
  if (k0 >= d.p) return;
  const double tiny = fp[d.inst * kFpCount + kFpTiny], perturb = fp[d.inst * kFpCount + kFpPerturb];
  const int kb = min(nb, d.p - k0);
  const int K1 = min(K0 + NB, d.p);
  const int p = d.p;
  const long long m = d.m;
  T *__restrict__ Fr = arena + d.off;
  T *__restrict__ Fi = Fr + m * m;
  constexpr int kWaves = NT / 64;
  // finished columns per load batch, fewer where the registers are full
  constexpr int kBatch = NT >= 1024 ? (R * sizeof(T) >= 24 ? 1 : (R * sizeof(T) > 8 ? 2 : kSubW))
                                    : (R * sizeof(T) > 48 ? 1 : (R * sizeof(T) > 16 ? 2 : kSubW));
  static_assert(R * NT <= kPanelMapRows, "rows beyond the row map");
  static_assert(kSubW % kBatch == 0, "batches tile the sub-panel");
  __shared__ int phys[kPanelMapRows];                          // logical row - k0 -> physical
  __shared__ T l11r[kLuNb][kLuNb + 1], l11i[kLuNb][kLuNb + 1];  // L[k0+i, k0+q]
  __shared__ T ur[kLuNb][kSubW], ui[kLuNb][kSubW];              // U[k0+i, c0+u]
  // per column, double-buffered by the column's parity: each wave's best
  // candidate (magnitude, row, the row's sub-panel values) and row j
  __shared__ double sv[2][kWaves];
  __shared__ int si[2][kWaves];
  __shared__ T cand_r[2][kWaves][kSubW], cand_i[2][kWaves][kSubW];
  __shared__ T jrow_r[2][kSubW], jrow_i[2][kSubW];
  __shared__ int spiv[kLuNb];  // pivots of the panel, stored to ipiv at the end
  __shared__ int smoved;
  __shared__ int moved[2 * kLuNb];
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  for (int r = tid; r < p - k0; r += NT) phys[r] = k0 + r;
  if (tid == 0) smoved = 0;
  __syncthreads();

  T ar[R][kSubW], ai[R][kSubW];
  for (int c0 = k0; c0 < k0 + kb; c0 += kSubW) {
    const int w = min(kSubW, k0 + kb - c0);
    const int c = c0 - k0;  // finished columns of the panel
    // (a) rows [k0, c0) of the sub-panel: U = L11^-1 X
    for (int e = tid; e < c * c; e += NT) {
      const int i = e % c, q = e / c;
      const long long idx = static_cast<long long>(k0 + q) * m + phys[i];
      l11r[i][q] = Fr[idx];
      l11i[i][q] = Fi[idx];
    }
    __syncthreads();
    if (wave < w && c > 0) {
      const long long col = static_cast<long long>(c0 + wave) * m;
      T xr = T(0), xi = T(0);
      if (lane < c) {
        xr = Fr[col + phys[lane]];
        xi = Fi[col + phys[lane]];
      }
      for (int q = 0; q + 1 < c; ++q) {
        const T br = readlane_t(xr, q), bi = readlane_t(xi, q);
        if (lane > q && lane < c) {
          const T lr = l11r[lane][q], li = l11i[lane][q];
          xr -= lr * br - li * bi;
          xi -= lr * bi + li * br;
        }
      }
      if (lane < c) {
        Fr[col + phys[lane]] = xr;
        Fi[col + phys[lane]] = xi;
        ur[lane][wave] = xr;
        ui[lane][wave] = xi;
      }
    }
    __syncthreads();
    // (b) rows [c0, p): X -= L[:, k0:c0] U[k0:c0, :], into registers.  The
    // finished columns are read kBatch at a time for all of a thread's rows
    // (c is a multiple of kSubW, hence of kBatch), double-buffered: the next
    // batch's loads are in flight during this batch's FMAs.  Rows outside
    // [c0, p) load a clamped in-front row and their results are never used.
    int prow[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int r = k0 + tid + s * NT;
      prow[s] = r >= c0 && r < p ? phys[r - k0] : k0;
#pragma unroll
      for (int u = 0; u < kSubW; ++u) {
        const long long cu = static_cast<long long>(c0 + min(u, w - 1)) * m;
        ar[s][u] = Fr[cu + prow[s]];
        ai[s][u] = Fi[cu + prow[s]];
      }
    }
    T nr[R][kBatch], ni[R][kBatch];
    auto load_batch = [&](int t0) {
#pragma unroll
      for (int tt = 0; tt < kBatch; ++tt) {
        const T *__restrict__ colr = Fr + static_cast<long long>(k0 + t0 + tt) * m;
        const T *__restrict__ coli = Fi + static_cast<long long>(k0 + t0 + tt) * m;
#pragma unroll
        for (int s = 0; s < R; ++s) {
          nr[s][tt] = colr[prow[s]];
          ni[s][tt] = coli[prow[s]];
        }
      }
    };
    if (c > 0) load_batch(0);
    for (int t0 = 0; t0 < c; t0 += kBatch) {
      T lr[R][kBatch], li[R][kBatch];
#pragma unroll
      for (int s = 0; s < R; ++s) {
#pragma unroll
        for (int tt = 0; tt < kBatch; ++tt) {
          lr[s][tt] = nr[s][tt];
          li[s][tt] = ni[s][tt];
        }
      }
      if (t0 + kBatch < c) load_batch(t0 + kBatch);
#pragma unroll
      for (int tt = 0; tt < kBatch; ++tt) {
#pragma unroll
        for (int u = 0; u < kSubW; ++u) {
          const T xr = ur[t0 + tt][u], xi = ui[t0 + tt][u];
#pragma unroll
          for (int s = 0; s < R; ++s) {
            ar[s][u] -= lr[s][tt] * xr - li[s][tt] * xi;
            ai[s][u] -= lr[s][tt] * xi + li[s][tt] * xr;
          }
        }
      }
    }
    // (c) the sub-panel's columns in registers, one barrier per column:
    // each wave reduces its pivot candidate by DPP and publishes it with its
    // row values, the owner of row j publishes row j; after the barrier
    // every wave reduces the candidates itself and takes the pivot row from
    // LDS.  The parity double buffer keeps a column's slots intact until all
    // waves have passed the next column's barrier.
#pragma unroll
    for (int u = 0; u < kSubW; ++u) {
      if (u < w) {
        const int j = c0 + u;
        const int b = j & 1;
        double best = -1.0;
        int bi = 0x7fffffff;
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int r = k0 + tid + s * NT;
          if (r >= j && r < p) {
            const double a = cabs1(static_cast<double>(ar[s][u]), static_cast<double>(ai[s][u]));
            if (a > best) {
              best = a;
              bi = r;
            }
          }
          if (r == j) {
#pragma unroll
            for (int q = 0; q < kSubW; ++q) {
              jrow_r[b][q] = ar[s][q];
              jrow_i[b][q] = ai[s][q];
            }
          }
        }
        const double wbest = wave_max(best);
        const int wbi = wave_min(best == wbest ? bi : 0x7fffffff);
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int r = k0 + tid + s * NT;
          if (r == wbi) {
#pragma unroll
            for (int q = 0; q < kSubW; ++q) {
              cand_r[b][wave][q] = ar[s][q];
              cand_i[b][wave][q] = ai[s][q];
            }
          }
        }
        if (lane == 0) {
          sv[b][wave] = wbest;
          si[b][wave] = wbi;
        }
        __syncthreads();
        const double gv = lane < kWaves ? sv[b][lane] : -1.0;
        const int gi = lane < kWaves ? si[b][lane] : 0x7fffffff;
        const double gbest = wave_max(gv);
        const int gpiv = wave_min(gv == gbest ? gi : 0x7fffffff);
        const bool small = !(gbest > tiny);
        const int piv = small ? j : gpiv;
        // the winning wave: the one lane whose candidate is the pivot row
        const unsigned long long hit = __ballot(lane < kWaves && gi == gpiv && gv == gbest);
        const int wwin = hit ? __ffsll(static_cast<long long>(hit)) - 1 : 0;
        T pr_[kSubW], pi_[kSubW];
#pragma unroll
        for (int q = 0; q < kSubW; ++q) {
          pr_[q] = small ? jrow_r[b][q] : cand_r[b][wwin][q];
          pi_[q] = small ? jrow_i[b][q] : cand_i[b][wwin][q];
        }
        if (tid == 0) {
          spiv[j - k0] = piv;
          if (piv != j) {
            const int t = phys[j - k0];
            phys[j - k0] = phys[piv - k0];
            phys[piv - k0] = t;
          }
          if (small) atomicAdd(nperturbed + d.inst, 1);
        }
        T pr = pr_[u], pi = pi_[u];
        if (small) {
          // static pivot perturbation: refinement recovers the accuracy
          pr = pr >= T(0) ? static_cast<T>(perturb) : static_cast<T>(-perturb);
          pi = T(0);
        }
        const double den = static_cast<double>(pr) * pr + static_cast<double>(pi) * pi;
        const T ir = static_cast<T>(pr / den), ii = static_cast<T>(-pi / den);
#pragma unroll
        for (int s = 0; s < R; ++s) {
          const int r = k0 + tid + s * NT;
          if (r == j) {
#pragma unroll
            for (int q = 0; q < kSubW; ++q) {
              ar[s][q] = pr_[q];
              ai[s][q] = pi_[q];
            }
            ar[s][u] = pr;
            ai[s][u] = pi;
          } else if (r == piv) {
#pragma unroll
            for (int q = 0; q < kSubW; ++q) {
              ar[s][q] = jrow_r[b][q];
              ai[s][q] = jrow_i[b][q];
            }
          }
          if (r > j && r < p) {
            const T xr = ar[s][u], xi = ai[s][u];
            const T lr = xr * ir - xi * ii, li = xr * ii + xi * ir;
            ar[s][u] = lr;
            ai[s][u] = li;
#pragma unroll
            for (int q = u + 1; q < kSubW; ++q) {
              const T yr = pr_[q], yi = pi_[q];
              ar[s][q] -= lr * yr - li * yi;
              ai[s][q] -= lr * yi + li * yr;
            }
          }
        }
      }
    }
    __syncthreads();  // the row map of this sub-panel's exchanges
    // (d) the sub-panel back to global memory through the row map
#pragma unroll
    for (int s = 0; s < R; ++s) {
      const int r = k0 + tid + s * NT;
      if (r >= c0 && r < p) {
        const long long pr = phys[r - k0];
#pragma unroll
        for (int u = 0; u < kSubW; ++u) {
          if (u < w) {
            Fr[static_cast<long long>(c0 + u) * m + pr] = ar[s][u];
            Fi[static_cast<long long>(c0 + u) * m + pr] = ai[s][u];
          }
        }
      }
    }
    __syncthreads();
  }
  // rows whose data moved: a subset of [k0, k0+kb) and the pivot rows
  if (tid < kb) ipiv[d.piv_off + k0 + tid] = spiv[tid];
  if (tid < 2 * kb) {
    const int cand = tid < kb ? k0 + tid : spiv[tid - kb];
    bool take = phys[cand - k0] != cand;
    if (tid >= kb) {
      if (cand < k0 + kb) take = false;  // listed by the first half
      for (int q = 0; q < tid - kb; ++q)
        if (spiv[q] == cand) take = false;
    }
    if (take) moved[atomicAdd(&smoved, 1)] = cand;
  }
  __syncthreads();
  // gather: column entries of the moved rows in every block column, in two
  // halves of the columns; within a half all loads precede the stores
  const int nm = smoved;
  constexpr int kHalf = kLuNbOuter / 2;
  constexpr int kGather = 2 * kLuNb * kHalf / NT;
  for (int h0 = K0; h0 < K1; h0 += kHalf) {
    const int pairs = nm * min(kHalf, K1 - h0);
    T gr[kGather], gi[kGather];
#pragma unroll
    for (int g = 0; g < kGather; ++g) {
      const int e = tid + g * NT;
      if (e < pairs) {
        const int i = moved[e % nm];
        const long long idx = static_cast<long long>(h0 + e / nm) * m + phys[i - k0];
        gr[g] = Fr[idx];
        gi[g] = Fi[idx];
      }
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < kGather; ++g) {
      const int e = tid + g * NT;
      if (e < pairs) {
        const long long idx = static_cast<long long>(h0 + e / nm) * m + moved[e % nm];
        Fr[idx] = gr[g];
        Fi[idx] = gi[g];
      }
    }
  }
}

// ------------------------------------------------------- MFMA complex GEMM
//
// C -= A B on split (re, im) column-major planes, batched over the fronts
// of a level.  A workgroup computes a 64 x 64 complex tile with four waves
// (2 x 2, 32 x 32 each = 2 x 2 MFMA 16 x 16 blocks, real and imaginary
// accumulators), K staged through LDS in chunks of 16 with the next chunk's
// global loads in flight during the MFMAs.
//
// The 16x16x4 MFMAs are issued as D^T = B^T A^T: lane l supplies
// B[k = l/16][j = l%16] as the first operand and A[i = l%16][k = l/16] as
// the second, which leaves each lane holding outputs of one row i = l%16,
// so each accumulator element stores 16 consecutive rows of a column.
// Output column of accumulator element v (lane l):
//   f64 16x16x4: 4 v + l/16        f32 16x16x4: 4 (l/16) + v
// (lane layout checked on gfx950 with value-coded operands).
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef double acc_t __attribute__((ext_vector_type(4)));
  static __device__ inline acc_t mma(double x, double y, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
  static __device__ inline int ocol(int l, int v) { return 4 * v + l / 16; }
  // the same as a function of the lane's group of 16 (g = l / 16)
  static __device__ constexpr int orow(int g, int v) { return 4 * v + g; }
};
template <> struct Mfma<float> {
  typedef float acc_t __attribute__((ext_vector_type(4)));
  static __device__ inline acc_t mma(float x, float y, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
  }
  static __device__ inline int ocol(int l, int v) { return 4 * (l / 16) + v; }
  static __device__ constexpr int orow(int g, int v) { return 4 * g + v; }
};

// ------------------------------------------- inverted pivot blocks (MFMA)
//
// The triangular solves with a panel's kb x kb pivot block are applied as
// multiplications by its inverse.  After the panel, k_diag_inv writes
// L_D^-1 (unit lower) and U_D^-1 (upper), zero-padded to kLuNb x kLuNb, to
// the panel's slot of the level's inverse workspace (slot dinv_off +
// k0 / kLuNb of the front: planes Lr, Li, Ur, Ui, column-major); the apply
// kernels below multiply by them on MFMA, in place.
constexpr int kInvSlot = kLuInvSlot;  // elements per panel slot

template <typename T>
__device__ inline T *inv_slot(T *dinv, const FrontDev &d, int k0) {
  return dinv + static_cast<long long>(d.dinv_off + k0 / kLuNb) * kInvSlot;
}

// One kLuNb x kLuNb plane pair of a slot (column-major) into LDS as
// Mr[i][q], Mi[i][q], by NT threads; each thread issues all of its loads
// before its first LDS store, so the loads are in flight together.
template <int NT, typename T>
__device__ inline void stage_slot(const T *slot, T (*Mr)[kLuNb + 1], T (*Mi)[kLuNb + 1]) {
  constexpr int kPer = kLuNb * kLuNb / NT;
  T vr[kPer], vi[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    vr[u] = slot[threadIdx.x + NT * u];
    vi[u] = slot[kLuNb * kLuNb + threadIdx.x + NT * u];
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int e = threadIdx.x + NT * u;
    Mr[e % kLuNb][e / kLuNb] = vr[u];
    Mi[e % kLuNb][e / kLuNb] = vi[u];
  }
}

// One front per workgroup: wave 0 computes L_D^-1, wave 1 U_D^-1, each by
// 2 x 2 blocks of 16:
//   L = [A 0; C B]:  L^-1 = [A^-1 0; -B^-1 (C A^-1)  B^-1]
//   U = [A C; 0 B]:  U^-1 = [A^-1  -A^-1 (C B^-1); 0 B^-1]
// Lanes 0..15 invert A and lanes 16..31 B (lane l: column l, by forward or
// back substitution over 16 rows), then the whole wave forms the
// off-diagonal block with two 16 x 16 x 16 complex products on MFMA.  The
// inverses are those of the leading kb x kb block, zero-padded: the zeros
// of D outside it carry through the products.
template <typename T>
__global__ __launch_bounds__(128) void k_diag_inv(const int *__restrict__ list, int nlist,
                                                  const FrontDev *__restrict__ fr, int k0, int nb,
                                                  const T *__restrict__ arena,
                                                  T *__restrict__ dinv) {
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int H = kLuNb / 2;
  static_assert(H == 16, "16 x 16 diagonal blocks");
  const FrontDev d = fr[list[blockIdx.x]];
  if (k0 >= d.p) return;
  const int kb = min(nb, d.p - k0);
  const long long m = d.m;
  const T *Fr = arena + d.off, *Fi = Fr + m * m;
  __shared__ T Dr[kLuNb][kLuNb + 1], Di[kLuNb][kLuNb + 1];  // D[i][q] = F[k0+i, k0+q]
  __shared__ T Xr[2][kLuNb][kLuNb + 1], Xi[2][kLuNb][kLuNb + 1];  // wave w's inverse
  __shared__ T Pr[2][H][H + 1], Pi[2][H][H + 1];                  // C A^-1 or C B^-1
  {
    // all loads issued before the first LDS store (clamped addresses, zeroed
    // outside the kb x kb block)
    constexpr int kPer = kLuNb * kLuNb / 128;  // 128 threads
    T vr[kPer], vi[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = threadIdx.x + 128 * u, i = e % kLuNb, q = e / kLuNb;
      const long long idx = static_cast<long long>(k0 + min(q, kb - 1)) * m + k0 + min(i, kb - 1);
      const bool ok = i < kb && q < kb;
      const T a = Fr[idx], b = Fi[idx];
      vr[u] = ok ? a : T(0);
      vi[u] = ok ? b : T(0);
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      const int e = threadIdx.x + 128 * u;
      Dr[e % kLuNb][e / kLuNb] = vr[u];
      Di[e % kLuNb][e / kLuNb] = vi[u];
    }
  }
  __syncthreads();
  const int l = threadIdx.x % 64, w = threadIdx.x / 64;
  if (l < kLuNb) {
    // column l of the diagonal block's inverse: rows o .. o + 15
    const int o = l < H ? 0 : H;
    const bool col_in = l < kb;
    T xr[H], xi[H];
    if (w == 0) {
      // unit lower: x_i = [o + i == l] - sum_{q < i} L[o+i][o+q] x_q
#pragma unroll
      for (int i = 0; i < H; ++i) {
        T ar = o + i == l ? T(1) : T(0), ai = T(0);
#pragma unroll
        for (int q = 0; q < i; ++q) {
          const T lr = Dr[o + i][o + q], li = Di[o + i][o + q];
          ar -= lr * xr[q] - li * xi[q];
          ai -= lr * xi[q] + li * xr[q];
        }
        const bool in = o + i < kb && col_in;
        xr[i] = in ? ar : T(0);
        xi[i] = in ? ai : T(0);
      }
    } else {
      // upper: x_i = ([o + i == l] - sum_{q > i} U[o+i][o+q] x_q) / U[o+i][o+i]
#pragma unroll
      for (int i = H - 1; i >= 0; --i) {
        T ar = o + i == l ? T(1) : T(0), ai = T(0);
#pragma unroll
        for (int q = i + 1; q < H; ++q) {
          const T ur = Dr[o + i][o + q], ui = Di[o + i][o + q];
          ar -= ur * xr[q] - ui * xi[q];
          ai -= ur * xi[q] + ui * xr[q];
        }
        xr[i] = T(0);
        xi[i] = T(0);
        if (o + i < kb && col_in) {
          const double dr = Dr[o + i][o + i], di = Di[o + i][o + i];
          const double den = dr * dr + di * di;
          xr[i] = static_cast<T>((ar * dr + ai * di) / den);
          xi[i] = static_cast<T>((ai * dr - ar * di) / den);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < H; ++i) {
      Xr[w][o + i][l] = xr[i];
      Xi[w][o + i][l] = xi[i];
      Xr[w][(o + i + H) % kLuNb][l] = T(0);  // the other off-diagonal rows
      Xi[w][(o + i + H) % kLuNb][l] = T(0);
    }
  }
  __syncthreads();
  // MFMA products: lane l supplies M[l%16][4s + g] and N[4s + g][l%16] and
  // receives rows orow(g, v) of column l%16
  const int r = l % 16, g = l / 16;
  // P = C X1:  L: C = D[16:32, 0:16], X1 = A^-1;  U: C = D[0:16, 16:32], X1 = B^-1
  {
    const int cr0 = w == 0 ? H : 0, cc0 = w == 0 ? 0 : H, x0 = w == 0 ? 0 : H;
    acc_t pr = acc_t{0, 0, 0, 0}, pi = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < H / 4; ++s) {
      const int k = 4 * s + g;
      const T ar = Dr[cr0 + r][cc0 + k], ai = Di[cr0 + r][cc0 + k];
      const T br = Xr[w][x0 + k][x0 + r], bi = Xi[w][x0 + k][x0 + r];
      pr = Mfma<T>::mma(ar, br, pr);
      pr = Mfma<T>::mma(ai, -bi, pr);
      pi = Mfma<T>::mma(ar, bi, pi);
      pi = Mfma<T>::mma(ai, br, pi);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Pr[w][Mfma<T>::orow(g, v)][r] = pr[v];
      Pi[w][Mfma<T>::orow(g, v)][r] = pi[v];
    }
  }
  __syncthreads();
  // off-diagonal block -X2 P:  L: X2 = B^-1, rows 16:32, columns 0:16;
  //                            U: X2 = A^-1, rows 0:16, columns 16:32
  {
    const int x0 = w == 0 ? H : 0, qr0 = w == 0 ? H : 0, qc0 = w == 0 ? 0 : H;
    acc_t qr = acc_t{0, 0, 0, 0}, qi = acc_t{0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < H / 4; ++s) {
      const int k = 4 * s + g;
      const T ar = Xr[w][x0 + r][x0 + k], ai = Xi[w][x0 + r][x0 + k];
      const T br = Pr[w][k][r], bi = Pi[w][k][r];
      qr = Mfma<T>::mma(ar, -br, qr);
      qr = Mfma<T>::mma(ai, bi, qr);
      qi = Mfma<T>::mma(ar, -bi, qi);
      qi = Mfma<T>::mma(ai, -br, qi);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      Xr[w][qr0 + Mfma<T>::orow(g, v)][qc0 + r] = qr[v];
      Xi[w][qr0 + Mfma<T>::orow(g, v)][qc0 + r] = qi[v];
    }
  }
  __syncthreads();
  // slot planes Lr, Li, Ur, Ui, column-major
  T *slot = inv_slot(dinv, d, k0);
#pragma unroll
  for (int u = 0; u < kLuNb * kLuNb / 128; ++u) {
    const int e = threadIdx.x + 128 * u, i = e % kLuNb, c = e / kLuNb;
    slot[e] = Xr[0][i][c];
    slot[kLuNb * kLuNb + e] = Xi[0][i][c];
    slot[2 * kLuNb * kLuNb + e] = Xr[1][i][c];
    slot[3 * kLuNb * kLuNb + e] = Xi[1][i][c];
  }
}

// X[k0:k0+kb, cols] <- M X (M kb x kb from the panel's slot):
//   mode 0: F columns [k0+kb, K1) of the current block, M = L_D^-1;
//   mode 2: Z columns [k0, p), M = U_D^-1.
// Four waves per workgroup, 16 columns per wave.  The MFMAs compute
// (M X)^T = X^T M^T: lane l supplies X[4s + l/16][l%16] and
// M[16 ib + l%16][4s + l/16], and holds the results of rows 16 ib + l%16.
// A wave reads all kb rows of its columns before storing any (the MFMA
// chain needs every k), so the update is in place.
template <typename T>
__global__ __launch_bounds__(256) void k_left_apply(const int *__restrict__ list, int nlist,
                                                    const FrontDev *__restrict__ fr, int k0, int nb,
                                                    int K0, int NB, int mode,
                                                    const T *__restrict__ dinv,
                                                    T *__restrict__ arena) {
  using acc_t = typename Mfma<T>::acc_t;
  const FrontDev d = fr[list[blockIdx.y]];
  if (k0 >= d.p) return;
  const int kb = min(nb, d.p - k0);
  const int K1 = min(K0 + NB, d.p);
  const long long m = d.m, pe = d.pe, mm = m * m, pp = pe * d.p;
  T *Fr = arena + d.off, *Fi = Fr + mm;
  T *Er = Fi + mm, *Ei = Er + pp;
  T *Zr = Ei + pp, *Zi = Zr + pp;
  const int ncols = mode == 0 ? K1 - (k0 + kb) : d.p - k0;
  const int col0 = blockIdx.x * 64;
  if (col0 >= ncols) return;
  __shared__ T Mr[kLuNb][kLuNb + 1], Mi[kLuNb][kLuNb + 1];
  stage_slot<256>(inv_slot(dinv, d, k0) + (mode == 2 ? 2 * kLuNb * kLuNb : 0), Mr, Mi);
  __syncthreads();
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  // row k0 of column j of the region, as (re, im) plane offsets
  auto column = [&](int j, T *&xr, T *&xi) {
    long long base;
    if (mode == 0) {
      base = static_cast<long long>(k0 + kb + j) * m + k0;
      xr = Fr + base;
      xi = Fi + base;
    } else {
      base = static_cast<long long>(k0 + j) * pe + k0;
      xr = Zr + base;
      xi = Zi + base;
    }
  };
  const int jl = col0 + wave * 16 + lane % 16;  // operand column of this lane
  T xr[kLuNb / 4], xi[kLuNb / 4];
  {
    T *cr, *ci;
    column(min(jl, ncols - 1), cr, ci);
#pragma unroll
    for (int s = 0; s < kLuNb / 4; ++s) {
      const int k = 4 * s + lane / 16;
      const bool ok = jl < ncols && k < kb;
      const int kk = min(k, kb - 1);  // clamped: addresses stay in the column's kb rows
      xr[s] = ok ? cr[kk] : T(0);
      xi[s] = ok ? ci[kk] : T(0);
    }
  }
  acc_t cr_[2], ci_[2];
#pragma unroll
  for (int ib = 0; ib < 2; ++ib) {
    cr_[ib] = acc_t{0, 0, 0, 0};
    ci_[ib] = acc_t{0, 0, 0, 0};
  }
#pragma unroll
  for (int s = 0; s < kLuNb / 4; ++s) {
#pragma unroll
    for (int ib = 0; ib < 2; ++ib) {
      const T mr = Mr[16 * ib + lane % 16][4 * s + lane / 16];
      const T mi = Mi[16 * ib + lane % 16][4 * s + lane / 16];
      cr_[ib] = Mfma<T>::mma(xr[s], mr, cr_[ib]);
      cr_[ib] = Mfma<T>::mma(-xi[s], mi, cr_[ib]);
      ci_[ib] = Mfma<T>::mma(xi[s], mr, ci_[ib]);
      ci_[ib] = Mfma<T>::mma(xr[s], mi, ci_[ib]);
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int j = col0 + wave * 16 + Mfma<T>::ocol(lane, v);
    if (j >= ncols) continue;
    T *cr, *ci;
    column(j, cr, ci);
#pragma unroll
    for (int ib = 0; ib < 2; ++ib) {
      const int i = 16 * ib + lane % 16;
      if (i < kb) {
        cr[i] = cr_[ib][v];
        ci[i] = ci_[ib][v];
      }
    }
  }
}

// U rows of a whole block [K0, K1) for the trailing F columns [K1, m) and
// all p columns of E: X <- L_blk^-1 X, where L_blk is the unit lower
// triangle of the block's rows and columns, in one sweep over its panels:
//   for panel j:  X_j <- L_D(j)^-1 X_j  (the slot's inverse),
//                 X_i -= L_ij X_j       for the later panels i,
// L_ij = F[rows of panel i, columns of panel j].  A workgroup holds the
// block's rows of 64 columns in registers for the whole sweep (four waves,
// 16 columns each), so each element is read and written once.
//
// The MFMAs compute M X with M (32 x 32, staged in LDS) as the first
// operand: lane l supplies M[16 ib + l%16][k] and X[k][l%16] for the k of
// its group g = l/16 in chunk s, and receives rows 16 ib + orow(g, v).  With
// k = 16 (s/4) + orow(g, s%4) for chunk s, a result lands in the register
// that held the same row of X, so results feed the next products without
// moving.  Rows and columns of each M are renumbered in LDS by
//   sigma(16 ib + orow(g, v)) = 8 g + 4 ib + v,
// under which lane group g holds the 8 consecutive rows 8 g .. 8 g + 7 of
// each panel: 64- or 32-byte vector loads and stores.
template <typename T>
__device__ inline int block_rows_pos(int r) {  // sigma^-1
  return 16 * ((r % 8) / 4) + Mfma<T>::orow(r / 8, r % 4);
}

template <typename T>
__global__ __launch_bounds__(256, 2) void k_block_rows(const int *__restrict__ list, int nlist,
                                                    const FrontDev *__restrict__ fr, int K0,
                                                    int NB, const T *__restrict__ dinv,
                                                    T *__restrict__ arena) {
  using acc_t = typename Mfma<T>::acc_t;
  typedef T v8 __attribute__((ext_vector_type(8)));
  constexpr int NP = kLuNbOuter / kLuNb;  // panels per block
  static_assert(NP == 4, "a block is four panels");
  const FrontDev d = fr[list[blockIdx.y]];
  if (K0 >= d.p) return;
  const int K1 = min(K0 + NB, d.p);
  const int nrows = K1 - K0, np = (nrows + kLuNb - 1) / kLuNb;
  const long long m = d.m, pe = d.pe, mm = m * m, pp = pe * d.p;
  T *Fr = arena + d.off, *Fi = Fr + mm;
  T *Er = Fi + mm, *Ei = Er + pp;
  const int nF = d.m - K1;
  const int ncols = nF + d.p;
  const int col0 = blockIdx.x * 64;
  if (col0 >= ncols) return;
  // LDS: stage 0 = L_D(j)^-1, stage t = L_{j+t, j}, renumbered by sigma
  __shared__ T Sr[NP][kLuNb][kLuNb + 1], Si[NP][kLuNb][kLuNb + 1];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64, g = lane / 16;
  const int jl = col0 + wave * 16 + lane % 16;  // the lane's column
  const bool live = jl < ncols;
  T *cr, *ci;  // row K0 of the lane's column (the last column for dead lanes)
  {
    const int j = min(jl, ncols - 1);
    const long long base =
        j < nF ? static_cast<long long>(K1 + j) * m + K0 : static_cast<long long>(j - nF) * pe + K0;
    cr = (j < nF ? Fr : Er) + base;
    ci = (j < nF ? Fi : Ei) + base;
  }
  // x[j][ib][v]: row 32 j + 8 g + 4 ib + v of the lane's column.  Every
  // load is issued unconditionally (rows past the block read row K0 and are
  // zeroed): a run of 8 rows starting below K1 stays inside the column,
  // whose length (m or pe) is a multiple of 32.
  acc_t xr[NP][2], xi[NP][2];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int r0 = 32 * j + 8 * g;
    const bool in = r0 < nrows;
    const v8 a = *reinterpret_cast<const v8 *>(cr + (in ? r0 : 0));
    const v8 b = *reinterpret_cast<const v8 *>(ci + (in ? r0 : 0));
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const bool ok = live && r0 + s < nrows;
      xr[j][s / 4][s % 4] = ok ? a[s] : T(0);
      xi[j][s / 4][s % 4] = ok ? b[s] : T(0);
    }
  }
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    if (j >= np) break;
    const int k0 = K0 + kLuNb * j;
    {
      // all of a thread's loads are issued before its first LDS store; the
      // blocks L_{j+t, j} past the last panel are loaded from row k0 (inside
      // the front: k0 + 32 <= m) and zeroed
      constexpr int kPer = kLuNb * kLuNb / 256;  // 256 threads
      const T *slot = inv_slot(dinv, d, k0);
      T vr[NP][kPer], vi[NP][kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = threadIdx.x + 256 * u, r = e % kLuNb, c = e / kLuNb;
        vr[0][u] = slot[e];
        vi[0][u] = slot[kLuNb * kLuNb + e];
#pragma unroll
        for (int t = 1; t < NP; ++t) {
          const int row = k0 + kLuNb * t + r;  // in the block's rows if < K1
          const bool ok = row < K1;
          const long long idx = static_cast<long long>(k0 + c) * m + (ok ? row : k0);
          const T a = Fr[idx], b = Fi[idx];
          vr[t][u] = ok ? a : T(0);
          vi[t][u] = ok ? b : T(0);
        }
      }
      __syncthreads();  // the previous panel's products have read the stages
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = threadIdx.x + 256 * u;
        const int pi = block_rows_pos<T>(e % kLuNb), pk = block_rows_pos<T>(e / kLuNb);
#pragma unroll
        for (int t = 0; t < NP; ++t) {
          Sr[t][pi][pk] = vr[t][u];
          Si[t][pi][pk] = vi[t][u];
        }
      }
    }
    __syncthreads();
    // X_j <- L_D(j)^-1 X_j
    acc_t ur[2], ui[2];
#pragma unroll
    for (int ib = 0; ib < 2; ++ib) {
      ur[ib] = acc_t{0, 0, 0, 0};
      ui[ib] = acc_t{0, 0, 0, 0};
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const int k = 16 * (s / 4) + Mfma<T>::orow(g, s % 4);
      const T br = xr[j][s / 4][s % 4], bi = xi[j][s / 4][s % 4];
#pragma unroll
      for (int ib = 0; ib < 2; ++ib) {
        const T ar = Sr[0][16 * ib + lane % 16][k], ai = Si[0][16 * ib + lane % 16][k];
        ur[ib] = Mfma<T>::mma(ar, br, ur[ib]);
        ur[ib] = Mfma<T>::mma(ai, -bi, ur[ib]);
        ui[ib] = Mfma<T>::mma(ar, bi, ui[ib]);
        ui[ib] = Mfma<T>::mma(ai, br, ui[ib]);
      }
    }
#pragma unroll
    for (int ib = 0; ib < 2; ++ib) {
      xr[j][ib] = ur[ib];
      xi[j][ib] = ui[ib];
    }
    // X_i -= L_ij X_j for the later panels
#pragma unroll
    for (int t = 1; t < NP; ++t) {
      const int i = j + t;
      if (i >= NP || i >= np) break;
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int k = 16 * (s / 4) + Mfma<T>::orow(g, s % 4);
        const T br = xr[j][s / 4][s % 4], bi = xi[j][s / 4][s % 4];
#pragma unroll
        for (int ib = 0; ib < 2; ++ib) {
          const T ar = Sr[t][16 * ib + lane % 16][k], ai = Si[t][16 * ib + lane % 16][k];
          xr[i][ib] = Mfma<T>::mma(ar, -br, xr[i][ib]);
          xr[i][ib] = Mfma<T>::mma(ai, bi, xr[i][ib]);
          xi[i][ib] = Mfma<T>::mma(ar, -bi, xi[i][ib]);
          xi[i][ib] = Mfma<T>::mma(ai, -br, xi[i][ib]);
        }
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int r0 = 32 * j + 8 * g;
    v8 a, b;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      a[s] = xr[j][s / 4][s % 4];
      b[s] = xi[j][s / 4][s % 4];
    }
    if (r0 + 8 <= nrows) {
      *reinterpret_cast<v8 *>(cr + r0) = a;
      *reinterpret_cast<v8 *>(ci + r0) = b;
    } else if (r0 < nrows) {
#pragma unroll
      for (int s = 0; s < 8; ++s)
        if (r0 + s < nrows) {
          cr[r0 + s] = a[s];
          ci[r0 + s] = b[s];
        }
    }
  }
}

// Z rows of a whole block [K0, K1), columns [K0, p): Z = U11^-1 by the
// block's panels from the last,
//   X_j <- U_jj^-1 X_j          back substitution,
//   X_i -= U_ij X_j             for the earlier panels i,
// with the block's rows of 64 columns in registers for the whole sweep,
// in the layout of k_block_rows (lane group g holds rows 8 g .. 8 g + 7 of
// each panel, column l%16).  The substitution scales each row by the
// reciprocal of U[r][r] (the diagonal of the slot's U_D^-1) and subtracts
// U[i][r] x_r from the rows above, eight rows per lane group at a time (see
// below): Z is used as an explicit inverse by the solves, so its panel rows
// are solved with U_jj itself rather than multiplied by the inverted pivot
// block, which is less accurate where U_jj is ill-conditioned.  The updates
// are MFMA products with the sigma-renumbered U_ij in LDS.  Columns
// [K0, k0) of panel j's rows are zero in Z, so the sweep may run over them.
template <typename T>
__global__ __launch_bounds__(256, 2) void k_block_z(const int *__restrict__ list, int nlist,
                                                    const FrontDev *__restrict__ fr, int K0,
                                                    int NB, const T *__restrict__ dinv,
                                                    T *__restrict__ arena) {
  using acc_t = typename Mfma<T>::acc_t;
  typedef T v8 __attribute__((ext_vector_type(8)));
  constexpr int NP = kLuNbOuter / kLuNb;
  static_assert(NP == 4, "a block is four panels");
  const FrontDev d = fr[list[blockIdx.y]];
  if (K0 >= d.p) return;
  const int K1 = min(K0 + NB, d.p);
  const int nrows = K1 - K0, np = (nrows + kLuNb - 1) / kLuNb;
  const long long m = d.m, pe = d.pe, mm = m * m, pp = pe * d.p;
  const T *Fr = arena + d.off, *Fi = Fr + mm;
  T *Zr = arena + d.off + 2 * mm + 2 * pp, *Zi = Zr + pp;
  const int ncols = d.p - K0;
  const int col0 = blockIdx.x * 64;
  if (col0 >= ncols) return;
  // stage 0: U_jj as it stands (row i, column r); stage t: U_{j-t, j}
  // renumbered by sigma; D: 1 / U[r][r]
  __shared__ T Sr[NP][kLuNb][kLuNb + 1], Si[NP][kLuNb][kLuNb + 1];
  __shared__ T Dr[kLuNb], Di[kLuNb];
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64, g = lane / 16;
  const int jl = col0 + wave * 16 + lane % 16;  // the lane's column - K0
  const bool live = jl < ncols;
  T *cr = Zr + static_cast<long long>(K0 + min(jl, ncols - 1)) * pe + K0;
  T *ci = Zi + static_cast<long long>(K0 + min(jl, ncols - 1)) * pe + K0;
  // x[j][ib][v]: row 32 j + 8 g + 4 ib + v; loads as in k_block_rows (a run
  // of 8 rows starting below K1 stays inside the column of pe rows)
  acc_t xr[NP][2], xi[NP][2];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int r0 = 32 * j + 8 * g;
    const bool in = r0 < nrows;
    const v8 a = *reinterpret_cast<const v8 *>(cr + (in ? r0 : 0));
    const v8 b = *reinterpret_cast<const v8 *>(ci + (in ? r0 : 0));
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const bool ok = live && r0 + s < nrows;
      xr[j][s / 4][s % 4] = ok ? a[s] : T(0);
      xi[j][s / 4][s % 4] = ok ? b[s] : T(0);
    }
  }
#pragma unroll
  for (int j = NP - 1; j >= 0; --j) {
    if (j >= np) continue;
    const int k0 = K0 + kLuNb * j;
    const int kb = min(kLuNb, K1 - k0);
    {
      // every load issued before the first LDS store; entries outside the
      // panel's kb columns (and blocks above the first panel) are zeroed,
      // their addresses clamped into the front
      constexpr int kPer = kLuNb * kLuNb / 256;  // 256 threads
      const T *slot = inv_slot(dinv, d, k0);
      T vr[NP][kPer], vi[NP][kPer];
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = threadIdx.x + 256 * u, r = e % kLuNb, c = e / kLuNb;
        const bool cin = c < kb;
#pragma unroll
        for (int t = 0; t < NP; ++t) {
          const int i = j - t;                                   // panel of the rows
          const bool ok = cin && i >= 0 && (t > 0 || r < kb);
          const int row = K0 + kLuNb * max(i, 0) + r;             // < K1 when ok
          const long long idx = static_cast<long long>(k0 + c) * m + (ok ? row : k0);
          const T a = Fr[idx], b = Fi[idx];
          vr[t][u] = ok ? a : T(0);
          vi[t][u] = ok ? b : T(0);
        }
      }
      const T dr = threadIdx.x < kLuNb ? slot[2 * kLuNb * kLuNb + threadIdx.x * (kLuNb + 1)] : T(0);
      const T di = threadIdx.x < kLuNb ? slot[3 * kLuNb * kLuNb + threadIdx.x * (kLuNb + 1)] : T(0);
      __syncthreads();  // the previous panel's steps have read the stages
#pragma unroll
      for (int u = 0; u < kPer; ++u) {
        const int e = threadIdx.x + 256 * u, r = e % kLuNb, c = e / kLuNb;
        Sr[0][r][c] = vr[0][u];
        Si[0][r][c] = vi[0][u];
        const int pi = block_rows_pos<T>(r), pk = block_rows_pos<T>(c);
#pragma unroll
        for (int t = 1; t < NP; ++t) {
          Sr[t][pi][pk] = vr[t][u];
          Si[t][pi][pk] = vi[t][u];
        }
      }
      if (threadIdx.x < kLuNb) {
        Dr[threadIdx.x] = dr;
        Di[threadIdx.x] = di;
      }
    }
    __syncthreads();
    // X_j <- U_jj^-1 X_j by back substitution over the panel's kb rows, in
    // four steps of 8 rows from the last: the lane group G holding rows
    // 8 G .. 8 G + 7 solves them in its own registers, passes them to the
    // other lane groups of the column by shuffles, and the groups above
    // subtract their products with U[i][8 G .. 8 G + 7]
#pragma unroll
    for (int G = 3; G >= 0; --G) {
      if (8 * G >= kb) continue;
      if (g == G) {
#pragma unroll
        for (int s = 7; s >= 0; --s) {
          const int r = 8 * G + s;
          if (r >= kb) continue;
          const T d_r = Dr[r], d_i = Di[r];
          const T x_r = xr[j][s / 4][s % 4], x_i = xi[j][s / 4][s % 4];
          const T yr = x_r * d_r - x_i * d_i, yi = x_r * d_i + x_i * d_r;
          xr[j][s / 4][s % 4] = yr;
          xi[j][s / 4][s % 4] = yi;
#pragma unroll
          for (int q = 0; q < s; ++q) {
            const T ur = Sr[0][8 * G + q][r], ui = Si[0][8 * G + q][r];
            xr[j][q / 4][q % 4] -= ur * yr - ui * yi;
            xi[j][q / 4][q % 4] -= ur * yi + ui * yr;
          }
        }
      }
      if (G == 0) break;
      const int src = (lane & 15) + 16 * G;
      T br[8], bi[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        br[s] = __shfl(xr[j][s / 4][s % 4], src);
        bi[s] = __shfl(xi[j][s / 4][s % 4], src);
      }
      if (g < G) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          T ar = xr[j][q / 4][q % 4], ai = xi[j][q / 4][q % 4];
#pragma unroll
          for (int s = 0; s < 8; ++s) {
            const T ur = Sr[0][8 * g + q][8 * G + s], ui = Si[0][8 * g + q][8 * G + s];
            ar -= ur * br[s] - ui * bi[s];
            ai -= ur * bi[s] + ui * br[s];
          }
          xr[j][q / 4][q % 4] = ar;
          xi[j][q / 4][q % 4] = ai;
        }
      }
    }
    // X_i -= U_ij X_j for the earlier panels
#pragma unroll
    for (int t = 1; t < NP; ++t) {
      const int i = j - t;
      if (i < 0) break;
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const int k = 16 * (s / 4) + Mfma<T>::orow(g, s % 4);
        const T br = xr[j][s / 4][s % 4], bi = xi[j][s / 4][s % 4];
#pragma unroll
        for (int ib = 0; ib < 2; ++ib) {
          const T ar = Sr[t][16 * ib + lane % 16][k], ai = Si[t][16 * ib + lane % 16][k];
          xr[i][ib] = Mfma<T>::mma(ar, -br, xr[i][ib]);
          xr[i][ib] = Mfma<T>::mma(ai, bi, xr[i][ib]);
          xi[i][ib] = Mfma<T>::mma(ar, -bi, xi[i][ib]);
          xi[i][ib] = Mfma<T>::mma(ai, -br, xi[i][ib]);
        }
      }
    }
  }
  if (!live) return;
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int r0 = 32 * j + 8 * g;
    v8 a, b;
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      a[s] = xr[j][s / 4][s % 4];
      b[s] = xi[j][s / 4][s % 4];
    }
    if (r0 + 8 <= nrows) {
      *reinterpret_cast<v8 *>(cr + r0) = a;
      *reinterpret_cast<v8 *>(ci + r0) = b;
    } else if (r0 < nrows) {
#pragma unroll
      for (int s = 0; s < 8; ++s)
        if (r0 + s < nrows) {
          cr[r0 + s] = a[s];
          ci[r0 + s] = b[s];
        }
    }
  }
}

// Rows [p, m) of the panel columns: X <- X U_D^-1.  Four waves per
// workgroup, 32 rows x 32 columns per wave; the MFMAs compute
// (X U)^T = U^T X^T: lane l supplies U[4s + l/16][16 jb + l%16] and
// X[16 rb + l%16][4s + l/16], and holds the results of rows 16 rb + l%16.
template <typename T>
__global__ __launch_bounds__(256) void k_right_apply(const int *__restrict__ list, int nlist,
                                                     const FrontDev *__restrict__ fr, int k0,
                                                     int nb, const T *__restrict__ dinv,
                                                     T *__restrict__ arena) {
  using acc_t = typename Mfma<T>::acc_t;
  const FrontDev d = fr[list[blockIdx.y]];
  if (k0 >= d.p || d.c == 0) return;
  const int kb = min(nb, d.p - k0);
  const int row0 = d.p + blockIdx.x * 128;
  if (row0 >= d.m) return;
  const long long m = d.m;
  T *Fr = arena + d.off, *Fi = Fr + m * m;
  __shared__ T Ur[kLuNb][kLuNb + 1], Ui[kLuNb][kLuNb + 1];  // U[k][j]
  stage_slot<256>(inv_slot(dinv, d, k0) + 2 * kLuNb * kLuNb, Ur, Ui);
  __syncthreads();
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  const int rw = row0 + wave * 32;
  T xr[2][kLuNb / 4], xi[2][kLuNb / 4];
#pragma unroll
  for (int rb = 0; rb < 2; ++rb) {
    const int r = rw + 16 * rb + lane % 16;
    const long long rr = min(r, d.m - 1);
#pragma unroll
    for (int s = 0; s < kLuNb / 4; ++s) {
      const int k = 4 * s + lane / 16;
      const long long idx = static_cast<long long>(k0 + min(k, kb - 1)) * m + rr;
      const bool ok = r < d.m && k < kb;
      xr[rb][s] = ok ? Fr[idx] : T(0);
      xi[rb][s] = ok ? Fi[idx] : T(0);
    }
  }
  acc_t cr_[2][2], ci_[2][2];
#pragma unroll
  for (int jb = 0; jb < 2; ++jb)
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      cr_[jb][rb] = acc_t{0, 0, 0, 0};
      ci_[jb][rb] = acc_t{0, 0, 0, 0};
    }
#pragma unroll
  for (int s = 0; s < kLuNb / 4; ++s) {
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
      const T ur = Ur[4 * s + lane / 16][16 * jb + lane % 16];
      const T ui = Ui[4 * s + lane / 16][16 * jb + lane % 16];
#pragma unroll
      for (int rb = 0; rb < 2; ++rb) {
        cr_[jb][rb] = Mfma<T>::mma(ur, xr[rb][s], cr_[jb][rb]);
        cr_[jb][rb] = Mfma<T>::mma(-ui, xi[rb][s], cr_[jb][rb]);
        ci_[jb][rb] = Mfma<T>::mma(ui, xr[rb][s], ci_[jb][rb]);
        ci_[jb][rb] = Mfma<T>::mma(ur, xi[rb][s], ci_[jb][rb]);
      }
    }
  }
#pragma unroll
  for (int jb = 0; jb < 2; ++jb)
#pragma unroll
    for (int rb = 0; rb < 2; ++rb) {
      const int r = rw + 16 * rb + lane % 16;
      if (r >= d.m) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int j = 16 * jb + Mfma<T>::ocol(lane, v);
        if (j < kb) {
          const long long idx = static_cast<long long>(k0 + j) * m + r;
          Fr[idx] = cr_[jb][rb][v];
          Fi[idx] = ci_[jb][rb][v];
        }
      }
    }
}

// ---------------------------------------------- row exchanges of a block
//
// The exchanges of block [K0, K1) (ipiv, in order) act on the rows [K0, p)
// of the trailing F columns and of E as one permutation.  k_block_perm
// composes them in an LDS row map and lists the rows that move (at most
// 2 (K1 - K0)) with the row each takes its values from; k_apply_perm moves
// them, a wave per column.  Fronts with more than kPanelMapRows rows below
// K0 take the exchanges one by one in k_block_swaps instead.
constexpr int kPermStride = kLuPermStride;  // count, destinations, sources

__global__ __launch_bounds__(256) void k_block_perm(const int *__restrict__ list, int nlist,
                                                    const FrontDev *__restrict__ fr, int K0, int NB,
                                                    const int *__restrict__ ipiv,
                                                    int *__restrict__ bperm) {
  const FrontDev d = fr[list[blockIdx.x]];
  int *out = bperm + static_cast<long long>(blockIdx.x) * kPermStride;
  const int n = d.p - K0;
  if (n <= 0 || n > kPanelMapRows) {
    if (threadIdx.x == 0) out[0] = 0;
    return;
  }
  const int K1 = min(K0 + NB, d.p);
  __shared__ int map[kPanelMapRows];  // row K0 + i holds the values of row K0 + map[i]
  __shared__ int pv[kLuNbOuter];
  __shared__ int cnt;
  for (int i = threadIdx.x; i < n; i += blockDim.x) map[i] = i;
  if (threadIdx.x < K1 - K0) pv[threadIdx.x] = ipiv[d.piv_off + K0 + threadIdx.x] - K0;
  if (threadIdx.x == 0) cnt = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int r = 0; r < K1 - K0; ++r) {
      const int q = pv[r];
      const int t = map[r];
      map[r] = map[q];
      map[q] = t;
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    if (map[i] != i) {
      const int k = atomicAdd(&cnt, 1);
      out[1 + k] = K0 + i;
      out[1 + 2 * kLuNbOuter + k] = K0 + map[i];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) out[0] = cnt;
}

template <typename T>
__device__ inline void trailing_column(const FrontDev &d, int K1, int col, T *arena, T *&xr,
                                       T *&xi) {
  const long long m = d.m, pe = d.pe, mm = m * m;
  const int nF = d.m - K1;
  if (col < nF) {
    xr = arena + d.off + static_cast<long long>(K1 + col) * m;
    xi = xr + mm;
  } else {
    xr = arena + d.off + 2 * mm + static_cast<long long>(col - nF) * pe;
    xi = xr + pe * d.p;
  }
}

// The listed row moves on F columns [K1, m) and E, or (c1 > c0) on F
// columns [c0, c1) only (an earlier block's L columns): a wave per column,
// lanes over the moved rows; all of a column's loads complete before its
// stores.
template <typename T>
__global__ __launch_bounds__(256) void k_apply_perm(const int *__restrict__ list, int nlist,
                                                    const FrontDev *__restrict__ fr, int K0, int NB,
                                                    const int *__restrict__ bperm,
                                                    T *__restrict__ arena, int c0, int c1) {
  const FrontDev d = fr[list[blockIdx.y]];
  if (K0 >= d.p) return;
  const int K1 = min(K0 + NB, d.p);
  const int *pl = bperm + static_cast<long long>(blockIdx.y) * kPermStride;
  const int n = pl[0];
  const int col = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
  const bool range = c1 > c0;
  if (n == 0 || col >= (range ? c1 - c0 : d.m - K1 + d.p)) return;
  T *Xr, *Xi;
  if (range) {
    Xr = arena + d.off + static_cast<long long>(c0 + col) * d.m;
    Xi = Xr + static_cast<long long>(d.m) * d.m;
  } else {
    trailing_column(d, K1, col, arena, Xr, Xi);
  }
  constexpr int kPer = 2 * kLuNbOuter / 64;
  T vr[kPer], vi[kPer];
#pragma unroll
  for (int g = 0; g < kPer; ++g) {
    const int e = lane + 64 * g;
    if (e < n) {
      const int src = pl[1 + 2 * kLuNbOuter + e];
      vr[g] = Xr[src];
      vi[g] = Xi[src];
    }
  }
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // every source read before any write
#pragma unroll
  for (int g = 0; g < kPer; ++g) {
    const int e = lane + 64 * g;
    if (e < n) {
      const int dst = pl[1 + e];
      Xr[dst] = vr[g];
      Xi[dst] = vi[g];
    }
  }
}

// Fallback for fronts whose rows [K0, p) exceed the row map: the exchanges
// of the block applied in order, a thread per trailing column (or, c1 > c0,
// per F column of [c0, c1)).
template <typename T>
__global__ __launch_bounds__(256) void k_block_swaps(const int *__restrict__ list, int nlist,
                                                     const FrontDev *__restrict__ fr, int K0, int NB,
                                                     const int *__restrict__ ipiv,
                                                     T *__restrict__ arena, int c0, int c1) {
  const FrontDev d = fr[list[blockIdx.y]];
  if (K0 >= d.p || d.p - K0 <= kPanelMapRows) return;
  const int K1 = min(K0 + NB, d.p);
  __shared__ int pv[kLuNbOuter];
  if (threadIdx.x < K1 - K0) pv[threadIdx.x] = ipiv[d.piv_off + K0 + threadIdx.x];
  __syncthreads();
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const bool range = c1 > c0;
  if (col >= (range ? c1 - c0 : d.m - K1 + d.p)) return;
  T *Xr, *Xi;
  if (range) {
    Xr = arena + d.off + static_cast<long long>(c0 + col) * d.m;
    Xi = Xr + static_cast<long long>(d.m) * d.m;
  } else {
    trailing_column(d, K1, col, arena, Xr, Xi);
  }
  for (int r = K0; r < K1; ++r) {
    const int q = pv[r - K0];
    if (q != r) {
      T t = Xr[r]; Xr[r] = Xr[q]; Xr[q] = t;
      t = Xi[r]; Xi[r] = Xi[q]; Xi[q] = t;
    }
  }
}

// TRAIL = TRAILHEAD (the next block's fully summed columns of F) + TRAILREST
// (the remaining columns of F and all of E); the split lets the next
// block's panels start while the rest of the update runs.  Z = U11^-1 is
// built block by block from the last: k_block_z finishes the block's rows
// [K0, K1), and ZBLK then updates all rows above the block at once
// (K = 128).
// With paired blocks (two-level blocking, LuPlan::pair_blocks), block A's
// update of the rows below it is deferred until block B = [K1, K1n) is
// factored: PAIRROWS applies it to B's rows of the columns right of B and
// of E (K = 128), and the pair's TRAILHEAD / TRAILREST / TRAIL then update
// everything below B with K = 256 (the two blocks' L columns and U rows are
// contiguous).
// (Regions 1 and 3 are unused: the block rows of the trailing columns and
// of Z are finished by the triangular sweeps k_block_rows and k_block_z.)
enum GemmRegion {
  kGemmInner = 0, kGemmTrail = 2, kGemmTrailHead = 4,
  kGemmTrailRest = 5, kGemmZBlk = 6, kGemmPairRows = 7
};

// Operands of one front's update.  Columns [0, N1) of B and C come from the
// first segment (rows [0, M) of C), columns [N1, N) from the second (rows
// [0, M2)); A is shared.
template <typename T> struct GemmView {
  const T *Ar, *Ai;
  long long lda;
  const T *Br, *Bi, *B2r, *B2i;
  long long ldb, ldb2;
  T *Cr, *Ci, *C2r, *C2i;
  long long ldc, ldc2;
  int M, M2, N, N1, K;
};

// Output rows (M for the first column segment, M2 for the second), columns
// N and depth K of one front's update; shared by the kernel and the host's
// flop accounting.
struct GemmShape {
  int M, M2, N, N1, K;
};
__host__ __device__ inline GemmShape gemm_shape(int region, const FrontDev &d, int k0, int nb,
                                                int K0, int NB) {
  GemmShape g{0, 0, 0, 0, 0};
  const bool trail = region == kGemmTrail || region == kGemmTrailHead ||
                     region == kGemmTrailRest || region == kGemmZBlk || region == kGemmPairRows;
  if (trail ? K0 >= d.p : k0 >= d.p) return g;
  const int kb = min(nb, d.p - k0);
  const int K1 = min(K0 + NB, d.p);           // end of the block (or pair) [K0, K1)
  const int K1n = min(K1 + kLuNbOuter, d.p);  // end of the next block
  switch (region) {
    case kGemmInner:
      g = {d.m - k0 - kb, d.m - k0 - kb, K1 - k0 - kb, K1 - k0 - kb, kb};
      break;
    case kGemmTrail:
      g = {d.m - K1, d.p - K1, d.m - K1 + d.p, d.m - K1, K1 - K0};
      break;
    case kGemmTrailHead:
      g = {d.m - K1, d.m - K1, K1n - K1, K1n - K1, K1 - K0};
      break;
    case kGemmTrailRest:
      g = {d.m - K1, d.p - K1, d.m - K1n + d.p, d.m - K1n, K1 - K0};
      break;
    case kGemmPairRows:
      g = {K1n - K1, K1n - K1, d.m - K1n + d.p, d.m - K1n, K1 - K0};
      break;
    case kGemmZBlk:
      g = {K0, K0, d.p - K0, d.p - K0, K1 - K0};
      break;
    default:
      break;
  }
  return g;
}

template <typename T>
__device__ inline bool gemm_view(int region, const FrontDev &d, int k0, int nb, int K0, int NB,
                                 T *arena, GemmView<T> &v) {
  const bool trail = region == kGemmTrail || region == kGemmTrailHead ||
                     region == kGemmTrailRest || region == kGemmZBlk || region == kGemmPairRows;
  if (trail ? K0 >= d.p : k0 >= d.p) return false;
  // p: the leading dimension of E and Z (counts below use d.p)
  const long long m = d.m, p = d.pe, mm = m * m, pp = p * d.p;
  T *Fr = arena + d.off, *Fi = Fr + mm;
  T *Er = Fi + mm, *Ei = Er + pp;
  T *Zr = Ei + pp, *Zi = Zr + pp;
  const int kb = min(nb, d.p - k0);
  const int K1 = min(K0 + NB, d.p);
  const int K1n = min(K1 + kLuNbOuter, d.p);
  switch (region) {
    case kGemmTrailHead: {
      const long long a = K0 * m + K1, b = K1 * m + K0, c = K1 * m + K1;
      v = {Fr + a, Fi + a, m, Fr + b, Fi + b, nullptr, nullptr, m, 0,
           Fr + c, Fi + c, nullptr, nullptr, m, 0,
           d.m - K1, d.m - K1, K1n - K1, K1n - K1, K1 - K0};
      break;
    }
    case kGemmTrailRest: {
      const long long a = K0 * m + K1, b = static_cast<long long>(K1n) * m + K0,
                      c = static_cast<long long>(K1n) * m + K1;
      v = {Fr + a, Fi + a, m, Fr + b, Fi + b, Er + K0, Ei + K0, m, p,
           Fr + c, Fi + c, Er + K1, Ei + K1, m, p,
           d.m - K1, d.p - K1, d.m - K1n + d.p, d.m - K1n, K1 - K0};
      break;
    }
    case kGemmPairRows: {  // TRAILREST restricted to the next block's rows [K1, K1n)
      const long long a = K0 * m + K1, b = static_cast<long long>(K1n) * m + K0,
                      c = static_cast<long long>(K1n) * m + K1;
      v = {Fr + a, Fi + a, m, Fr + b, Fi + b, Er + K0, Ei + K0, m, p,
           Fr + c, Fi + c, Er + K1, Ei + K1, m, p,
           K1n - K1, K1n - K1, d.m - K1n + d.p, d.m - K1n, K1 - K0};
      break;
    }
    case kGemmInner: {
      const long long a = k0 * m + k0 + kb, b = (k0 + kb) * m + k0, c = (k0 + kb) * m + k0 + kb;
      v = {Fr + a, Fi + a, m, Fr + b, Fi + b, nullptr, nullptr, m, 0,
           Fr + c, Fi + c, nullptr, nullptr, m, 0,
           d.m - k0 - kb, d.m - k0 - kb, K1 - k0 - kb, K1 - k0 - kb, kb};
      break;
    }
    case kGemmTrail: {
      const long long a = K0 * m + K1, b = K1 * m + K0, c = K1 * m + K1;
      v = {Fr + a, Fi + a, m, Fr + b, Fi + b, Er + K0, Ei + K0, m, p,
           Fr + c, Fi + c, Er + K1, Ei + K1, m, p,
           d.m - K1, d.p - K1, d.m - K1 + d.p, d.m - K1, K1 - K0};
      break;
    }
    case kGemmZBlk: {
      const long long a = K0 * m, b = K0 * p + K0, c = K0 * p;
      v = {Fr + a, Fi + a, m, Zr + b, Zi + b, nullptr, nullptr, p, 0,
           Zr + c, Zi + c, nullptr, nullptr, p, 0,
           K0, K0, d.p - K0, d.p - K0, K1 - K0};
      break;
    }
    default:
      return false;
  }
  return v.M > 0 && v.N > 0 && v.K > 0;
}

constexpr int kXcds = 8;       // accelerator complex dies of an MI355X
constexpr int kGemmGroup = 8;  // row blocks per tile group
constexpr int GKC = 16;  // K chunk staged in LDS

// Tile shapes (chosen per level by the autotuner): WMW x WNW waves, each
// owning WMB x WNB MFMA 16x16 blocks, so a workgroup computes a
// (16 WMW WMB) x (16 WNW WNB) complex tile.
//   shape 0: 64 x 64, 256 threads, K chunk 16, 3M form, two waves per SIMD;
//   shape 1: 64 x 64, 256 threads, K chunk 4, 3M form, registers held to
//            three waves per SIMD (a quarter of shape 0's LDS);
//   shape 2: 64 x 64, 256 threads, K chunk 8, 3M form, operands loaded
//            straight into a ring of LDS stages (k_gemm_lds; FP64 only,
//            FP32 fronts use shape 1's tile);
//   shape 3: 64 x 64, 256 threads, K chunk 8, 3M form, two waves per SIMD.
// The shorter K chunks trade more barriers for more resident workgroups,
// which hide the first chunk's and C's load latency of each tile.
enum { kShape64x64 = 0, kShape64x64k4 = 1, kShapeLds = 2, kShape64x64k8 = 3, kShapeCount = 4 };
static_assert(kShapeCount == kGemmShapes, "shape count");

// Output tile (row0, col0) of this workgroup in its front's update v.  The
// tiles holding output are enumerated as two blocks: every row tile of the
// column tiles that start in the first column segment, then the row tiles
// above M2 of the column tiles in the second segment (the E columns of a
// trailing update hold output in rows [0, p - K1) only; a tile below would
// run its whole K loop to store nothing).  False for a workgroup past the
// front's tiles.
template <typename T>
__device__ inline bool gemm_tile(const GemmView<T> &v, int TM, int TN, int order, int &row0,
                                 int &col0) {
  const int tiles_m = (v.M + TM - 1) / TM, tiles_n = (v.N + TN - 1) / TN;
  const int split = min((v.N1 + TN - 1) / TN, tiles_n);
  const int tiles_m2 = min(max(v.M2, 0) + TM - 1, v.M + TM - 1) / TM;
  const int na = tiles_m * split, ntiles = na + tiles_m2 * (tiles_n - split);
  int tile;
  if (order == kOrderLinear) {
    tile = blockIdx.x;  // row tiles fastest
  } else {
    // XCD-aware order: workgroups are dispatched round-robin over the 8
    // XCDs, so XCD x (= blockIdx.x % 8, gridDim.x being a multiple of 8)
    // receives a contiguous eighth of this front's tiles (rotated by front);
    // within it, tiles go kGemmGroup row tiles at a time so that
    // co-resident workgroups share A and B tiles in that XCD's L2.
    const int per = (ntiles + kXcds - 1) / kXcds;
    const int q = static_cast<int>(blockIdx.x / kXcds);
    if (q >= per) return false;
    const int slot = (static_cast<int>(blockIdx.x % kXcds) - static_cast<int>(blockIdx.y % kXcds) +
                      kXcds) % kXcds;
    tile = slot * per + q;
  }
  if (tile >= ntiles) return false;
  int rows = tiles_m, cols = split, coff = 0;
  if (tile >= na) {
    tile -= na;
    rows = tiles_m2;
    cols = tiles_n - split;
    coff = split;
  }
  int tm, tn;
  if (order == kOrderLinear) {
    tm = tile % rows;
    tn = tile / rows;
  } else {
    const int in_group = kGemmGroup * cols;
    const int first_m = (tile / in_group) * kGemmGroup;
    const int group_m = min(rows - first_m, kGemmGroup);
    tm = first_m + (tile % in_group) % group_m;
    tn = (tile % in_group) / group_m;
  }
  row0 = tm * TM;
  col0 = (tn + coff) * TN;
  return true;
}

template <typename T, int WMW, int WNW, int WMB, int WNB, int KC = GKC, int MINW = 1, bool M3 = false>
__global__ __launch_bounds__(64 * WMW * WNW, MINW) void k_gemm(const int *__restrict__ list,
                                                               int nlist,
                                                               const FrontDev *__restrict__ fr,
                                                               int region, int k0, int nb,
                                                               int K0, int NB, int order,
                                                               T *__restrict__ arena) {
  constexpr int GKC = KC;  // K chunk of this instance
  constexpr int THREADS = 64 * WMW * WNW;
  constexpr int TM = 16 * WMW * WMB, TN = 16 * WNW * WNB;
  // LDS layouts: A as [k][row] with row stride TM + 16 (the four k rows of a
  // wave's reads fall in different banks); B as [column][k] with stride
  // GKC + 1, so the stash writes (consecutive k of a column) are contiguous
  // and the operand reads (16 columns at one k) hit distinct banks
  constexpr int LDA_ = TM + 16, LDB_ = GKC + 1;
  constexpr int AE = TM * GKC / THREADS, BE = TN * GKC / THREADS;
  static_assert(AE * THREADS == TM * GKC && BE * THREADS == TN * GKC, "tile / thread mismatch");
  const int f = list[blockIdx.y];
  const FrontDev d = fr[f];
  GemmView<T> v;
  if (!gemm_view(region, d, k0, nb, K0, NB, arena, v)) return;
  int row0, col0;
  if (!gemm_tile(v, TM, TN, order, row0, col0)) return;
  // two LDS buffers: the next K chunk is stored while the current one feeds
  // the MFMAs, so each chunk needs one barrier
  __shared__ T sAr[2][GKC][LDA_], sAi[2][GKC][LDA_], sBr[2][TN][LDB_], sBi[2][TN][LDB_];
  const int t = threadIdx.x, l = t % 64, w = t / 64;
  const int wm = w % WMW, wn = w / WMW;
  // global -> register staging (element e = t + s * THREADS): A (row e % TM,
  // k e / TM), B (k e % GKC, column e / GKC).  Loads are unconditional from
  // clamped in-bounds addresses; out-of-range entries become zero.
  const T *bcol_r[BE], *bcol_i[BE];
  bool bcol_ok[BE];
#pragma unroll
  for (int s = 0; s < BE; ++s) {
    const int col = col0 + (t + s * THREADS) / GKC;
    bcol_ok[s] = col < v.N;
    if (!bcol_ok[s]) {
      bcol_r[s] = v.Br;
      bcol_i[s] = v.Bi;
    } else if (col < v.N1) {
      bcol_r[s] = v.Br + col * v.ldb;
      bcol_i[s] = v.Bi + col * v.ldb;
    } else {
      bcol_r[s] = v.B2r + (col - v.N1) * v.ldb2;
      bcol_i[s] = v.B2i + (col - v.N1) * v.ldb2;
    }
  }
  // staged global loads: chunk c + 1 is requested while chunk c feeds the
  // MFMAs.  The loads come from clamped in-bounds addresses and stay raw
  // until the stash, which zeroes the out-of-range entries: a select right
  // after the loads would make the wave wait for them before the chunk's
  // MFMAs.  (A second register set, requesting chunk c + 2, measured slower.)
  T ga_r[AE], ga_i[AE], gb_r[BE], gb_i[BE];
  bool ga_ok[AE], gb_ok[BE];
  auto fetch = [&](int kc) {
#pragma unroll
    for (int s = 0; s < AE; ++s) {
      const int e = t + s * THREADS, r = e % TM, k = e / TM;
      const bool oka = row0 + r < v.M && kc + k < v.K;
      const long long ia = oka ? static_cast<long long>(kc + k) * v.lda + row0 + r : 0;
      ga_r[s] = v.Ar[ia];
      ga_i[s] = v.Ai[ia];
      ga_ok[s] = oka;
    }
#pragma unroll
    for (int s = 0; s < BE; ++s) {
      const int kbb = kc + (t + s * THREADS) % GKC;
      const bool okb = bcol_ok[s] && kbb < v.K;
      const int ib = okb ? kbb : 0;
      gb_r[s] = bcol_r[s][ib];
      gb_i[s] = bcol_i[s][ib];
      gb_ok[s] = okb;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int s = 0; s < AE; ++s) {
      const int e = t + s * THREADS;
      sAr[buf][e / TM][e % TM] = ga_ok[s] ? ga_r[s] : T(0);
      sAi[buf][e / TM][e % TM] = ga_ok[s] ? ga_i[s] : T(0);
    }
#pragma unroll
    for (int s = 0; s < BE; ++s) {
      const int e = t + s * THREADS;
      sBr[buf][e / GKC][e % GKC] = gb_ok[s] ? -gb_r[s] : T(0);
      sBi[buf][e / GKC][e % GKC] = gb_ok[s] ? -gb_i[s] : T(0);
    }
  };
  // B is staged negated (Bn = -B) and the accumulators start from C, so the
  // tile ends with plain stores; the C loads are issued with the first
  // chunk's loads and share their latency.  Complex products:
  //  * four real MFMAs:  cr = Cr + sum (Ar Bnr - Ai Bni),  ci = Ci + sum (Ar Bni + Ai Bnr);
  //  * three (M3, the 3M form):  cr = x = Cr + sum Ar Bnr,  cy = y = -sum Ai Bni,
  //    ci = z = Cr + Ci + sum (Ar + Ai)(Bnr + Bni), and C - AB = (x + y) + i (z - x + y).
  //    The 3M form is normwise stable (its error bound is a small multiple
  //    of the four-product form's) and needs a third accumulator set.
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int NY = M3 ? WMB : 1;
  acc_t cr[WMB][WNB], ci[WMB][WNB], cy[NY][WNB];
#pragma unroll
  for (int a = 0; a < WMB; ++a) {
    const int row = row0 + wm * 16 * WMB + a * 16 + l % 16;
#pragma unroll
    for (int b = 0; b < WNB; ++b) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int col = col0 + wn * 16 * WNB + b * 16 + Mfma<T>::ocol(l, e);
        T xr = T(0), xi = T(0);
        if (col < v.N1) {
          if (row < v.M) {
            const long long ic = static_cast<long long>(col) * v.ldc + row;
            xr = v.Cr[ic];
            xi = v.Ci[ic];
          }
        } else if (col < v.N && row < v.M2) {
          const long long ic = static_cast<long long>(col - v.N1) * v.ldc2 + row;
          xr = v.C2r[ic];
          xi = v.C2i[ic];
        }
        cr[a][b][e] = xr;
        if constexpr (M3) {
          ci[a][b][e] = xr + xi;
          cy[a][b][e] = T(0);
        } else {
          ci[a][b][e] = xi;
        }
      }
    }
  }
  auto compute = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < GKC; kk += 4) {
      const int kr = kk + l / 16;
      T a_r[WMB], a_i[WMB], b_r[WNB], b_i[WNB];
#pragma unroll
      for (int q = 0; q < WMB; ++q) {
        a_r[q] = sAr[buf][kr][wm * 16 * WMB + q * 16 + l % 16];
        a_i[q] = sAi[buf][kr][wm * 16 * WMB + q * 16 + l % 16];
      }
#pragma unroll
      for (int q = 0; q < WNB; ++q) {
        b_r[q] = sBr[buf][wn * 16 * WNB + q * 16 + l % 16][kr];
        b_i[q] = sBi[buf][wn * 16 * WNB + q * 16 + l % 16][kr];
      }
      if constexpr (M3) {
        T a_s[WMB], b_s[WNB];
#pragma unroll
        for (int q = 0; q < WMB; ++q) a_s[q] = a_r[q] + a_i[q];
#pragma unroll
        for (int q = 0; q < WNB; ++q) b_s[q] = b_r[q] + b_i[q];
#pragma unroll
        for (int a = 0; a < WMB; ++a)
#pragma unroll
          for (int b = 0; b < WNB; ++b) {
            cr[a][b] = Mfma<T>::mma(b_r[b], a_r[a], cr[a][b]);
            cy[a][b] = Mfma<T>::mma(b_i[b], -a_i[a], cy[a][b]);
            ci[a][b] = Mfma<T>::mma(b_s[b], a_s[a], ci[a][b]);
          }
      } else {
#pragma unroll
        for (int a = 0; a < WMB; ++a)
#pragma unroll
          for (int b = 0; b < WNB; ++b) {
            cr[a][b] = Mfma<T>::mma(b_r[b], a_r[a], cr[a][b]);
            cr[a][b] = Mfma<T>::mma(b_i[b], -a_i[a], cr[a][b]);
            ci[a][b] = Mfma<T>::mma(b_i[b], a_r[a], ci[a][b]);
            ci[a][b] = Mfma<T>::mma(b_r[b], a_i[a], ci[a][b]);
          }
      }
    }
  };
  const int nch = (v.K + GKC - 1) / GKC;
  fetch(0);
  stash(0);
  __syncthreads();
  int buf = 0;
  for (int c = 0; c < nch; ++c) {
    const bool more = c + 1 < nch;
    if (more) fetch((c + 1) * GKC);
    compute(buf);
    if (more) stash(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int a = 0; a < WMB; ++a) {
    const int row = row0 + wm * 16 * WMB + a * 16 + l % 16;
#pragma unroll
    for (int b = 0; b < WNB; ++b) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int col = col0 + wn * 16 * WNB + b * 16 + Mfma<T>::ocol(l, e);
        if (col >= v.N) continue;
        T re = cr[a][b][e], im = ci[a][b][e];
        if constexpr (M3) {
          const T y = cy[a][b][e];
          im = (im - re) + y;
          re = re + y;
        }
        if (col < v.N1) {
          if (row >= v.M) continue;
          const long long ic = static_cast<long long>(col) * v.ldc + row;
          v.Cr[ic] = re;
          v.Ci[ic] = im;
        } else {
          if (row >= v.M2) continue;
          const long long ic = static_cast<long long>(col - v.N1) * v.ldc2 + row;
          v.C2r[ic] = re;
          v.C2i[ic] = im;
        }
      }
    }
  }
}

// ----------------------------------------- frontal GEMM, direct LDS loads
//
// Tile shape 2 (FP64): the 64 x 64 3M tile of shape 3 (2 x 2 waves, each
// 2 x 2 MFMA blocks, K chunk 8), fed by global -> LDS loads that write
// LDS without passing through registers (global_load_lds_dwordx4: each lane
// moves 16 bytes, a wave's 1 KB lands contiguously).  The chunks rotate
// through a ring of S LDS stages (kGemmLdsStages = 2: chunk c + 1 is in
// flight while chunk c feeds the MFMAs; with three, chunk c + 2 as well),
// and a chunk costs one barrier and no VALU work to stage.  Per chunk each
// wave issues four loads: A columns k = 2w, 2w + 1
// (lanes 0-31 the real plane's 64 rows, two per lane; lanes 32-63 the
// imaginary), and the real and imaginary planes of B's column group w
// (16 columns x 8 k; lane l: column l % 16, k pair l / 16).  Stage layout:
//   A  [k (8)][plane (2)][row (64)], 16 doubles of padding per k;
//   B  per plane [column group (4)][k pair (4)][column (16)][2].
// Loads never stray outside the front's planes: a row pair starting beyond
// M, or a column beyond N, is redirected to row pair 0 or the last column
// (such rows and columns only reach outputs that are not stored), and a row
// pair (M - 1, M) reads the element after its column's last row, which is
// the next column's first (or the next plane's).  A K tail (K mod 8) is
// staged through registers with zeros past K, as in k_gemm.
//
// Each stage's A and B planes are arrays of their own, addressed with
// compile-time stage indices: the compiler then tells the MFMA operand reads
// of one stage from the loads in flight into the others and waits only for
// those it must (the waits the ring needs are written out below).
constexpr int kLdsKc = 8, kLdsAk = 2 * 64 + 16;
constexpr int kLdsASize = kLdsKc * kLdsAk, kLdsBSize = 4 * 128;  // doubles per stage (B: per plane)

// s_waitcnt immediate that waits until at most n vector-memory operations
// of this wave are outstanding (other counters untouched)
constexpr int wait_vm_imm(int n) { return (n & 0xF) | ((n >> 4) << 14) | (0x7 << 4) | (0xF << 8); }

// One global -> LDS load of 16 bytes per lane (the wave's 1 KB lands at l,
// wave-uniform, + 16 x lane).  A device function of its own: the builtin
// exists for the device only.
__device__ inline void lds_load16(const double *g, double *l) {
  __builtin_amdgcn_global_load_lds(g, l, 16, 0, 0);
}

template <int MINW, int S>
__global__ __launch_bounds__(256, MINW) void k_gemm_lds(const int *__restrict__ list, int nlist,
                                                      const FrontDev *__restrict__ fr, int region,
                                                      int k0, int nb, int K0, int NB, int order,
                                                      double *__restrict__ arena) {
  constexpr int TM = 64, TN = 64, KC = kLdsKc, AK = kLdsAk;
  static_assert(S == 2 || S == 3, "the stage arrays below are written out for two or three stages");
  const int f = list[blockIdx.y];
  const FrontDev d = fr[f];
  GemmView<double> v;
  if (!gemm_view(region, d, k0, nb, K0, NB, arena, v)) return;
  int row0, col0;
  if (!gemm_tile(v, TM, TN, order, row0, col0)) return;
  __shared__ __attribute__((aligned(16))) double sa0[kLdsASize], sbr0[kLdsBSize], sbi0[kLdsBSize];
  __shared__ __attribute__((aligned(16))) double sa1[kLdsASize], sbr1[kLdsBSize], sbi1[kLdsBSize];
  __shared__ __attribute__((aligned(16))) double sa2[S > 2 ? kLdsASize : 2], sbr2[S > 2 ? kLdsBSize : 2],
      sbi2[S > 2 ? kLdsBSize : 2];
  auto stage_a = [&](int st) -> double * { return st == 0 ? sa0 : st == 1 ? sa1 : sa2; };
  auto stage_br = [&](int st) -> double * { return st == 0 ? sbr0 : st == 1 ? sbr1 : sbr2; };
  auto stage_bi = [&](int st) -> double * { return st == 0 ? sbi0 : st == 1 ? sbi1 : sbi2; };
  const int t = threadIdx.x, l = t % 64, w = t / 64, wm = w % 2, wn = w / 2;
  // load sources: A row pair of this lane (plane l / 32) and B columns
  const int arow = row0 + 2 * (l % 32);
  const double *a_src = (l < 32 ? v.Ar : v.Ai) + (arow < v.M ? arow : 0);
  const int bcol = min(col0 + 16 * w + l % 16, v.N - 1);
  const double *b_src[2];
  if (bcol < v.N1) {
    b_src[0] = v.Br + bcol * v.ldb;
    b_src[1] = v.Bi + bcol * v.ldb;
  } else {
    b_src[0] = v.B2r + (bcol - v.N1) * v.ldb2;
    b_src[1] = v.B2i + (bcol - v.N1) * v.ldb2;
  }
  const int bk = 2 * (l / 16);
  auto issue = [&](int st, int kc) {
#pragma unroll
    for (int u = 0; u < 2; ++u)
      lds_load16(a_src + static_cast<long long>(kc + 2 * w + u) * v.lda, stage_a(st) + (2 * w + u) * AK);
    lds_load16(b_src[0] + kc + bk, stage_br(st) + 128 * w);
    lds_load16(b_src[1] + kc + bk, stage_bi(st) + 128 * w);
  };
  const int nfull = v.K / KC, ktail = v.K - nfull * KC;
#pragma unroll
  for (int st = 0; st < S - 1; ++st)
    if (st < nfull) issue(st, st * KC);
  // accumulators of the 3M form with B not negated: x = P1 - Cr,
  // y = P2, z = P3 - Cr - Ci, where P1 = sum Ar Br, P2 = sum Ai Bi and
  // P3 = sum (Ar + Ai)(Br + Bi); then C - AB = (y - x) + i (x + y - z)
  typedef Mfma<double>::acc_t acc_t;
  acc_t cx[2][2], cy[2][2], cz[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int row = row0 + wm * 32 + a * 16 + l % 16;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int col = col0 + wn * 32 + b * 16 + Mfma<double>::ocol(l, e);
        double xr = 0.0, xi = 0.0;
        if (col < v.N1) {
          if (row < v.M) {
            const long long ic = static_cast<long long>(col) * v.ldc + row;
            xr = v.Cr[ic];
            xi = v.Ci[ic];
          }
        } else if (col < v.N && row < v.M2) {
          const long long ic = static_cast<long long>(col - v.N1) * v.ldc2 + row;
          xr = v.C2r[ic];
          xi = v.C2i[ic];
        }
        cx[a][b][e] = -xr;
        cy[a][b][e] = 0.0;
        cz[a][b][e] = -(xr + xi);
      }
    }
  }
  auto compute = [&](const double *A, const double *Br, const double *Bi) {
#pragma unroll
    for (int kk = 0; kk < KC; kk += 4) {
      const int k = kk + l / 16;
      double a_r[2], a_i[2], a_s[2], b_r[2], b_i[2], b_s[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int row = wm * 32 + q * 16 + l % 16;
        a_r[q] = A[k * AK + row];
        a_i[q] = A[k * AK + 64 + row];
        a_s[q] = a_r[q] + a_i[q];
        const int off = (wn * 2 + q) * 128 + (k / 2) * 32 + (l % 16) * 2 + (k % 2);
        b_r[q] = Br[off];
        b_i[q] = Bi[off];
        b_s[q] = b_r[q] + b_i[q];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
          cx[a][b] = Mfma<double>::mma(b_r[b], a_r[a], cx[a][b]);
          cy[a][b] = Mfma<double>::mma(b_i[b], a_i[a], cy[a][b]);
          cz[a][b] = Mfma<double>::mma(b_s[b], a_s[a], cz[a][b]);
        }
    }
  };
  for (int c0 = 0; c0 < nfull; c0 += S) {
#pragma unroll
    for (int st = 0; st < S; ++st) {
      const int c = c0 + st;
      if (c < nfull) {
        // chunk c's loads are done once at most the later chunk's (c + 1,
        // four loads, issued with three stages) are in flight
        if (S == 3 && c + 1 < nfull) __builtin_amdgcn_s_waitcnt(wait_vm_imm(4));
        else __builtin_amdgcn_s_waitcnt(wait_vm_imm(0));
        // a plain barrier: each wave has waited for its own loads of chunk
        // c, and consumed its reads of the stage refilled next (chunk
        // c - 1's) before arriving
        asm volatile("" ::: "memory");
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (c + S - 1 < nfull) issue((st + S - 1) % S, (c + S - 1) * KC);
        compute(stage_a(st), stage_br(st), stage_bi(st));
      }
    }
  }
  if (ktail > 0) {
    // the stage after the last full chunk's; no loads are in flight
    const int st = nfull % S;
    double *sa = stage_a(st), *sbr = stage_br(st), *sbi = stage_bi(st);
    const int kc = nfull * KC;
    // A: 8 k x 2 planes x 64 rows; B: 2 planes x 64 columns x 8 k (four each per thread)
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int e = t + 256 * s, r = e % 64, pl = (e / 64) % 2, k = e / 128;
      double x = 0.0;
      if (k < ktail && row0 + r < v.M)
        x = (pl ? v.Ai : v.Ar)[static_cast<long long>(kc + k) * v.lda + row0 + r];
      sa[k * AK + pl * 64 + r] = x;
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int e = t + 256 * s, k = e % 8, col = (e / 8) % 64, pl = e / 512;
      const int gc = col0 + col;
      double x = 0.0;
      if (k < ktail && gc < v.N) {
        if (gc < v.N1) x = (pl ? v.Bi : v.Br)[static_cast<long long>(gc) * v.ldb + kc + k];
        else x = (pl ? v.B2i : v.B2r)[static_cast<long long>(gc - v.N1) * v.ldb2 + kc + k];
      }
      (pl ? sbi : sbr)[(col / 16) * 128 + (k / 2) * 32 + (col % 16) * 2 + (k % 2)] = x;
    }
    __syncthreads();
    compute(sa, sbr, sbi);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int row = row0 + wm * 32 + a * 16 + l % 16;
#pragma unroll
    for (int b = 0; b < 2; ++b) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int col = col0 + wn * 32 + b * 16 + Mfma<double>::ocol(l, e);
        if (col >= v.N) continue;
        const double x = cx[a][b][e], y = cy[a][b][e];
        const double re = y - x, im = (x + y) - cz[a][b][e];
        if (col < v.N1) {
          if (row >= v.M) continue;
          const long long ic = static_cast<long long>(col) * v.ldc + row;
          v.Cr[ic] = re;
          v.Ci[ic] = im;
        } else {
          if (row >= v.M2) continue;
          const long long ic = static_cast<long long>(col - v.N1) * v.ldc2 + row;
          v.C2r[ic] = re;
          v.C2i[ic] = im;
        }
      }
    }
  }
}

// Product form and waves-per-SIMD bound of each tile shape.  The 64x64
// tiles use the 3M form; its third accumulator set spills at four waves per
// SIMD (K chunk 4) and three (K chunk 8), so their bounds are three and two,
// and two for K chunk 16.  Measured on the 200x100 factorization (ms of GEMM
// per factorization, XCD-grouped order): K chunk 4: 3M at three waves 106.8,
// four products at four waves 116.1; K chunk 8: 3M at two 110.5, four
// products at three 118.6; K chunk 16: 3M at two 110.5, four products at
// one 126.8.  The LDS-ring tile (shape 2): against 101.6 ms for the best
// register-staged tile, three stages (52 KB of LDS per workgroup) with the
// loads written as inline assembly 92.0 at two waves per SIMD and 98.9 at
// three (spilling); on another GPU, three stages 93.0 (inline assembly, two
// waves) and 92.8 (builtin loads, three waves), two stages (35 KB) 89.9 at
// three waves (two compile to the same code).
constexpr int kGemmK4Waves = 3, kGemmK8Waves = 2, kGemmK16Waves = 2, kGemmLdsWaves = 3, kGemmLdsStages = 2;
constexpr bool kGemmK4M3 = true, kGemmK8M3 = true, kGemmK16M3 = true;

// Launch one batched GEMM with the given tile shape; rows/cols are upper
// bounds of M and N over the level's fronts.
template <typename T>
hipError_t launch_gemm(int shape, int order, long long rows, long long cols, int count,
                       hipStream_t st, const int *list, const FrontDev *fr, int region, int k0,
                       int nb, int K0, int NB, T *arena) {
  const int tm = 64, tn = 64, threads = 256;
  const long long tr = rows <= 0 ? 1 : (rows + tm - 1) / tm, tc = cols <= 0 ? 1 : (cols + tn - 1) / tn;
  unsigned tiles = static_cast<unsigned>(tr * tc);
  if (order != kOrderLinear) tiles = (tiles + kXcds - 1) / kXcds * kXcds;  // XCD order: multiple of 8
  const dim3 grid(tiles, count), block(threads);
  switch (shape) {
    case kShape64x64k4:
      hipLaunchKernelGGL((k_gemm<T, 2, 2, 2, 2, 4, kGemmK4Waves, kGemmK4M3>), grid, block, 0, st,
                         list, count, fr, region, k0, nb, K0, NB, order, arena);
      break;
    case kShapeLds:
      if constexpr (sizeof(T) == 8)
        hipLaunchKernelGGL((k_gemm_lds<kGemmLdsWaves, kGemmLdsStages>), grid, block, 0, st, list, count, fr, region,
                           k0, nb, K0, NB, order, arena);
      else
        hipLaunchKernelGGL((k_gemm<T, 2, 2, 2, 2, 4, kGemmK4Waves, kGemmK4M3>), grid, block, 0, st,
                           list, count, fr, region, k0, nb, K0, NB, order, arena);
      break;
    case kShape64x64k8:
      hipLaunchKernelGGL((k_gemm<T, 2, 2, 2, 2, 8, kGemmK8Waves, kGemmK8M3>), grid, block, 0, st,
                         list, count, fr, region, k0, nb, K0, NB, order, arena);
      break;
    default:
      hipLaunchKernelGGL((k_gemm<T, 2, 2, 2, 2, GKC, kGemmK16Waves, kGemmK16M3>), grid, block, 0, st, list,
                         count, fr, region, k0, nb, K0, NB, order, arena);
      break;
  }
  return hipGetLastError();
}

// Panel LU of one level: the register-resident kernel (kRegsThreads threads,
// the fewest rows per thread that cover the level's largest front: pmax
// bounds p - k0) when the fronts have at most kPanelMapRows fully summed
// rows, else, or when `streaming` is set, the streaming kernel.
constexpr int kRegsThreads = 512;

template <typename T, int R>
hipError_t launch_panel_regs(int rows, dim3 grid, hipStream_t st, const int *list, int count,
                             const FrontDev *fr, int k0, int nb, int K0, int NB, const double *fp,
                             T *arena, int *ipiv, int *nperturbed) {
  if constexpr (R * kRegsThreads > kPanelMapRows) {
    return hipErrorInvalidValue;  // excluded by the caller
  } else {
    if (rows > R)
      return launch_panel_regs<T, R + 1>(rows, grid, st, list, count, fr, k0, nb, K0, NB, fp,
                                         arena, ipiv, nperturbed);
    hipLaunchKernelGGL((k_panel_regs<T, R, kRegsThreads>), grid, dim3(kRegsThreads), 0, st, list,
                       count, fr, k0, nb, K0, NB, fp, arena, ipiv, nperturbed);
    return hipGetLastError();
  }
}

template <typename T>
hipError_t launch_panel(int pmax, bool streaming, int count, hipStream_t st, const int *list,
                        const FrontDev *fr, int k0, int nb, int K0, int NB, const double *fp,
                        T *arena, int *ipiv, int *nperturbed) {
  const dim3 grid(count);
  if (!streaming && pmax <= kPanelMapRows)
    return launch_panel_regs<T, 1>((pmax + kRegsThreads - 1) / kRegsThreads, grid, st, list,
                                   count, fr, k0, nb, K0, NB, fp, arena, ipiv, nperturbed);
  hipLaunchKernelGGL(k_panel<T>, grid, dim3(kPanelThreads), 0, st, list, count, fr, k0, nb, K0,
                     NB, fp, arena, ipiv, nperturbed);
  return hipGetLastError();
}

// ------------------------------------------------------------------ solves
//
// Vectors are split complex planes in elimination (dof) order: position
// 5*pos[node] + dof.  The per-front workspace W (m entries per front)
// collects the right-hand side of the front's pivots (W[0:p]) and the
// contributions passing through its update rows (W[p:m]):
//
//   forward, levels bottom-up   op 0  Y_f = E_f W_f[0:p]
//                               op 1  W_parent[rel_f] += W_f[p:m] - L21_f Y_f
//   backward, levels top-down   op 2  Y_f -= U12_f x[update rows of f]
//                               op 3  x_f  = Z_f Y_f
//
// k_solve_init sets W = [b; 0], Y = 0 and x = 0 for every front in one
// launch; all GEMV outputs are accumulated with FP64 atomics.

__global__ void k_solve_init(const FrontDev *__restrict__ fr, const double *br, const double *bi,
                             double *__restrict__ Wr, double *__restrict__ Wi,
                             double *__restrict__ Yr, double *__restrict__ Yi, double *xr,
                             double *xi) {
  const FrontDev d = fr[blockIdx.x];
  for (int k = threadIdx.x; k < d.m; k += blockDim.x) {
    const bool piv = k < d.p;
    // b may alias x: each dof is read and then cleared by the same thread
    const double vr = piv ? br[d.dof0 + k] : 0.0, vi = piv ? bi[d.dof0 + k] : 0.0;
    Wr[d.woff + k] = vr;
    Wi[d.woff + k] = vi;
    if (piv) {
      Yr[d.dof0 + k] = 0.0;
      Yi[d.dof0 + k] = 0.0;
      xr[d.dof0 + k] = 0.0;
      xi[d.dof0 + k] = 0.0;
    }
  }
}

constexpr int GVR = 64, GVW = 4;  // rows per workgroup, waves per workgroup
constexpr long long kGemvWideMin = 1024;  // workgroups of the wide form that fill the GPU

// One batched GEMV y (+)= add + sign * A x.  xmap / ymap, when set, gather x
// and scatter y through node maps (entry j -> 5 * map[j / 5] + j % 5).
template <typename T> struct GemvView {
  const T *Ar, *Ai;
  long long lda;
  int M, N;
  const double *xr, *xi;
  const int *xmap;
  double *yr, *yi;
  const int *ymap;
  const double *addr, *addi;
  double sign;
  bool upper;  // A is upper triangular (skip tiles below the diagonal)
  int rshift;  // row tiles start at row -rshift (so that they start on a 128-byte line)
};

struct SolveMaps {
  const int *parent, *rel, *upd_pos;
  const long long *rel_off, *upd_off;
};

template <typename T>
__device__ inline bool gemv_view(int op, const FrontDev *fr, int f, const FrontDev &d,
                                 const SolveMaps &mp, const T *arena, double *Wr, double *Wi,
                                 double *Yr, double *Yi, double *Xr, double *Xi, GemvView<T> &v) {
  const long long mm = static_cast<long long>(d.m) * d.m, pp = static_cast<long long>(d.pe) * d.p;
  const T *Fr = arena + d.off, *Fi = Fr + mm;
  const T *Er = Fi + mm, *Ei = Er + pp;
  const T *Zr = Ei + pp, *Zi = Zr + pp;
  switch (op) {
    case 0:
      v = {Er, Ei, d.pe, d.p, d.p, Wr + d.woff, Wi + d.woff, nullptr, Yr + d.dof0, Yi + d.dof0,
           nullptr, nullptr, nullptr, 1.0, false, 0};
      return d.p > 0;
    case 1: {
      if (d.c == 0 || d.p == 0) return false;
      const long long pw = fr[mp.parent[f]].woff;
      // L21 starts at row p of F's line-aligned columns
      constexpr int line = 128 / static_cast<int>(sizeof(T));
      v = {Fr + d.p, Fi + d.p, d.m, d.c, d.p, Yr + d.dof0, Yi + d.dof0, nullptr, Wr + pw, Wi + pw,
           mp.rel + mp.rel_off[f], Wr + d.woff + d.p, Wi + d.woff + d.p, -1.0, false, d.p % line};
      return true;
    }
    case 2: {
      if (d.c == 0 || d.p == 0) return false;
      const long long a = static_cast<long long>(d.p) * d.m;
      v = {Fr + a, Fi + a, d.m, d.p, d.c, Xr, Xi, mp.upd_pos + mp.upd_off[f], Yr + d.dof0,
           Yi + d.dof0, nullptr, nullptr, nullptr, -1.0, false, 0};
      return true;
    }
    default:
      v = {Zr, Zi, d.pe, d.p, d.p, Yr + d.dof0, Yi + d.dof0, nullptr, Xr + d.dof0, Xi + d.dof0,
           nullptr, nullptr, nullptr, 1.0, true, 0};
      return d.p > 0;
  }
}

// One right-hand side of a solve: the front workspace W, the accumulated
// update Y and the solution x (split planes).
struct RhsSet {
  double *Wr, *Wi, *Yr, *Yi, *Xr, *Xi;
};

// One workgroup: GVR output rows x (GVW * GVC) reduction columns; each wave
// takes GVC of the columns and the four partial sums are combined in LDS
// before one FP64 atomic per row.  GVC = 64 for levels with enough work to
// fill the GPU, 16 (one batch of loads per wave, four times the workgroups)
// for the small ones, whose launches are latency-bound.  NR right-hand
// sides (1 or 2) share every load of A; each one's sums are formed in the
// same order as with one.  Factor entries of type T are widened to FP64;
// sums are FP64.  A wave steps through its columns 32 at a time with FP32
// entries and 16 at a time with FP64 ones, so both keep 256 B of factor
// loads in flight per lane: with 16 FP32 columns per step the solve was
// bound by the bytes in flight, at the FP64 solve's time for half the bytes
// (200x100: 1.43 ms per FP32 solve before, 0.91 ms after; FP64 1.43 ms,
// and 1.47 ms with 32 FP64 columns per step at half the occupancy).
template <typename T, int GVC, int NR>
__global__ __launch_bounds__(GVR * GVW) void k_gemv(const int *__restrict__ list, int nlist,
                                                    const FrontDev *__restrict__ fr, SolveMaps mp,
                                                    int op, const T *__restrict__ arena,
                                                    RhsSet rhs0, RhsSet rhs1) {
  static_assert(NR == 1 || NR == 2, "one or two right-hand sides");
  const int f = list[blockIdx.y];
  const FrontDev d = fr[f];
  const RhsSet rs[2] = {rhs0, rhs1};
  GemvView<T> v[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r)
    if (!gemv_view(op, fr, f, d, mp, arena, rs[r].Wr, rs[r].Wi, rs[r].Yr, rs[r].Yi, rs[r].Xr,
                   rs[r].Xi, v[r]))
      return;
  const GemvView<T> &V = v[0];  // A, its shape and the maps are shared
  const int lane = threadIdx.x % GVR, wave = threadIdx.x / GVR;
  const int r0 = blockIdx.x * GVR - V.rshift;
  const int row = r0 + lane;
  const int cw = blockIdx.z * GVW * GVC;  // first column of this workgroup
  if (r0 >= V.M || cw >= V.N) return;
  if (V.upper && cw + GVW * GVC <= r0) return;  // every entry of the tile is below the diagonal
  const int c0 = cw + wave * GVC;
  const int c1 = min(V.N, c0 + GVC);
  __shared__ double sxr[NR][GVW * GVC], sxi[NR][GVW * GVC];
  __shared__ double pr[NR][GVW][GVR], pi[NR][GVW][GVR];
  const bool inside = row >= 0 && row < V.M;
  const bool active = inside && !(V.upper && c1 <= r0);
  const T *ar = V.Ar + (inside ? row : 0), *ai = V.Ai + (inside ? row : 0);
  // KS columns per step: every lane keeps 2 * KS factor loads in flight,
  // 256 B with FP64 entries (KS = 16) and, with KS = 32, also with FP32
  // entries, so both precisions hold the same bytes in flight per wave.
  // Entries stay in their stored type in registers and are widened to FP64
  // where they are used; sums are FP64.
  constexpr int KS = (sizeof(T) == sizeof(float) && GVC % 32 == 0) ? 32 : 16;
  static_assert(GVC % KS == 0, "a wave's columns are whole steps");
  // the first KS columns of A are requested before x is staged, so the two
  // loads overlap
  T a_r[KS], a_i[KS];
  const bool first = active && c0 + KS <= c1;
  if (first) {
#pragma unroll
    for (int u = 0; u < KS; ++u) {
      const long long idx = static_cast<long long>(c0 + u) * V.lda;
      a_r[u] = ar[idx];
      a_i[u] = ai[idx];
    }
  }
  for (int k = threadIdx.x; k < GVW * GVC; k += GVR * GVW) {
    const int col = cw + k;
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      double xr = 0.0, xi = 0.0;
      if (col < V.N) {
        const int g = V.xmap ? 5 * V.xmap[col / 5] + col % 5 : col;
        xr = v[r].xr[g];
        xi = v[r].xi[g];
      }
      sxr[r][k] = xr;
      sxi[r][k] = xi;
    }
  }
  __syncthreads();
  double accr[NR], acci[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    accr[r] = 0.0;
    acci[r] = 0.0;
  }
  if (active) {
    int k = c0;
    for (; k + KS <= c1; k += KS) {
      if (k != c0 || !first) {
#pragma unroll
        for (int u = 0; u < KS; ++u) {
          const long long idx = static_cast<long long>(k + u) * V.lda;
          a_r[u] = ar[idx];
          a_i[u] = ai[idx];
        }
      }
#pragma unroll
      for (int u = 0; u < KS; ++u) {
        const double br = static_cast<double>(a_r[u]), bi = static_cast<double>(a_i[u]);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
          const double x_r = sxr[r][k + u - cw], x_i = sxi[r][k + u - cw];
          accr[r] += br * x_r - bi * x_i;
          acci[r] += br * x_i + bi * x_r;
        }
      }
    }
    // the last c1 - k < KS columns (front orders are multiples of 5, so
    // most waves end with some): every load is issued before the first use,
    // and the sums run over the columns in the same order as above
    const int rem = c1 - k;
    if (rem > 0) {
#pragma unroll
      for (int u = 0; u < KS; ++u) {
        if (u < rem) {
          const long long idx = static_cast<long long>(k + u) * V.lda;
          a_r[u] = ar[idx];
          a_i[u] = ai[idx];
        }
      }
#pragma unroll
      for (int u = 0; u < KS; ++u) {
        if (u < rem) {
          const double br = static_cast<double>(a_r[u]), bi = static_cast<double>(a_i[u]);
#pragma unroll
          for (int r = 0; r < NR; ++r) {
            const double x_r = sxr[r][k + u - cw], x_i = sxi[r][k + u - cw];
            accr[r] += br * x_r - bi * x_i;
            acci[r] += br * x_i + bi * x_r;
          }
        }
      }
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    pr[r][wave][lane] = accr[r];
    pi[r][wave][lane] = acci[r];
  }
  __syncthreads();
  if (wave == 0 && inside) {
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      double sr = 0.0, si = 0.0;
      for (int w = 0; w < GVW; ++w) {
        sr += pr[r][w][lane];
        si += pi[r][w][lane];
      }
      sr *= v[r].sign;
      si *= v[r].sign;
      if (v[r].addr && blockIdx.z == 0) {
        sr += v[r].addr[row];
        si += v[r].addi[row];
      }
      const int o = v[r].ymap ? 5 * v[r].ymap[row / 5] + row % 5 : row;
      atomicAdd(v[r].yr + o, sr);
      atomicAdd(v[r].yi + o, si);
    }
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

// The K of one plan `part` for the factorization: fbase is where the part's
// fronts start in the front table the factorization runs on (0 for the
// part's own plan, else its place in a batch plan), fp its parameter block.
struct ScatterJob {
  const LuPlan *part;
  LuInputs in;
  int fbase;
  const double *fp;
};

// E = I, Z = I and the pivot counters, then level by level: the level's
// fronts built from their children, K of every job added, and the blocked
// LU of the fronts (P: a context's plan or a batch plan).
template <typename T>
static hipError_t factor_levels(const LuPlan &P, const std::vector<ScatterJob> &jobs,
                                hipStream_t s, FactorTiming *tm) {
  hipError_t e;
  std::vector<hipEvent_t> events;  // start/stop pairs around GEMM launches when timing
  std::vector<int> event_level;    // level of each pair
  std::vector<int> event_region;   // update region of each pair
  if (tm) tm->level_gemm_flops.assign(P.levels.size(), 0.0);
  T *arena = reinterpret_cast<T *>(P.arena);
  hipLaunchKernelGGL(k_init_identity<T>, dim3(64, P.nsn), dim3(256), 0, s, P.d_front, P.nsn, arena);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_fill, dim3(1), dim3(64), 0, s, reinterpret_cast<double *>(P.d_counters),
                     static_cast<long long>((P.instances + 1) / 2), 0.0);
  const int nb = P.nb, NB = kLuNbOuter;
  T *dinv = reinterpret_cast<T *>(P.dinv);
  for (size_t lev = 0; lev < P.levels.size(); ++lev) {
    const LevelPlan &L = P.levels[lev];
    const int *list = P.d_level_list + L.list_off;
    {
      const int nrc = static_cast<int>(cdivu(L.mmax, kBuildRows));
      hipLaunchKernelGGL(k_build_fronts<T>, dim3(nrc * cdivu(L.mmax, kBuildCols), L.count),
                         dim3(256), 0, s, list, L.count, P.d_front, P.d_build, P.d_inv, nrc, arena);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      for (const ScatterJob &J : jobs) {
        if (lev >= J.part->levels.size()) continue;
        const LevelPlan &Lp = J.part->levels[lev];
        if (Lp.ablk_count == 0) continue;
        const long long threads = Lp.ablk_count * 25;
        hipLaunchKernelGGL(k_assemble_fronts<T>, dim3(static_cast<unsigned>((threads + 255) / 256)),
                           dim3(256), 0, s, Lp.ablk_count, J.part->d_ablk + Lp.ablk_off, J.in.per,
                           J.in.avals, J.in.colidx, J.in.bvals, J.fp, J.part->d_amap, P.d_front,
                           J.fbase, arena);
        if ((e = hipGetLastError()) != hipSuccess) return e;
      }
    }
    const int order = lev < P.gemm_order.size() ? P.gemm_order[lev] : kOrderLinear;
    const int shape = lev < P.gemm_shape.size() ? P.gemm_shape[lev] : kShape64x64;
    // one batched GEMM of `region`; W: the K-block width of the trailing
    // regions (NB, or 2 NB for a pair of blocks)
    auto gemm = [&](long long rows, long long cols, int region, int k0, int K0, int W,
                    hipStream_t st) -> hipError_t {
      hipEvent_t e0 = nullptr, e1 = nullptr;
      if (tm) {
        hipError_t r;
        if ((r = hipEventCreate(&e0)) != hipSuccess) return r;
        if ((r = hipEventCreate(&e1)) != hipSuccess) return r;
        events.push_back(e0);
        events.push_back(e1);
        event_level.push_back(static_cast<int>(lev));
        event_region.push_back(region);
        if ((r = hipEventRecord(e0, st)) != hipSuccess) return r;
        for (int q = 0; q < L.count; ++q) {
          // algorithmic flops: the front's own order, not the rows and
          // columns that round it up to whole 128-byte lines
          FrontDev dq = P.h_front[P.h_level_list[L.list_off + q]];
          dq.m = dq.p + dq.c;
          const GemmShape g = gemm_shape(region, dq, k0, nb, K0, W);
          if (g.M > 0 && g.N > 0 && g.K > 0) {
            const double fl = 8.0 * g.K * (static_cast<double>(g.M) * g.N1 +
                                           static_cast<double>(g.M2) * (g.N - g.N1));
            tm->gemm_flops += fl;
            tm->level_gemm_flops[lev] += fl;
            tm->region_gemm_flops[region] += fl;
          }
        }
        ++tm->gemm_launches;
      }
      hipError_t r = launch_gemm<T>(shape, order, rows, cols, L.count, st, list, P.d_front,
                                    region, k0, nb, K0, W, arena);
      if (r != hipSuccess) return r;
      if (tm) return hipEventRecord(e1, st);
      return hipSuccess;
    };
    // panels of block [K0, K0 + NB): LU of each panel, its L21 rows, and the
    // block columns right of it (touches only the block's columns and ipiv)
    auto panel_phase = [&](int K0, hipStream_t st) -> hipError_t {
      hipError_t r;
      const int Kend = std::min(K0 + NB, L.pmax);
      for (int k0 = K0; k0 < Kend; k0 += nb) {
        if ((r = launch_panel<T>(L.pmax, P.panel_streaming, L.count, st, list, P.d_front, k0,
                                 nb, K0, NB, P.d_fparams, arena, P.d_ipiv,
                                 P.d_counters)) != hipSuccess)
          return r;
        hipLaunchKernelGGL(k_diag_inv<T>, dim3(L.count), dim3(128), 0, st, list, L.count,
                           P.d_front, k0, nb, arena, dinv);
        if (L.cmax > 0)
          hipLaunchKernelGGL(k_right_apply<T>, dim3(cdivu(L.cmax, 128), L.count), dim3(256), 0,
                             st, list, L.count, P.d_front, k0, nb, dinv, arena);
        if (k0 + nb < Kend) {
          hipLaunchKernelGGL(k_left_apply<T>, dim3(cdivu(NB - nb, 64), L.count), dim3(256), 0,
                             st, list, L.count, P.d_front, k0, nb, K0, NB, 0, dinv, arena);
          if ((r = gemm(L.mmax, NB - nb, kGemmInner, k0, K0, NB, st)) != hipSuccess) return r;
        }
        if ((r = hipGetLastError()) != hipSuccess) return r;
      }
      return hipSuccess;
    };
    // the row exchanges of block [K0, K0 + NB) on the trailing F columns and
    // E (composed first), or, c1 > c0, on F columns [c0, c1) with the
    // permutation composed by the previous call for the same block
    auto exchanges = [&](int K0, int c0, int c1) -> hipError_t {
      const bool range = c1 > c0;
      const long long ncols = range ? c1 - c0 : L.mmax + L.pmax;
      if (!range)
        hipLaunchKernelGGL(k_block_perm, dim3(L.count), dim3(256), 0, s, list, L.count,
                           P.d_front, K0, NB, P.d_ipiv, P.d_bperm);
      hipLaunchKernelGGL(k_apply_perm<T>, dim3(cdivu(ncols, 4), L.count), dim3(256), 0, s,
                         list, L.count, P.d_front, K0, NB, P.d_bperm, arena, c0, c1);
      if (L.pmax > kPanelMapRows)
        hipLaunchKernelGGL(k_block_swaps<T>, dim3(cdivu(ncols, 256), L.count), dim3(256), 0, s,
                           list, L.count, P.d_front, K0, NB, P.d_ipiv, arena, c0, c1);
      return hipGetLastError();
    };
    // U rows of block [K0, K0 + NB) for the trailing F columns and E
    auto u_rows = [&](int K0) -> hipError_t {
      hipLaunchKernelGGL(k_block_rows<T>, dim3(cdivu(L.mmax + L.pmax, 64), L.count), dim3(256),
                         0, s, list, L.count, P.d_front, K0, NB, dinv, arena);
      return hipGetLastError();
    };
    // Look-ahead: once the TRAILHEAD update has produced the next block's
    // columns, TRAILREST (later columns and E) runs on the side stream while
    // the main stream factors the next block's panels.  The two touch
    // disjoint columns; the main stream waits for TRAILREST before the next
    // block's row exchanges and U rows.  The side stream is confined to all
    // but kSideReservedCus compute units (LuPlan::side), so the panel chain
    // -- a dozen short launches per panel, one workgroup per front -- always
    // finds free CUs instead of waiting for TRAILREST's workgroups to retire.
    // A timed factorization runs the same launches in order on one stream,
    // so each kernel's duration is its own.
    //
    // Paired blocks (P.pair_blocks): blocks A = [K0, K0 + NB) and
    // B = [K0 + NB, K0 + 2 NB) share one update of the rows below B
    // (K = 2 NB).  After A's exchanges and U rows, TRAILHEAD brings B's
    // columns up to date and B is factored (its panels are not overlapped);
    // B's exchanges then also permute A's L columns, so that the deferred
    // update of A sees the rows in B's order, PAIRROWS applies A's update to
    // B's rows, and B's U rows follow.  Rows below B get both blocks' update
    // in the pair's TRAILHEAD / TRAILREST / TRAIL.
    const bool lookahead = tm == nullptr;
    if ((e = panel_phase(0, s)) != hipSuccess) return e;
    for (int K0 = 0; K0 < L.pmax;) {
      const bool pair = P.pair_blocks && K0 + NB < L.pmax;
      const int W = pair ? 2 * NB : NB;
      if ((e = exchanges(K0, 0, 0)) != hipSuccess) return e;
      if ((e = u_rows(K0)) != hipSuccess) return e;
      if (pair) {
        const int KB = K0 + NB;
        if ((e = gemm(L.mmax, NB, kGemmTrailHead, K0, K0, NB, s)) != hipSuccess) return e;
        if ((e = panel_phase(KB, s)) != hipSuccess) return e;
        if ((e = exchanges(KB, 0, 0)) != hipSuccess) return e;
        if ((e = exchanges(KB, K0, KB)) != hipSuccess) return e;
        if ((e = gemm(NB, L.mmax + L.pmax, kGemmPairRows, K0, K0, NB, s)) != hipSuccess) return e;
        if ((e = u_rows(KB)) != hipSuccess) return e;
      }
      const bool next = K0 + W < L.pmax;
      if (next) {
        if ((e = gemm(L.mmax, NB, kGemmTrailHead, K0, K0, W, s)) != hipSuccess) return e;
        if (lookahead) {
          if ((e = hipEventRecord(P.ev_head, s)) != hipSuccess) return e;
          if ((e = hipStreamWaitEvent(P.side, P.ev_head, 0)) != hipSuccess) return e;
          if ((e = gemm(L.mmax, L.mmax + L.pmax, kGemmTrailRest, K0, K0, W, P.side)) !=
              hipSuccess)
            return e;
          if ((e = hipEventRecord(P.ev_panel, P.side)) != hipSuccess) return e;
          if ((e = panel_phase(K0 + W, s)) != hipSuccess) return e;
          if ((e = hipStreamWaitEvent(s, P.ev_panel, 0)) != hipSuccess) return e;
        } else {
          if ((e = panel_phase(K0 + W, s)) != hipSuccess) return e;
          if ((e = gemm(L.mmax, L.mmax + L.pmax, kGemmTrailRest, K0, K0, W, s)) != hipSuccess)
            return e;
        }
      } else {
        if ((e = gemm(L.mmax, L.mmax + L.pmax, kGemmTrail, K0, K0, W, s)) != hipSuccess) return e;
      }
      if ((e = hipGetLastError()) != hipSuccess) return e;
      K0 += W;
    }
    // Z = U11^-1 by backward block-Jordan steps: the block's rows in one
    // sweep (k_block_z), then ZBLK for every row above the block
    for (int K0 = ((L.pmax - 1) / NB) * NB; K0 >= 0; K0 -= NB) {
      hipLaunchKernelGGL(k_block_z<T>, dim3(cdivu(L.pmax - K0, 64), L.count), dim3(256), 0, s,
                         list, L.count, P.d_front, K0, NB, dinv, arena);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      if (K0 > 0)
        if ((e = gemm(L.pmax, L.pmax, kGemmZBlk, K0, K0, NB, s)) != hipSuccess) return e;
    }
  }
  if (tm) {
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    tm->level_gemm_ms.assign(P.levels.size(), 0.0);
    for (size_t q = 0; q + 1 < events.size(); q += 2) {
      float ms = 0.f;
      if ((e = hipEventElapsedTime(&ms, events[q], events[q + 1])) != hipSuccess) return e;
      tm->gemm_ms += ms;
      tm->level_gemm_ms[event_level[q / 2]] += ms;
      tm->region_gemm_ms[event_region[q / 2]] += ms;
    }
    for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
  }
  return hipSuccess;
}

template <typename T>
static hipError_t factor_impl(const LuPlan &P, const LuInputs &in, hipStream_t s,
                              FactorTiming *tm) {
  return factor_levels<T>(P, {ScatterJob{&P, in, 0, P.d_fparams}}, s, tm);
}

hipError_t lu_factor(const LuPlan &P, const LuInputs &in, bool fp32, hipStream_t s,
                     FactorTiming *tm) {
  if (!tm) {
    const hipError_t e = lu_plan_side(P);
    if (e != hipSuccess) return e;
  }
  return fp32 ? factor_impl<float>(P, in, s, tm) : factor_impl<double>(P, in, s, tm);
}

hipError_t lu_set_sigma(const LuPlan &P, double sig_re, double sig_im, hipStream_t s, int inst) {
  if (inst < 0 || inst >= P.instances) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_set_sigma, dim3(1), dim3(1), 0, s, P.d_fparams + inst * kFpCount, sig_re,
                     sig_im);
  return hipGetLastError();
}

// Norms of A_bc and B_bc into `partials`, then the pivot threshold and
// perturbation into the parameter block fp (its shift already set).
static hipError_t factor_params_at(double *fp, const LuInputs &in, bool fp32, double *partials,
                                   hipStream_t s) {
  const long long nodes = in.nnzb / in.per;
  hipLaunchKernelGGL(k_sumsq_flat, dim3(kParamBlocks), dim3(256), 0, s, in.avals, in.nnzb * 50,
                     partials);
  hipLaunchKernelGGL(k_sumsq_flat, dim3(kParamBlocks), dim3(256), 0, s, in.bvals, nodes * 50,
                     partials + kParamBlocks);
  hipLaunchKernelGGL(k_factor_params, dim3(1), dim3(1), 0, s, partials, 5.0 * nodes, fp32 ? 1 : 0,
                     fp);
  return hipGetLastError();
}

hipError_t lu_factor_params(const LuPlan &P, const LuInputs &in, bool fp32, double *partials,
                            hipStream_t s) {
  return factor_params_at(P.d_fparams, in, fp32, partials, s);
}

template <typename T>
static hipError_t factor_batch_impl(const LuPlan &U, const std::vector<BatchPart> &parts,
                                    const std::vector<int> &fbase,
                                    const std::vector<LuInputs> &in,
                                    const std::vector<double *> &partials, bool fp32,
                                    hipStream_t s, FactorTiming *tm) {
  std::vector<ScatterJob> jobs;
  for (size_t b = 0; b < parts.size(); ++b) {
    double *fp = U.d_fparams + b * kFpCount;
    hipError_t e = factor_params_at(fp, in[b], fp32, partials[b], s);
    if (e != hipSuccess) return e;
    jobs.push_back({parts[b].plan, in[b], fbase[b], fp});
  }
  return factor_levels<T>(U, jobs, s, tm);
}

hipError_t lu_factor_batch(const LuPlan &U, const std::vector<BatchPart> &parts,
                           const std::vector<int> &fbase, const std::vector<LuInputs> &in,
                           const std::vector<double *> &partials, bool fp32, hipStream_t s,
                           FactorTiming *tm) {
  const size_t nb = parts.size();
  if (nb == 0 || static_cast<int>(nb) != U.instances || fbase.size() != nb || in.size() != nb ||
      partials.size() != nb)
    return hipErrorInvalidValue;
  return fp32 ? factor_batch_impl<float>(U, parts, fbase, in, partials, true, s, tm)
              : factor_batch_impl<double>(U, parts, fbase, in, partials, false, s, tm);
}

// x = K^-1 b for vectors in elimination order (split planes).  b may alias x.
// x_r = K^-1 b_r for NR right-hand sides in one pass over the factors
// (b[2 r], b[2 r + 1]: real and imaginary planes of right-hand side r).
template <typename T, int NR>
static hipError_t solve_impl(const LuPlan &P, const double *const *b, double *const *x,
                             hipStream_t s) {
  const T *arena = reinterpret_cast<const T *>(P.arena);
  RhsSet rs[2];
  for (int r = 0; r < NR; ++r) {
    double *W = r == 0 ? P.W : P.W2, *Y = r == 0 ? P.Y : P.Y2;
    rs[r] = {W, W + P.w_len, Y, Y + P.n, x[2 * r], x[2 * r + 1]};
    hipLaunchKernelGGL(k_solve_init, dim3(P.nsn), dim3(256), 0, s, P.d_front, b[2 * r],
                       b[2 * r + 1], rs[r].Wr, rs[r].Wi, rs[r].Yr, rs[r].Yi, rs[r].Xr, rs[r].Xi);
  }
  if (NR == 1) rs[1] = rs[0];
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const SolveMaps mp{P.d_parent, P.d_rel, P.d_upd_pos, P.d_rel_off, P.d_upd_off};
  auto gemv = [&](const LevelPlan &L, int op, int rows, int cols) {
    const long long wide = static_cast<long long>(cdivu(rows, GVR)) * L.count * cdivu(cols, 64 * GVW);
    if (wide >= kGemvWideMin)
      hipLaunchKernelGGL((k_gemv<T, 64, NR>), dim3(cdivu(rows, GVR), L.count, cdivu(cols, 64 * GVW)),
                         dim3(GVR * GVW), 0, s, P.d_level_list + L.list_off, L.count, P.d_front,
                         mp, op, arena, rs[0], rs[1]);
    else
      hipLaunchKernelGGL((k_gemv<T, 16, NR>), dim3(cdivu(rows, GVR), L.count, cdivu(cols, 16 * GVW)),
                         dim3(GVR * GVW), 0, s, P.d_level_list + L.list_off, L.count, P.d_front,
                         mp, op, arena, rs[0], rs[1]);
  };
  for (size_t lev = 0; lev < P.levels.size(); ++lev) {
    const LevelPlan &L = P.levels[lev];
    gemv(L, 0, L.pmax, L.pmax);
    // op 1's row tiles start up to one line before row 0 (GemvView::rshift)
    if (L.cmax > 0) gemv(L, 1, L.cmax + 128 / static_cast<int>(sizeof(T)), L.pmax);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  for (int lev = static_cast<int>(P.levels.size()) - 1; lev >= 0; --lev) {
    const LevelPlan &L = P.levels[lev];
    if (L.cmax > 0) gemv(L, 2, L.pmax, L.cmax);
    gemv(L, 3, L.pmax, L.pmax);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t lu_solve(const LuPlan &P, const double *br, const double *bi, double *xr, double *xi,
                    bool fp32, hipStream_t s) {
  const double *b[2] = {br, bi};
  double *x[2] = {xr, xi};
  return fp32 ? solve_impl<float, 1>(P, b, x, s) : solve_impl<double, 1>(P, b, x, s);
}

hipError_t lu_solve2(const LuPlan &P, const double *b1r, const double *b1i, const double *b2r,
                     const double *b2i, double *x1r, double *x1i, double *x2r, double *x2i,
                     bool fp32, hipStream_t s) {
  const double *b[4] = {b1r, b1i, b2r, b2i};
  double *x[4] = {x1r, x1i, x2r, x2i};
  return fp32 ? solve_impl<float, 2>(P, b, x, s) : solve_impl<double, 2>(P, b, x, s);
}

}  // namespace tailor
