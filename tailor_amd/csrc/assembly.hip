// assembly.hip -- gfx950 kernels of the BiGlobal operator assembly.
//
//   k_metrics_first        first inverse metrics          Metrics.cpp:56-225
//   k_metrics2_derivatives second metrics + base-flow     Metrics.cpp:227-352,
//                          derivatives                    BaseFlowDerivatives.cpp:51-251
//   k_characteristics      far-field eigenbasis, R^-1,    BoundaryConditions.cpp:92-280
//                          P = R^-1 Gamma^-1
//   k_operator             LNS point coefficients ->      LNSCoefficients.cpp:173-404,
//                          Fourier -> curvilinear ->      StreamwiseFourier.cpp:16-44,
//                          77 BSR blocks of A_bc (= -L    CurvilinearTransform.cpp:62-101,
//                          with wall / far-field rows)    GlobalOperator.cpp:285-378,
//                          and B_bc                       BoundaryConditions.cpp:354-548
//
// The operator kernel is the write-bound one: 616 MB of complex128 blocks
// at 200x100.  One workgroup owns NPW consecutive block rows, which are
// contiguous in the BSR value array (77 blocks x 25 entries x 16 B = 30.8 KB
// per row).  Its threads first build the per-node coefficient blocks in LDS
// (six threads per node for the physical blocks, 25 per node for the Fourier
// and curvilinear entries), then stream the rows out with one 16-byte
// non-temporal store per lane per step, so each wave writes 1 KB contiguous.
//
// This file is compiled with -ffp-contract=off and evaluates every sum in
// the reference's order, so the values match a sequential CPU evaluation of
// the same formulas bit for bit.
#include <cfloat>
#include <cmath>

#include "internal.hpp"

namespace tailor {
namespace {

constexpr double kEps = 2.220446049250313e-16;
constexpr int kThreads = 256;

__device__ inline int wrap_mod(int idx, int n) {
  int r = idx % n;
  return r < 0 ? r + n : r;
}

__device__ inline int wrap_count(int idx, int n) {
  return idx >= 0 ? idx / n : -((-idx + n - 1) / n);
}

__device__ inline bool finite(double x) { return isfinite(x); }

__device__ void atomic_min_d(double *addr, double v) {
  auto *a = reinterpret_cast<unsigned long long *>(addr);
  unsigned long long old = *a;
  while (v < __longlong_as_double(static_cast<long long>(old))) {
    unsigned long long prev = atomicCAS(a, old, static_cast<unsigned long long>(__double_as_longlong(v)));
    if (prev == old) break;
    old = prev;
  }
}

__device__ void atomic_max_d(double *addr, double v) {
  auto *a = reinterpret_cast<unsigned long long *>(addr);
  unsigned long long old = *a;
  while (v > __longlong_as_double(static_cast<long long>(old))) {
    unsigned long long prev = atomicCAS(a, old, static_cast<unsigned long long>(__double_as_longlong(v)));
    if (prev == old) break;
    old = prev;
  }
}

// Block-wide reduction of (or-flags, min, max, max) then one atomic each.
__device__ void reduce_publish(unsigned flags, double mn, double mx, double mx2,
                               unsigned *gflags, double *gmin, double *gmax, double *gmax2) {
  __shared__ unsigned sf[kThreads];
  __shared__ double s1[kThreads], s2[kThreads], s3[kThreads];
  const int t = threadIdx.x;
  sf[t] = flags;
  s1[t] = mn;
  s2[t] = mx;
  s3[t] = mx2;
  __syncthreads();
  for (int w = kThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
      sf[t] |= sf[t + w];
      s1[t] = fmin(s1[t], s1[t + w]);
      s2[t] = fmax(s2[t], s2[t + w]);
      s3[t] = fmax(s3[t], s3[t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    if (sf[0]) atomicOr(gflags, sf[0]);
    if (gmin) atomic_min_d(gmin, s1[0]);
    if (gmax) atomic_max_d(gmax, s2[0]);
    if (gmax2) atomic_max_d(gmax2, s3[0]);
  }
}

// The streamwise wavenumber and spanwise period of the next assembly into
// the case's parameter block (launched before an assembly, or before the
// replay of a captured one).
__global__ void k_case_scalars(double *__restrict__ recipe, double alpha_re, double alpha_im,
                               double period) {
  recipe[kRecipeAlphaRe] = alpha_re;
  recipe[kRecipeAlphaIm] = alpha_im;
  recipe[kRecipePeriod] = period;
}

__global__ void k_reset_diag(DeviceDiag *d) {
  d->flags = 0;
  d->char_status = 0;
  d->incoming = 0;
  d->neutral = 0;
  d->min_incoming = 5;
  d->max_incoming = 0;
  d->min_jac = DBL_MAX;
  d->max_jac = -DBL_MAX;
  d->max_grid_imag = 0.0;
  d->max_base_imag = 0.0;
  d->max_char_imag = 0.0;
  d->max_cond = 0.0;
}

// ---------------------------------------------------------------------------
// First inverse metrics, one thread per node (Metrics.cpp:91-187).
__global__ __launch_bounds__(kThreads) void k_metrics_first(
    CaseDims d, const int *__restrict__ xoff, const double *__restrict__ xw,
    const int *__restrict__ eoff, const double *__restrict__ ew, const double *__restrict__ recipe,
    const double *__restrict__ grid, double *__restrict__ metrics, DeviceDiag *diag) {
  const long long node = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x;
  unsigned flags = 0;
  double jmin = DBL_MAX, jmax = -DBL_MAX, imax = 0.0;
  const double period = recipe[kRecipePeriod];
  if (node < d.nodes) {
    const int i = static_cast<int>(node % d.ny), j = static_cast<int>(node / d.ny);
    const double *w1x = xw + static_cast<size_t>(d.ny) * d.ssx + static_cast<size_t>(i) * d.ssx;
    const double *w1e = ew + static_cast<size_t>(d.nz) * d.sse + static_cast<size_t>(j) * d.sse;
    double yXi = 0.0, yEta = 0.0, zXi = 0.0, zEta = 0.0;
    for (int s = 0; s < d.ssx; ++s) {
      const int col = i + xoff[i * d.ssx + s];
      const double *g = grid + 6 * (static_cast<size_t>(j) * d.ny + col);
      const double yr = g[2], yi = g[3], zr = g[4], zi = g[5];
      if (!finite(yr) || !finite(yi) || !finite(zr) || !finite(zi)) flags |= kGridNonFinite;
      imax = fmax(imax, fmax(fabs(yi), fabs(zi)));
      const double w = w1x[s];
      yXi += w * yr;
      zXi += w * zr;
    }
    for (int t = 0; t < d.sse; ++t) {
      const int row = j + eoff[j * d.sse + t];
      const int jw = wrap_mod(row, d.nz);
      const double *g = grid + 6 * (static_cast<size_t>(jw) * d.ny + i);
      const double yr = g[2], yi = g[3];
      const double zr = g[4] + static_cast<double>(wrap_count(row, d.nz)) * period;
      const double zi = g[5];
      if (!finite(yr) || !finite(yi) || !finite(zr) || !finite(zi)) flags |= kGridNonFinite;
      imax = fmax(imax, fmax(fabs(yi), fabs(zi)));
      const double w = w1e[t];
      yEta += w * yr;
      zEta += w * zr;
    }
    if (!finite(yXi) || !finite(yEta) || !finite(zXi) || !finite(zEta)) {
      flags |= kGridNonFinite;
    } else {
      const double p1 = yXi * zEta, p2 = yEta * zXi;
      const double J = p1 - p2;
      const double scale = fmax(1.0, fmax(fabs(p1), fabs(p2)));
      const double tol = 100.0 * kEps * scale;
      if (!finite(J)) {
        flags |= kGridNonFinite;
      } else {
        jmin = J;
        jmax = J;
        if (J <= tol) {
          flags |= kJacobian;
        } else {
          double *m = metrics + 10 * node;
          const double m0 = zEta / J, m1 = -yEta / J, m2 = -zXi / J, m3 = yXi / J;
          if (!finite(m0) || !finite(m1) || !finite(m2) || !finite(m3)) flags |= kGridNonFinite;
          m[0] = m0;
          m[1] = m1;
          m[2] = m2;
          m[3] = m3;
        }
      }
    }
  }
  reduce_publish(flags, jmin, jmax, imax, &diag->flags, &diag->min_jac, &diag->max_jac,
                 &diag->max_grid_imag);
}

// ---------------------------------------------------------------------------
// Second metrics (chain rule of first metrics, Metrics.cpp:255-337) and the 25
// physical base-flow derivatives (BaseFlowDerivatives.cpp:85-218).
//
// One workgroup owns a DTX x DTY tile of nodes and stages the tile plus its
// stencil halo (hx nodes each side in xi, clipped at the walls; hy rows each
// side in eta, wrapped periodically) of the base flow and of the first
// metrics in LDS.  Thread (node, field) then evaluates its 11 + 7 + 77 taps
// from LDS; the threads of field 0 first produce the node's second metrics.
constexpr int DTX = 16, DTY = 8, DNODES = DTX * DTY, DTHREADS = 5 * DNODES;

__global__ __launch_bounds__(DTHREADS) void k_metrics2_derivatives(
    CaseDims d, const int *__restrict__ xoff, const double *__restrict__ xw,
    const int *__restrict__ eoff, const double *__restrict__ ew,
    const double *__restrict__ base, double *__restrict__ metrics, double *__restrict__ deriv,
    DeviceDiag *diag) {
  extern __shared__ double lds[];
  __shared__ double m2[DNODES][6];
  __shared__ unsigned s_flags;
  __shared__ unsigned long long s_imax;
  const int tid = threadIdx.x;
  const int i0 = blockIdx.x * DTX, j0 = blockIdx.y * DTY;
  const int xlo = max(0, i0 - d.hx), xhi = min(d.ny, i0 + DTX + d.hx);
  const int EX = xhi - xlo, EY = DTY + 2 * d.hy;
  const int plane = EX * EY;
  double *sb = lds;                 // [5][EY][EX] base flow (real parts)
  double *sm = lds + 5 * plane;     // [4][EY][EX] first metrics
  if (tid == 0) {
    s_flags = 0;
    s_imax = 0;
  }
  __syncthreads();
  unsigned flags = 0;
  double imax = 0.0;
  for (int idx = tid; idx < plane; idx += DTHREADS) {
    const int ly = idx / EX, lx = idx - EX * (idx / EX);
    const int gy = wrap_mod(j0 - d.hy + ly, d.nz);
    const size_t node = static_cast<size_t>(gy) * d.ny + xlo + lx;
    const double *b = base + 10 * node;
    for (int f = 0; f < 5; ++f) {
      const double re = b[2 * f], im = b[2 * f + 1];
      if (!finite(re) || !finite(im)) flags |= kBaseNonFinite;
      imax = fmax(imax, fabs(im));
      sb[f * plane + idx] = re;
    }
    const double *mm = metrics + 10 * node;
    for (int c = 0; c < 4; ++c) sm[c * plane + idx] = mm[c];
  }
  __syncthreads();

  const int ln = tid % DNODES, f = tid / DNODES;
  const int i = i0 + ln % DTX, j = j0 + ln / DTX;
  const bool active = i < d.ny && j < d.nz;
  const double *w1x = xw + static_cast<size_t>(d.ny) * d.ssx + static_cast<size_t>(i) * d.ssx;
  const double *w2x = w1x + static_cast<size_t>(d.ny) * d.ssx;
  const double *w1e = ew + static_cast<size_t>(d.nz) * d.sse + static_cast<size_t>(j) * d.sse;
  const double *w2e = w1e + static_cast<size_t>(d.nz) * d.sse;
  const int *ox = xoff + i * d.ssx;
  const int *oe = eoff + j * d.sse;
  const int cy = j - j0 + d.hy, cx = i - xlo;  // local coordinates of the node
  if (active && f == 0) {
    double dXi[4] = {0.0, 0.0, 0.0, 0.0}, dEta[4] = {0.0, 0.0, 0.0, 0.0};
    for (int s = 0; s < d.ssx; ++s) {
      const int l = cy * EX + cx + ox[s];
      const double w = w1x[s];
      for (int c = 0; c < 4; ++c) dXi[c] += w * sm[c * plane + l];
    }
    for (int t = 0; t < d.sse; ++t) {
      const int l = (cy + oe[t]) * EX + cx;
      const double w = w1e[t];
      for (int c = 0; c < 4; ++c) dEta[c] += w * sm[c * plane + l];
    }
    const int l0 = cy * EX + cx;
    const double xiY = sm[l0], xiZ = sm[plane + l0], etaY = sm[2 * plane + l0],
                 etaZ = sm[3 * plane + l0];
    double v[6];
    v[0] = xiY * dXi[0] + etaY * dEta[0];
    v[1] = xiZ * dXi[1] + etaZ * dEta[1];
    v[2] = xiZ * dXi[0] + etaZ * dEta[0];
    v[3] = xiY * dXi[2] + etaY * dEta[2];
    v[4] = xiZ * dXi[3] + etaZ * dEta[3];
    v[5] = xiZ * dXi[2] + etaZ * dEta[2];
    double *m = metrics + 10 * (static_cast<size_t>(j) * d.ny + i);
    for (int c = 0; c < 6; ++c) {
      if (!finite(v[c])) flags |= kSecondNonFinite;
      m2[ln][c] = v[c];
      m[4 + c] = v[c];
    }
    const double rho = sb[l0], T = sb[4 * plane + l0];
    if (finite(rho) && finite(T) && (rho <= 0.0 || T <= 0.0)) flags |= kBasePositive;
  }
  __syncthreads();
  if (active) {
    const double *sf = sb + f * plane;
    double vXi = 0.0, vEta = 0.0, vXiXi = 0.0, vEtaEta = 0.0, vXiEta = 0.0;
    for (int s = 0; s < d.ssx; ++s) {
      const double b = sf[cy * EX + cx + ox[s]];
      vXi += w1x[s] * b;
      vXiXi += w2x[s] * b;
    }
    for (int t = 0; t < d.sse; ++t) {
      const double b = sf[(cy + oe[t]) * EX + cx];
      vEta += w1e[t] * b;
      vEtaEta += w2e[t] * b;
    }
    for (int t = 0; t < d.sse; ++t) {
      const double we = w1e[t];
      const double *row = sf + (cy + oe[t]) * EX + cx;
      for (int s = 0; s < d.ssx; ++s) vXiEta += we * w1x[s] * row[ox[s]];
    }
    const int l0 = cy * EX + cx;
    const double xiY = sm[l0], xiZ = sm[plane + l0], etaY = sm[2 * plane + l0],
                 etaZ = sm[3 * plane + l0];
    const double xiYY = m2[ln][0], xiZZ = m2[ln][1], xiYZ = m2[ln][2];
    const double etaYY = m2[ln][3], etaZZ = m2[ln][4], etaYZ = m2[ln][5];
    const double vY = xiY * vXi + etaY * vEta;
    const double vZ = xiZ * vXi + etaZ * vEta;
    const double vYY = xiY * xiY * vXiXi + 2.0 * xiY * etaY * vXiEta + etaY * etaY * vEtaEta +
                       xiYY * vXi + etaYY * vEta;
    const double vZZ = xiZ * xiZ * vXiXi + 2.0 * xiZ * etaZ * vXiEta + etaZ * etaZ * vEtaEta +
                       xiZZ * vXi + etaZZ * vEta;
    const double vYZ = xiY * xiZ * vXiXi + (xiY * etaZ + etaY * xiZ) * vXiEta +
                       etaY * etaZ * vEtaEta + xiYZ * vXi + etaYZ * vEta;
    if (!finite(vY) || !finite(vZ) || !finite(vYY) || !finite(vZZ) || !finite(vYZ)) {
      flags |= kDerivNonFinite;
    } else {
      double *dd = deriv + 25 * (static_cast<size_t>(j) * d.ny + i);
      dd[0 * 5 + f] = vY;
      dd[1 * 5 + f] = vZ;
      dd[2 * 5 + f] = vYY;
      dd[3 * 5 + f] = vZZ;
      dd[4 * 5 + f] = vYZ;
    }
  }
  if (flags) atomicOr(&s_flags, flags);
  if (imax > 0.0) atomicMax(&s_imax, static_cast<unsigned long long>(__double_as_longlong(imax)));
  __syncthreads();
  if (tid == 0) {
    if (s_flags) atomicOr(&diag->flags, s_flags);
    if (s_imax) atomic_max_d(&diag->max_base_imag, __longlong_as_double(static_cast<long long>(s_imax)));
  }
}

// ---------------------------------------------------------------------------
// Pointwise physics.  Base-flow point: value + y/z derivatives; the
// x-derivatives are zero in the locally parallel BiGlobal model
// (LNSCoefficients.cpp:75-110), so the x terms of the reference formulas
// drop out exactly.
typedef double double2v __attribute__((ext_vector_type(2)));

struct Point {
  double rho, U, V, W, T;
  double rhoy, Uy, Vy, Wy, Ty;
  double rhoz, Uz, Vz, Wz, Tz;
  double Uyy, Vyy, Wyy, Tyy;
  double Uzz, Vzz, Wzz, Tzz;
  double Vyz, Wyz;
};

struct Transport {
  double Re, Pr, Pe, g2, mu, muT, muTT;
};

// Returns the CoefficientStatus of makeBaseFlowPoint (0 = success).
__device__ int load_point(const double *__restrict__ base, const double *__restrict__ deriv,
                          long long node, Point &p) {
  const double *b = base + 10 * node;
  const double *dd = deriv + 25 * node;
  double v[5];
  for (int f = 0; f < 5; ++f) {
    if (!finite(b[2 * f]) || !finite(b[2 * f + 1]) || b[2 * f + 1] != 0.0) return 1;
    v[f] = b[2 * f];
  }
  for (int k = 0; k < 25; ++k)
    if (!finite(dd[k])) return 1;
  p.rho = v[0]; p.U = v[1]; p.V = v[2]; p.W = v[3]; p.T = v[4];
  p.rhoy = dd[0]; p.Uy = dd[1]; p.Vy = dd[2]; p.Wy = dd[3]; p.Ty = dd[4];
  p.rhoz = dd[5]; p.Uz = dd[6]; p.Vz = dd[7]; p.Wz = dd[8]; p.Tz = dd[9];
  p.Uyy = dd[11]; p.Vyy = dd[12]; p.Wyy = dd[13]; p.Tyy = dd[14];
  p.Uzz = dd[16]; p.Vzz = dd[17]; p.Wzz = dd[18]; p.Tzz = dd[19];
  p.Vyz = dd[22]; p.Wyz = dd[23];
  if (p.rho <= 0.0) return 2;
  if (p.T <= 0.0) return 3;
  return 0;
}

__device__ Transport transport(const double *recipe, double T) {
  Transport t;
  const double Ma = recipe[1], gam = recipe[3];
  const double s = recipe[5] / recipe[4];
  t.Re = recipe[0];
  t.Pr = recipe[2];
  t.Pe = 1.0 / (gam * Ma * Ma);
  t.g2 = 1.0 / ((gam - 1.0) * Ma * Ma);
  t.mu = T * sqrt(T) * (1.0 + s) / (T + s);
  t.muT = t.mu * (1.5 / T - 1.0 / (T + s));
  t.muTT = t.muT * (1.5 / T - 1.0 / (T + s)) - t.mu * (1.5 / (T * T) - 1.0 / ((T + s) * (T + s)));
  return t;
}

constexpr double k13 = 1.0 / 3.0, k23 = 2.0 / 3.0, k43 = 4.0 / 3.0;

// Physical blocks, grouped so that six threads of a node share the work:
// part 0: Gamma + the six viscous blocks, 1: A, 2: B, 3: C, 4: D.
// out points at 11 row-major 5x5 blocks (Gamma, A, B, C, D, Vxx, Vxy, Vxz,
// Vyy, Vyz, Vzz); each part only writes its own blocks.
__device__ void physical_part(int part, const Point &p, const Transport &q, double *out) {
  const double rho = p.rho, U = p.U, V = p.V, W = p.W, T = p.T;
  const double Re = q.Re, Pr = q.Pr, Pe = q.Pe, g2 = q.g2, mu = q.mu, muT = q.muT, muTT = q.muTT;
  // x-derivatives of the locally parallel base flow
  const double rhox = 0.0, Ux = 0.0, Vx = 0.0, Wx = 0.0, Tx = 0.0;
  const double Uxx = 0.0, Vxx = 0.0, Wxx = 0.0, Txx = 0.0, Uxy = 0.0, Vxy = 0.0, Uxz = 0.0,
               Wxz = 0.0;
  const double mux = muT * Tx, muy = muT * p.Ty, muz = muT * p.Tz;
  (void)Vxy;
  auto Z = [&](int blk) {
    for (int k = 0; k < 25; ++k) out[blk * 25 + k] = 0.0;
  };
#define SETB(blk, r, c, v) out[(blk) * 25 + (r) * 5 + (c)] = (v)
  switch (part) {
    case 0: {
      Z(0);
      for (int b = 5; b < 11; ++b) Z(b);
      SETB(0, 0, 0, 1.0);
      SETB(0, 1, 1, rho);
      SETB(0, 2, 2, rho);
      SETB(0, 3, 3, rho);
      SETB(0, 4, 0, -Pe * T);
      SETB(0, 4, 4, rho * g2 - Pe * rho);
      const double m43 = k43 * mu / Re, m1 = mu / Re, mk = mu * g2 / (Re * Pr), m13 = k13 * mu / Re;
      SETB(5, 1, 1, m43); SETB(5, 2, 2, m1); SETB(5, 3, 3, m1); SETB(5, 4, 4, mk);
      SETB(8, 1, 1, m1); SETB(8, 2, 2, m43); SETB(8, 3, 3, m1); SETB(8, 4, 4, mk);
      SETB(10, 1, 1, m1); SETB(10, 2, 2, m1); SETB(10, 3, 3, m43); SETB(10, 4, 4, mk);
      SETB(6, 1, 2, m13); SETB(6, 2, 1, m13);
      SETB(7, 1, 3, m13); SETB(7, 3, 1, m13);
      SETB(9, 2, 3, m13); SETB(9, 3, 2, m13);
      break;
    }
    case 1: {
      const double Vy = p.Vy, Wz = p.Wz, Uy = p.Uy, Uz = p.Uz;
      Z(1);
      SETB(1, 0, 0, U);
      SETB(1, 0, 1, rho);
      SETB(1, 1, 0, Pe * T);
      SETB(1, 1, 1, rho * U - k43 * mux / Re);
      SETB(1, 1, 2, -muy / Re);
      SETB(1, 1, 3, -muz / Re);
      SETB(1, 1, 4, Pe * rho - muT / Re * (k43 * Ux - k23 * Vy - k23 * Wz));
      SETB(1, 2, 1, k23 * muy / Re);
      SETB(1, 2, 2, rho * U - mux / Re);
      SETB(1, 2, 4, -muT * (Uy + Vx) / Re);
      SETB(1, 3, 1, k23 * muz / Re);
      SETB(1, 3, 3, rho * U - mux / Re);
      SETB(1, 3, 4, -muT * (Wx + Uz) / Re);
      SETB(1, 4, 0, -Pe * U * T);
      SETB(1, 4, 1, -2.0 * mu * (k43 * Ux - k23 * Vy - k23 * Wz) / Re);
      SETB(1, 4, 2, -2.0 * mu * (Uy + Vx) / Re);
      SETB(1, 4, 3, -2.0 * mu * (Wx + Uz) / Re);
      SETB(1, 4, 4, rho * U * g2 - rho * U * Pe - 2.0 * Tx * muT * g2 / (Re * Pr));
      break;
    }
    case 2: {
      const double Vy = p.Vy, Wz = p.Wz, Uy = p.Uy, Vz = p.Vz, Wy = p.Wy, Ty = p.Ty;
      Z(2);
      SETB(2, 0, 0, V);
      SETB(2, 0, 2, rho);
      SETB(2, 1, 1, rho * V - muy / Re);
      SETB(2, 1, 2, k23 * mux / Re);
      SETB(2, 1, 4, -muT * (Uy + Vx) / Re);
      SETB(2, 2, 0, Pe * T);
      SETB(2, 2, 1, -mux / Re);
      SETB(2, 2, 2, rho * V - k43 * muy / Re);
      SETB(2, 2, 3, -muz / Re);
      SETB(2, 2, 4, Pe * rho - muT / Re * (-k23 * Ux + k43 * Vy - k23 * Wz));
      SETB(2, 3, 2, k23 * muz / Re);
      SETB(2, 3, 3, rho * V - muy / Re);
      SETB(2, 3, 4, -muT * (Vz + Wy) / Re);
      SETB(2, 4, 0, -Pe * V * T);
      SETB(2, 4, 1, -2.0 * mu * (Uy + Vx) / Re);
      SETB(2, 4, 2, -2.0 * mu * (-k23 * Ux + k43 * Vy - k23 * Wz) / Re);
      SETB(2, 4, 3, -2.0 * mu * (Vz + Wy) / Re);
      SETB(2, 4, 4, rho * V * g2 - rho * V * Pe - 2.0 * Ty * muT * g2 / (Re * Pr));
      break;
    }
    case 3: {
      const double Vy = p.Vy, Wz = p.Wz, Uz = p.Uz, Vz = p.Vz, Wy = p.Wy, Tz = p.Tz;
      Z(3);
      SETB(3, 0, 0, W);
      SETB(3, 0, 3, rho);
      SETB(3, 1, 1, rho * W - muz / Re);
      SETB(3, 1, 3, k23 * mux / Re);
      SETB(3, 1, 4, -muT * (Wx + Uz) / Re);
      SETB(3, 2, 2, rho * W - muz / Re);
      SETB(3, 2, 3, k23 * muy / Re);
      SETB(3, 2, 4, -muT * (Vz + Wy) / Re);
      SETB(3, 3, 0, Pe * T);
      SETB(3, 3, 1, -mux / Re);
      SETB(3, 3, 2, -muy / Re);
      SETB(3, 3, 3, rho * W - k43 * muz / Re);
      SETB(3, 3, 4, Pe * rho - muT / Re * (-k23 * Ux - k23 * Vy + k43 * Wz));
      SETB(3, 4, 0, -Pe * W * T);
      SETB(3, 4, 1, -2.0 * mu * (Wx + Uz) / Re);
      SETB(3, 4, 2, -2.0 * mu * (Vz + Wy) / Re);
      SETB(3, 4, 3, -2.0 * mu * (-k23 * Ux - k23 * Vy + k43 * Wz) / Re);
      SETB(3, 4, 4, rho * W * g2 - rho * W * Pe - 2.0 * Tz * muT * g2 / (Re * Pr));
      break;
    }
    case 4: {
      const double rhoy = p.rhoy, rhoz = p.rhoz, Uy = p.Uy, Uz = p.Uz, Vy = p.Vy, Vz = p.Vz,
                   Wy = p.Wy, Wz = p.Wz, Ty = p.Ty, Tz = p.Tz;
      const double Uyy = p.Uyy, Uzz = p.Uzz, Vyy = p.Vyy, Vzz = p.Vzz, Wyy = p.Wyy, Wzz = p.Wzz;
      const double Tyy = p.Tyy, Tzz = p.Tzz, Vyz = p.Vyz, Wyz = p.Wyz;
      Z(4);
      SETB(4, 0, 0, Ux + Vy + Wz);
      SETB(4, 0, 1, rhox);
      SETB(4, 0, 2, rhoy);
      SETB(4, 0, 3, rhoz);
      SETB(4, 1, 0, U * Ux + V * Uy + W * Uz + Pe * Tx);
      SETB(4, 1, 1, rho * Ux);
      SETB(4, 1, 2, rho * Uy);
      SETB(4, 1, 3, rho * Uz);
      SETB(4, 1, 4,
           Pe * rhox - (muTT * Tx * k23 * (2.0 * Ux - Vy - Wz) + muTT * Ty * (Uy + Vx) +
                        muTT * Tz * (Wx + Uz) +
                        muT * (k43 * Uxx + Uyy + Uzz + k13 * Vxy + k13 * Wxz)) /
                           Re);
      SETB(4, 2, 0, U * Vx + V * Vy + W * Vz + Pe * Ty);
      SETB(4, 2, 1, rho * Vx);
      SETB(4, 2, 2, rho * Vy);
      SETB(4, 2, 3, rho * Vz);
      SETB(4, 2, 4,
           Pe * rhoy - (muTT * Tx * (Uy + Vx) + muTT * Ty * k23 * (-Ux + 2.0 * Vy - Wz) +
                        muTT * Tz * (Vz + Wy) +
                        muT * (Vxx + k43 * Vyy + Vzz + k13 * Uxy + k13 * Wyz)) /
                           Re);
      SETB(4, 3, 0, U * Wx + V * Wy + W * Wz + Pe * Tz);
      SETB(4, 3, 1, rho * Wx);
      SETB(4, 3, 2, rho * Wy);
      SETB(4, 3, 3, rho * Wz);
      SETB(4, 3, 4,
           Pe * rhoz - (muTT * Tx * (Wx + Uz) + muTT * Ty * (Vz + Wy) +
                        muTT * Tz * k23 * (-Ux - Vy + 2.0 * Wz) +
                        muT * (Wxx + Wyy + k43 * Wzz + k13 * Uxz + k13 * Vyz)) /
                           Re);
      SETB(4, 4, 0, (g2 - Pe) * (U * Tx + V * Ty + W * Tz));
      SETB(4, 4, 1, rho * Tx * g2 - Pe * (rho * Tx + T * rhox));
      SETB(4, 4, 2, rho * Ty * g2 - Pe * (rho * Ty + T * rhoy));
      SETB(4, 4, 3, rho * Tz * g2 - Pe * (rho * Tz + T * rhoz));
      SETB(4, 4, 4,
           -Pe * (U * rhox + V * rhoy + W * rhoz) - (Txx + Tyy + Tzz) * muT * g2 / (Re * Pr) -
               (Tx * Tx + Ty * Ty + Tz * Tz) * muTT * g2 / (Re * Pr) -
               muT *
                   (Uy * Uy + Vx * Vx + 2.0 * Uy * Vx + Uz * Uz + Wx * Wx + 2.0 * Uz * Wx +
                    Vz * Vz + Wy * Wy + 2.0 * Vz * Wy) /
                   Re -
               muT * k43 * (Ux * Ux + Vy * Vy + Wz * Wz - Ux * Wz - Ux * Vy - Vy * Wz) / Re);
      break;
    }
    default:
      break;
  }
#undef SETB
}

// ---------------------------------------------------------------------------
// Far-field characteristic data, one thread per eta index
// (BoundaryConditions.cpp:152-280).  Writes per node 50 doubles of
// P = R^-1 Gamma^-1, 50 of R^-1, and the incoming mask; statuses follow
// CharacteristicStatus (1 invalid input, 2 singular metric, 6 singular
// eigenvectors, 7 ill-conditioned).
//
// The normal flux eigenproblem (n_y Bc + n_z Cc) r = c Gamma r has speeds
// u_n - a, u_n (three times), u_n + a, a = sqrt(T)/Ma.  The eigenvectors
// used as columns of R: acoustic [rho, 0, -+n_y a, -+n_z a, (gamma-1) T],
// streamwise e_1, tangential [0, 0, n_z, -n_y, 0], entropy [rho, 0, 0, 0, -T].

// Pivoted Gauss-Jordan inverse of a 5x5 (BoundaryConditions.cpp:92-140).
// The closed-form eigenbasis and Gamma are real, and for real operands the
// reference's complex steps (|z| by hypot, division, multiply-subtract)
// reduce exactly to these real ones, so the inverse is computed in real
// arithmetic: fully unrolled, register resident.
__device__ __forceinline__ bool invert5(double (&a)[25], double (&inv)[25]) {
#pragma unroll
  for (int k = 0; k < 25; ++k) inv[k] = 0.0;
#pragma unroll
  for (int k = 0; k < 5; ++k) inv[k * 5 + k] = 1.0;
#pragma unroll
  for (int pc = 0; pc < 5; ++pc) {
    int pr = pc;
    double pm = fabs(a[pc * 5 + pc]);
#pragma unroll
    for (int r = pc + 1; r < 5; ++r) {
      const double mg = fabs(a[r * 5 + pc]);
      if (mg > pm) {
        pm = mg;
        pr = r;
      }
    }
    if (!finite(pm) || pm <= 100.0 * kEps) return false;
#pragma unroll
    for (int r = pc + 1; r < 5; ++r) {
      if (r == pr) {
#pragma unroll
        for (int c = 0; c < 5; ++c) {
          double t = a[r * 5 + c];
          a[r * 5 + c] = a[pc * 5 + c];
          a[pc * 5 + c] = t;
          t = inv[r * 5 + c];
          inv[r * 5 + c] = inv[pc * 5 + c];
          inv[pc * 5 + c] = t;
        }
      }
    }
    const double piv = a[pc * 5 + pc];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      a[pc * 5 + c] = a[pc * 5 + c] / piv;
      inv[pc * 5 + c] = inv[pc * 5 + c] / piv;
    }
#pragma unroll
    for (int r = 0; r < 5; ++r) {
      if (r == pc) continue;
      const double f = a[r * 5 + pc];
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        a[r * 5 + c] -= f * a[pc * 5 + c];
        inv[r * 5 + c] -= f * inv[pc * 5 + c];
      }
    }
  }
  return true;
}

__global__ __launch_bounds__(64) void k_characteristics(CaseDims d, const double *__restrict__ recipe,
                                  const double *__restrict__ base,
                                  const double *__restrict__ deriv,
                                  const double *__restrict__ metrics, double *__restrict__ bcdata,
                                  DeviceDiag *diag) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d.nz) return;
  const long long node = static_cast<long long>(j) * d.ny + (d.ny - 1);
  double *out = bcdata + 106 * static_cast<size_t>(j);  // P[50], Rinv[50], incoming[5], status
  Point p;
  int st = load_point(base, deriv, node, p) ? 1 : 0;
  const double *m = metrics + 10 * node;
#pragma unroll
  for (int k = 0; k < 10; ++k)
    if (!finite(m[k])) st = st ? st : 1;
  const double Ma = recipe[1], gam = recipe[3];
  const double Pe = 1.0 / (gam * Ma * Ma), g2 = 1.0 / ((gam - 1.0) * Ma * Ma);
  double nm = 0.0;
  if (!st) {
    nm = sqrt(m[0] * m[0] + m[1] * m[1]);
    if (!finite(nm) || nm <= 100.0 * kEps) st = 2;
  }
  double R[25], Ri[25], G[25], Gi[25], tmp[25];
  double c[5] = {0, 0, 0, 0, 0};
  double cond = 0.0;
  if (!st) {
    const double ny = m[0] / nm, nz = m[1] / nm;
    const double un = ny * p.V + nz * p.W;
    const double a = sqrt(p.T) / Ma;
    c[0] = un - a; c[1] = un; c[2] = un; c[3] = un; c[4] = un + a;
#pragma unroll
    for (int k = 0; k < 25; ++k) {
      R[k] = 0.0;
      G[k] = 0.0;
    }
    R[0 * 5 + 0] = p.rho; R[2 * 5 + 0] = -ny * a; R[3 * 5 + 0] = -nz * a;
    R[4 * 5 + 0] = (gam - 1.0) * p.T;
    R[1 * 5 + 1] = 1.0;
    R[2 * 5 + 2] = nz; R[3 * 5 + 2] = -ny;
    R[0 * 5 + 3] = p.rho; R[4 * 5 + 3] = -p.T;
    R[0 * 5 + 4] = p.rho; R[2 * 5 + 4] = ny * a; R[3 * 5 + 4] = nz * a;
    R[4 * 5 + 4] = (gam - 1.0) * p.T;
#pragma unroll
    for (int k = 0; k < 25; ++k) tmp[k] = R[k];
    G[0] = 1.0; G[6] = p.rho; G[12] = p.rho; G[18] = p.rho;
    G[20] = -Pe * p.T; G[24] = p.rho * g2 - Pe * p.rho;
    if (!invert5(tmp, Ri)) st = 6;
    if (!st && !invert5(G, Gi)) st = 6;
    if (!st) {
      double nR = 0.0, nRi = 0.0;
#pragma unroll
      for (int r = 0; r < 5; ++r) {
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int cc = 0; cc < 5; ++cc) {
          s1 += fabs(R[r * 5 + cc]);
          s2 += fabs(Ri[r * 5 + cc]);
        }
        nR = fmax(nR, s1);
        nRi = fmax(nRi, s2);
      }
      cond = nR * nRi;
      if (!finite(cond) || cond > 1.0e12) st = 7;
    }
  }
  if (st) {
    out[105] = st;
    atomicOr(&diag->flags, kCharacteristic);
    atomicMax(&diag->char_status, st);
    return;
  }
  double maxSpeed = 1.0;
#pragma unroll
  for (int k = 0; k < 5; ++k) maxSpeed = fmax(maxSpeed, fabs(c[k]));
  const double tol = 1000.0 * kEps * maxSpeed;
  int inc = 0, neu = 0;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const bool in = c[k] < -tol;
    out[100 + k] = in ? 1.0 : 0.0;
    inc += in;
    neu += fabs(c[k]) <= tol;
  }
#pragma unroll
  for (int r = 0; r < 5; ++r)
#pragma unroll
    for (int cc = 0; cc < 5; ++cc) {
      double sr = 0.0;
#pragma unroll
      for (int k = 0; k < 5; ++k) sr += Ri[r * 5 + k] * Gi[k * 5 + cc];
      out[2 * (r * 5 + cc)] = sr;
      out[2 * (r * 5 + cc) + 1] = 0.0;
      out[50 + 2 * (r * 5 + cc)] = Ri[r * 5 + cc];
      out[50 + 2 * (r * 5 + cc) + 1] = 0.0;
    }
  out[105] = 0.0;
  atomicAdd(&diag->incoming, inc);
  atomicAdd(&diag->neutral, neu);
  atomicMin(&diag->min_incoming, inc);
  atomicMax(&diag->max_incoming, inc);
  atomic_max_d(&diag->max_cond, cond);
}

// ---------------------------------------------------------------------------
// The operator writer.

// Per-node coefficient record written by k_coefficients and read by the
// writer (doubles): K0, Kxi, Keta (25 complex each), Vxixi, Vxieta, Vetaeta
// (25 real each) and Gamma (25 real).
constexpr int kCoefDoubles = 250;
// nodes per k_coefficients workgroup (16 or 32 measured the same at 200x100)
constexpr int kCoefNodes = 8;

// Physical LNS blocks (LNSCoefficients.cpp:173-404), their finiteness, and
// the Fourier (StreamwiseFourier.cpp:24-41) and curvilinear
// (CurvilinearTransform.cpp:68-96) combinations of every node; status 0 on
// success, nonzero for the reference's NonFinite/NonPositive outcomes.
__global__ __launch_bounds__(kThreads) void k_coefficients(
    CaseDims d, const double *__restrict__ recipe, const double *__restrict__ base,
    const double *__restrict__ deriv, const double *__restrict__ metrics,
    double *__restrict__ coef, int *__restrict__ status) {
  __shared__ double phys[kCoefNodes][11 * 25];
  __shared__ int st[kCoefNodes];
  const int tid = threadIdx.x;
  const long long first = static_cast<long long>(blockIdx.x) * kCoefNodes;
  if (tid < kCoefNodes) st[tid] = 0;
  __syncthreads();
  // six threads per node evaluate the physical blocks; every thread that
  // writes blocks checks their finiteness (NonFiniteOutput of evaluate).  All
  // six see the same point status s0: a failed point is recorded by one of
  // them, and only nodes with s0 == 0 reach the atomic.
  for (int t6 = tid; t6 < 6 * kCoefNodes; t6 += kThreads) {
    const int n = t6 / 6, part = t6 - 6 * (t6 / 6);
    const long long node = first + n;
    if (node < d.nodes) {
      Point p;
      const int s0 = load_point(base, deriv, node, p);
      if (part == 5) {
        if (s0 != 0) st[n] = s0;
      } else if (s0 == 0) {
        const Transport q = transport(recipe, p.T);
        physical_part(part, p, q, phys[n]);
        // part 0 wrote Gamma and the six viscous blocks, part k block k
        const int b0 = part == 0 ? 0 : part, b1 = part == 0 ? 1 : part + 1;
        bool ok = true;
        for (int b = b0; b < b1; ++b)
          for (int k = 0; k < 25; ++k) ok = ok && finite(phys[n][b * 25 + k]);
        if (part == 0)
          for (int b = 5; b < 11; ++b)
            for (int k = 0; k < 25; ++k) ok = ok && finite(phys[n][b * 25 + k]);
        if (!ok) atomicMax(&st[n], 4);
      }
    }
  }
  __syncthreads();
  // 25 threads per node: Fourier and curvilinear entries
  for (int t25 = tid; t25 < 25 * kCoefNodes; t25 += kThreads) {
    const int n = t25 / 25, e = t25 - 25 * (t25 / 25);
    const long long node = first + n;
    if (node < d.nodes && st[n] == 0) {
      const double alpha_re = recipe[kRecipeAlphaRe], alpha_im = recipe[kRecipeAlphaIm];
      const double ia_re = 0.0 * alpha_re - 1.0 * alpha_im;
      const double ia_im = 0.0 * alpha_im + 1.0 * alpha_re;
      const double a2_re = alpha_re * alpha_re - alpha_im * alpha_im;
      const double a2_im = alpha_re * alpha_im + alpha_im * alpha_re;
      const double *ph = phys[n];
      const double A = ph[25 + e], B = ph[50 + e], C = ph[75 + e], D = ph[100 + e];
      const double Vxx = ph[125 + e], Vxy = ph[150 + e], Vxz = ph[175 + e];
      const double Vyy = ph[200 + e], Vyz = ph[225 + e], Vzz = ph[250 + e];
      const double K0r = D + ia_re * A + a2_re * Vxx;
      const double K0i = 0.0 + ia_im * A + a2_im * Vxx;
      const double Kyr = B - ia_re * Vxy, Kyi = 0.0 - ia_im * Vxy;
      const double Kzr = C - ia_re * Vxz, Kzi = 0.0 - ia_im * Vxz;
      const double *m = metrics + 10 * node;
      const double xiY = m[0], xiZ = m[1], etaY = m[2], etaZ = m[3];
      const double xiYY = m[4], xiZZ = m[5], xiYZ = m[6], etaYY = m[7], etaZZ = m[8], etaYZ = m[9];
      bool ok = finite(K0r) && finite(K0i) && finite(Kyr) && finite(Kyi) && finite(Kzr) &&
                finite(Kzi);
      for (int k = 0; k < 10; ++k) ok = ok && finite(m[k]);
      const double Kxr = xiY * Kyr + xiZ * Kzr - xiYY * Vyy - xiZZ * Vzz - xiYZ * Vyz;
      const double Kxi = xiY * Kyi + xiZ * Kzi;
      const double Ker = etaY * Kyr + etaZ * Kzr - etaYY * Vyy - etaZZ * Vzz - etaYZ * Vyz;
      const double Kei = etaY * Kyi + etaZ * Kzi;
      const double Vxixi = xiY * xiY * Vyy + xiZ * xiZ * Vzz + xiY * xiZ * Vyz;
      const double Vxieta =
          2.0 * xiY * etaY * Vyy + 2.0 * xiZ * etaZ * Vzz + (xiY * etaZ + etaY * xiZ) * Vyz;
      const double Veta = etaY * etaY * Vyy + etaZ * etaZ * Vzz + etaY * etaZ * Vyz;
      ok = ok && finite(Kxr) && finite(Kxi) && finite(Ker) && finite(Kei) && finite(Vxixi) &&
           finite(Vxieta) && finite(Veta);
      double *c = coef + kCoefDoubles * node;
      c[2 * e] = K0r;
      c[2 * e + 1] = K0i;
      c[50 + 2 * e] = Kxr;
      c[50 + 2 * e + 1] = Kxi;
      c[100 + 2 * e] = Ker;
      c[100 + 2 * e + 1] = Kei;
      c[150 + e] = Vxixi;
      c[175 + e] = Vxieta;
      c[200 + e] = Veta;
      c[225 + e] = ph[e];  // Gamma
      if (!ok) atomicMax(&st[n], 4);
    }
  }
  __syncthreads();
  for (int n = tid; n < kCoefNodes; n += kThreads)
    if (first + n < d.nodes) status[first + n] = st[n];
}

constexpr int kOperatorRows = 8;  // block rows (waves) per writer workgroup

template <int SSX, int SSE>
struct NodeRec {
  double gamma[25];
  dcx K0[25], Kxi[25], Keta[25];
  double Vxx[25], Vxe[25], Vee[25];  // Vxixi, Vxieta, Vetaeta (real)
  double w1x[SSX], mw2x[SSX], w1e[SSE], mw2e[SSE];
  dcx P[25], Ri[25];
  int slot_of_rank[SSE];
  int incoming[5];
  int t0, s0, kind, status;
};

// The operator writer: workgroup = NPW consecutive block rows, one wave per
// row.  The node's coefficient record is staged in LDS; the row is then
// streamed out with each lane owning one block entry.
// SSX x SSE: the stencil sizes; with RT they are only the array bounds and
// the case's own sizes (d.ssx <= SSX, d.sse <= SSE) are used at run time.
template <int SSX, int SSE, int NPW, bool RT, bool BYPASS>
__global__ __launch_bounds__(64 * NPW) void k_operator(
    CaseDims d, const int *__restrict__ xoff, const double *__restrict__ xw,
    const int *__restrict__ eoff, const double *__restrict__ ew, const double *__restrict__ coef,
    const int *__restrict__ status, const double *__restrict__ bcdata, int constrained,
    double *__restrict__ vals, double *__restrict__ mvals, DeviceDiag *diag) {
  const int NSX = RT ? d.ssx : SSX, NSE = RT ? d.sse : SSE;
  const int PER = NSX * NSE;
  __shared__ NodeRec<SSX, SSE> rec[NPW];
  __shared__ int bad;
  const int tid = threadIdx.x;
  const long long first = static_cast<long long>(blockIdx.x) * NPW;
  if (tid == 0) bad = 0;
  __syncthreads();
  // node metadata, one item per thread: per node the status and kind, each
  // xi-stencil slot s, each eta-stencil slot t (with its rank among the
  // node's eta neighbours), and the far-field P, Ri and incoming entries.
  // (One thread per node doing all of it measured 163 us for the writer at
  // 200x100 against 151.5 us spread like this.)
  {
    const int W = 1 + NSX + NSE + 25 + 5;
    for (int k = tid; k < NPW * W; k += 64 * NPW) {
      const int n = k / W, q = k - W * (k / W);
      const long long node = first + n;
      if (node >= d.nodes) continue;
      NodeRec<SSX, SSE> &R = rec[n];
      const int i = static_cast<int>(node % d.ny), j = static_cast<int>(node / d.ny);
      const int kind = i == 0 ? 1 : (i == d.ny - 1 ? 2 : 0);
      if (q == 0) {
        R.status = status[node];
        if (R.status != 0) atomicOr(&bad, 1);
        R.kind = kind;
      } else if (q <= NSX) {
        const int sx = q - 1;
        const double *w1x = xw + static_cast<size_t>(d.ny) * NSX + static_cast<size_t>(i) * NSX;
        const double *w2x = w1x + static_cast<size_t>(d.ny) * NSX;
        R.w1x[sx] = w1x[sx];
        R.mw2x[sx] = -w2x[sx];
        if (xoff[i * NSX + sx] == 0) R.s0 = sx;
      } else if (q <= NSX + NSE) {
        const int t = q - 1 - NSX;
        const double *w1e = ew + static_cast<size_t>(d.nz) * NSE + static_cast<size_t>(j) * NSE;
        const double *w2e = w1e + static_cast<size_t>(d.nz) * NSE;
        R.w1e[t] = w1e[t];
        R.mw2e[t] = -w2e[t];
        const int off = eoff[j * NSE + t];
        if (off == 0) R.t0 = t;
        const int ct = wrap_mod(j + off, d.nz);
        int rank = 0;
        for (int u = 0; u < NSE; ++u) rank += wrap_mod(j + eoff[j * NSE + u], d.nz) < ct;
        R.slot_of_rank[rank] = t;
      } else if (kind == 2 && constrained) {
        const double *bc = bcdata + 106 * static_cast<size_t>(j);
        if (q <= NSX + NSE + 25) {
          const int e = q - 1 - NSX - NSE;
          R.P[e] = dcx{bc[2 * e], bc[2 * e + 1]};
          R.Ri[e] = dcx{bc[50 + 2 * e], bc[50 + 2 * e + 1]};
        } else {
          const int r = q - 1 - NSX - NSE - 25;
          R.incoming[r] = bc[100 + r] != 0.0;
        }
      }
    }
  }
  // coefficient records: NPW consecutive nodes are contiguous in `coef`
  {
    const long long nn = min(static_cast<long long>(NPW), d.nodes - first);
    const double *src = coef + kCoefDoubles * first;
    for (int k = tid; k < nn * kCoefDoubles; k += 64 * NPW) {
      const int n = k / kCoefDoubles, q = k - kCoefDoubles * n;
      NodeRec<SSX, SSE> &R = rec[n];
      const double v = src[k];
      if (q < 150) {
        dcx *dst = q < 50 ? R.K0 : (q < 100 ? R.Kxi : R.Keta);
        const int qq = q - 50 * (q / 50);
        if (qq & 1) dst[qq >> 1].im = v;
        else dst[qq >> 1].re = v;
      } else if (q < 175) {
        R.Vxx[q - 150] = v;
      } else if (q < 200) {
        R.Vxe[q - 175] = v;
      } else if (q < 225) {
        R.Vee[q - 200] = v;
      } else {
        R.gamma[q - 225] = v;
      }
    }
  }
  __syncthreads();
  if (bad) {
    if (tid == 0) atomicOr(&diag->flags, kCoefficient);
    return;
  }

  // The workgroup's block rows are one contiguous span of nrows * PER * 25
  // complex entries; with NPW rows per workgroup of PER * 400 bytes each the
  // span starts on a 128-byte line whenever NPW * PER * 25 is a multiple of 8
  // (NPW = 8).  Thread k of the workgroup writes entries k, k + 64 * NPW, ...
  // of the span, so each store instruction of a wave writes 64 x 16 B = 1 KB
  // of consecutive bytes (a wave per row, 50 lanes storing 800 B at a time
  // from registers, stored at 4.3 TB/s against 6.5 TB/s for a plain fill).
  // Each entry reads its row's coefficients from the record in LDS.
  {
    const int per_row = PER * 25;
    const long long nrows = min(static_cast<long long>(NPW), d.nodes - first);
    const int span = static_cast<int>(nrows) * per_row;
    double2v *out = reinterpret_cast<double2v *>(vals) + first * per_row;
    // L entry e = (r, c) of block (t, s) of a row, accumulation order K0,
    // xi, eta, mixed (GlobalOperator.cpp:323-365)
    auto l_entry = [&](const NodeRec<SSX, SSE> &R, int t, int s, int e, double &vr,
                       double &vi) {
      vr = 0.0;
      vi = 0.0;
      if (t == R.t0) {
        if (s == R.s0) {
          vr = vr + R.K0[e].re;
          vi = vi + R.K0[e].im;
        }
        const double wv = R.w1x[s];
        double ar = R.Kxi[e].re * wv;
        const double ai = R.Kxi[e].im * wv;
        ar = ar + R.Vxx[e] * R.mw2x[s];
        vr = vr + ar;
        vi = vi + ai;
      }
      if (s == R.s0) {
        const double wv = R.w1e[t];
        double ar = R.Keta[e].re * wv;
        const double ai = R.Keta[e].im * wv;
        ar = ar + R.Vee[e] * R.mw2e[t];
        vr = vr + ar;
        vi = vi + ai;
      }
      vr = vr + R.Vxe[e] * (-R.w1e[t] * R.w1x[s]);
    };
    for (int k = tid; k < span; k += 64 * NPW) {
      const int w = k / per_row, rem = k - per_row * w;
      const int b = rem / 25, e = rem - 25 * b;
      const int r = e / 5, c = e - 5 * r;
      const NodeRec<SSX, SSE> &R = rec[w];
      const int tr = b / NSX, s = b - NSX * tr;
      const int t = R.slot_of_rank[tr];
      double vr, vi;
      l_entry(R, t, s, e, vr, vi);
      double2v val;
      if (!constrained) {
        val = double2v{vr, vi};
      } else {
        const double ar = vr * -1.0, ai = vi * -1.0;  // A_bc = -L
        const bool diagblk = t == R.t0 && s == R.s0;
        if (R.kind == 1 && r >= 1) {
          // isothermal no-slip wall: u, v, w, T rows become identity rows
          val = double2v{(diagblk && r == c) ? 1.0 : 0.0, 0.0};
        } else if (R.kind == 2) {
          // far field: row r = P[r,:] (rows of -L) unless mode r is incoming
          if (R.incoming[r]) {
            const dcx ri = R.Ri[r * 5 + c];
            val = diagblk ? double2v{ri.re, ri.im} : double2v{0.0, 0.0};
          } else {
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < 5; ++q) {
              const dcx pk = R.P[r * 5 + q];
              if (pk.re == 0.0 && pk.im == 0.0) continue;
              double xr, xi;
              l_entry(R, t, s, q * 5 + c, xr, xi);
              xr = xr * -1.0;
              xi = xi * -1.0;
              sr += pk.re * xr - pk.im * xi;
              si += pk.re * xi + pk.im * xr;
            }
            val = double2v{sr, si};
          }
        } else {
          val = double2v{ar, ai};
        }
      }
      // (a run-time choice between the two stores is merged by the
      // compiler into one cached store: the choice is a template argument)
      if constexpr (BYPASS)
        __builtin_nontemporal_store(val, &out[k]);
      else
        out[k] = val;
    }
  }

  // the block-diagonal M_Gamma / B_bc
  if (tid < 25 * NPW) {
    const int n = tid / 25, e = tid - 25 * (tid / 25);
    const long long node = first + n;
    if (node < d.nodes) {
      const NodeRec<SSX, SSE> &R = rec[n];
      const int r = e / 5, c = e - 5 * (e / 5);
      dcx val{R.gamma[e], 0.0};
      if (constrained) {
        if (R.kind == 1 && r >= 1) {
          val = dcx{0.0, 0.0};
        } else if (R.kind == 2) {
          if (R.incoming[r]) {
            val = dcx{0.0, 0.0};
          } else {
            double sr = 0.0, si = 0.0;
            for (int k = 0; k < 5; ++k) {
              const dcx pk = R.P[r * 5 + k];
              if (pk.re == 0.0 && pk.im == 0.0) continue;
              const double br = R.gamma[k * 5 + c], bi = 0.0;
              sr += pk.re * br - pk.im * bi;
              si += pk.re * bi + pk.im * br;
            }
            val = dcx{sr, si};
          }
        }
      }
      reinterpret_cast<dcx *>(mvals)[node * 25 + e] = val;
    }
  }
}

// Generic-size variant dispatch table (RT: run-time stencil sizes up to
// SSX x SSE).
template <int SSX, int SSE, bool RT = false>
hipError_t launch_op(tailor_ctx *ctx, bool constrained, hipStream_t st) {
  constexpr int NPW = kOperatorRows;
  const CaseDims &d = ctx->dims;
  const unsigned grid = static_cast<unsigned>((d.nodes + NPW - 1) / NPW);
  // Operators larger than the 256 MB last-level cache bypass it: cached
  // stores only evict what the next launches read (200x100, 616 MB: writer
  // 136.8 -> 130.9 us, whole assembly 265 -> 253 us), while a 100x50 operator
  // (154 MB) stays faster cached (36.4 against 41.6 us).
  const bool bypass = static_cast<size_t>(d.nnzb) * 400 > (static_cast<size_t>(256) << 20);
  auto kernel = bypass ? k_operator<SSX, SSE, NPW, RT, true> : k_operator<SSX, SSE, NPW, RT, false>;
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(64 * NPW), 0, st, d,
                     ctx->d_xoff, ctx->d_xw, ctx->d_eoff, ctx->d_ew, ctx->d_coef,
                     ctx->d_coef_status, ctx->d_bcdata, constrained ? 1 : 0, ctx->d_vals,
                     ctx->d_mvals, ctx->d_diag);
  return hipGetLastError();
}

// Per-component sums of squares: rows x ncomp doubles with row stride
// `stride`; each block writes ncomp partial sums.
__global__ __launch_bounds__(kThreads) void k_sumsq(const double *__restrict__ data, long long rows,
                                                    int ncomp, long long stride,
                                                    double *__restrict__ partials) {
  __shared__ double red[kThreads];
  for (int c = 0; c < ncomp; ++c) {
    double acc = 0.0;
    for (long long r = static_cast<long long>(blockIdx.x) * kThreads + threadIdx.x; r < rows;
         r += static_cast<long long>(gridDim.x) * kThreads) {
      const double x = data[r * stride + c];
      acc += x * x;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
      if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) partials[static_cast<size_t>(blockIdx.x) * ncomp + c] = red[0];
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_case_scalars(tailor_ctx *ctx, hipStream_t s) {
  hipLaunchKernelGGL(k_case_scalars, dim3(1), dim3(1), 0, s, ctx->d_recipe, ctx->alpha_re,
                     ctx->alpha_im, ctx->period);
  return hipGetLastError();
}

hipError_t launch_reset_diag(tailor_ctx *ctx, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_diag, dim3(1), dim3(1), 0, s, ctx->d_diag);
  return hipGetLastError();
}

hipError_t launch_metrics(tailor_ctx *ctx, hipStream_t s) {
  const CaseDims &d = ctx->dims;
  const unsigned grid = static_cast<unsigned>((d.nodes + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_metrics_first, dim3(grid), dim3(kThreads), 0, s, d, ctx->d_xoff,
                     ctx->d_xw, ctx->d_eoff, ctx->d_ew, ctx->d_recipe, ctx->d_grid, ctx->d_metrics,
                     ctx->d_diag);
  return hipGetLastError();
}

hipError_t launch_derivatives(tailor_ctx *ctx, hipStream_t s) {
  const CaseDims &d = ctx->dims;
  const dim3 grid((d.ny + DTX - 1) / DTX, (d.nz + DTY - 1) / DTY);
  const size_t lds = sizeof(double) * 9 * static_cast<size_t>(DTX + 2 * d.hx) * (DTY + 2 * d.hy);
  hipLaunchKernelGGL(k_metrics2_derivatives, grid, dim3(DTHREADS), lds, s, d,
                     ctx->d_xoff, ctx->d_xw, ctx->d_eoff, ctx->d_ew, ctx->d_base, ctx->d_metrics,
                     ctx->d_deriv, ctx->d_diag);
  return hipGetLastError();
}

hipError_t launch_characteristics(tailor_ctx *ctx, hipStream_t s) {
  const CaseDims &d = ctx->dims;
  const unsigned g = static_cast<unsigned>((d.nz + 63) / 64);
  hipLaunchKernelGGL(k_characteristics, dim3(g), dim3(64), 0, s, d, ctx->d_recipe,
                     ctx->d_base, ctx->d_deriv, ctx->d_metrics, ctx->d_bcdata, ctx->d_diag);
  return hipGetLastError();
}

hipError_t launch_coefficients(tailor_ctx *ctx, hipStream_t s) {
  const CaseDims &d = ctx->dims;
  const unsigned grid = static_cast<unsigned>((d.nodes + kCoefNodes - 1) / kCoefNodes);
  hipLaunchKernelGGL(k_coefficients, dim3(grid), dim3(kThreads), 0, s, d, ctx->d_recipe,
                     ctx->d_base, ctx->d_deriv, ctx->d_metrics, ctx->d_coef, ctx->d_coef_status);
  return hipGetLastError();
}

hipError_t launch_operator(tailor_ctx *ctx, bool constrained, hipStream_t s) {
  const CaseDims &d = ctx->dims;
  if (d.ssx == 11 && d.sse == 7) return launch_op<11, 7>(ctx, constrained, s);
  if (d.ssx == 5 && d.sse == 3) return launch_op<5, 3>(ctx, constrained, s);
  if (d.ssx == 5 && d.sse == 5) return launch_op<5, 5>(ctx, constrained, s);
  if (d.ssx == 7 && d.sse == 5) return launch_op<7, 5>(ctx, constrained, s);
  if (d.ssx == 9 && d.sse == 7) return launch_op<9, 7>(ctx, constrained, s);
  if (d.ssx == 11 && d.sse == 5) return launch_op<11, 5>(ctx, constrained, s);
  if (d.ssx == 7 && d.sse == 7) return launch_op<7, 7>(ctx, constrained, s);
  // any other stencil pair within the run-time instance's bounds (the loader
  // rejects wider ones, and the derivative tile's LDS bound is tighter)
  if (d.ssx <= kOperatorMaxSsx && d.sse <= kOperatorMaxSse)
    return launch_op<kOperatorMaxSsx, kOperatorMaxSse, true>(ctx, constrained, s);
  return hipErrorInvalidValue;
}

hipError_t launch_sumsq(tailor_ctx *ctx, const double *data, long long rows, int ncomp,
                        long long stride, double *host_out) {
  const int blocks = kPartialBlocks;
  if (ncomp > kPartialComps) return hipErrorInvalidValue;
  double *partials = ctx->d_partials;
  hipLaunchKernelGGL(k_sumsq, dim3(blocks), dim3(kThreads), 0, ctx->stream, data, rows, ncomp,
                     stride, partials);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  std::vector<double> h(static_cast<size_t>(blocks) * ncomp);
  e = hipMemcpyAsync(h.data(), partials, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                     ctx->stream);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return e;
  for (int c = 0; c < ncomp; ++c) {
    double s = 0.0;
    for (int b = 0; b < blocks; ++b) s += h[static_cast<size_t>(b) * ncomp + c];
    host_out[c] = s;
  }
  return hipSuccess;
}

}  // namespace tailor
