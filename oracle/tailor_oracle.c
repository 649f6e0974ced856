/*
 * tailor_oracle.c -- CPU FP64 restatement of the reference's assemble path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the GPU path is
 * compared with.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links or calls it.
 *
 * Every function restates one piece of the reference (Bzhouha/Tailor,
 * src/CPlusPlus) in plain C, in the same floating-point operation order, so
 * that the GPU kernels (which follow the same order, compiled without FP
 * contraction) can be compared entry by entry:
 *
 *   tor_metrics           Metrics.cpp:56-225 (first order), :227-352 (second)
 *   tor_derivatives       BaseFlowDerivatives.cpp:51-251
 *   tor_lns_physical      LNSCoefficients.cpp:173-404 (+ :112-116 Sutherland)
 *   tor_lns_inviscid      LNSCoefficients.cpp:118-171
 *   tor_fourier           StreamwiseFourier.cpp:16-44
 *   tor_curvilinear       CurvilinearTransform.cpp:62-101
 *   tor_assemble          GlobalOperator.cpp:209-378 (BSR layout, values)
 *   tor_characteristics   BoundaryConditions.cpp:92-140, :152-280, with a
 *                         closed-form eigenbasis in place of LAPACK zggev
 *   tor_apply_bc          BoundaryConditions.cpp:354-548
 *
 * Complex numbers are (re, im) pairs of doubles.  Products of a complex by a
 * real-valued PetscScalar reduce to component-wise real products, as they do
 * for finite operands in the reference.
 *
 * Parity status per function (how each is pinned) is listed in the oracle
 * tests: Fortran gold (LNS, transforms), matrix-free action (assembly),
 * boundary-row invariants + LAPACK eigen-decomposition (characteristics),
 * closed-form metrics of analytic mappings (metrics, derivatives).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>


typedef struct {
  int n;          /* node count */
  int ss;         /* stencil size q+1 */
  const long *idx;   /* (n, ss) */
  const long *off;   /* (n, ss) */
  const double *w0, *w1, *w2;  /* (n, ss) */
} orule;


static const double kEps = 2.220446049250313e-16;

/* ---------------------------------------------------------------- metrics */

static long periodic_wrap(long index, long n) {
  if (index >= 0) return index / n;
  return -((-index + n - 1) / n);
}

static long wrap_index(long index, long n) {
  long r = index % n;
  return r < 0 ? r + n : r;
}

/* grid: (N,3) complex interleaved -> metrics (N,10).  diag[0..2] = min J,
 * max J, max |Im grid|.  Returns 0 or an error code:
 * 2 non-finite, 3 non-positive Jacobian, 4 complex grid. */
int tor_metrics(int ny, int nz, const orule *xr, const orule *er,
                double period, const double *grid, double *metrics,
                double *diag) {
  const long n = (long)ny * nz;
  int fail_nf = 0, fail_jac = 0;
  double minJ = 1.79769313486231570e308, maxJ = -1.79769313486231570e308;
  double maxIm = 0.0;
  for (long j = 0; j < nz; ++j) {
    for (long i = 0; i < ny; ++i) {
      double yXi = 0.0, yEta = 0.0, zXi = 0.0, zEta = 0.0;
      for (int s = 0; s < xr->ss; ++s) {
        long col = i + xr->off[i * xr->ss + s];
        const double *g = grid + 6 * (j * ny + col);
        double yr = g[2], yi = g[3], zr = g[4], zi = g[5];
        if (!isfinite(yr) || !isfinite(yi) || !isfinite(zr) || !isfinite(zi)) fail_nf = 1;
        if (fabs(yi) > maxIm) maxIm = fabs(yi);
        if (fabs(zi) > maxIm) maxIm = fabs(zi);
        double w = xr->w1[i * xr->ss + s];
        yXi += w * yr;
        zXi += w * zr;
      }
      for (int t = 0; t < er->ss; ++t) {
        long row = j + er->off[j * er->ss + t];
        long jw = wrap_index(row, nz);
        const double *g = grid + 6 * (jw * ny + i);
        double yr = g[2], yi = g[3];
        double zr = g[4] + (double)periodic_wrap(row, nz) * period, zi = g[5];
        if (!isfinite(yr) || !isfinite(yi) || !isfinite(zr) || !isfinite(zi)) fail_nf = 1;
        if (fabs(yi) > maxIm) maxIm = fabs(yi);
        if (fabs(zi) > maxIm) maxIm = fabs(zi);
        double w = er->w1[j * er->ss + t];
        yEta += w * yr;
        zEta += w * zr;
      }
      double *m = metrics + 10 * (j * ny + i);
      if (!isfinite(yXi) || !isfinite(yEta) || !isfinite(zXi) || !isfinite(zEta)) {
        fail_nf = 1;
        continue;
      }
      double p1 = yXi * zEta, p2 = yEta * zXi;
      double J = p1 - p2;
      double scale = 1.0;
      if (fabs(p1) > scale) scale = fabs(p1);
      if (fabs(p2) > scale) scale = fabs(p2);
      double tolJ = 100.0 * kEps * scale;
      if (!isfinite(J)) { fail_nf = 1; continue; }
      if (J < minJ) minJ = J;
      if (J > maxJ) maxJ = J;
      if (J <= tolJ) { fail_jac = 1; continue; }
      m[0] = zEta / J;
      m[1] = -yEta / J;
      m[2] = -zXi / J;
      m[3] = yXi / J;
      for (int c = 0; c < 4; ++c)
        if (!isfinite(m[c])) fail_nf = 1;
    }
  }
  diag[0] = minJ; diag[1] = maxJ; diag[2] = maxIm;
  if (fail_nf) return 2;
  if (fail_jac) return 3;
  if (maxIm > 100.0 * kEps) return 4;
  /* second order: chain rule on the synchronized first-order metrics.
   * Values go to a scratch array first so that every node differentiates
   * first-order values only. */
  {
    double *tmp = (double *)malloc(sizeof(double) * 6 * n);
    if (!tmp) return 9;
    for (long j = 0; j < nz; ++j) {
      for (long i = 0; i < ny; ++i) {
        double dXi[4] = {0, 0, 0, 0}, dEta[4] = {0, 0, 0, 0};
        for (int s = 0; s < xr->ss; ++s) {
          long col = i + xr->off[i * xr->ss + s];
          double w = xr->w1[i * xr->ss + s];
          const double *mm = metrics + 10 * (j * ny + col);
          for (int c = 0; c < 4; ++c) dXi[c] += w * mm[c];
        }
        for (int t = 0; t < er->ss; ++t) {
          long jw = wrap_index(j + er->off[j * er->ss + t], nz);
          double w = er->w1[j * er->ss + t];
          const double *mm = metrics + 10 * (jw * ny + i);
          for (int c = 0; c < 4; ++c) dEta[c] += w * mm[c];
        }
        const double *m = metrics + 10 * (j * ny + i);
        double xiY = m[0], xiZ = m[1], etaY = m[2], etaZ = m[3];
        double *v = tmp + 6 * (j * ny + i);
        v[0] = xiY * dXi[0] + etaY * dEta[0];
        v[1] = xiZ * dXi[1] + etaZ * dEta[1];
        v[2] = xiZ * dXi[0] + etaZ * dEta[0];
        v[3] = xiY * dXi[2] + etaY * dEta[2];
        v[4] = xiZ * dXi[3] + etaZ * dEta[3];
        v[5] = xiZ * dXi[2] + etaZ * dEta[2];
        for (int c = 0; c < 6; ++c)
          if (!isfinite(v[c])) fail_nf = 1;
      }
    }
    for (long k = 0; k < n; ++k)
      for (int c = 0; c < 6; ++c) metrics[10 * k + 4 + c] = tmp[6 * k + c];
    free(tmp);
  }
  return fail_nf ? 2 : 0;
}

/* ------------------------------------------------------ base-flow derivs */

/* baseflow (N,5) complex -> deriv (N,25) derivative-major [Dy,Dz,Dyy,Dzz,Dyz]
 * x [rho,u,v,w,T].  diag[0] = max |Im baseflow|.  Returns 0, 2 (non-finite
 * values), 3 (rho or T <= 0), 5 (non-finite derivative), 4 (complex). */
int tor_derivatives(int ny, int nz, const orule *xr, const orule *er,
                    const double *baseflow, const double *metrics,
                    double *deriv, double *diag) {
  int f_nf = 0, f_pos = 0, f_dnf = 0;
  double maxIm = 0.0;
  for (long j = 0; j < nz; ++j) {
    for (long i = 0; i < ny; ++i) {
      const double *b0 = baseflow + 10 * (j * ny + i);
      double rho_r = b0[0], rho_i = b0[1], T_r = b0[8], T_i = b0[9];
      if (!isfinite(rho_r) || !isfinite(rho_i) || !isfinite(T_r) || !isfinite(T_i))
        f_nf = 1;
      else if (rho_r <= 0.0 || T_r <= 0.0)
        f_pos = 1;
      const double *m = metrics + 10 * (j * ny + i);
      for (int f = 0; f < 5; ++f) {
        double vXi = 0, vEta = 0, vXiXi = 0, vEtaEta = 0, vXiEta = 0;
        for (int s = 0; s < xr->ss; ++s) {
          long col = i + xr->off[i * xr->ss + s];
          const double *b = baseflow + 10 * (j * ny + col) + 2 * f;
          if (!isfinite(b[0]) || !isfinite(b[1])) f_nf = 1;
          if (fabs(b[1]) > maxIm) maxIm = fabs(b[1]);
          vXi += xr->w1[i * xr->ss + s] * b[0];
          vXiXi += xr->w2[i * xr->ss + s] * b[0];
        }
        for (int t = 0; t < er->ss; ++t) {
          long jw = wrap_index(j + er->off[j * er->ss + t], nz);
          const double *b = baseflow + 10 * (jw * ny + i) + 2 * f;
          if (!isfinite(b[0]) || !isfinite(b[1])) f_nf = 1;
          if (fabs(b[1]) > maxIm) maxIm = fabs(b[1]);
          vEta += er->w1[j * er->ss + t] * b[0];
          vEtaEta += er->w2[j * er->ss + t] * b[0];
        }
        for (int t = 0; t < er->ss; ++t) {
          long jw = wrap_index(j + er->off[j * er->ss + t], nz);
          double we = er->w1[j * er->ss + t];
          for (int s = 0; s < xr->ss; ++s) {
            long col = i + xr->off[i * xr->ss + s];
            const double *b = baseflow + 10 * (jw * ny + col) + 2 * f;
            if (!isfinite(b[0]) || !isfinite(b[1])) f_nf = 1;
            if (fabs(b[1]) > maxIm) maxIm = fabs(b[1]);
            vXiEta += we * xr->w1[i * xr->ss + s] * b[0];
          }
        }
        double xiY = m[0], xiZ = m[1], etaY = m[2], etaZ = m[3];
        double xiYY = m[4], xiZZ = m[5], xiYZ = m[6];
        double etaYY = m[7], etaZZ = m[8], etaYZ = m[9];
        double v[5];
        v[0] = xiY * vXi + etaY * vEta;
        v[1] = xiZ * vXi + etaZ * vEta;
        v[2] = xiY * xiY * vXiXi + 2.0 * xiY * etaY * vXiEta +
               etaY * etaY * vEtaEta + xiYY * vXi + etaYY * vEta;
        v[3] = xiZ * xiZ * vXiXi + 2.0 * xiZ * etaZ * vXiEta +
               etaZ * etaZ * vEtaEta + xiZZ * vXi + etaZZ * vEta;
        v[4] = xiY * xiZ * vXiXi + (xiY * etaZ + etaY * xiZ) * vXiEta +
               etaY * etaZ * vEtaEta + xiYZ * vXi + etaYZ * vEta;
        int bad = 0;
        for (int d = 0; d < 5; ++d)
          if (!isfinite(v[d])) bad = 1;
        if (bad) { f_dnf = 1; continue; }
        for (int d = 0; d < 5; ++d) deriv[25 * (j * ny + i) + d * 5 + f] = v[d];
      }
    }
  }
  diag[0] = maxIm;
  if (f_nf) return 2;
  if (f_pos) return 3;
  if (f_dnf) return 5;
  if (maxIm > 100.0 * kEps) return 4;
  return 0;
}

/* -------------------------------------------------------- LNS physics */

/* recipe: [Re, Ma, Pr, gamma, Tref, S].
 * pt: 10 tuples of 5 = value, dx, dy, dz, dxx, dyy, dzz, dxy, dxz, dyz.
 * out: 11 blocks (Gamma, A, B, C, D, Vxx, Vxy, Vxz, Vyy, Vyz, Vzz), 5x5
 * row-major real.  Returns 0, 1 non-finite input, 2 rho<=0, 3 T<=0,
 * 4 non-finite output (CoefficientStatus order). */
int tor_lns_physical(const double *recipe, const double *pt, double *out) {
  for (int k = 0; k < 50; ++k)
    if (!isfinite(pt[k])) return 1;
  const double Re = recipe[0], Ma = recipe[1], Pr = recipe[2], gam = recipe[3];
  const double s = recipe[5] / recipe[4];
  const double oneThird = 1.0 / 3.0, twoThirds = 2.0 / 3.0, fourThirds = 4.0 / 3.0;
  const double *val = pt, *dx = pt + 5, *dy = pt + 10, *dz = pt + 15;
  const double *dxx = pt + 20, *dyy = pt + 25, *dzz = pt + 30, *dxy = pt + 35,
               *dxz = pt + 40, *dyz = pt + 45;
  const double rho = val[0], U = val[1], V = val[2], W = val[3], T = val[4];
  if (rho <= 0.0) return 2;
  if (T <= 0.0) return 3;
  const double rhox = dx[0], Ux = dx[1], Vx = dx[2], Wx = dx[3], Tx = dx[4];
  const double rhoy = dy[0], Uy = dy[1], Vy = dy[2], Wy = dy[3], Ty = dy[4];
  const double rhoz = dz[0], Uz = dz[1], Vz = dz[2], Wz = dz[3], Tz = dz[4];
  const double Uxx = dxx[1], Vxx = dxx[2], Wxx = dxx[3], Txx = dxx[4];
  const double Uyy = dyy[1], Vyy = dyy[2], Wyy = dyy[3], Tyy = dyy[4];
  const double Uzz = dzz[1], Vzz = dzz[2], Wzz = dzz[3], Tzz = dzz[4];
  const double Uxy = dxy[1], Vxy = dxy[2];
  const double Uxz = dxz[1], Wxz = dxz[3];
  const double Vyz = dyz[2], Wyz = dyz[3];
  (void)Vxy;

  const double Pe = 1.0 / (gam * Ma * Ma);
  const double g2 = 1.0 / ((gam - 1.0) * Ma * Ma);
  const double mu = T * sqrt(T) * (1.0 + s) / (T + s);
  const double muT = mu * (1.5 / T - 1.0 / (T + s));
  const double muTT = muT * (1.5 / T - 1.0 / (T + s)) -
                      mu * (1.5 / (T * T) - 1.0 / ((T + s) * (T + s)));
  const double mux = muT * Tx, muy = muT * Ty, muz = muT * Tz;

  memset(out, 0, sizeof(double) * 11 * 25);
  double *G = out, *A = out + 25, *B = out + 50, *C = out + 75, *D = out + 100;
  double *MVxx = out + 125, *MVxy = out + 150, *MVxz = out + 175;
  double *MVyy = out + 200, *MVyz = out + 225, *MVzz = out + 250;
#define E(M, r, c) M[(r) * 5 + (c)]
  E(G, 0, 0) = 1.0;
  E(G, 1, 1) = rho;
  E(G, 2, 2) = rho;
  E(G, 3, 3) = rho;
  E(G, 4, 0) = -Pe * T;
  E(G, 4, 4) = rho * g2 - Pe * rho;

  E(A, 0, 0) = U;
  E(A, 0, 1) = rho;
  E(A, 1, 0) = Pe * T;
  E(A, 1, 1) = rho * U - fourThirds * mux / Re;
  E(A, 1, 2) = -muy / Re;
  E(A, 1, 3) = -muz / Re;
  E(A, 1, 4) = Pe * rho - muT / Re * (fourThirds * Ux - twoThirds * Vy - twoThirds * Wz);
  E(A, 2, 1) = twoThirds * muy / Re;
  E(A, 2, 2) = rho * U - mux / Re;
  E(A, 2, 4) = -muT * (Uy + Vx) / Re;
  E(A, 3, 1) = twoThirds * muz / Re;
  E(A, 3, 3) = rho * U - mux / Re;
  E(A, 3, 4) = -muT * (Wx + Uz) / Re;
  E(A, 4, 0) = -Pe * U * T;
  E(A, 4, 1) = -2.0 * mu * (fourThirds * Ux - twoThirds * Vy - twoThirds * Wz) / Re;
  E(A, 4, 2) = -2.0 * mu * (Uy + Vx) / Re;
  E(A, 4, 3) = -2.0 * mu * (Wx + Uz) / Re;
  E(A, 4, 4) = rho * U * g2 - rho * U * Pe - 2.0 * Tx * muT * g2 / (Re * Pr);

  E(B, 0, 0) = V;
  E(B, 0, 2) = rho;
  E(B, 1, 1) = rho * V - muy / Re;
  E(B, 1, 2) = twoThirds * mux / Re;
  E(B, 1, 4) = -muT * (Uy + Vx) / Re;
  E(B, 2, 0) = Pe * T;
  E(B, 2, 1) = -mux / Re;
  E(B, 2, 2) = rho * V - fourThirds * muy / Re;
  E(B, 2, 3) = -muz / Re;
  E(B, 2, 4) = Pe * rho - muT / Re * (-twoThirds * Ux + fourThirds * Vy - twoThirds * Wz);
  E(B, 3, 2) = twoThirds * muz / Re;
  E(B, 3, 3) = rho * V - muy / Re;
  E(B, 3, 4) = -muT * (Vz + Wy) / Re;
  E(B, 4, 0) = -Pe * V * T;
  E(B, 4, 1) = -2.0 * mu * (Uy + Vx) / Re;
  E(B, 4, 2) = -2.0 * mu * (-twoThirds * Ux + fourThirds * Vy - twoThirds * Wz) / Re;
  E(B, 4, 3) = -2.0 * mu * (Vz + Wy) / Re;
  E(B, 4, 4) = rho * V * g2 - rho * V * Pe - 2.0 * Ty * muT * g2 / (Re * Pr);

  E(C, 0, 0) = W;
  E(C, 0, 3) = rho;
  E(C, 1, 1) = rho * W - muz / Re;
  E(C, 1, 3) = twoThirds * mux / Re;
  E(C, 1, 4) = -muT * (Wx + Uz) / Re;
  E(C, 2, 2) = rho * W - muz / Re;
  E(C, 2, 3) = twoThirds * muy / Re;
  E(C, 2, 4) = -muT * (Vz + Wy) / Re;
  E(C, 3, 0) = Pe * T;
  E(C, 3, 1) = -mux / Re;
  E(C, 3, 2) = -muy / Re;
  E(C, 3, 3) = rho * W - fourThirds * muz / Re;
  E(C, 3, 4) = Pe * rho - muT / Re * (-twoThirds * Ux - twoThirds * Vy + fourThirds * Wz);
  E(C, 4, 0) = -Pe * W * T;
  E(C, 4, 1) = -2.0 * mu * (Wx + Uz) / Re;
  E(C, 4, 2) = -2.0 * mu * (Vz + Wy) / Re;
  E(C, 4, 3) = -2.0 * mu * (-twoThirds * Ux - twoThirds * Vy + fourThirds * Wz) / Re;
  E(C, 4, 4) = rho * W * g2 - rho * W * Pe - 2.0 * Tz * muT * g2 / (Re * Pr);

  E(D, 0, 0) = Ux + Vy + Wz;
  E(D, 0, 1) = rhox;
  E(D, 0, 2) = rhoy;
  E(D, 0, 3) = rhoz;
  E(D, 1, 0) = U * Ux + V * Uy + W * Uz + Pe * Tx;
  E(D, 1, 1) = rho * Ux;
  E(D, 1, 2) = rho * Uy;
  E(D, 1, 3) = rho * Uz;
  E(D, 1, 4) = Pe * rhox -
               (muTT * Tx * twoThirds * (2.0 * Ux - Vy - Wz) + muTT * Ty * (Uy + Vx) +
                muTT * Tz * (Wx + Uz) +
                muT * (fourThirds * Uxx + Uyy + Uzz + oneThird * dxy[2] + oneThird * Wxz)) /
                   Re;
  E(D, 2, 0) = U * Vx + V * Vy + W * Vz + Pe * Ty;
  E(D, 2, 1) = rho * Vx;
  E(D, 2, 2) = rho * Vy;
  E(D, 2, 3) = rho * Vz;
  E(D, 2, 4) = Pe * rhoy -
               (muTT * Tx * (Uy + Vx) + muTT * Ty * twoThirds * (-Ux + 2.0 * Vy - Wz) +
                muTT * Tz * (Vz + Wy) +
                muT * (Vxx + fourThirds * Vyy + Vzz + oneThird * Uxy + oneThird * Wyz)) /
                   Re;
  E(D, 3, 0) = U * Wx + V * Wy + W * Wz + Pe * Tz;
  E(D, 3, 1) = rho * Wx;
  E(D, 3, 2) = rho * Wy;
  E(D, 3, 3) = rho * Wz;
  E(D, 3, 4) = Pe * rhoz -
               (muTT * Tx * (Wx + Uz) + muTT * Ty * (Vz + Wy) +
                muTT * Tz * twoThirds * (-Ux - Vy + 2.0 * Wz) +
                muT * (Wxx + Wyy + fourThirds * Wzz + oneThird * Uxz + oneThird * Vyz)) /
                   Re;
  E(D, 4, 0) = (g2 - Pe) * (U * Tx + V * Ty + W * Tz);
  E(D, 4, 1) = rho * Tx * g2 - Pe * (rho * Tx + T * rhox);
  E(D, 4, 2) = rho * Ty * g2 - Pe * (rho * Ty + T * rhoy);
  E(D, 4, 3) = rho * Tz * g2 - Pe * (rho * Tz + T * rhoz);
  E(D, 4, 4) = -Pe * (U * rhox + V * rhoy + W * rhoz) -
               (Txx + Tyy + Tzz) * muT * g2 / (Re * Pr) -
               (Tx * Tx + Ty * Ty + Tz * Tz) * muTT * g2 / (Re * Pr) -
               muT *
                   (Uy * Uy + Vx * Vx + 2.0 * Uy * Vx + Uz * Uz + Wx * Wx + 2.0 * Uz * Wx +
                    Vz * Vz + Wy * Wy + 2.0 * Vz * Wy) /
                   Re -
               muT * fourThirds * (Ux * Ux + Vy * Vy + Wz * Wz - Ux * Wz - Ux * Vy - Vy * Wz) /
                   Re;

  E(MVxx, 1, 1) = fourThirds * mu / Re;
  E(MVxx, 2, 2) = mu / Re;
  E(MVxx, 3, 3) = mu / Re;
  E(MVxx, 4, 4) = mu * g2 / (Re * Pr);
  E(MVyy, 1, 1) = mu / Re;
  E(MVyy, 2, 2) = fourThirds * mu / Re;
  E(MVyy, 3, 3) = mu / Re;
  E(MVyy, 4, 4) = mu * g2 / (Re * Pr);
  E(MVzz, 1, 1) = mu / Re;
  E(MVzz, 2, 2) = mu / Re;
  E(MVzz, 3, 3) = fourThirds * mu / Re;
  E(MVzz, 4, 4) = mu * g2 / (Re * Pr);
  E(MVxy, 1, 2) = oneThird * mu / Re;
  E(MVxy, 2, 1) = oneThird * mu / Re;
  E(MVxz, 1, 3) = oneThird * mu / Re;
  E(MVxz, 3, 1) = oneThird * mu / Re;
  E(MVyz, 2, 3) = oneThird * mu / Re;
  E(MVyz, 3, 2) = oneThird * mu / Re;
#undef E
  for (int k = 0; k < 275; ++k)
    if (!isfinite(out[k])) return 4;
  return 0;
}

/* Gamma, Bc, Cc (LNSCoefficients.cpp:118-171).  val = [rho,U,V,W,T]. */
int tor_lns_inviscid(const double *recipe, const double *val, double *out) {
  for (int k = 0; k < 5; ++k)
    if (!isfinite(val[k])) return 1;
  const double rho = val[0], V = val[2], W = val[3], T = val[4];
  if (rho <= 0.0) return 2;
  if (T <= 0.0) return 3;
  const double gam = recipe[3], Ma = recipe[1];
  const double Pe = 1.0 / (gam * Ma * Ma);
  const double g2 = 1.0 / ((gam - 1.0) * Ma * Ma);
  memset(out, 0, sizeof(double) * 75);
  double *G = out, *Bc = out + 25, *Cc = out + 50;
#define E(M, r, c) M[(r) * 5 + (c)]
  E(G, 0, 0) = 1.0;
  E(G, 1, 1) = rho;
  E(G, 2, 2) = rho;
  E(G, 3, 3) = rho;
  E(G, 4, 0) = -Pe * T;
  E(G, 4, 4) = rho * g2 - Pe * rho;
  E(Bc, 0, 0) = V;
  E(Bc, 0, 2) = rho;
  E(Bc, 1, 1) = rho * V;
  E(Bc, 2, 0) = Pe * T;
  E(Bc, 2, 2) = rho * V;
  E(Bc, 2, 4) = Pe * rho;
  E(Bc, 3, 3) = rho * V;
  E(Bc, 4, 0) = -Pe * V * T;
  E(Bc, 4, 4) = rho * V * g2 - rho * V * Pe;
  E(Cc, 0, 0) = W;
  E(Cc, 0, 3) = rho;
  E(Cc, 1, 1) = rho * W;
  E(Cc, 2, 2) = rho * W;
  E(Cc, 3, 0) = Pe * T;
  E(Cc, 3, 3) = rho * W;
  E(Cc, 3, 4) = Pe * rho;
  E(Cc, 4, 0) = -Pe * W * T;
  E(Cc, 4, 4) = rho * W * g2 - rho * W * Pe;
#undef E
  for (int k = 0; k < 75; ++k)
    if (!isfinite(out[k])) return 4;
  return 0;
}

/* ------------------------------------------------------- transforms */

/* phys: 11 real blocks; out: 7 complex blocks (Gamma, K0, Ky, Kz, Vyy, Vyz,
 * Vzz), (re, im) interleaved, 25 entries each. */
int tor_fourier(const double *phys, double alpha_re, double alpha_im, double *out) {
  if (!isfinite(alpha_re) || !isfinite(alpha_im)) return 1;
  for (int k = 0; k < 275; ++k)
    if (!isfinite(phys[k])) return 1;
  /* i*alpha and alpha^2 as complex products (StreamwiseFourier.cpp:30-31) */
  const double ia_re = 0.0 * alpha_re - 1.0 * alpha_im;
  const double ia_im = 0.0 * alpha_im + 1.0 * alpha_re;
  const double a2_re = alpha_re * alpha_re - alpha_im * alpha_im;
  const double a2_im = alpha_re * alpha_im + alpha_im * alpha_re;
  const double *G = phys, *A = phys + 25, *B = phys + 50, *C = phys + 75, *D = phys + 100;
  const double *Vxx = phys + 125, *Vxy = phys + 150, *Vxz = phys + 175;
  const double *Vyy = phys + 200, *Vyz = phys + 225, *Vzz = phys + 250;
  for (int e = 0; e < 25; ++e) {
    double *o = out;
    o[2 * e] = G[e]; o[2 * e + 1] = 0.0;
    o = out + 50;
    /* D + (ia*A) + (a2*Vxx) with real A, Vxx */
    o[2 * e] = D[e] + ia_re * A[e] + a2_re * Vxx[e];
    o[2 * e + 1] = 0.0 + ia_im * A[e] + a2_im * Vxx[e];
    o = out + 100;
    o[2 * e] = B[e] - ia_re * Vxy[e];
    o[2 * e + 1] = 0.0 - ia_im * Vxy[e];
    o = out + 150;
    o[2 * e] = C[e] - ia_re * Vxz[e];
    o[2 * e + 1] = 0.0 - ia_im * Vxz[e];
    o = out + 200; o[2 * e] = Vyy[e]; o[2 * e + 1] = 0.0;
    o = out + 250; o[2 * e] = Vyz[e]; o[2 * e + 1] = 0.0;
    o = out + 300; o[2 * e] = Vzz[e]; o[2 * e + 1] = 0.0;
  }
  for (int k = 0; k < 350; ++k)
    if (!isfinite(out[k])) return 4;
  return 0;
}

/* fourier: 7 complex blocks; m: 10 metrics; out: 7 complex blocks (Gamma,
 * K0, Kxi, Keta, Vxixi, Vxieta, Vetaeta). */
int tor_curvilinear(const double *fourier, const double *m, double *out) {
  for (int k = 0; k < 350; ++k)
    if (!isfinite(fourier[k])) return 1;
  for (int k = 0; k < 10; ++k)
    if (!isfinite(m[k])) return 1;
  const double xiY = m[0], xiZ = m[1], etaY = m[2], etaZ = m[3];
  const double xiYY = m[4], xiZZ = m[5], xiYZ = m[6];
  const double etaYY = m[7], etaZZ = m[8], etaYZ = m[9];
  for (int e = 0; e < 25; ++e) {
    for (int p = 0; p < 2; ++p) {
      const int k = 2 * e + p;
      const double Gm = fourier[k], K0 = fourier[50 + k], Ky = fourier[100 + k],
                   Kz = fourier[150 + k], Vyy = fourier[200 + k], Vyz = fourier[250 + k],
                   Vzz = fourier[300 + k];
      out[k] = Gm;
      out[50 + k] = K0;
      out[100 + k] = xiY * Ky + xiZ * Kz - xiYY * Vyy - xiZZ * Vzz - xiYZ * Vyz;
      out[150 + k] = etaY * Ky + etaZ * Kz - etaYY * Vyy - etaZZ * Vzz - etaYZ * Vyz;
      out[200 + k] = xiY * xiY * Vyy + xiZ * xiZ * Vzz + xiY * xiZ * Vyz;
      out[250 + k] = 2.0 * xiY * etaY * Vyy + 2.0 * xiZ * etaZ * Vzz +
                     (xiY * etaZ + etaY * xiZ) * Vyz;
      out[300 + k] = etaY * etaY * Vyy + etaZ * etaZ * Vzz + etaY * etaZ * Vyz;
    }
  }
  for (int k = 0; k < 350; ++k)
    if (!isfinite(out[k])) return 4;
  return 0;
}

/* Base-flow point of one node: value + y/z derivatives, x-derivatives zero
 * (makeBaseFlowPoint, LNSCoefficients.cpp:75-110). */
static int node_point(const double *baseflow, const double *deriv, long node, double *pt) {
  memset(pt, 0, sizeof(double) * 50);
  const double *b = baseflow + 10 * node;
  const double *d = deriv + 25 * node;
  for (int f = 0; f < 5; ++f) {
    if (!isfinite(b[2 * f]) || !isfinite(b[2 * f + 1]) || b[2 * f + 1] != 0.0) return 1;
    pt[f] = b[2 * f];
    pt[10 + f] = d[0 * 5 + f];  /* dy */
    pt[15 + f] = d[1 * 5 + f];  /* dz */
    pt[25 + f] = d[2 * 5 + f];  /* dyy */
    pt[30 + f] = d[3 * 5 + f];  /* dzz */
    pt[45 + f] = d[4 * 5 + f];  /* dyz */
  }
  for (int k = 0; k < 50; ++k)
    if (!isfinite(pt[k])) return 1;
  if (pt[0] <= 0.0) return 2;
  if (pt[4] <= 0.0) return 3;
  return 0;
}

/* Full node chain: 7 curvilinear complex blocks. */
int tor_node_coefficients(const double *recipe, double alpha_re, double alpha_im,
                          const double *baseflow, const double *deriv,
                          const double *metrics, long node, double *curv) {
  double pt[50], phys[275], four[350];
  int st = node_point(baseflow, deriv, node, pt);
  if (!st) st = tor_lns_physical(recipe, pt, phys);
  if (!st) st = tor_fourier(phys, alpha_re, alpha_im, four);
  if (!st) st = tor_curvilinear(four, metrics + 10 * node, curv);
  return st;
}

/* ------------------------------------------------------------ assembly */

/* Rank of eta slot t of row j in ascending wrapped-column order. */
static void eta_sorted_rank(const orule *er, int nz, long j, int *rank) {
  int ss = er->ss;
  for (int t = 0; t < ss; ++t) {
    long ct = wrap_index(j + er->off[j * ss + t], nz);
    int r = 0;
    for (int u = 0; u < ss; ++u) {
      long cu = wrap_index(j + er->off[j * ss + u], nz);
      if (cu < ct) ++r;
    }
    rank[t] = r;
  }
}

/* BSR of L (row pointers, sorted block columns, 5x5 row-major complex
 * blocks) and the block-diagonal M.  rowptr has N+1 entries, colidx and
 * lvals nnzb = N*ssx*sse entries (25 complex each), mvals N blocks.
 * Returns 0 or the first failing node's CoefficientStatus. */
int tor_assemble(int ny, int nz, const orule *xr, const orule *er,
                 const double *recipe, double alpha_re, double alpha_im,
                 const double *baseflow, const double *deriv, const double *metrics,
                 int *rowptr, int *colidx, double *lvals, double *mvals) {
  const int ssx = xr->ss, sse = er->ss, per = ssx * sse;
  double curv[350];
  int rank[64];
  for (long j = 0; j < nz; ++j) {
    eta_sorted_rank(er, nz, j, rank);
    for (long i = 0; i < ny; ++i) {
      const long row = j * ny + i;
      rowptr[row] = (int)(row * per);
      int st = tor_node_coefficients(recipe, alpha_re, alpha_im, baseflow, deriv,
                                     metrics, row, curv);
      if (st) return st;
      const double *Gm = curv, *K0 = curv + 50, *Kxi = curv + 100, *Keta = curv + 150,
                   *Vxx = curv + 200, *Vxe = curv + 250, *Vee = curv + 300;
      /* column indices and zeroed values for this row */
      for (int t = 0; t < sse; ++t) {
        long jc = wrap_index(j + er->off[j * sse + t], nz);
        for (int s = 0; s < ssx; ++s) {
          long ic = i + xr->off[i * ssx + s];
          colidx[row * per + rank[t] * ssx + s] = (int)(jc * ny + ic);
        }
      }
      double *rv = lvals + (size_t)row * per * 50;
      memset(rv, 0, sizeof(double) * per * 50);
      /* M(row,row) = Gamma (INSERT) */
      memcpy(mvals + row * 50, Gm, sizeof(double) * 50);
      /* ADD order: K0, xi slots, eta slots, mixed (GlobalOperator.cpp:323-365) */
      int s0 = -1, t0 = -1;
      for (int s = 0; s < ssx; ++s)
        if (xr->off[i * ssx + s] == 0) s0 = s;
      for (int t = 0; t < sse; ++t)
        if (er->off[j * sse + t] == 0) t0 = t;
      {
        double *blk = rv + (rank[t0] * ssx + s0) * 50;
        for (int k = 0; k < 50; ++k) blk[k] += K0[k];
      }
      for (int s = 0; s < ssx; ++s) {
        const double w1 = xr->w1[i * ssx + s], mw2 = -xr->w2[i * ssx + s];
        double *blk = rv + (rank[t0] * ssx + s) * 50;
        for (int k = 0; k < 50; ++k) {
          double v = Kxi[k] * w1;
          v = v + Vxx[k] * mw2;
          blk[k] += v;
        }
      }
      for (int t = 0; t < sse; ++t) {
        const double w1 = er->w1[j * sse + t], mw2 = -er->w2[j * sse + t];
        double *blk = rv + (rank[t] * ssx + s0) * 50;
        for (int k = 0; k < 50; ++k) {
          double v = Keta[k] * w1;
          v = v + Vee[k] * mw2;
          blk[k] += v;
        }
      }
      for (int t = 0; t < sse; ++t) {
        const double we = er->w1[j * sse + t];
        for (int s = 0; s < ssx; ++s) {
          const double scale = -we * xr->w1[i * ssx + s];
          double *blk = rv + (rank[t] * ssx + s) * 50;
          for (int k = 0; k < 50; ++k) blk[k] += Vxe[k] * scale;
        }
      }
    }
  }
  rowptr[(long)ny * nz] = (int)((long)ny * nz * per);
  return 0;
}

/* ------------------------------------------------- boundary conditions */

static double cabs2(double re, double im) { return hypot(re, im); }

/* Complex division by Smith's method (R. L. Smith, CACM 5 (1962) 435). */
static void cdiv(double a, double b, double c, double d, double *qr, double *qi) {
  if (fabs(c) >= fabs(d)) {
    const double r = d / c, den = c + d * r;
    *qr = (a + b * r) / den;
    *qi = (b - a * r) / den;
  } else {
    const double r = c / d, den = c * r + d;
    *qr = (a * r + b) / den;
    *qi = (b * r - a) / den;
  }
}

/* Pivoted Gauss-Jordan inverse of a 5x5 complex matrix
 * (BoundaryConditions.cpp:92-140).  a, inv: 25 complex interleaved. */
int tor_invert5(const double *a_in, double *inv) {
  double a[50];
  memcpy(a, a_in, sizeof(a));
  memset(inv, 0, sizeof(double) * 50);
  for (int d = 0; d < 5; ++d) inv[2 * (d * 5 + d)] = 1.0;
  for (int pc = 0; pc < 5; ++pc) {
    int pr = pc;
    double pm = cabs2(a[2 * (pr * 5 + pc)], a[2 * (pr * 5 + pc) + 1]);
    for (int r = pc + 1; r < 5; ++r) {
      double mg = cabs2(a[2 * (r * 5 + pc)], a[2 * (r * 5 + pc) + 1]);
      if (mg > pm) { pm = mg; pr = r; }
    }
    if (!isfinite(pm) || pm <= 100.0 * kEps) return 0;
    if (pr != pc) {
      for (int c = 0; c < 5; ++c) {
        for (int p = 0; p < 2; ++p) {
          double t = a[2 * (pr * 5 + c) + p];
          a[2 * (pr * 5 + c) + p] = a[2 * (pc * 5 + c) + p];
          a[2 * (pc * 5 + c) + p] = t;
          t = inv[2 * (pr * 5 + c) + p];
          inv[2 * (pr * 5 + c) + p] = inv[2 * (pc * 5 + c) + p];
          inv[2 * (pc * 5 + c) + p] = t;
        }
      }
    }
    const double pre = a[2 * (pc * 5 + pc)], pim = a[2 * (pc * 5 + pc) + 1];
    for (int c = 0; c < 5; ++c) {
      for (int w = 0; w < 2; ++w) {
        double *m = w ? inv : a;
        double xr = m[2 * (pc * 5 + c)], xi = m[2 * (pc * 5 + c) + 1];
        double qr, qi;
        cdiv(xr, xi, pre, pim, &qr, &qi);
        m[2 * (pc * 5 + c)] = qr;
        m[2 * (pc * 5 + c) + 1] = qi;
      }
    }
    for (int r = 0; r < 5; ++r) {
      if (r == pc) continue;
      const double fr = a[2 * (r * 5 + pc)], fi = a[2 * (r * 5 + pc) + 1];
      for (int c = 0; c < 5; ++c) {
        for (int w = 0; w < 2; ++w) {
          double *m = w ? inv : a;
          double yr = m[2 * (pc * 5 + c)], yi = m[2 * (pc * 5 + c) + 1];
          m[2 * (r * 5 + c)] -= fr * yr - fi * yi;
          m[2 * (r * 5 + c) + 1] -= fr * yi + fi * yr;
        }
      }
    }
  }
  return 1;
}

/* Characteristic data of one far-field node.
 * out (complex 5x5 interleaved): R[50], Rinv[50], P[50]; speeds[5] (real);
 * flags[5]: bit0 incoming, bit1 neutral; extra[2]: max |Im c|, cond_inf(R).
 * Returns 0 or a CharacteristicStatus code (1 invalid input, 2 singular
 * metric, 6 singular eigenvectors, 7 ill-conditioned).
 *
 * The normal inviscid system (n_y Bc + n_z Cc) r = c Gamma r has the closed
 * form eigenvalues c = u_n - a, u_n (x3), u_n + a with a = sqrt(T)/Ma.  The
 * basis used here: acoustic [rho, 0, -/+ n_y a, -/+ n_z a, (gamma-1) T],
 * streamwise e_1, tangential [0, 0, n_z, -n_y, 0], entropy [rho,0,0,0,-T].
 * zggev returns some basis of the triple eigenvalue; any two bases differ by
 * an invertible transform inside that eigenspace, which changes the replaced
 * rows by a left transform and leaves the eigenproblem unchanged. */
int tor_characteristics(const double *recipe, const double *val, const double *m,
                        double *R, double *Rinv, double *P, double *speeds,
                        int *flags, double *extra) {
  double inv3[75];
  if (tor_lns_inviscid(recipe, val, inv3)) return 1;
  for (int k = 0; k < 10; ++k)
    if (!isfinite(m[k])) return 1;
  const double nm = sqrt(m[0] * m[0] + m[1] * m[1]);
  if (!isfinite(nm) || nm <= 100.0 * kEps) return 2;
  const double ny = m[0] / nm, nz = m[1] / nm;
  const double rho = val[0], V = val[2], W = val[3], T = val[4];
  const double gam = recipe[3], Ma = recipe[1];
  const double un = ny * V + nz * W;
  const double a = sqrt(T) / Ma;
  double c[5] = {un - a, un, un, un, un + a};
  double Rr[25] = {0};
  /* columns = modes */
  Rr[0 * 5 + 0] = rho; Rr[2 * 5 + 0] = -ny * a; Rr[3 * 5 + 0] = -nz * a; Rr[4 * 5 + 0] = (gam - 1.0) * T;
  Rr[1 * 5 + 1] = 1.0;
  Rr[2 * 5 + 2] = nz; Rr[3 * 5 + 2] = -ny;
  Rr[0 * 5 + 3] = rho; Rr[4 * 5 + 3] = -T;
  Rr[0 * 5 + 4] = rho; Rr[2 * 5 + 4] = ny * a; Rr[3 * 5 + 4] = nz * a; Rr[4 * 5 + 4] = (gam - 1.0) * T;
  for (int k = 0; k < 25; ++k) { R[2 * k] = Rr[k]; R[2 * k + 1] = 0.0; }
  double maxSpeed = 1.0;
  for (int k = 0; k < 5; ++k) {
    speeds[k] = c[k];
    if (fabs(c[k]) > maxSpeed) maxSpeed = fabs(c[k]);
  }
  extra[0] = 0.0;
  if (!tor_invert5(R, Rinv)) return 6;
  double G[50], Gi[50];
  for (int k = 0; k < 25; ++k) { G[2 * k] = inv3[k]; G[2 * k + 1] = 0.0; }
  if (!tor_invert5(G, Gi)) return 6;
  for (int r = 0; r < 5; ++r) {
    for (int cc = 0; cc < 5; ++cc) {
      double sr = 0.0, si = 0.0;
      for (int k = 0; k < 5; ++k) {
        double ar = Rinv[2 * (r * 5 + k)], ai = Rinv[2 * (r * 5 + k) + 1];
        double br = Gi[2 * (k * 5 + cc)], bi = Gi[2 * (k * 5 + cc) + 1];
        sr += ar * br - ai * bi;
        si += ar * bi + ai * br;
      }
      P[2 * (r * 5 + cc)] = sr;
      P[2 * (r * 5 + cc) + 1] = si;
    }
  }
  double nR = 0.0, nRi = 0.0;
  for (int r = 0; r < 5; ++r) {
    double s1 = 0.0, s2 = 0.0;
    for (int cc = 0; cc < 5; ++cc) {
      s1 += cabs2(R[2 * (r * 5 + cc)], R[2 * (r * 5 + cc) + 1]);
      s2 += cabs2(Rinv[2 * (r * 5 + cc)], Rinv[2 * (r * 5 + cc) + 1]);
    }
    if (s1 > nR) nR = s1;
    if (s2 > nRi) nRi = s2;
  }
  extra[1] = nR * nRi;
  if (!isfinite(extra[1]) || extra[1] > 1.0e12) return 7;
  const double tol = 1000.0 * kEps * maxSpeed;
  for (int k = 0; k < 5; ++k) {
    flags[k] = 0;
    if (c[k] < -tol) flags[k] |= 1;
    if (fabs(c[k]) <= tol) flags[k] |= 2;
  }
  return 0;
}

/* A_bc = -L, B_bc = M with wall and far-field rows replaced.
 * avals has the layout of lvals; bvals that of mvals.  diag (8 doubles):
 * wall nodes, far-field nodes, incoming, neutral, min incoming/node,
 * max incoming/node, max |Im c|, max cond(R).  Returns 0 or the maximal
 * characteristic failure status. */
int tor_apply_bc(int ny, int nz, const orule *xr, const orule *er,
                 const double *recipe, const double *baseflow, const double *deriv,
                 const double *metrics, const int *colidx, const double *lvals,
                 const double *mvals, double *avals, double *bvals, double *diag) {
  const int per = xr->ss * er->ss;
  const long n = (long)ny * nz;
  for (long k = 0; k < n * per * 50; ++k) avals[k] = lvals[k] * -1.0;
  memcpy(bvals, mvals, sizeof(double) * n * 50);
  int fail = 0, inc_total = 0, neu_total = 0, minInc = 5, maxInc = 0;
  double maxIm = 0.0, maxCond = 0.0;
  for (long j = 0; j < nz; ++j) {
    /* wall i = 0: rows 1..4 -> identity in A, zero in B */
    {
      const long row = j * ny + 0;
      double *rv = avals + (size_t)row * per * 50;
      for (int b = 0; b < per; ++b) {
        const int diagblk = colidx[row * per + b] == row;
        for (int r = 1; r < 5; ++r)
          for (int cc = 0; cc < 5; ++cc) {
            rv[b * 50 + 2 * (r * 5 + cc)] = (diagblk && r == cc) ? 1.0 : 0.0;
            rv[b * 50 + 2 * (r * 5 + cc) + 1] = 0.0;
          }
      }
      for (int r = 1; r < 5; ++r)
        for (int cc = 0; cc < 5; ++cc) {
          bvals[row * 50 + 2 * (r * 5 + cc)] = 0.0;
          bvals[row * 50 + 2 * (r * 5 + cc) + 1] = 0.0;
        }
    }
    /* far field i = Ny-1 */
    {
      const long row = j * ny + (ny - 1);
      double pt[50];
      int st = node_point(baseflow, deriv, row, pt);
      double R[50], Ri[50], P[50], sp[5], ex[2];
      int fl[5];
      if (!st) st = tor_characteristics(recipe, pt, metrics + 10 * row, R, Ri, P, sp, fl, ex);
      else st = 1;
      if (st) { if (st > fail) fail = st; continue; }
      int inc = 0;
      for (int k = 0; k < 5; ++k) {
        inc += fl[k] & 1;
        neu_total += (fl[k] >> 1) & 1;
      }
      inc_total += inc;
      if (inc < minInc) minInc = inc;
      if (inc > maxInc) maxInc = inc;
      if (ex[0] > maxIm) maxIm = ex[0];
      if (ex[1] > maxCond) maxCond = ex[1];
      double *rv = avals + (size_t)row * per * 50;
      const double *old = lvals + (size_t)row * per * 50;  /* -L read as old*-1 */
      double *brow = bvals + row * 50;
      double oldB[50];
      memcpy(oldB, mvals + row * 50, sizeof(oldB));
      double *newA = (double *)malloc(sizeof(double) * per * 50);
      double newB[50];
      memset(newB, 0, sizeof(newB));
      for (int mo = 0; mo < 5; ++mo) {
        if (fl[mo] & 1) {
          for (int b = 0; b < per; ++b) {
            const int diagblk = colidx[row * per + b] == row;
            for (int cc = 0; cc < 5; ++cc) {
              newA[b * 50 + 2 * (mo * 5 + cc)] = diagblk ? Ri[2 * (mo * 5 + cc)] : 0.0;
              newA[b * 50 + 2 * (mo * 5 + cc) + 1] = diagblk ? Ri[2 * (mo * 5 + cc) + 1] : 0.0;
            }
          }
          for (int cc = 0; cc < 5; ++cc) {
            newB[2 * (mo * 5 + cc)] = 0.0;
            newB[2 * (mo * 5 + cc) + 1] = 0.0;
          }
          continue;
        }
        for (int b = 0; b < per; ++b) {
          for (int cc = 0; cc < 5; ++cc) {
            double sr = 0.0, si = 0.0;
            for (int s = 0; s < 5; ++s) {
              const double pr = P[2 * (mo * 5 + s)], pi = P[2 * (mo * 5 + s) + 1];
              if (pr == 0.0 && pi == 0.0) continue;
              const double ar = old[b * 50 + 2 * (s * 5 + cc)] * -1.0;
              const double ai = old[b * 50 + 2 * (s * 5 + cc) + 1] * -1.0;
              sr += pr * ar - pi * ai;
              si += pr * ai + pi * ar;
            }
            newA[b * 50 + 2 * (mo * 5 + cc)] = sr;
            newA[b * 50 + 2 * (mo * 5 + cc) + 1] = si;
          }
        }
        for (int cc = 0; cc < 5; ++cc) {
          double sr = 0.0, si = 0.0;
          for (int s = 0; s < 5; ++s) {
            const double pr = P[2 * (mo * 5 + s)], pi = P[2 * (mo * 5 + s) + 1];
            if (pr == 0.0 && pi == 0.0) continue;
            const double br = oldB[2 * (s * 5 + cc)], bi = oldB[2 * (s * 5 + cc) + 1];
            sr += pr * br - pi * bi;
            si += pr * bi + pi * br;
          }
          newB[2 * (mo * 5 + cc)] = sr;
          newB[2 * (mo * 5 + cc) + 1] = si;
        }
      }
      memcpy(rv, newA, sizeof(double) * per * 50);
      memcpy(brow, newB, sizeof(newB));
      free(newA);
    }
  }
  diag[0] = nz; diag[1] = nz; diag[2] = inc_total; diag[3] = neu_total;
  diag[4] = minInc; diag[5] = maxInc; diag[6] = maxIm; diag[7] = maxCond;
  return fail;
}
