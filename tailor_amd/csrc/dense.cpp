// dense.cpp -- see dense.hpp.
#include "dense.hpp"

#include <algorithm>
#include <cmath>

namespace tailor {

namespace {

constexpr double kEps = 2.220446049250313e-16;

// Givens rotation with real cosine: [c s; -conj(s) c] [f; g] = [r; 0].
void givens(cplx f, cplx g, double &c, cplx &s, cplx &r) {
  const double af = std::abs(f), ag = std::abs(g);
  if (ag == 0.0) {
    c = 1.0;
    s = 0.0;
    r = f;
    return;
  }
  if (af == 0.0) {
    c = 0.0;
    s = std::conj(g) / ag;
    r = ag;
    return;
  }
  const double nrm = std::hypot(af, ag);
  c = af / nrm;
  const cplx phase = f / af;
  s = phase * std::conj(g) / nrm;
  r = phase * nrm;
}

// rows i, k of A (columns c0..c1-1): [x; y] <- [c s; -conj(s) c] [x; y]
void rot_rows(CMat &A, int i, int k, double c, cplx s, int c0, int c1) {
  for (int j = c0; j < c1; ++j) {
    const cplx x = A(i, j), y = A(k, j);
    A(i, j) = c * x + s * y;
    A(k, j) = c * y - std::conj(s) * x;
  }
}

// columns i, k of A (rows r0..r1-1): multiply on the right by the adjoint
// of the row rotation, keeping similarity.
void rot_cols(CMat &A, int i, int k, double c, cplx s, int r0, int r1) {
  for (int j = r0; j < r1; ++j) {
    const cplx x = A(j, i), y = A(j, k);
    A(j, i) = c * x + std::conj(s) * y;
    A(j, k) = c * y - s * x;
  }
}

void hessenberg(CMat &A, CMat &Q) {
  const int n = A.n;
  for (int k = 0; k < n - 2; ++k) {
    double alpha2 = 0.0;
    for (int i = k + 1; i < n; ++i) alpha2 += std::norm(A(i, k));
    const double alpha = std::sqrt(alpha2);
    if (alpha == 0.0) continue;
    std::vector<cplx> v(n, 0.0);
    const cplx x0 = A(k + 1, k);
    const cplx ph = std::abs(x0) == 0.0 ? cplx(1.0) : x0 / std::abs(x0);
    v[k + 1] = x0 + ph * alpha;
    for (int i = k + 2; i < n; ++i) v[i] = A(i, k);
    double vn2 = 0.0;
    for (int i = k + 1; i < n; ++i) vn2 += std::norm(v[i]);
    if (vn2 == 0.0) continue;
    // A <- H A H with H = I - 2 v v^H / (v^H v)
    for (int j = 0; j < n; ++j) {
      cplx s = 0.0;
      for (int i = k + 1; i < n; ++i) s += std::conj(v[i]) * A(i, j);
      s *= 2.0 / vn2;
      for (int i = k + 1; i < n; ++i) A(i, j) -= v[i] * s;
    }
    for (int i = 0; i < n; ++i) {
      cplx s = 0.0;
      for (int j = k + 1; j < n; ++j) s += A(i, j) * v[j];
      s *= 2.0 / vn2;
      for (int j = k + 1; j < n; ++j) A(i, j) -= s * std::conj(v[j]);
    }
    for (int i = 0; i < n; ++i) {
      cplx s = 0.0;
      for (int j = k + 1; j < n; ++j) s += Q(i, j) * v[j];
      s *= 2.0 / vn2;
      for (int j = k + 1; j < n; ++j) Q(i, j) -= s * std::conj(v[j]);
    }
    for (int i = k + 2; i < n; ++i) A(i, k) = 0.0;
  }
}

}  // namespace

bool schur_decompose(CMat &A, CMat &Q) {
  const int n = A.n;
  Q = CMat(n, n);
  for (int i = 0; i < n; ++i) Q(i, i) = 1.0;
  if (n <= 1) return true;
  hessenberg(A, Q);
  int hi = n - 1, iter = 0, total = 0;
  while (hi > 0) {
    int l = hi;
    for (; l > 0; --l) {
      const double s = std::abs(A(l, l)) + std::abs(A(l - 1, l - 1));
      if (std::abs(A(l, l - 1)) <= kEps * (s == 0.0 ? 1.0 : s)) {
        A(l, l - 1) = 0.0;
        break;
      }
    }
    if (l == hi) {
      --hi;
      iter = 0;
      continue;
    }
    if (++total > 100 * n) return false;
    ++iter;
    cplx mu;
    if (iter % 11 == 10) {
      mu = A(hi, hi) + std::abs(A(hi, hi - 1));  // exceptional shift
    } else {
      const cplx a = A(hi - 1, hi - 1), b = A(hi - 1, hi), c = A(hi, hi - 1), d = A(hi, hi);
      const cplx tr = a + d, det = a * d - b * c;
      const cplx disc = std::sqrt(tr * tr / 4.0 - det);
      const cplx m1 = tr / 2.0 + disc, m2 = tr / 2.0 - disc;
      mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
    }
    cplx x = A(l, l) - mu, y = A(l + 1, l);
    for (int k = l; k < hi; ++k) {
      double c;
      cplx s, r;
      givens(x, y, c, s, r);
      rot_rows(A, k, k + 1, c, s, std::max(l, k - 1), n);
      rot_cols(A, k, k + 1, c, s, 0, std::min(k + 3, hi + 1));
      rot_cols(Q, k, k + 1, c, s, 0, n);
      if (k + 1 < hi) {
        x = A(k + 1, k);
        y = A(k + 2, k);
      }
    }
  }
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) A(i, j) = 0.0;
  return true;
}

void schur_reorder(CMat &T, CMat &Q, const std::vector<bool> &select) {
  const int n = T.n;
  int target = 0;
  for (int k = 0; k < n; ++k) {
    if (!select[k]) continue;
    // bubble T(k,k) up to position `target`
    for (int j = k - 1; j >= target; --j) {
      const cplx t11 = T(j, j), t22 = T(j + 1, j + 1);
      double c;
      cplx s, r;
      givens(T(j, j + 1), t22 - t11, c, s, r);
      rot_rows(T, j, j + 1, c, s, j + 2, n);
      rot_cols(T, j, j + 1, c, s, 0, j);
      T(j, j) = t22;
      T(j + 1, j + 1) = t11;
      rot_cols(Q, j, j + 1, c, s, 0, n);
    }
    ++target;
  }
}

std::vector<cplx> triangular_eigenvector(const CMat &T, int i) {
  const int n = T.n;
  std::vector<cplx> y(n, 0.0);
  y[i] = 1.0;
  const cplx lam = T(i, i);
  double tnorm = 0.0;
  for (int k = 0; k <= i; ++k) tnorm = std::max(tnorm, std::abs(T(k, k)));
  const double small = kEps * std::max(tnorm, 1e-300);
  for (int k = i - 1; k >= 0; --k) {
    cplx s = 0.0;
    for (int j = k + 1; j <= i; ++j) s += T(k, j) * y[j];
    cplx den = T(k, k) - lam;
    if (std::abs(den) < small) den = small;
    y[k] = -s / den;
  }
  double nrm = 0.0;
  for (const cplx &v : y) nrm += std::norm(v);
  nrm = std::sqrt(nrm);
  for (cplx &v : y) v /= nrm;
  return y;
}

bool dense_solve(CMat A, std::vector<cplx> &b) {
  const int n = A.n;
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(A(i, k)) > std::abs(A(p, k))) p = i;
    if (std::abs(A(p, k)) == 0.0) return false;
    if (p != k) {
      for (int j = 0; j < n; ++j) std::swap(A(k, j), A(p, j));
      std::swap(b[k], b[p]);
    }
    for (int i = k + 1; i < n; ++i) {
      const cplx f = A(i, k) / A(k, k);
      for (int j = k; j < n; ++j) A(i, j) -= f * A(k, j);
      b[i] -= f * b[k];
    }
  }
  for (int k = n - 1; k >= 0; --k) {
    cplx s = b[k];
    for (int j = k + 1; j < n; ++j) s -= A(k, j) * b[j];
    b[k] = s / A(k, k);
  }
  return true;
}

}  // namespace tailor
