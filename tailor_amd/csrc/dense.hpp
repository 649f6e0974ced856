// dense.hpp -- small dense complex linear algebra for the Krylov-Schur
// projected problem (dimension ncv, a few tens).  Host code.
//
// Published methods: Householder reduction to Hessenberg form, the implicit
// single-shift complex QR algorithm with Wilkinson shifts, and reordering of
// the complex Schur form by adjacent swaps (Golub & Van Loan, "Matrix
// Computations", 4th ed., sections 7.4-7.6).
#pragma once

#include <complex>
#include <vector>

namespace tailor {

using cplx = std::complex<double>;

// Column-major n x n matrix.
struct CMat {
  int n = 0, cols = 0;
  std::vector<cplx> a;
  CMat() = default;
  CMat(int rows, int c) : n(rows), cols(c), a(static_cast<size_t>(rows) * c) {}
  cplx &operator()(int i, int j) { return a[static_cast<size_t>(j) * n + i]; }
  const cplx &operator()(int i, int j) const { return a[static_cast<size_t>(j) * n + i]; }
};

// A = Q T Q^H with T upper triangular.  A is overwritten by T, Q returned.
// Returns false if the QR iteration did not converge.
bool schur_decompose(CMat &A, CMat &Q);

// Reorder the Schur form so that the eigenvalues marked in `select` come
// first (in their current relative order); updates T and Q.
void schur_reorder(CMat &T, CMat &Q, const std::vector<bool> &select);

// Eigenvector y of upper triangular T for eigenvalue T(i,i): y(i) = 1,
// y(k) = 0 for k > i, unit 2-norm.
std::vector<cplx> triangular_eigenvector(const CMat &T, int i);

// Dense complex LU solve (partial pivoting), small systems.
bool dense_solve(CMat A, std::vector<cplx> &b);

}  // namespace tailor
