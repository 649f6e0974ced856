"""Pin the oracle's shift-invert Krylov-Schur (oracle/eigen.py).

* The known 10x10 generalized problem of test_eigen_solver.cpp:17-136:
  target omega = -0.65-0.45i -> lambda_s = -0.45+0.65i; the two nearest
  eigenvalues are -0.5+0.7i and -1.2-0.4i within 2e-10, residual <= 2e-10.
* Dense LAPACK brute force on a small BiGlobal grid: the 8 eigenvalues
  nearest the target within 1e-8 relative.
"""
import numpy as np
import scipy.linalg as sla

from conftest import small_case
from oracle import eigen


def test_target_conversion():
    assert eigen.target_lambda(-0.65 - 0.45j) == complex(-0.45, 0.65)


def test_known_generalized_problem():
    A, B = eigen.known_problem()
    sigma = eigen.target_lambda(-0.65 - 0.45j)
    lams, vecs, res, _ = eigen.krylov_schur(A, B, sigma, nev=2, tol=1e-12, maxit=200, ncv=6)
    got = sorted(lams[:2], key=lambda z: abs(z - sigma))
    assert abs(got[0] - (-0.5 + 0.7j)) <= 2e-10
    assert abs(got[1] - (-1.2 - 0.4j)) <= 2e-10
    assert np.all(res[:2] <= 2e-10)


def test_known_problem_eigenvalues_are_the_diagonal():
    A, B = eigen.known_problem()
    w = sla.eig(A, B, right=False)
    expected = np.diag(A) / np.diag(B)
    assert np.allclose(np.sort_complex(w), np.sort_complex(expected), atol=1e-13)


def test_biglobal_small_grid_vs_dense(oracle_lib):
    c = small_case(24, 12)
    a = oracle_lib.assemble(c)
    n = c.ny * c.nz
    A = oracle_lib.to_scipy(a.rowptr, a.colidx, a.A, n)
    B = oracle_lib.block_diag_to_scipy(a.B)
    sigma = 0.0 + 0.0j
    lams, _, res, _ = eigen.krylov_schur(A, B, sigma, nev=8, tol=1e-12)
    w = sla.eig(A.toarray(), B.toarray(), right=False)
    w = w[np.isfinite(w)]
    ref = w[np.argsort(np.abs(w - sigma))][:8]
    for lam in ref:
        assert np.min(np.abs(lams - lam)) <= 1e-8 * max(1, abs(lam))
    assert np.all(res[:8] <= 1e-10)
