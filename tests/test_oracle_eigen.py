"""Pin the oracle's shift-invert Krylov-Schur (oracle/eigen.py).

* The known 10x10 generalized problem of test_eigen_solver.cpp:17-136:
  target omega = -0.65-0.45i -> lambda_s = -0.45+0.65i; the two nearest
  eigenvalues are -0.5+0.7i and -1.2-0.4i within 2e-10, residual <= 2e-10.
* Dense LAPACK brute force on a small BiGlobal grid: the 8 eigenvalues
  nearest the target within 1e-8 relative.
"""
import numpy as np
import pytest
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


@pytest.fixture(scope="module")
def dense_24x12(oracle_lib):
    """A_bc, B_bc of the 24x12 case and the 8 finite eigenvalues nearest 0
    from dense LAPACK (zggev), computed once for the tests below."""
    c = small_case(24, 12)
    a = oracle_lib.assemble(c)
    n = c.ny * c.nz
    A = oracle_lib.to_scipy(a.rowptr, a.colidx, a.A, n)
    B = oracle_lib.block_diag_to_scipy(a.B)
    w = sla.eig(A.toarray(), B.toarray(), right=False)
    w = w[np.isfinite(w)]
    return c, A, B, w[np.argsort(np.abs(w))][:8]


def test_biglobal_small_grid_vs_dense(dense_24x12):
    c, A, B, ref = dense_24x12
    sigma = 0.0 + 0.0j
    lams, _, res, _ = eigen.krylov_schur(A, B, sigma, nev=8, tol=1e-12)
    for lam in ref:
        assert np.min(np.abs(lams - lam)) <= 1e-8 * max(1, abs(lam))
    assert np.all(res[:8] <= 1e-10)



@pytest.mark.parametrize("shape,omega", [((24, 12), 0j), ((40, 16), 0.02 + 0j)])
def test_multifrontal_lu_matches_superlu(oracle_lib, shape, omega):
    """oracle/multifrontal.py (the CPU baseline's LU: nested dissection,
    dense frontal LU with partial pivoting, extend-add) against SuperLU on
    the same K = A_bc - lambda_s B_bc: solutions agree and both residuals are
    at rounding level."""
    import scipy.sparse.linalg as spla
    from oracle import multifrontal
    c = small_case(*shape)
    a = oracle_lib.assemble(c)
    n = c.ny * c.nz
    sigma = eigen.target_lambda(omega)
    per = len(a.colidx) // n
    kvals = a.A.copy()
    kvals[a.colidx == np.repeat(np.arange(n), per)] -= sigma * a.B
    lu = multifrontal.MultifrontalLU(c, a.rowptr, a.colidx, kvals)
    K = oracle_lib.to_scipy(a.rowptr, a.colidx, kvals, n).tocsc()
    ref = spla.splu(K)
    rng = np.random.default_rng(7)
    for _ in range(3):
        b = rng.standard_normal(5 * n) + 1j * rng.standard_normal(5 * n)
        x, xr = lu.solve(b), ref.solve(b)
        assert np.linalg.norm(K @ x - b) <= 1e-12 * np.linalg.norm(b) * max(1.0, abs(K).max())
        assert np.linalg.norm(x - xr) <= 1e-9 * np.linalg.norm(xr)


def test_multifrontal_solve_case_vs_dense(dense_24x12):
    """The CPU baseline end to end (multifrontal LU + Krylov-Schur) against
    dense LAPACK on the 24x12 grid: the 8 eigenvalues nearest the target."""
    from oracle import multifrontal
    c, _, _, ref = dense_24x12
    out = multifrontal.solve_case(c, nev=8, tol=1e-10)
    for lam in ref:
        assert np.min(np.abs(out["lambdas"] - lam)) <= 1e-8 * max(1, abs(lam))
    assert np.all(out["residuals"][:8] <= 1e-10)
