"""Golden eigenvalues of the synthetic BiGlobal cases (fixture generator).

Writes ``eig_<ny>x<nz>.json`` beside this file: the eigenvalues of
A_bc q = lambda B_bc q nearest the target lambda_s = Im(omega_s) - i Re(omega_s)
(EigenSolver.cpp:46-48) of

* the 100x50 case (alpha 2.43, omega_s 0, Re 1000, Ma 0.5, seed 0), and
* the 200x100 case (alpha 2.43, omega_s 0.02, otherwise the same),

computed on the CPU, independently of the GPU path and of the oracle's own
Krylov-Schur:

* A_bc and B_bc come from the C oracle (oracle/tailor_oracle.c), whose
  entries the GPU assembly reproduces bit for bit;
* K = A_bc - lambda_s B_bc is factored by SuperLU (scipy, its default
  COLAMD column ordering and partial pivoting);
* ARPACK (scipy.sparse.linalg.eigs, implicitly restarted Arnoldi) finds the
  12 largest |theta| of the shift-invert operator K^-1 B_bc, lambda =
  lambda_s + 1/theta, and the 12 are sorted by |lambda - lambda_s|.

Each stored eigenvalue carries its true residual ||A q - lambda B q|| / ||q||
on the oracle matrices, and the file records the distance gap between the
8th and 9th eigenvalue, so that a test can tell a selection tie from an
error.  The reference ships no golden spectrum for any case: these values
pin the GPU solver against the same operator, not against the reference's
own output.

Run from the repository root (minutes for 100x50; the 200x100 case needs
tens of minutes and about 20 GB of memory):

    python tests/golden/make_eig_gold.py            # both cases
    python tests/golden/make_eig_gold.py 100x50     # one case
"""
from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np
import scipy.sparse.linalg as spla

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from oracle import eigen  # noqa: E402
from tailor_amd import case as cs  # noqa: E402

CASES = {
    "100x50": dict(ny=100, nz=50, omega=0j),
    "200x100": dict(ny=200, nz=100, omega=0.02 + 0j),
}
KEEP = 12


def golden(name: str) -> dict:
    spec = CASES[name]
    case = cs.synthetic_case(spec["ny"], spec["nz"], seed=0, omega_shift=spec["omega"])
    t0 = time.perf_counter()
    a = oracle.assemble(case)
    A = oracle.to_scipy(a.rowptr, a.colidx, a.A, case.nodes).tocsr()
    B = oracle.block_diag_to_scipy(a.B).tocsr()
    sigma = eigen.target_lambda(case.omega_shift)
    K = (A - sigma * B).tocsc()
    t1 = time.perf_counter()
    lu = spla.splu(K)
    t2 = time.perf_counter()
    print(f"{name}: assembled {t1 - t0:.1f} s, SuperLU {t2 - t1:.1f} s, "
          f"{lu.L.nnz + lu.U.nnz} factor entries", flush=True)

    def op(v):
        return lu.solve(B @ v)

    n = A.shape[0]
    Op = spla.LinearOperator((n, n), matvec=op, dtype=np.complex128)
    v0 = np.ones(n, dtype=np.complex128)
    theta, vecs = spla.eigs(Op, k=KEEP, which="LM", tol=1e-15, ncv=48, v0=v0, maxiter=5000)
    t3 = time.perf_counter()
    lam = sigma + 1.0 / theta
    idx = np.argsort(np.abs(lam - sigma), kind="stable")
    lam, vecs = lam[idx], vecs[:, idx]
    res = np.array([np.linalg.norm(A @ v - l * (B @ v)) / np.linalg.norm(v)
                    for l, v in zip(lam, vecs.T)])
    print(f"{name}: ARPACK {t3 - t2:.1f} s, max residual {res.max():.2e}", flush=True)
    dist = np.abs(lam - sigma)
    return {
        "case": f"synthetic_case({spec['ny']}, {spec['nz']}, seed=0), alpha 2.43, Re 1000, "
                f"Ma 0.5, Pr 0.72, gamma 1.4, Tref 288.15, S 110.4",
        "ny": spec["ny"], "nz": spec["nz"],
        "omega_shift": [spec["omega"].real, spec["omega"].imag],
        "lambda_shift": [sigma.real, sigma.imag],
        "method": "oracle-assembled A_bc, B_bc; SuperLU (COLAMD) of A_bc - lambda_s B_bc; "
                  "ARPACK shift-invert, 12 eigenvalues nearest lambda_s",
        "lambda": [[float(l.real), float(l.imag)] for l in lam],
        "residual": [float(r) for r in res],
        "distance": [float(d) for d in dist],
        "gap_8_9": float(dist[8] - dist[7]),
        "factor_entries": int(lu.L.nnz + lu.U.nnz),
    }


def main(argv) -> int:
    names = argv[1:] or list(CASES)
    for name in names:
        g = golden(name)
        out = HERE / f"eig_{name}.json"
        out.write_text(json.dumps(g, indent=1) + "\n")
        print(f"wrote {out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
