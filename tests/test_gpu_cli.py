"""The ``tailor`` solver executable end to end on the GPU.

A synthetic 24x12 case is written as a schema-v2 FD-q cache plus a YAML case
file, solved through :func:`tailor_amd.cli.solver_main`, and its schema-v1
output is checked the way the reference's test_eigen_output.py checks its own
(ordering and schema attributes, omega = i lambda, real residuals within the
tolerance, unit-norm modes whose largest entry is real and positive), and
against the eigenvalues of the same solve through the context API and dense
LAPACK.  Rule corruptions that only the device-side loader sees (an index out
of range, a non-finite weight) fail with the messages of Prepare.cpp:376-401.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import small_case
import tailor_amd as ta
from tailor_amd import cli, h5

pytestmark = pytest.mark.gpu


def _case_dir(tmp_path, c):
    cfg = cli.write_config(tmp_path / "case" / "config.yaml", c)
    cli.write_cache(c, tmp_path / "case" / "data" / "FD-q" / "fdq_case_qy10_qz6.h5")
    return cfg


def test_solver_executable_writes_the_schema_v1_file(tmp_path, capsys, oracle_lib):
    c = small_case()
    cfg = _case_dir(tmp_path, c)
    print("tailor -c <24x12 synthetic case>", flush=True)
    assert cli.solver_main(["-c", str(cfg), "-eps_type", "krylovschur"]) == 0
    out = capsys.readouterr().out
    assert "Prepared case" in out and "eigensolver: requested=8" in out
    assert "  output: " in out and "modes)" in out

    path = tmp_path / "case" / "results" / "eigenmodes.h5"
    assert path.is_file() and not path.with_name("eigenmodes.h5.tmp").exists()
    f = h5.read(path)
    assert f.attrs["ordering"] == "k_j_dof" and f.attrs["schema_version"] == 1
    assert f.attrs["mode_dof"] == "[rho,u,v,w,T]"
    assert f.attrs["lambda_convention"] == "lambda=-i*omega; A_bc*q=lambda*B_bc*q"
    assert f.attrs["requested_modes"] == 8
    conv = f.attrs["converged_modes"]
    assert conv >= 8 and (f.attrs["Ny"], f.attrs["Nz"]) == (24, 12)
    np.testing.assert_array_equal(h5.as_complex(f["grid"]), c.grid)
    np.testing.assert_array_equal(h5.as_complex(f["baseflow"]), c.baseflow)
    lam = h5.as_complex(f["spectrum/lambda"])
    np.testing.assert_allclose(h5.as_complex(f["spectrum/omega"]), 1j * lam, rtol=1e-13, atol=1e-13)
    res = h5.as_complex(f["spectrum/residual"])
    assert np.max(np.abs(res.imag)) == 0.0 and np.max(res.real) <= 2e-10

    a = oracle_lib.assemble(c)
    n = c.ny * c.nz
    A = oracle_lib.to_scipy(a.rowptr, a.colidx, a.A, n).toarray()
    B = oracle_lib.block_diag_to_scipy(a.B).toarray()
    w = sla.eig(A, B, right=False)
    w = w[np.isfinite(w)]
    ref = w[np.argsort(np.abs(w))[:8]]
    for k in range(conv):
        q = h5.as_complex(f[f"modes/mode_{k:03d}"]).reshape(-1)
        assert h5.as_complex(f[f"modes/mode_{k:03d}"]).shape == (n, 5)
        assert abs(np.linalg.norm(q) - 1.0) <= 2e-13
        p = q[np.argmax(np.abs(q))]
        assert abs(p.imag) <= 2e-13 and p.real > 0
        r = np.linalg.norm(A @ q - lam[k] * (B @ q))
        assert r <= 1e-10 * max(1.0, abs(lam[k])), (k, r)
    for lr in ref:
        assert np.min(np.abs(lam - lr)) <= 1e-8 * max(1.0, abs(lr)), lr

    with ta.Context(0) as ctx:
        ctx.load_case(c)
        ctx.assemble(diagnostics=False)
        ctx.solve(nev=8, omega=c.omega_shift, tol=1e-10, maxit=1000)
        direct = np.array([p["lambda"] for p in ctx.eigenpairs()])
    np.testing.assert_allclose(lam, direct, rtol=1e-12, atol=1e-12)


def test_assemble_only_skips_the_solve(tmp_path, capsys):
    cfg = _case_dir(tmp_path, small_case())
    assert cli.solver_main(["-c", str(cfg), "-tailor_assemble_only"]) == 0
    out = capsys.readouterr().out
    assert "assemble-only: boundary matrices validated; EPSSolve skipped" in out
    assert "||A_bc||F:" in out
    assert not (tmp_path / "case" / "results").exists()


@pytest.mark.parametrize("name,edit,message", [
    ("out_of_range_index", lambda r: r["discretization/y/stencil_indices"].data.__setitem__(
        (0, 0), r.attrs["Ny"]), "out-of-range index"),
    ("non_finite_weight", lambda r: r["discretization/y/weights/d1"].data.__setitem__(
        (0, 0), np.nan), "weight is not finite"),
])
def test_rule_corruptions_fail_in_the_loader(tmp_path, capsys, name, edit, message):
    cfg = _case_dir(tmp_path, small_case())
    cache = cli.parse_config(cfg).input_file
    root = h5.read(cache)
    for path in ("discretization/y/stencil_indices", "discretization/y/weights/d1"):
        root[path].data = root[path].data.copy()
    edit(root)
    h5.write(cache, root)
    print(f"tailor -c <corrupted cache: {name}> -tailor_assemble_only", flush=True)
    assert cli.solver_main(["-c", str(cfg), "-tailor_assemble_only"]) != 0
    assert message in capsys.readouterr().err


def test_stored_relative_errors_meet_the_tolerance_when_lambda_is_small(tmp_path, capsys):
    """/spectrum/residual holds SLEPc's relative error ||Aq - lambda Bq|| /
    (|lambda| ||q||) (EigenSolver.cpp:77-82); with |lambda| << 1 it is the
    stricter of the two measures, and every stored value must meet tol."""
    c = small_case(alpha=0.3 + 0j)
    cfg = _case_dir(tmp_path, c)
    print("tailor -c <24x12 synthetic case, alpha 0.3>", flush=True)
    assert cli.solver_main(["-c", str(cfg)]) == 0
    f = h5.read(tmp_path / "case" / "results" / "eigenmodes.h5")
    lam = h5.as_complex(f["spectrum/lambda"])
    rel = h5.as_complex(f["spectrum/residual"]).real
    assert np.min(np.abs(lam)) < 0.05  # the relative error is > 20x the residual here
    assert np.max(rel) <= 1e-10, rel
    out = capsys.readouterr().out
    max_rel = float(out.split("max relative error: ")[1].split()[0])
    assert max_rel <= 1e-10


def test_unconverged_solve_fails_before_any_output(tmp_path, capsys):
    """Fewer than nev pairs after maxit restarts: the reference's
    'SLEPc failed to converge' error (EigenSolver.cpp:84-88), a nonzero exit
    and no output file."""
    cfg = _case_dir(tmp_path, small_case())
    print("tailor -c <24x12 synthetic case> -eps_max_it 1 -eps_ncv 10", flush=True)
    rc = cli.solver_main(["-c", str(cfg), "-eps_max_it", "1", "-eps_ncv", "10"])
    err = capsys.readouterr().err
    assert rc != 0
    assert "SLEPc failed to converge (reason -1, " in err and "after 1 iterations" in err
    assert not (tmp_path / "case" / "results").exists()
