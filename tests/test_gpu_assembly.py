"""GPU assembly parity: libtailor (HIP, gfx950) vs the CPU oracle.

The bar: assembled block entries within 1e-12 relative.  The kernels follow the oracle's operation order without FP
contraction, so most entries are expected to be bit-identical; the test
reports the bit-exact fraction and enforces the 1e-12 bar.
"""
import numpy as np
import pytest

from conftest import small_case
import tailor_amd as ta
from tailor_amd import case as cs

pytestmark = pytest.mark.gpu


def rel_err(got, ref):
    return np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref)))


def assemble_both(c, oracle_lib):
    ref = oracle_lib.assemble(c)
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        diag = ctx.assemble(diagnostics=True)
        m, d = ctx.export_fields()
        rp, ci, A = ctx.export_bsr(ta.OP_A)
        _, _, B = ctx.export_bsr(ta.OP_B)
        ctx.assemble_unconstrained()
        _, _, L = ctx.export_bsr(ta.OP_L)
        _, _, M = ctx.export_bsr(ta.OP_M)
    return ref, diag, dict(metrics=m, deriv=d, rowptr=rp, colidx=ci, A=A, B=B, L=L, M=M)


@pytest.mark.parametrize("shape", [(24, 12), (100, 50)])
def test_assembly_matches_oracle(oracle_lib, shape):
    ny, nz = shape
    c = small_case(ny, nz) if ny < 100 else cs.synthetic_case(100, 50)
    ref, diag, got = assemble_both(c, oracle_lib)
    print(f"case {ny}x{nz}", flush=True)
    assert np.array_equal(got["rowptr"], ref.rowptr)
    assert np.array_equal(got["colidx"], ref.colidx)
    assert rel_err(got["metrics"], ref.metrics) <= 1e-13
    assert rel_err(got["deriv"], ref.deriv) <= 1e-13
    for key in ("L", "M", "A", "B"):
        g, r = got[key], getattr(ref, key)
        err = rel_err(g, r)
        exact = np.mean(g == r)
        print(f"{key}: max rel err {err:.3e}, bit-exact fraction {exact:.6f}")
        assert err <= 1e-12, key
    assert diag["wall_nodes"] == nz and diag["farfield_nodes"] == nz
    assert diag["incoming_modes"] == int(ref.bc_diag[2])
    assert diag["neutral_modes"] == int(ref.bc_diag[3])
    assert abs(diag["a_fro"] - np.linalg.norm(ref.A)) <= 1e-12 * np.linalg.norm(ref.A)
    assert abs(diag["spatial_fro"] - np.linalg.norm(ref.L)) <= 1e-12 * np.linalg.norm(ref.L)
    assert abs(diag["b_fro"] - np.linalg.norm(ref.B)) <= 1e-12 * np.linalg.norm(ref.B)


def test_full_size_assembly_matches_oracle(oracle_lib):
    c = cs.synthetic_case(200, 100, omega_shift=0.02 + 0j)
    ref = oracle_lib.assemble(c)
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        ctx.assemble(diagnostics=False)
        _, _, A = ctx.export_bsr(ta.OP_A)
        _, _, B = ctx.export_bsr(ta.OP_B)
    assert A.shape[0] == 200 * 100 * 77
    assert rel_err(A, ref.A) <= 1e-12
    assert rel_err(B, ref.B) <= 1e-12
    print(f"200x100 bit-exact fraction A {np.mean(A == ref.A):.6f}")


def test_alpha_change_reassembles(oracle_lib):
    c = small_case(24, 12)
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        ctx.set_alpha(0.7 + 0.1j)
        ctx.assemble(diagnostics=False)
        _, _, A = ctx.export_bsr(ta.OP_A)
    ref = oracle_lib.assemble(c.with_alpha(0.7 + 0.1j))
    assert rel_err(A, ref.A) <= 1e-12


def test_uniform_flow_counts():
    c = cs.uniform_case(24, 8, q_y=4, q_z=2)
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        d = ctx.assemble()
    assert (d["incoming_modes"], d["neutral_modes"]) == (8, 24)
    assert (d["min_incoming_per_node"], d["max_incoming_per_node"]) == (1, 1)


@pytest.mark.parametrize("mutation,status,message", [
    ("nan_grid", ta.TailorError, "Grid coordinates or computed first-order metrics are not finite"),
    ("complex_grid", ta.TailorError, "Grid y/z coordinates must be real"),
    ("degenerate_grid", ta.TailorError, "Grid Jacobian must be positive and nonsingular"),
    ("negative_T", ta.TailorError, "density and temperature must both be positive"),
    ("complex_base", ta.TailorError, "Base-flow values must be real"),
])
def test_invalid_cases_fail_with_reference_messages(mutation, status, message):
    c = small_case(24, 12)
    if mutation == "nan_grid":
        c.grid = c.grid.copy(); c.grid[5, 1] = np.nan
    elif mutation == "complex_grid":
        c.grid = c.grid.copy(); c.grid[:, 1] += 1e-6j
    elif mutation == "degenerate_grid":
        c.grid = c.grid.copy(); c.grid[:, 1:] = 0.0
    elif mutation == "negative_T":
        c.baseflow = c.baseflow.copy(); c.baseflow[7, 4] = -1.0
    elif mutation == "complex_base":
        c.baseflow = c.baseflow.copy(); c.baseflow[:, 1] += 1e-6j
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        with pytest.raises(status) as e:
            ctx.assemble(diagnostics=False)
    assert message in str(e.value)


def test_physics_validation_message():
    c = small_case(24, 12)
    c.physics = cs.Physics(reynolds=-1.0)
    with ta.Context(0) as ctx:
        with pytest.raises(ta.TailorError) as e:
            ctx.load_case(c)
    assert "outside their valid ranges" in str(e.value)


def test_derivatives_closed_form_on_skewed_grid():
    """The GPU's metrics and base-flow derivatives against the closed form of
    a skewed affine grid with base-flow fields P(y) H(eta)
    (conftest.skewed_product_case), with the bounds the oracle meets."""
    from conftest import skewed_product_case
    c, exact, metrics = skewed_product_case()
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        ctx.assemble(diagnostics=False)
        m, d = ctx.export_fields()
    assert np.max(np.abs(m[:, :4] - metrics[:4])) <= 1e-13
    assert np.max(np.abs(m[:, 4:])) <= 1e-10
    scale = np.max(np.abs(exact), axis=0)
    err = np.max(np.abs(d - exact), axis=0) / np.maximum(scale, 1e-3)
    print(f"closed-form derivative error {np.max(err):.3e}", flush=True)
    assert np.max(err) <= 1e-9


@pytest.mark.parametrize("q_y,q_z", [(8, 4), (12, 2), (6, 8), (4, 10)])
def test_any_stencil_pair_matches_oracle(oracle_lib, q_y, q_z):
    """Q-Value pairs other than the writer's specialised stencil sizes (the
    reference accepts any q >= 2, Parser.cpp:110-112) run through its
    run-time-sized instance: A_bc, B_bc, L and M equal the oracle's."""
    c = small_case(24, 12, q_y=q_y, q_z=q_z)
    print(f"stencils {q_y + 1} x {q_z + 1}", flush=True)
    ref, diag, got = assemble_both(c, oracle_lib)
    assert np.array_equal(got["colidx"], ref.colidx)
    for key in ("L", "M", "A", "B"):
        assert rel_err(got[key], getattr(ref, key)) <= 1e-12, key
    assert diag["incoming_modes"] == int(ref.bc_diag[2])


def test_solve_with_a_generic_stencil_pair(oracle_lib):
    """The eigen-solve on a 9 x 5 stencil case: the 8 eigenvalues nearest the
    target equal ARPACK's (SuperLU shift-invert) on the oracle matrices."""
    import scipy.sparse.linalg as spla
    c = small_case(24, 12, q_y=8, q_z=4)
    ref = oracle_lib.assemble(c)
    n = c.ny * c.nz
    A = oracle_lib.to_scipy(ref.rowptr, ref.colidx, ref.A, n).tocsc()
    B = oracle_lib.block_diag_to_scipy(ref.B).tocsc()
    lu = spla.splu(A.tocsc())
    op = spla.LinearOperator(A.shape, matvec=lambda v: lu.solve(B @ v), dtype=np.complex128)
    theta = spla.eigs(op, k=10, which="LM", tol=1e-14, return_eigenvectors=False)
    lam_ref = 1.0 / theta
    lam_ref = lam_ref[np.argsort(np.abs(lam_ref))][:8]
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        ctx.assemble(diagnostics=False)
        d = ctx.solve(nev=8, omega=0j, tol=1e-10)
        pairs = ctx.eigenpairs()[:8]
    lam = np.array([p["lambda"] for p in pairs])
    assert d["converged"] >= 8
    for l in lam_ref:
        assert np.min(np.abs(lam - l)) <= 1e-8 * abs(l), l
