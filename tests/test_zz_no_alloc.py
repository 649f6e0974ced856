"""tailor_linear_solve_noalloc (no device allocation) against
tailor_linear_solve, both called through the C ABI.

Two contexts are solved in a row, the second with the smaller case, so the
second call reuses the device's scratch array after a larger call: for FP64
and complex-FP32 factors and 0 or 2 refinement steps the device-buffer
solution must equal the host-buffer entry point's (same kernels, same order;
the triangular solves accumulate with atomics, so equality is to rounding)
and solve K x = b.  Sizes above the scratch limit and bad arguments fail
with an error and leave the context usable.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import small_case
import tailor_amd as ta

pytestmark = pytest.mark.gpu


def noalloc_solve(ctx, b, steps):
    import torch
    bt = torch.from_numpy(np.ascontiguousarray(b).view(np.float64).copy()).to("cuda:0")
    xt = torch.full_like(bt, float("nan"))
    torch.cuda.synchronize()  # the library's stream does not wait for torch's
    st = ctx.lib.tailor_linear_solve_noalloc(ctx.ptr, C.c_void_p(bt.data_ptr()),
                                             C.c_void_p(xt.data_ptr()), steps)
    if st:
        raise ta.TailorError(st, ctx.lib.tailor_last_error(ctx.ptr).decode())
    return xt.cpu().numpy().view(np.complex128)


def operator(oracle_lib, c, sigma):
    a = oracle_lib.assemble(c)
    n = c.ny * c.nz
    A = oracle_lib.to_scipy(a.rowptr, a.colidx, a.A, n).tocsr()
    B = oracle_lib.block_diag_to_scipy(a.B).tocsr()
    return (A - sigma * B).tocsr()


@pytest.mark.parametrize("precision", [ta.PRECISION_FP64, ta.PRECISION_MIXED])
def test_noalloc_solve_matches_linear_solve(oracle_lib, precision):
    rng = np.random.default_rng(11)
    ctxs = []
    try:
        # the second call is the smaller case
        for (ny, nz), omega in (((40, 16), 0.02 + 0j), ((24, 12), 0j)):
            c = small_case(ny, nz)
            ctx = ta.Context(0)
            ctxs.append(ctx)
            ctx.load_case(c)
            ctx.assemble(diagnostics=False)
            ctx.set_precision(precision)
            sigma = complex(omega.imag, -omega.real)
            ctx.factor(sigma)
            K = operator(oracle_lib, c, sigma)
            n = 5 * c.nodes
            b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
            for steps in (0, 2):
                print(f"noalloc solve {ny}x{nz}, precision {precision}, {steps} steps", flush=True)
                x = noalloc_solve(ctx, b, steps)
                xref, _, done = ctx.linear_solve(b, refine_max=steps, refine_tol=0.0)
                assert done == steps
                assert np.all(np.isfinite(x))
                diff = np.linalg.norm(x - xref) / np.linalg.norm(xref)
                res = np.linalg.norm(K @ x - b) / np.linalg.norm(b)
                print(f"  |x - x_ref| / |x_ref| = {diff:.2e}, residual {res:.2e}", flush=True)
                assert diff <= 1e-12
                if precision == ta.PRECISION_FP64:
                    assert res <= 1e-10
                elif steps == 2:
                    assert res <= 1e-8
    finally:
        for ctx in ctxs:
            ctx.close()


def test_noalloc_solve_rejects_oversized_cases_and_bad_arguments():
    from tailor_amd import case as cs
    big = cs.synthetic_case(100, 530)  # 265,000 unknowns > 262,144
    with ta.Context(0) as ctx:
        ctx.load_case(big)
        print("noalloc solve with 265,000 unknowns (expects a size error)", flush=True)
        st = ctx.lib.tailor_linear_solve_noalloc(ctx.ptr, C.c_void_p(1), C.c_void_p(1), 0)
        assert st == 3  # TAILOR_ERR_ARG_OUTOFRANGE, before any device access
        assert "scratch limit of 262144" in ctx.lib.tailor_last_error(ctx.ptr).decode()
    c = small_case(24, 12)
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        ctx.assemble(diagnostics=False)
        st = ctx.lib.tailor_linear_solve_noalloc(ctx.ptr, C.c_void_p(1), C.c_void_p(1), 0)
        assert st == 8  # not factored yet
        ctx.factor(0j)
        for steps in (-1, 65):
            assert ctx.lib.tailor_linear_solve_noalloc(ctx.ptr, C.c_void_p(1), C.c_void_p(1),
                                                       steps) == 3
        assert ctx.lib.tailor_linear_solve_noalloc(ctx.ptr, None, None, 0) == 7
        # still usable after the rejected calls
        b = np.ones(5 * c.nodes, dtype=np.complex128)
        x = noalloc_solve(ctx, b, 1)
        xref, _, _ = ctx.linear_solve(b, refine_max=1, refine_tol=0.0)
        assert np.linalg.norm(x - xref) <= 1e-12 * np.linalg.norm(xref)
