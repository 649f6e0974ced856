"""Factorizations replayed from HIP graphs (tailor_factor_graph,
tailor_factor_batch) against the eager factorization (tailor_factor).

Every replay is checked: the graph is replayed at several shifts, in both
front precisions, with eager factorizations, solves and synchronizations of
the same contexts between the replays, and each replay's solution of
K x = b must equal the eager factorization's at the same shift (the same
kernels in the same order; the solves accumulate with atomics, so equality
is to rounding).  A batch factors contexts of two grid sizes in one graph
launch each (every context's graph on its own stream), three times, the
last after one context changed precision (its graph is captured again).
"""
import numpy as np
import pytest

from conftest import small_case
import tailor_amd as ta

pytestmark = pytest.mark.gpu


def solve(ctx, b):
    x, _, _ = ctx.linear_solve(b, refine_max=0)
    return x


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("precision", [ta.PRECISION_FP64, ta.PRECISION_MIXED])
def test_graph_replays_match_eager_factorizations(precision):
    c = small_case(40, 16)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(5 * c.nodes) + 1j * rng.standard_normal(5 * c.nodes)
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        ctx.assemble(diagnostics=False)
        ctx.set_precision(precision)
        for k, sigma in enumerate([-0.02j, 0.1 - 0.3j, -0.02j, 0.05 + 0.01j]):
            print(f"replay {k}: precision {precision}, sigma {sigma}", flush=True)
            de = ctx.factor(sigma)
            xe = solve(ctx, b)
            dg = ctx.factor_graph(sigma)
            xg = solve(ctx, b)
            assert dg["perturbed_pivots"] == de["perturbed_pivots"]
            assert rel(xg, xe) <= 1e-12, (k, rel(xg, xe))
            assert np.all(np.isfinite(xg))


def test_eigenpairs_after_a_graph_factorization():
    c = small_case(24, 12)
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        ctx.assemble(diagnostics=False)
        ctx.solve(nev=8, omega=c.omega_shift, tol=1e-10)
        eager = np.array([p["lambda"] for p in ctx.eigenpairs()[:8]])
        ctx.assemble(diagnostics=False)  # drops the factorization
        ctx.factor_graph(c.lambda_shift)
        d = ctx.solve(nev=8, omega=c.omega_shift, tol=1e-10)  # no refactorization
        graph = np.array([p["lambda"] for p in ctx.eigenpairs()[:8]])
    assert d["converged"] >= 8
    for lam in eager:
        assert np.min(np.abs(graph - lam)) <= 1e-10 * abs(lam)


def test_batch_factorization_matches_eager():
    cases = [small_case(40, 16), small_case(24, 12), small_case(40, 16, alpha=1.7 + 0j, seed=5)]
    rng = np.random.default_rng(4)
    bs = [rng.standard_normal(5 * c.nodes) + 1j * rng.standard_normal(5 * c.nodes) for c in cases]
    ctxs = [ta.Context(0) for _ in cases]
    try:
        for ctx, c in zip(ctxs, cases):
            ctx.load_case(c)
            ctx.assemble(diagnostics=False)
        rounds = [[-0.02j, 0.0j, 0.1 - 0.2j], [0.03 - 0.01j, -0.02j, 0.0j], [-0.02j, -0.02j, -0.02j]]
        for r, sigmas in enumerate(rounds):
            if r == 2:
                ctxs[1].set_precision(ta.PRECISION_MIXED)  # new graph for one context
            eager = []
            for ctx, b, s in zip(ctxs, bs, sigmas):
                ctx.factor(s)
                eager.append(solve(ctx, b))
            print(f"batch {r}: {len(ctxs)} contexts, shifts {sigmas}", flush=True)
            ms = ta.factor_batch(ctxs, sigmas)
            assert ms > 0
            for k, (ctx, b) in enumerate(zip(ctxs, bs)):
                xb = solve(ctx, b)
                assert rel(xb, eager[k]) <= 1e-12, (r, k, rel(xb, eager[k]))
        with pytest.raises(ta.TailorError):
            ta.factor_batch([ctxs[0], ctxs[0]], [0j, 0j])
    finally:
        for ctx in ctxs:
            ctx.close()
