"""Joined factorizations (tailor_batch_*): the fronts of several contexts
factored by one launch sequence over their joined elimination levels,
eagerly and replayed from the batch's HIP graph, against each context's own
factorization (tailor_factor) at the same shift.

Every replay is checked.  Contexts of one grid size run the same kernels on
their fronts in a batch as alone, so their solutions agree to the rounding
of the solves' atomic accumulation; contexts of different sizes share level
launches whose kernel choice follows the largest front of the level, so
theirs agree to the rounding of the factorization.  The eigenpairs computed
after a batch factorization are those of the context's own factorization.
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


def contexts_for(cases):
    ctxs = [ta.Context(0) for _ in cases]
    for ctx, c in zip(ctxs, cases):
        ctx.load_case(c)
        ctx.assemble(diagnostics=False)
    return ctxs


def own_solutions(ctxs, bs, sigmas):
    out = []
    for ctx, b, s in zip(ctxs, bs, sigmas):
        ctx.factor(s)
        out.append(solve(ctx, b))
    return out


@pytest.mark.parametrize("precision", [ta.PRECISION_FP64, ta.PRECISION_MIXED])
def test_same_size_batch_matches_own_factorizations(precision):
    cases = [small_case(40, 16, alpha=a, seed=s) for a, s in ((2.43, 0), (1.1, 3), (0.7, 9))]
    rng = np.random.default_rng(11)
    bs = [rng.standard_normal(5 * c.nodes) + 1j * rng.standard_normal(5 * c.nodes) for c in cases]
    ctxs = contexts_for(cases)
    try:
        for ctx in ctxs:
            ctx.set_precision(precision)
            ctx.set_gemm_config(3)
        with ta.Batch(ctxs) as batch:
            batch.set_gemm_config(3)
            rounds = [([-0.02j, 0.0j, 0.1 - 0.2j], True), ([0.03 - 0.01j, -0.02j, 0.0j], True),
                      ([-0.02j, 0.05j, 0.01 - 0.01j], False), ([0.0j, 0.0j, -0.02j], True)]
            for r, (sigmas, graph) in enumerate(rounds):
                own = own_solutions(ctxs, bs, sigmas)
                print(f"batch round {r}: precision {precision}, graph {graph}, shifts {sigmas}",
                      flush=True)
                ms = batch.factor(sigmas, graph=graph)
                assert ms > 0
                for k, (ctx, b) in enumerate(zip(ctxs, bs)):
                    xb = solve(ctx, b)
                    assert np.all(np.isfinite(xb))
                    assert rel(xb, own[k]) <= 1e-12, (r, k, rel(xb, own[k]))
    finally:
        for ctx in ctxs:
            ctx.close()


def test_mixed_size_batch_matches_own_factorizations():
    cases = [small_case(40, 16), small_case(24, 12, alpha=1.3), small_case(64, 32, alpha=0.9, seed=4)]
    rng = np.random.default_rng(12)
    bs = [rng.standard_normal(5 * c.nodes) + 1j * rng.standard_normal(5 * c.nodes) for c in cases]
    ctxs = contexts_for(cases)
    try:
        with ta.Batch(ctxs) as batch:
            for r, sigmas in enumerate([[-0.02j, 0.1 - 0.2j, 0.0j], [0.0j, -0.02j, 0.02 - 0.05j]]):
                own = own_solutions(ctxs, bs, sigmas)
                print(f"mixed-size batch round {r}: shifts {sigmas}", flush=True)
                batch.factor(sigmas, graph=True)
                for k, (ctx, b) in enumerate(zip(ctxs, bs)):
                    xb = solve(ctx, b)
                    assert rel(xb, own[k]) <= 1e-10, (r, k, rel(xb, own[k]))
            tune = batch.autotune(-0.02j)
            assert tune["levels"] >= 1 and len(tune["config"]) == tune["levels"]
            sigmas = [0.01j, -0.02j, 0.0j]
            own = own_solutions(ctxs, bs, sigmas)
            batch.factor(sigmas, graph=True)  # captured again for the tuned tiles
            for k, (ctx, b) in enumerate(zip(ctxs, bs)):
                assert rel(solve(ctx, b), own[k]) <= 1e-10
    finally:
        for ctx in ctxs:
            ctx.close()


def test_eigenpairs_after_a_batch_factorization():
    cases = [small_case(24, 12), small_case(24, 12, alpha=1.9 + 0j, seed=2)]
    ctxs = contexts_for(cases)
    try:
        own = []
        for ctx, c in zip(ctxs, cases):
            ctx.solve(nev=8, omega=c.omega_shift, tol=1e-10)
            own.append(np.array([p["lambda"] for p in ctx.eigenpairs()[:8]]))
            ctx.assemble(diagnostics=False)  # drops the factorization
        with ta.Batch(ctxs) as batch:
            batch.factor([c.lambda_shift for c in cases])
            for k, (ctx, c) in enumerate(zip(ctxs, cases)):
                d = ctx.solve(nev=8, omega=c.omega_shift, tol=1e-10)  # no refactorization
                lam = np.array([p["lambda"] for p in ctx.eigenpairs()[:8]])
                assert d["converged"] >= 8
                for x in own[k]:
                    assert np.min(np.abs(lam - x)) <= 1e-10 * abs(x), (k, x)
    finally:
        for ctx in ctxs:
            ctx.close()


def test_batch_rejects_changed_structures_and_mixed_precisions():
    cases = [small_case(24, 12), small_case(24, 12, alpha=1.5 + 0j)]
    ctxs = contexts_for(cases)
    try:
        with pytest.raises(ta.TailorError):
            ta.Batch([ctxs[0], ctxs[0]])
        with ta.Batch(ctxs) as batch:
            ctxs[1].set_precision(ta.PRECISION_MIXED)
            with pytest.raises(ta.TailorError, match="one front precision"):
                batch.factor([0j, 0j])
            ctxs[1].set_precision(ta.PRECISION_FP64)
            with pytest.raises(ValueError):
                batch.factor([0j])
            ctxs[1].load_case(small_case(40, 16))
            ctxs[1].assemble(diagnostics=False)
            with pytest.raises(ta.TailorError, match="another structure"):
                batch.factor([0j, 0j])
    finally:
        for ctx in ctxs:
            ctx.close()


def test_mixed_size_mixed_precision_sweep_through_joined_factorizations():
    """Config 3 at small scale through joined factorizations: instances of
    both grid sizes with their own base flows in groups of two (each group
    one 100x50 and one 200x100 instance, factored by one launch sequence
    replayed from the batch's graph), complex FP32 fronts with every solve
    refined to 1e-12.  Every pair meets the tolerance and equals the plain
    FP64 sweep's eigenvalue within 1e-8."""
    from tailor_amd import sweep as sw
    insts = sw.sweep_instances(4, seed=1, sizes=[(100, 50), (200, 100)], per_instance_seeds=True)
    cases = sw.prepare(insts)
    print("mixed-precision sweep: 2 x 100x50 + 2 x 200x100, joined pairs", flush=True)
    runner = sw.Sweeper(ncv=96, batch=2)
    mixed = runner.run(cases, "mixed")
    runner.close()
    plain_runner = sw.Sweeper(ncv=96)
    plain = plain_runner.run(cases, "plain")
    plain_runner.close()
    print(f"mixed {mixed.totals()}\nplain {plain.totals()}", flush=True)
    assert mixed.good_pairs == 8 * len(insts) and plain.good_pairs == 8 * len(insts)
    assert mixed.totals()["refinement_steps"] >= mixed.totals()["solves"]
    by = {id(r.instance): r for r in plain.records}
    for a in mixed.records:
        b = by[id(a.instance)]
        for w in b.omegas:
            assert np.min(np.abs(a.omegas - w)) <= 1e-8 * abs(w), (a.instance, w)


def test_batch_assembly_and_factorization_in_one_graph():
    """tailor_batch_assemble_factor: every context's assembly and the joined
    factorization replayed from one graph, across wavenumber changes (the
    wavenumber is read from each case's parameter block, written before each
    replay).  A_bc and B_bc equal a separate context's own assembly bit for
    bit, and the solutions equal the own factorization's."""
    cases = [small_case(40, 16, alpha=2.43), small_case(40, 16, alpha=1.2, seed=6),
             small_case(24, 12, alpha=0.8)]
    rng = np.random.default_rng(13)
    bs = [rng.standard_normal(5 * c.nodes) + 1j * rng.standard_normal(5 * c.nodes) for c in cases]
    ctxs = [ta.Context(0) for _ in cases]
    refs = [ta.Context(0) for _ in cases]
    try:
        for ctx, ref, c in zip(ctxs, refs, cases):
            ctx.load_case(c)
            ref.load_case(c)
        with ta.Batch(ctxs) as batch:
            for r, (alphas, sigmas) in enumerate([([2.43, 1.2, 0.8], [-0.02j, 0.0j, 0.05 - 0.1j]),
                                                  ([1.9, 0.6, 2.7], [0.0j, -0.02j, -0.02j]),
                                                  ([1.9, 0.6, 2.7], [0.01j, 0.0j, 0.0j])]):
                for ctx, ref, a in zip(ctxs, refs, alphas):
                    ctx.set_alpha(a)
                    ref.set_alpha(a)
                    ref.assemble(diagnostics=False)
                print(f"assemble+factor round {r}: alphas {alphas}, shifts {sigmas}", flush=True)
                batch.assemble_factor(sigmas, graph=True)
                for k, (ctx, ref, b, s) in enumerate(zip(ctxs, refs, bs, sigmas)):
                    for op in (ta.OP_A, ta.OP_B):
                        assert np.array_equal(ctx.export_bsr(op)[2], ref.export_bsr(op)[2]), (r, k, op)
                    xb = solve(ctx, b)
                    ref.factor(s)
                    assert rel(xb, solve(ref, b)) <= 1e-12, (r, k)
    finally:
        for ctx in ctxs + refs:
            ctx.close()


def test_batch_assembly_reports_invalid_cases():
    good, bad = small_case(24, 12), small_case(24, 12, alpha=1.4)
    bad.baseflow = bad.baseflow.copy()
    bad.baseflow[7, 4] = -1.0
    ctxs = [ta.Context(0), ta.Context(0)]
    try:
        ctxs[0].load_case(good)
        ctxs[1].load_case(bad)
        with ta.Batch(ctxs) as batch:
            with pytest.raises(ta.TailorError, match="context 1 of the batch: Base-flow density "
                                                     "and temperature must both be positive"):
                batch.assemble_factor([0j, 0j])
    finally:
        for ctx in ctxs:
            ctx.close()
