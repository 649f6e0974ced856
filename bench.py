"""Benchmark of the BiGlobal assemble-and-solve path on one MI355X.

Headline (`value`): eigenpairs/s of a wavenumber sweep on the 200x100 case
(FD-q q_y=10 bounded x q_z=6 periodic, 100,000 complex unknowns, Re=1500,
Ma=0.7, alpha ~ U[0.5, 3] from seed 1, a base flow per instance from seeds
1000 + k) in the learned-shift form: the shift network proposes each
instance's shift centre, instances run in wavenumber order with warm-started
Krylov-Schur, passes end as soon as the wanted pairs' Ritz estimates meet
tol, the fixed-shift search is the fallback, and eight contexts pipeline the
sweep (the next seven instances' assembly and factorization run while the
current instance's Krylov-Schur runs).  A step is one sweep instance
(assembly of A_bc, B_bc, sparse LU of A_bc - sigma B_bc, Krylov-Schur for 8
eigenpairs); an eigenpair counts if its residual ||A q - lambda B q|| / ||q||
and its relative error (the same over |lambda|) are <= 1e-10.

The same instances also run
* in the plain form as the reference runs it (complex FP64, fixed
  omega_s = 0.02, cold start, ncv = max(2 nev, nev + 15) = 23, one instance
  after another);
* in the plain form at its fastest subspace size, through the same
  eight-context pipeline and through joined groups of four (each group's
  assembly and joined factorization replayed from one HIP graph); the
  headline ratio is taken against the strongest plain form;
* with the network switched off (fixed shift, same warm starts, early pass
  ends and pipeline): the learned form's rate over this one is what the
  network itself gains.

Also reported: the plain single instance at alpha = 2.43 (Re=1000, Ma=0.5)
at several subspace sizes, the mixed-precision form on that operator
(per-eigenvalue error against the FP64 result, refinement steps per solve),
a sweep over both grid sizes with per-instance base flows in mixed
precision through joined groups (and through the pipeline beside it), the
factorization throughput of joined groups against one factorization after
another, assembly, triangular-solve and SpMV HBM GB/s, the roofline of the
frontal-update GEMM (the kernel with the largest share of a sweep instance)
from HIP events on the library stream, and the CPU oracle on the 100x50 case
beside the GPU on the same case.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

# The sweeps drive two HIP streams per context (the factorization's main and
# look-ahead streams).  With HIP's default of four hardware queues per
# process, unrelated streams share queues and wait for each other: measured
# at 200x100 with three contexts, 199 ms per instance at 4 queues against
# 185 ms at 8 or 16.  24 covers the eight pipelined contexts' streams and
# the one-context plain form's.  Read by HIP when it initialises, so set
# here, before anything touches the GPU (also where the environment sets the
# default of 4).
os.environ["GPU_MAX_HW_QUEUES"] = "24"


import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
NEV, TOL, OMEGA = 8, 1e-10, 0.02 + 0j
REFERENCE_NCV = max(2 * NEV, NEV + 15)  # SLEPc's default subspace (EigenSolver.cpp:55-56)
PLAIN_NCV_CANDIDATES = (REFERENCE_NCV, 32, 48, 64, 96)
LEARNED_NCV = 40  # subspace of the learned form: the wanted cluster converges in one pass
LEARNED_CHECK_FROM = 28  # learned form: convergence also tested every second step from step 28
PIPELINE_CONTEXTS = 8  # Krylov-Schur of one instance, factorization of the next seven
MIXED_SIZE_INSTANCES = 4  # sweep over both grid sizes in mixed precision
JOINED_GROUP = 4  # instances per joined assembly + factorization (tailor_batch)


_T0 = time.perf_counter()


def log(msg: str) -> None:
    """A progress line on stderr as each phase ends (stdout carries only the
    JSON result)."""
    print(f"bench [{time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def algorithmic_assembly_bytes(ny: int, nz: int, ssx: int, sse: int) -> int:
    """A_bc and B_bc written (16 B x 25 per block), grid + base flow read
    (8 complex per node = 128 B) and the FD-q weights."""
    nodes = ny * nz
    nnzb_a = nodes * ssx * sse
    return 16 * 25 * (nnzb_a + nodes) + 128 * nodes + 3 * 8 * (ny * ssx + nz * sse)


def measured_traffic(kernel: str):
    """HBM bytes per dispatch of `kernel` from the committed counter runs
    (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a 1 GiB copy), or None."""
    path = ROOT / "profiles" / "hbm_traffic_200x100.json"
    if not path.exists():
        return None
    k = json.loads(path.read_text()).get("kernels", {}).get(kernel)
    return k["traffic_bytes_per_dispatch"] if k else None


def traced_gemm_window():
    """Mean k_gemm duration inside tailor_profile_factor's window from the
    committed kernel-trace summary of this benchmark (rocprofv3), or None."""
    import re
    path = ROOT / "profiles" / "bench_sweep_200x100_kernel_stats.md"
    if not path.exists():
        return None
    m = re.search(r"Frontal GEMM \(k_gemm and k_gemm_lds, all tile shapes\): (\d+) launches, "
                  r"([0-9.]+) ms, "
                  r"mean ([0-9.]+) us per launch", path.read_text())
    if not m:
        return None
    return {"launches": int(m.group(1)), "ms": float(m.group(2)), "us_per_launch": float(m.group(3)),
            "source": str(path.relative_to(ROOT))}


def host_facts() -> dict:
    """The host the CPU baseline runs on."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"nproc": usable, "cpu_model": model, "mpi_ranks": 1}


def cpu_baseline(ncv: int) -> dict:
    """The CPU oracle (C assembly + multifrontal LU on LAPACK/BLAS + numpy
    Krylov-Schur, the same algorithm as the GPU path) on the 100x50 case of
    the same generator (omega_s = 0, the reference's settings: nev 8, tol
    1e-10, ncv 23), with the assembly, factorization and Krylov-Schur times
    apart.  PETSc, SLEPc and MUMPS are not installed on the GPU host, so the
    reference's own CPU path cannot run beside it."""
    import oracle
    from oracle import eigen, multifrontal
    from tailor_amd import case as cs
    oracle.build()
    c = cs.synthetic_case(100, 50, omega_shift=0j)
    n = c.nodes
    sigma = eigen.target_lambda(c.omega_shift)
    t0 = time.perf_counter()
    a = oracle.assemble(c)
    t1 = time.perf_counter()
    per = len(a.colidx) // n
    kvals = a.A.copy()
    kvals[a.colidx == np.repeat(np.arange(n), per)] -= sigma * a.B
    lu = multifrontal.MultifrontalLU(c, a.rowptr, a.colidx, kvals)
    t2 = time.perf_counter()
    A = oracle.to_scipy(a.rowptr, a.colidx, a.A, n)
    B = oracle.block_diag_to_scipy(a.B)
    lams, _, res, restarts = eigen.krylov_schur(A, B, sigma, nev=NEV, tol=TOL, ncv=ncv,
                                                solver=lu.solve, true_residual=True)
    t3 = time.perf_counter()
    rel = res[:NEV] / np.maximum(np.abs(lams[:NEV]), 1e-300)
    good = int(np.sum((res[:NEV] <= TOL) & (rel <= TOL)))
    threads = multifrontal.blas_threads()
    facts = host_facts()
    return {"value": round(good / (t3 - t0), 4), "unit": "eigenpairs/s", "cores": threads,
            "kind": "port",
            "sample": f"one assemble + factor + Krylov-Schur of the 100x50 case (25,000 unknowns, "
                      f"omega_s = 0, nev 8, tol 1e-10, ncv {ncv}): {t3 - t0:.1f} s, {good} of "
                      f"{NEV} pairs at tol, {restarts} restarts",
            "assembly_s": round(t1 - t0, 3), "factor_s": round(t2 - t1, 3),
            "krylov_s": round(t3 - t2, 3), "blas_threads": threads, **facts,
            "lambdas": [[float(l.real), float(l.imag)] for l in lams[:NEV]]}


def gpu_small_case(ta, cs, ncv: int) -> dict:
    """The plain FP64 GPU form on the CPU baseline's 100x50 case."""
    c = cs.synthetic_case(100, 50, omega_shift=0j)
    with ta.Context(0) as ctx:
        ctx.load_case(c)
        ctx.assemble(diagnostics=False)
        ctx.solve(nev=NEV, omega=c.omega_shift, tol=TOL, ncv=ncv)  # plan + warm-up
        t0 = time.perf_counter()
        ctx.assemble(diagnostics=False)
        d = ctx.solve(nev=NEV, omega=c.omega_shift, tol=TOL, ncv=ncv)
        el = time.perf_counter() - t0
        pairs = ctx.eigenpairs()[:NEV]
    good = sum(1 for p in pairs if p["residual"] <= TOL and p["relative_error"] <= TOL)
    return {"eigenpairs_per_s": round(good / el, 4), "seconds": round(el, 4), "ncv": ncv,
            "factor_ms": round(d["factor_ms"], 2), "krylov_ms": round(d["solve_ms"], 2),
            "restarts": d["iterations"], "solves": d["solves"],
            "lambdas": [[p["lambda"].real, p["lambda"].imag] for p in pairs]}


def plain_step(ta, cs) -> dict:
    """The plain FP64 form on one instance (alpha 2.43, omega_s 0.02) at
    several subspace sizes, the kernels' rooflines on its operator, and the
    mixed-precision form on the same operator."""
    case = cs.synthetic_case(200, 100, omega_shift=OMEGA)
    ctx = ta.Context(0)
    ctx.load_case(case)
    ctx.assemble(diagnostics=False)
    ctx.time_assembly(3)
    write_ms, asm_ms = ctx.time_assembly(20)
    asm_bytes = algorithmic_assembly_bytes(200, 100, 11, 7)

    def step(ncv):
        t0 = time.perf_counter()
        ctx.assemble(diagnostics=False)
        d = ctx.solve(nev=NEV, omega=OMEGA, tol=TOL, ncv=ncv, raise_on_failure=False)
        el = time.perf_counter() - t0
        pairs = ctx.eigenpairs()[:NEV]
        good = sum(1 for p in pairs if p["residual"] <= TOL and p["relative_error"] <= TOL)
        return el, good, d, pairs

    tune = ctx.autotune(case.lambda_shift)  # frontal GEMM tile shape and order per level
    log("plain instance: autotuned")
    step(REFERENCE_NCV)  # warm-up
    by_ncv = {}
    for ncv in PLAIN_NCV_CANDIDATES:
        el, good, d, pairs = step(ncv)
        by_ncv[ncv] = {"seconds": round(el, 4), "eigenpairs_per_s": round(good / el, 4),
                       "restarts": d["iterations"], "solves": d["solves"],
                       "factor_ms": round(d["factor_ms"], 2), "krylov_ms": round(d["solve_ms"], 2)}
    best_ncv = max(by_ncv, key=lambda k: by_ncv[k]["eigenpairs_per_s"])
    log(f"plain instance: subspace sizes {list(by_ncv)} timed, best {best_ncv}")
    el, good, d, pairs = step(REFERENCE_NCV)
    lam64 = np.array([p["lambda"] for p in pairs])
    solve_ms, solve_bytes = ctx.time_linear_solve(20)
    spmv_ms, spmv_bytes = ctx.time_operator(50)
    prof = ctx.profile_factor(case.lambda_shift)
    log("plain instance: solve, SpMV and frontal GEMM timed")
    # mixed precision on the same operator and shift
    ctx.set_precision(ta.PRECISION_MIXED)
    t0 = time.perf_counter()
    dm = ctx.solve(nev=NEV, omega=OMEGA, tol=TOL, ncv=REFERENCE_NCV, raise_on_failure=False)
    mixed_s = time.perf_counter() - t0
    pm = ctx.eigenpairs()[:NEV]
    lam32 = np.array([p["lambda"] for p in pm])
    errs = [float(np.min(np.abs(lam32 - l)) / abs(l)) for l in lam64]
    solve32_ms, solve32_bytes = ctx.time_linear_solve(20)
    ctx.close()
    log("plain instance: mixed precision done")
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9
    return {
        "plain": {"eigenpairs_per_s": round(good / el, 4), "seconds": round(el, 4),
                  "ncv": REFERENCE_NCV, "factor_ms": round(d["factor_ms"], 2),
                  "krylov_ms": round(d["solve_ms"], 2), "solves": d["solves"],
                  "restarts": d["iterations"], "max_residual": d["max_residual"],
                  "max_relative_error": d["max_relative_error"],
                  "factor_TFLOPs_lu_flops": round(d["factor_flops"] / (d["factor_ms"] * 1e-3) / 1e12, 3),
                  "by_ncv": by_ncv, "best_ncv": best_ncv},
        "assembly": {"GBps_writer": round(gbs(asm_bytes, write_ms), 2),
                     "GBps_full": round(gbs(asm_bytes, asm_ms), 2),
                     "frac_hbm_writer": round(gbs(asm_bytes, write_ms) / HBM_PEAK_GBS, 4),
                     "frac_hbm_full": round(gbs(asm_bytes, asm_ms) / HBM_PEAK_GBS, 4),
                     "writer_ms": round(write_ms, 5), "full_ms": round(asm_ms, 5),
                     "algorithmic_bytes": asm_bytes},
        "solve_sweep": {"kernel": "k_gemv (triangular solves), FP64 factors",
                        "ms_per_solve": round(solve_ms, 4), "bytes": int(solve_bytes),
                        "GBps": round(gbs(solve_bytes, solve_ms), 2),
                        "frac_hbm": round(gbs(solve_bytes, solve_ms) / HBM_PEAK_GBS, 4),
                        "fp32_factors": {"ms_per_solve": round(solve32_ms, 4),
                                         "bytes": int(solve32_bytes),
                                         "GBps": round(gbs(solve32_bytes, solve32_ms), 2),
                                         "frac_hbm": round(gbs(solve32_bytes, solve32_ms) / HBM_PEAK_GBS, 4)}},
        "spmv": {"kernel": "k_kmul (y = (A_bc - sigma B_bc) x)", "ms": round(spmv_ms, 5),
                 "bytes": int(spmv_bytes), "GBps": round(gbs(spmv_bytes, spmv_ms), 2),
                 "frac_hbm": round(gbs(spmv_bytes, spmv_ms) / HBM_PEAK_GBS, 4)},
        "gemm": prof,
        "tune": tune,
        "mixed": {"seconds": round(mixed_s, 3), "factor_ms": round(dm["factor_ms"], 2),
                  "solves": dm["solves"], "refinement_steps": dm["refinement_steps"],
                  "refinement_steps_per_solve": round(dm["refinement_steps"] / max(1, dm["solves"]), 2),
                  "converged": dm["converged"], "max_residual": dm["max_residual"],
                  "eig_rel_err_vs_fp64": [float(f"{e:.3e}") for e in errs],
                  "fp64_seconds_same_operator": round(el, 3), "ncv": REFERENCE_NCV},
    }


def without_assembly(totals: dict, assembly_ms: float) -> dict:
    """Add the sweep's rate without assembly: its wall time less one full
    assembly (A_bc and B_bc, timed separately on the same case) per
    instance."""
    sec = totals["seconds"] - totals["instances"] * assembly_ms * 1e-3
    out = dict(totals)
    out["assembly_ms_per_instance"] = round(assembly_ms, 4)
    out["eigenpairs_per_s_without_assembly"] = round(totals["good_pairs"] / sec, 4) if sec > 0 else None
    return out


def headline_instances(ta, steps: int, warmup: int):
    """The headline sweep's instances (200x100, Re 1500, Ma 0.7, alpha ~
    U[0.5, 3] from seed 1, base-flow seeds 1000 + k), prepared on the host,
    and the network's shift centres for them."""
    from tailor_amd import sweep as sw
    from tailor_amd.shift_model import ShiftModel
    from tailor_amd.train_shift_model import MODEL
    if not MODEL.exists():
        raise SystemExit(f"{MODEL} is missing: run `python -m tailor_amd.train_shift_model`")
    model = ShiftModel(0).load(MODEL)
    insts = sw.sweep_instances(warmup + steps, seed=1, sizes=[(200, 100)], reynolds=1500.0,
                               mach=0.7, per_instance_seeds=True)
    cases = sw.prepare(insts, OMEGA)
    alphas = np.array([i.alpha for i in insts])
    shifts = model.shift_centres(alphas, np.full(len(insts), 1500.0), np.full(len(insts), 0.7))
    warm = list(range(warmup))
    timed = list(range(warmup, warmup + steps))
    return insts, cases, shifts, warm, timed


def record_dicts(result) -> list:
    return [{"seed": r.instance.seed, "alpha": r.instance.alpha, "size": f"{r.instance.ny}x{r.instance.nz}",
             "good": r.good, "seconds": r.seconds, "solves": r.solves,
             "omegas": [[w.real, w.imag] for w in r.omegas],
             "residuals": [float(x) for x in r.residuals]} for r in result.records]


def pair_overlap(a: list, b: list, rtol=1e-8) -> dict:
    """Per instance: how many of sweep a's converged pairs sweep b also found
    (same eigenvalue within rtol), and the largest difference among them."""
    by = {(r["seed"], r["alpha"]): r for r in b}
    counts, worst = [], 0.0
    for r in a:
        p = by[(r["seed"], r["alpha"])]
        good_b = np.array([complex(*w) for w, x in zip(p["omegas"], p["residuals"]) if x <= TOL])
        n = 0
        for w, x in zip(r["omegas"], r["residuals"]):
            w = complex(*w)
            if x > TOL or len(good_b) == 0:
                continue
            d = float(np.min(np.abs(good_b - w)) / max(abs(w), 1e-300))
            if d <= rtol:
                n += 1
                worst = max(worst, d)
        counts.append(n)
    return {"matched_per_instance": counts, "matched_total": int(sum(counts)),
            "max_rel_diff": worst}


def phase_pipeline(ta, steps, warmup, best_ncv, learned_ncv, contexts) -> dict:
    """The learned form, the network switched off and the plain form at its
    fastest subspace, all through the same pipeline of contexts."""
    from tailor_amd import sweep as sw
    insts, cases, shifts, warm, timed = headline_instances(ta, steps, warmup)
    pipe = sw.Sweeper(ncv=learned_ncv, omega_s=OMEGA, contexts=contexts,
                      check_from=LEARNED_CHECK_FROM)
    # warm-up instances (untimed): build the plans, tune, and let the first
    # one choose the factorization precision of the learned form by
    # measurement
    choice = {}
    precision = ta.PRECISION_FP64
    if warm:
        t = {}
        for prec in (ta.PRECISION_FP64, ta.PRECISION_MIXED):
            r = pipe.run(cases, "learned", shifts=shifts, precision=prec, order=warm[:1])
            t[prec] = r.seconds
        precision = min(t, key=t.get)
        choice = {"fp64_s": round(t[ta.PRECISION_FP64], 3), "mixed_s": round(t[ta.PRECISION_MIXED], 3),
                  "chosen": "mixed" if precision == ta.PRECISION_MIXED else "fp64"}
    log(f"pipeline: contexts planned and tuned, precision {choice.get('chosen', 'fp64')}")
    order = sorted(timed, key=lambda k: insts[k].alpha)
    learned = pipe.run(cases, "learned", shifts=shifts, precision=precision, order=order)
    log(f"learned sweep: {learned.eigenpairs_per_s:.2f} eigenpairs/s")
    # the network switched off: the same pipeline, warm starts and early
    # pass ends at the fixed shift
    fixed = pipe.run(cases, "fixed", precision=precision, order=order)
    log(f"network-off sweep: {fixed.eigenpairs_per_s:.2f} eigenpairs/s")
    # the plain form at its fastest subspace through the same pipeline
    pipe.ncv, pipe.check_from = best_ncv, 0
    plain_best = pipe.run(cases, "plain", precision=ta.PRECISION_FP64, order=timed)
    log(f"plain sweep (ncv {best_ncv}, pipelined): {plain_best.eigenpairs_per_s:.2f} eigenpairs/s")
    pipe.close()
    lr, pr = record_dicts(learned), record_dicts(plain_best)
    return {"learned": learned.totals(), "fixed": fixed.totals(), "plain_best": plain_best.totals(),
            "choice": choice, "overlap": pair_overlap(lr, pr),
            "autotune": {f"{k[0]}x{k[1]}": v for k, v in pipe.tuned.items()}}


def phase_reference(ta, steps, warmup) -> dict:
    """The plain form as the reference runs it: ncv 23, one instance after
    another, on one context."""
    from tailor_amd import sweep as sw
    insts, cases, shifts, warm, timed = headline_instances(ta, steps, warmup)
    ref = sw.Sweeper(ncv=REFERENCE_NCV, omega_s=OMEGA)
    ref.run(cases, "plain", precision=ta.PRECISION_FP64, order=warm[:1])  # plan, tune, warm-up
    plain_ref = ref.run(cases, "plain", precision=ta.PRECISION_FP64, order=timed)
    ref.close()
    log(f"plain sweep (reference settings): {plain_ref.eigenpairs_per_s:.2f} eigenpairs/s")
    return {"plain_ref": plain_ref.totals()}


def phase_joined_plain(ta, steps, warmup, best_ncv, threads=True) -> dict:
    """The plain form at its fastest subspace through joined groups of
    JOINED_GROUP instances (assembly and joined factorization of a group
    replayed from one HIP graph, the group's Krylov-Schur runs side by side
    while the next group is factored)."""
    from tailor_amd import sweep as sw
    insts, cases, shifts, warm, timed = headline_instances(ta, steps, warmup)
    runner = sw.Sweeper(ncv=best_ncv, omega_s=OMEGA, batch=JOINED_GROUP, threads=threads)
    runner.run(cases, "plain", precision=ta.PRECISION_FP64, order=warm[:1])  # plans, graphs
    res = runner.run(cases, "plain", precision=ta.PRECISION_FP64, order=timed)
    runner.close()
    log(f"plain sweep (ncv {best_ncv}, joined groups of {JOINED_GROUP}): "
        f"{res.eigenpairs_per_s:.2f} eigenpairs/s")
    return {"plain_joined": res.totals()}


def phase_mixed_size(ta, count, ncv, contexts, joined=True) -> dict:
    """Wavenumber sweep over both grid sizes (half 100x50, half 200x100),
    base flows from seeds 1000 + k, complex FP32 fronts with FP64
    refinement of every solve to 1e-12, fixed shift: through joined groups
    of JOINED_GROUP instances of a size (assembly and joined factorization
    of a group replayed from one HIP graph), or (joined=False) through the
    pipeline of contexts.  Each size's half runs as its own timed sweep, so
    the per-size rates are wall-clock rates; the whole sweep's rate is their
    combination."""
    from tailor_amd import sweep as sw
    c3 = sw.sweep_instances(count + 2, seed=1, sizes=[(100, 50), (200, 100)])
    cases = sw.prepare(c3, OMEGA)
    runner = (sw.Sweeper(ncv=ncv, omega_s=OMEGA, batch=JOINED_GROUP, threads=contexts > 1)
              if joined else sw.Sweeper(ncv=ncv, omega_s=OMEGA, contexts=contexts))
    runner.run(cases, "mixed", order=[0, 1])  # plans and tuning of both sizes
    per_size, good, secs = {}, 0, 0.0
    totals = {"instances": 0, "factorizations": 0, "restarts": 0, "solves": 0,
              "refinement_steps": 0, "fallbacks": 0}
    for ny, nz in ((100, 50), (200, 100)):
        idx = [k for k in range(2, count + 2) if (c3[k].ny, c3[k].nz) == (ny, nz)]
        res = runner.run(cases, "mixed", order=idx)
        t = res.totals()
        per_size[f"{ny}x{nz}"] = {"instances": t["instances"], "good_pairs": t["good_pairs"],
                                  "seconds": t["seconds"], "eigenpairs_per_s": t["eigenpairs_per_s"],
                                  "solves": t["solves"], "refinement_steps": t["refinement_steps"]}
        good += t["good_pairs"]
        secs += t["seconds"]
        for key in totals:
            totals[key] += t[key]
        log(f"mixed-precision sweep {ny}x{nz} ({'joined groups' if joined else 'pipeline'}): "
            f"{t['eigenpairs_per_s']:.2f} eigenpairs/s")
    runner.close()
    how = ({"joined_group": JOINED_GROUP, "graph": "assembly + joined factorization per group"}
           if joined else {"contexts": contexts})
    return {"mixed_sweep": {**totals, "good_pairs": good, "seconds": round(secs, 4),
                            "eigenpairs_per_s": round(good / secs, 4) if secs > 0 else None,
                            "ncv": ncv, **how, "per_size": per_size}}


def phase_batch(ta, count: int = 4) -> dict:
    """Factorization throughput of several 200x100 instances (headline
    alphas, per-instance base flows, their own shifts): one after another on
    their contexts, every context's graph on its own stream at once
    (tailor_factor_batch), and joined into one launch sequence over their
    joined elimination levels replayed from a HIP graph (tailor_batch); then
    the same joined launch over two 200x100 and two 100x50 instances.  A
    solve on each context after the joined factorization is compared with
    the same solve after its own factorization."""
    from tailor_amd import sweep as sw
    insts = sw.sweep_instances(count, seed=1, sizes=[(200, 100)], reynolds=1500.0, mach=0.7)
    mixed = sw.sweep_instances(count, seed=1, sizes=[(200, 100), (100, 50)], reynolds=1500.0,
                               mach=0.7)
    out = {}
    rng = np.random.default_rng(5)
    for name, group in (("200x100", insts), ("200x100+100x50", mixed)):
        cases = sw.prepare(group, OMEGA)
        # targets omega_s = 0.02 + 0.001 k: sigma = Im(omega_s) - i Re(omega_s)
        sig = [complex(0.0, -(0.02 + 0.001 * k)) for k in range(count)]
        ctxs = []
        for _inst, c in cases:
            ctx = ta.Context(0)
            ctx.load_case(c)
            ctx.assemble(diagnostics=False)
            ctx.autotune(sig[0])
            ctxs.append(ctx)
        bs = [rng.standard_normal(5 * c.nodes) + 1j * rng.standard_normal(5 * c.nodes)
              for _i, c in cases]
        t = 0.0
        for _rep in range(2):  # the second pass is timed
            t0 = time.perf_counter()
            for ctx, s_ in zip(ctxs, sig):
                ctx.factor(s_)
            t = time.perf_counter() - t0
        own = [ctx.linear_solve(b)[0] for ctx, b in zip(ctxs, bs)]
        conc = [ta.factor_batch(ctxs, sig) for _ in range(3)][-1]
        batch = ta.Batch(ctxs)
        eager = batch.factor(sig, graph=False)
        joined = [batch.factor(sig, graph=True) for _ in range(3)][-1]
        diffs = [float(np.linalg.norm(ctx.linear_solve(b)[0] - x) / np.linalg.norm(x))
                 for ctx, b, x in zip(ctxs, bs, own)]
        batch.close()
        for ctx in ctxs:
            ctx.close()
        out[name] = {"instances": count, "own_ms_per_instance": round(1e3 * t / count, 2),
                     "concurrent_graphs_ms_per_instance": round(conc / count, 2),
                     "joined_eager_ms_per_instance": round(eager / count, 2),
                     "joined_graph_ms_per_instance": round(joined / count, 2),
                     "solve_rel_diff_vs_own": max(diffs)}
        log(f"joined factorization {name}: {joined / count:.1f} ms per instance "
            f"(own {1e3 * t / count:.1f}, concurrent graphs {conc / count:.1f})")
    return {"batch": out}


def run_phase(name: str, args, extra=()) -> dict:
    """One measurement phase in a child process (a fresh HIP runtime: the
    runtime keeps the hardware queues of every stream a process has created,
    and with many queues in use each kernel launch waits longer; measured at
    200x100, an unrelated factorization took 177 ms in a fresh process and
    207 ms after 24 streams had been created, also once they were destroyed)."""
    import subprocess
    cmd = [sys.executable, str(Path(__file__).resolve()), "--phase", name,
           "--steps", str(args.steps), "--warmup", str(max(1, args.warmup)),
           "--learned-ncv", str(args.learned_ncv), "--contexts", str(args.contexts), *extra]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
    if r.returncode != 0:
        raise SystemExit(f"bench.py: phase {name} failed with status {r.returncode}")
    return json.loads(r.stdout.strip().splitlines()[-1])


def phase_main(args) -> int:
    """Run one phase in this process and print its result as JSON."""
    import tailor_amd as ta
    from tailor_amd import case as cs
    if args.phase == "single":
        peaks = ta.measure_peaks(0)
        log(f"measured ceilings: FP64 MFMA {peaks['fp64_mfma_tflops']:.1f} TFLOP/s, "
            f"HBM read {peaks['read_gbs']:.0f} GB/s")
        single = plain_step(ta, cs)
        small = gpu_small_case(ta, cs, REFERENCE_NCV)
        log("100x50 plain instance done")
        out = {"single": single, "small": small, "peaks": peaks}
    elif args.phase == "pipeline":
        out = phase_pipeline(ta, args.steps, args.warmup, args.best_ncv, args.learned_ncv,
                             args.contexts)
    elif args.phase == "reference":
        out = phase_reference(ta, args.steps, args.warmup)
    elif args.phase == "batch":
        out = phase_batch(ta)
    elif args.phase == "joined":
        out = phase_joined_plain(ta, args.steps, args.warmup, args.best_ncv,
                                 threads=args.contexts > 1)
    elif args.phase == "mixed_size":
        out = phase_mixed_size(ta, args.mixed_count, args.best_ncv, args.contexts, joined=True)
    elif args.phase == "mixed_size_pipeline":
        out = phase_mixed_size(ta, args.mixed_count, args.best_ncv, args.contexts, joined=False)
    else:
        raise SystemExit(f"unknown phase {args.phase}")
    print(json.dumps(out))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64,
                    help="timed sweep instances (the learned-shift sweep is specified at 64)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed warm-up instances")
    ap.add_argument("--learned-ncv", type=int, default=LEARNED_NCV)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--contexts", type=int, default=PIPELINE_CONTEXTS,
                    help="contexts of the pipelined sweeps (1: no worker threads anywhere, "
                         "also in the joined sweeps; for kernel traces)")
    ap.add_argument("--phase", help=argparse.SUPPRESS)
    ap.add_argument("--best-ncv", type=int, default=REFERENCE_NCV, help=argparse.SUPPRESS)
    ap.add_argument("--mixed-count", type=int, default=MIXED_SIZE_INSTANCES, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU sweep would shard whole alpha instances across GPUs and "
              "gather only eigenvalues and residuals)", file=sys.stderr)
        return 2
    if args.steps < 1:
        print("bench.py: --steps must be at least 1", file=sys.stderr)
        return 2
    if args.phase:
        return phase_main(args)

    # every GPU phase in a child process; this process never touches the GPU
    sp = run_phase("single", args)
    single, small, peaks = sp["single"], sp["small"], sp["peaks"]
    mfma_meas, read_meas = peaks["fp64_mfma_tflops"], peaks["read_gbs"]
    best_ncv = single["plain"]["best_ncv"]
    pipe = run_phase("pipeline", args, ("--best-ncv", str(best_ncv)))
    ref = run_phase("reference", args)
    mixed_count = max(MIXED_SIZE_INSTANCES, 2 * (args.steps // 2))  # 64 at the default
    joined = run_phase("joined", args, ("--best-ncv", str(best_ncv)))
    msz_args = ("--best-ncv", str(best_ncv), "--mixed-count", str(mixed_count),
                "--contexts", str(min(args.contexts, 4)))
    msz = run_phase("mixed_size", args, msz_args)
    msz_pipe = run_phase("mixed_size_pipeline", args, msz_args)
    bt = run_phase("batch", args)
    L, F, PB, PR = pipe["learned"], pipe["fixed"], pipe["plain_best"], ref["plain_ref"]
    PJ = joined["plain_joined"]
    rate = lambda t: t["eigenpairs_per_s"]
    strongest = max(rate(PB), rate(PR), rate(PJ))
    g = single["gemm"]
    gemm_tf = g["gemm_flops"] / (g["gemm_ms"] * 1e-3) / 1e12
    asm_ms = single["assembly"]["full_ms"]
    out = {
        "metric": "eigenpairs/s at resid<=1e-10 (learned-shift sweep; plain FP64 alongside)",
        "value": round(rate(L), 4),
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(1, args.warmup),
        "ms_per_step": round(L["seconds"] * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if pipe["choice"].get("chosen", "fp64") == "fp64" else "c64 factors + c128",
        "data": "synthetic (erf boundary layer on a tanh-clustered, wall-bumped grid; base flow "
                "per instance from seeds 1000 + k)",
        "config": {"workload": "BiGlobal wavenumber sweep, 200x100 FD-q grid (q_y=10, q_z=6), "
                               "100,000 unknowns, Re=1500, Ma=0.7, alpha ~ U[0.5,3] seed 1, nev=8, "
                               "tol=1e-10; learned shift centres, warm-started Krylov-Schur, "
                               "fixed-shift fallback, " + (f"{args.contexts}-context pipeline"
                                                          if args.contexts > 1 else "one context"),
                   "nev": NEV, "ncv_learned": args.learned_ncv, "ncv_plain_reference": REFERENCE_NCV,
                   "ncv_plain_best": best_ncv, "pipeline_contexts": args.contexts,
                   "learned_check_from": LEARNED_CHECK_FROM, "instances": args.steps,
                   "phases": "each GPU phase in its own child process",
                   "precision_choice": pipe["choice"], "gemm_autotune": pipe["autotune"]},
        "learned_sweep": without_assembly(L, asm_ms),
        "network_off_sweep": without_assembly(F, asm_ms),
        "plain_sweep_best": without_assembly(PB, asm_ms),
        "plain_sweep_best_joined": without_assembly(PJ, asm_ms),
        "plain_sweep_reference": without_assembly(PR, asm_ms),
        "speedup_vs_strongest_plain": round(rate(L) / max(strongest, 1e-300), 3),
        "speedup_vs_reference_plain": round(rate(L) / max(rate(PR), 1e-300), 3),
        "network_gain": round(rate(L) / max(rate(F), 1e-300), 3),
        "learned_pairs_found_by_plain": pipe["overlap"],
        "roofline": {"bound": "mfma", "kernel": "k_gemm / k_gemm_lds (frontal updates, MFMA f64 16x16x4)",
                     "achieved": round(gemm_tf, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(gemm_tf / FP64_PEAK_TFLOPS, 4),
                     "measured_peak": round(mfma_meas, 2),
                     "frac_of_measured_peak": round(gemm_tf / mfma_meas, 4),
                     "measured_peak_note": "sustained dense FP64 MFMA rate of this GPU "
                                           "(tailor_measure_peaks: v_mfma_f64_16x16x4f64 back "
                                           "to back, two waves per SIMD, every CU); the frontal "
                                           "GEMM's 3M form issues 6 of the 8 counted flops",
                     "traffic": measured_traffic("k_gemm<double>"),
                     "traffic_note": "mean HBM bytes per k_gemm dispatch (counter runs in profiles/)",
                     "timing": "HIP events around every k_gemm launch of one serialized "
                               "factorization on the library stream (tailor_profile_factor)",
                     "gemm_ms_per_factorization": round(g["gemm_ms"], 2),
                     "gemm_us_per_launch": round(g["gemm_ms"] * 1e3 / max(g["gemm_launches"], 1), 2),
                     "gemm_flops_per_factorization": g["gemm_flops"],
                     "gemm_launches": g["gemm_launches"],
                     "tile_config_per_level": single["tune"]["config"],
                     "gemm_ms_per_config": single["tune"]["gemm_ms"],
                     "factor_ms": round(g["factor_ms"], 2)},
        "roofline_traced": None,
        "plain_instance": single["plain"],
        "measured_ceilings": {"fp64_mfma_tflops": round(mfma_meas, 2),
                              "hbm_read_gbs": round(read_meas, 1),
                              "published": {"fp64_mfma_tflops": FP64_PEAK_TFLOPS,
                                            "hbm_gbs": HBM_PEAK_GBS}},
        "assembly": {**single["assembly"],
                     "frac_of_measured_read_full": round(single["assembly"]["GBps_full"] / read_meas, 4)},
        "solve_sweep": {**single["solve_sweep"],
                        "frac_of_measured_read": round(single["solve_sweep"]["GBps"] / read_meas, 4)},
        "spmv": {**single["spmv"], "frac_of_measured_read": round(single["spmv"]["GBps"] / read_meas, 4)},
        "mixed_precision": single["mixed"],
        "mixed_size_sweep": msz["mixed_sweep"],
        "mixed_size_sweep_pipeline": msz_pipe["mixed_sweep"],
        "joined_factorization": bt["batch"],
        "gpu_100x50_plain": {k: v for k, v in small.items() if k != "lambdas"},
    }
    tw = traced_gemm_window()
    if tw and tw["launches"] == g["gemm_launches"]:
        tf = g["gemm_flops"] / (tw["ms"] * 1e-3) / 1e12
        out["roofline_traced"] = {**tw, "achieved": round(tf, 3),
                                  "frac": round(tf / FP64_PEAK_TFLOPS, 4),
                                  "frac_of_measured_peak": round(tf / mfma_meas, 4),
                                  "note": "the same factorization's k_gemm launches timed by a "
                                          "rocprofv3 kernel trace of this benchmark"}
    if not args.no_cpu_baseline:
        cb = cpu_baseline(REFERENCE_NCV)
        log("CPU baseline done")
        gl = np.array([complex(*x) for x in small["lambdas"]])
        cl = np.array([complex(*x) for x in cb.pop("lambdas")])
        cb["max_rel_diff_vs_gpu"] = float(max(np.min(np.abs(gl - l)) / abs(l) for l in cl)) \
            if len(gl) and len(cl) else None
        out["cpu_baseline"] = cb
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
