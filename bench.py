"""Benchmark of the BiGlobal assemble-and-solve path on one MI355X.

Headline (`value`): eigenpairs/s of a wavenumber sweep on the 200x100 case
(FD-q q_y=10 bounded x q_z=6 periodic, 100,000 complex unknowns, Re=1500,
Ma=0.7, alpha ~ U[0.5, 3] from seed 1) in the learned-shift form: the shift
network proposes each instance's shift centre, instances run in wavenumber
order with warm-started Krylov-Schur, and the fixed-shift search is the
fallback; eight contexts pipeline the sweep (the next seven instances'
assembly and factorization run while the current instance's Krylov-Schur
runs).  A step is one sweep instance (assembly of A_bc, B_bc, sparse LU
of A_bc - sigma B_bc, Krylov-Schur for 8 eigenpairs); an eigenpair counts if
its true residual ||A q - lambda B q|| / ||q|| <= 1e-10.  The same instances
are also run in the plain form (complex FP64, fixed omega_s = 0.02, cold
start) and the ratio is reported.

Also reported: the plain single-instance step at alpha = 2.43 (Re=1000,
Ma=0.5), the mixed-precision form on that operator (per-eigenvalue error
against the FP64 result, refinement steps per solve), a small sweep over
both grid sizes with per-instance base flows in mixed precision, assembly
HBM GB/s, and the roofline of the frontal-update GEMM (the kernel with the
largest share of a sweep instance) from HIP events on the library stream.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

# The learned sweep drives two HIP streams per context (the factorization's
# main and look-ahead streams).  With HIP's default of four hardware queues
# per process, unrelated streams share queues and wait for each other:
# measured at 200x100 with three contexts, 199 ms per instance at 4 queues
# against 185 ms at 8 or 16.  24 covers the eight learned contexts' streams
# and the plain form's.  Read by HIP when it initialises, so set here,
# before anything touches the GPU (also where the environment sets the
# default of 4).
os.environ["GPU_MAX_HW_QUEUES"] = "24"


import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
NEV, TOL, OMEGA = 8, 1e-10, 0.02 + 0j
NCV = 64  # Krylov subspace size (the -eps_ncv of the reference's options database)
LEARNED_NCV = 40  # subspace of the learned form: the wanted cluster converges in one pass
LEARNED_CHECK_FROM = 28  # learned form: convergence also tested every second step from step 28
PIPELINE_CONTEXTS = 8  # learned sweep: Krylov-Schur of one instance, factorization of the next seven


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


def cpu_baseline() -> dict:
    """The CPU oracle (C assembly + multifrontal LU on LAPACK/BLAS + numpy
    Krylov-Schur) on the 100x50 grid of the same generator (the 200x100
    case needs minutes of host time)."""
    import oracle
    from oracle import multifrontal
    from tailor_amd import case as cs
    oracle.build()
    c = cs.synthetic_case(100, 50, omega_shift=0j)
    t0 = time.perf_counter()
    res = multifrontal.solve_case(c, nev=NEV, tol=TOL, ncv=NCV)
    el = time.perf_counter() - t0
    good = int(np.sum(res["residuals"][:NEV] <= TOL))
    threads = multifrontal.blas_threads()
    return {"value": round(good / el, 4), "unit": "eigenpairs/s", "cores": threads,
            "kind": "port",
            "sample": f"one full assemble+factor+Krylov-Schur of the 100x50 case (25,000 unknowns, "
                      f"omega_s=0), {el:.1f} s, {good} of {NEV} pairs at residual <= 1e-10, "
                      f"{res['restarts']} restarts; BLAS threads {threads}"}


def plain_step(ta, cs, steps: int, ncv: int) -> dict:
    """The plain FP64 form on one instance (alpha 2.43, omega_s 0.02)."""
    case = cs.synthetic_case(200, 100, omega_shift=OMEGA)
    ctx = ta.Context(0)
    ctx.load_case(case)
    ctx.assemble(diagnostics=False)
    ctx.time_assembly(3)
    write_ms, asm_ms = ctx.time_assembly(20)
    asm_bytes = algorithmic_assembly_bytes(200, 100, 11, 7)

    def step():
        ctx.assemble(diagnostics=False)
        d = ctx.solve(nev=NEV, omega=OMEGA, tol=TOL, ncv=ncv, refine_max=0, raise_on_failure=False)
        pairs = ctx.eigenpairs()[:NEV]
        return sum(1 for p in pairs if p["residual"] <= TOL), d, pairs

    tune = ctx.autotune(case.lambda_shift)  # frontal GEMM tile order per level
    step()
    good, t0 = 0, time.perf_counter()
    for _ in range(steps):
        g, d, pairs = step()
        good += g
    el = time.perf_counter() - t0
    lam64 = np.array([p["lambda"] for p in pairs])
    solve_ms, solve_bytes = ctx.time_linear_solve(20)
    prof = ctx.profile_factor(case.lambda_shift)
    # mixed precision on the same operator and shift
    ctx.set_precision(ta.PRECISION_MIXED)
    t0 = time.perf_counter()
    dm = ctx.solve(nev=NEV, omega=OMEGA, tol=TOL, ncv=ncv, refine_max=20, refine_tol=1e-12,
                   raise_on_failure=False)
    mixed_s = time.perf_counter() - t0
    pm = ctx.eigenpairs()[:NEV]
    lam32 = np.array([p["lambda"] for p in pm])
    errs = [float(np.min(np.abs(lam32 - l)) / abs(l)) for l in lam64]
    ctx.close()
    return {
        "plain": {"eigenpairs_per_s": round(good / el, 4), "ms_per_step": round(el * 1e3 / steps, 2),
                  "steps": steps, "factor_ms": round(d["factor_ms"], 2),
                  "krylov_ms": round(d["solve_ms"], 2), "solves": d["solves"],
                  "restarts": d["iterations"], "max_residual": d["max_residual"],
                  "factor_TFLOPs_lu_flops": round(d["factor_flops"] / (d["factor_ms"] * 1e-3) / 1e12, 3)},
        "assembly": {"GBps_writer": round(asm_bytes / (write_ms * 1e-3) / 1e9, 2),
                     "GBps_full": round(asm_bytes / (asm_ms * 1e-3) / 1e9, 2),
                     "frac_hbm_writer": round(asm_bytes / (write_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "writer_ms": round(write_ms, 5), "full_ms": round(asm_ms, 5),
                     "algorithmic_bytes": asm_bytes},
        "solve_sweep": {"kernel": "k_gemv (triangular solves)", "ms_per_solve": round(solve_ms, 4),
                        "bytes": int(solve_bytes),
                        "GBps": round(solve_bytes / (solve_ms * 1e-3) / 1e9, 2),
                        "frac_hbm": round(solve_bytes / (solve_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "gemm": prof,
        "tune": tune,
        "mixed": {"seconds": round(mixed_s, 3), "factor_ms": round(dm["factor_ms"], 2),
                  "solves": dm["solves"], "refinement_steps": dm["refinement_steps"],
                  "refinement_steps_per_solve": round(dm["refinement_steps"] / max(1, dm["solves"]), 2),
                  "converged": dm["converged"], "max_residual": dm["max_residual"],
                  "eig_rel_err_vs_fp64": [float(f"{e:.3e}") for e in errs],
                  "fp64_seconds_same_operator": round(el / steps, 3)},
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


def sweeps(ta, steps: int, warmup: int, ncv: int, learned_ncv: int, contexts: int) -> dict:
    from tailor_amd import sweep as sw
    from tailor_amd.shift_model import ShiftModel
    from tailor_amd.train_shift_model import MODEL
    if not MODEL.exists():
        raise SystemExit(f"{MODEL} is missing: run `python -m tailor_amd.train_shift_model`")
    model = ShiftModel(0).load(MODEL)
    insts = sw.sweep_instances(warmup + steps, seed=1, sizes=[(200, 100)], reynolds=1500.0,
                               mach=0.7, per_instance_seeds=False)
    cases = sw.prepare(insts, OMEGA)
    alphas = np.array([i.alpha for i in insts])
    shifts = model.shift_centres(alphas, np.full(len(insts), 1500.0), np.full(len(insts), 0.7))
    runner = sw.Sweeper(ncv=learned_ncv, omega_s=OMEGA, contexts=contexts,
                        check_from=LEARNED_CHECK_FROM)
    plain_runner = sw.Sweeper(ncv=ncv, omega_s=OMEGA)
    # warm-up instances (untimed): build the plan, and let the first one
    # choose the factorization precision by measurement
    warm = list(range(warmup))
    choice = {}
    precision = ta.PRECISION_FP64
    if warm:
        t = {}
        for prec in (ta.PRECISION_FP64, ta.PRECISION_MIXED):
            r = runner.run(cases, "learned", shifts=shifts, precision=prec, order=warm[:1])
            t[prec] = r.seconds
        precision = min(t, key=t.get)
        choice = {"fp64_s": round(t[ta.PRECISION_FP64], 3), "mixed_s": round(t[ta.PRECISION_MIXED], 3),
                  "chosen": "mixed" if precision == ta.PRECISION_MIXED else "fp64"}
    timed = list(range(warmup, warmup + steps))
    order = sorted(timed, key=lambda k: insts[k].alpha)
    learned = runner.run(cases, "learned", shifts=shifts, precision=precision, order=order)
    runner.close()
    plain_runner.run(cases, "plain", precision=ta.PRECISION_FP64, order=warm[:1])  # warm-up
    plain = plain_runner.run(cases, "plain", precision=ta.PRECISION_FP64, order=timed)
    # eigenvalues found by both forms must agree
    lam_diff = 0.0
    by_inst = {id(r.instance): r for r in plain.records}
    for r in learned.records:
        p = by_inst[id(r.instance)]
        for w in r.omegas[r.residuals <= TOL]:
            d = np.min(np.abs(p.omegas - w)) / max(abs(w), 1e-300)
            if d < 1e-6:
                lam_diff = max(lam_diff, float(d))
    # config 3: both grid sizes, per-instance base flows, mixed precision
    c3 = sw.sweep_instances(4, seed=1, sizes=[(100, 50), (200, 100)])
    c3_cases = sw.prepare(c3, OMEGA)
    mixed_sweep = plain_runner.run(c3_cases, "mixed")
    plain_runner.close()
    return {"learned": learned, "plain": plain, "choice": choice, "common_pair_rel_diff": lam_diff,
            "mixed_sweep": mixed_sweep, "shifts": shifts,
            "autotune": {f"{k[0]}x{k[1]}": v for k, v in runner.tuned.items()}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64,
                    help="timed sweep instances (the learned-shift sweep is specified at 64)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed warm-up instances")
    ap.add_argument("--ncv", type=int, default=NCV, help="Krylov subspace of the plain form")
    ap.add_argument("--learned-ncv", type=int, default=LEARNED_NCV)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--contexts", type=int, default=PIPELINE_CONTEXTS,
                    help="contexts of the learned sweep's pipeline (1: no worker threads)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(a multi-GPU sweep would shard whole alpha instances across GPUs and "
              "gather only eigenvalues and residuals)", file=sys.stderr)
        return 2
    if args.steps < 1:
        print("bench.py: --steps must be at least 1", file=sys.stderr)
        return 2

    import tailor_amd as ta
    from tailor_amd import case as cs

    single = plain_step(ta, cs, 2, args.ncv)
    sw = sweeps(ta, args.steps, max(1, args.warmup), args.ncv, args.learned_ncv, args.contexts)
    L, P = sw["learned"], sw["plain"]
    g = single["gemm"]
    gemm_tf = g["gemm_flops"] / (g["gemm_ms"] * 1e-3) / 1e12
    out = {
        "metric": "eigenpairs/s at resid<=1e-10 (learned-shift sweep; plain FP64 alongside)",
        "value": round(L.eigenpairs_per_s, 4),
        "unit": "eigenpairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": max(1, args.warmup),
        "ms_per_step": round(L.seconds * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128" if sw["choice"].get("chosen", "fp64") == "fp64" else "c64 factors + c128",
        "data": "synthetic (erf boundary layer on tanh-clustered wall-bumped grid, seed 0)",
        "config": {"workload": "BiGlobal wavenumber sweep, 200x100 FD-q grid (q_y=10, q_z=6), "
                               "100,000 unknowns, Re=1500, Ma=0.7, alpha ~ U[0.5,3] seed 1, nev=8, "
                               "tol=1e-10; learned shift centres, warm-started Krylov-Schur, "
                               "fixed-shift fallback, " + (f"{args.contexts}-context pipeline"
                                                          if args.contexts > 1 else "one context"),
                   "nev": NEV, "ncv_learned": args.learned_ncv, "ncv_plain": args.ncv,
                   "pipeline_contexts": args.contexts,
                   "learned_check_from": LEARNED_CHECK_FROM,
                   "instances": args.steps,
                   "precision_choice": sw["choice"], "gemm_autotune": sw["autotune"]},
        "learned_sweep": without_assembly(L.totals(), single["assembly"]["full_ms"]),
        "plain_sweep": without_assembly(P.totals(), single["assembly"]["full_ms"]),
        "speedup_vs_plain_sweep": round(L.eigenpairs_per_s / max(P.eigenpairs_per_s, 1e-300), 3),
        "common_pair_max_rel_diff": sw["common_pair_rel_diff"],
        "roofline": {"bound": "mfma", "kernel": "k_gemm (frontal updates, MFMA f64 16x16x4)",
                     "achieved": round(gemm_tf, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(gemm_tf / FP64_PEAK_TFLOPS, 4),
                     "traffic": measured_traffic("k_gemm<double>"),
                     "traffic_note": "mean HBM bytes per k_gemm dispatch (counter runs in profiles/)",
                     "gemm_ms_per_factorization": round(g["gemm_ms"], 2),
                     "gemm_flops_per_factorization": g["gemm_flops"],
                     "gemm_launches": g["gemm_launches"],
                     "tile_config_per_level": single["tune"]["config"],
                     "gemm_ms_per_config": single["tune"]["gemm_ms"],
                     "factor_ms": round(g["factor_ms"], 2)},
        "plain_instance": single["plain"],
        "assembly": single["assembly"],
        "solve_sweep": single["solve_sweep"],
        "mixed_precision": single["mixed"],
        "mixed_size_sweep": sw["mixed_sweep"].totals(),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
