"""Benchmark of the BiGlobal assemble path on one MI355X.

Workload (N=1): the 200x100 synthetic case (FD-q q_y=10 bounded x q_z=6
periodic, 100,000 complex unknowns, 1,540,000 5x5 blocks), alpha = 2.43.
A step assembles A_bc and B_bc from the device-resident grid and base flow:
metrics, base-flow derivatives, far-field characteristics and the fused
LNS/Fourier/curvilinear/BSR/boundary-row writer.

Prints one JSON line (see the keys below).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(ny: int, nz: int, ssx: int, sse: int) -> int:
    """Bytes one assembly must move: A_bc and B_bc written (16 B x 25 per
    block), grid + base flow read (8 complex per node = 128 B) and the FD-q
    weights (3 orders, both directions)."""
    nodes = ny * nz
    nnzb_a = nodes * ssx * sse
    return 16 * 25 * (nnzb_a + nodes) + 128 * nodes + 3 * 8 * (ny * ssx + nz * sse)


def cpu_baseline(case, seconds: float = 12.0) -> dict:
    """The C oracle (single thread) assembling the same case, repeated for
    about `seconds`."""
    import oracle
    oracle.build()
    oracle.assemble(case)  # warm-up
    n = 0
    t0 = time.perf_counter()
    while True:
        oracle.assemble(case)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    per = el / n
    nbytes = algorithmic_bytes(case.ny, case.nz, case.xi_rule.stencil_size, case.eta_rule.stencil_size)
    return {"value": nbytes / per / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{n} full assemblies (metrics, derivatives, L/M, A_bc/B_bc) of the "
                      f"{case.ny}x{case.nz} case, {per * 1e3:.1f} ms each"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the alpha sweep shards by instance across GPUs; see DESIGN.md)", file=sys.stderr)
        return 2

    import tailor_amd as ta
    from tailor_amd import case as cs

    case = cs.synthetic_case(200, 100, omega_shift=0.02 + 0j)
    ctx = ta.Context(0)
    ctx.load_case(case)
    ctx.assemble(diagnostics=False)
    # warmup + timed loop of whole assembly steps (kernel timing with HIP
    # events on the library's stream)
    ctx.time_assembly(max(1, args.warmup))
    write_ms, full_ms = ctx.time_assembly(args.steps)
    nbytes = algorithmic_bytes(200, 100, 11, 7)
    achieved = nbytes / (write_ms * 1e-3) / 1e9
    value = nbytes / (full_ms * 1e-3) / 1e9
    out = {
        "metric": "assembly HBM GB/s",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(full_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (erf boundary layer on tanh-clustered wall-bumped grid, seed 0)",
        "config": {"workload": "BiGlobal assembly of A_bc, B_bc, 200x100 FD-q grid (q_y=10, "
                               "q_z=6), 100,000 unknowns, 1,540,000 5x5 blocks, alpha=2.43",
                   "algorithmic_bytes": nbytes},
        "roofline": {"bound": "hbm", "kernel": "k_operator", "achieved": round(achieved, 2),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_ms": round(write_ms, 5)},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case)
    ctx.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
