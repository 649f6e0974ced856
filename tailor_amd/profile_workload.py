"""A short, fixed workload for hardware-counter profiles.

    python -m tailor_amd.profile_workload [--ny 200 --nz 100 --solves 4 --gemm-config 3]

Runs on the synthetic case: one assembly of A_bc and B_bc, one FP64
factorization of A_bc - sigma B_bc (sigma = lambda_s for omega_s = 0.02) with
the frontal GEMM tiles forced to one configuration (default 3: 64x64 tiles,
K chunk 4, XCD-grouped order, the autotuner's choice on most levels at
200x100; -1 keeps the untuned default),
`--solves` triangular-solve pairs, and one device-to-device copy of exactly
1 GiB (a kernel whose HBM traffic is known, used to scale the FETCH_SIZE /
WRITE_SIZE counters of the other kernels).  Prints the algorithmic bytes or
flops of each kernel family so that counter readings can be compared with
them.
"""
from __future__ import annotations

import argparse
import sys

import numpy as np


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--ny", type=int, default=200)
    ap.add_argument("--nz", type=int, default=100)
    ap.add_argument("--solves", type=int, default=4)
    ap.add_argument("--gemm-config", type=int, default=3)
    a = ap.parse_args(argv)
    import torch
    import tailor_amd as ta
    from tailor_amd import case as cs
    c = cs.synthetic_case(a.ny, a.nz, omega_shift=0.02)
    ctx = ta.Context(0)
    ctx.load_case(c)
    ctx.assemble(diagnostics=False)
    if a.gemm_config >= 0:
        ctx.set_gemm_config(a.gemm_config)
    fd = ctx.factor(c.lambda_shift)
    ms, nbytes = ctx.time_linear_solve(a.solves)
    nodes = a.ny * a.nz
    asm = 16 * 25 * (nodes * 77 + nodes) + 128 * nodes + 3 * 8 * (a.ny * 11 + a.nz * 7)
    print(f"assembly algorithmic bytes {asm}")
    print(f"factorization LU flops {fd['flops']:.6e}, arena bytes {fd['arena_bytes']}")
    print(f"solve algorithmic bytes {nbytes:.6e} per solve, {ms:.4f} ms per solve "
          f"({a.solves + 1} solves incl. warm-up)")
    ctx.close()
    n = (1 << 30) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print("device copy: 1073741824 bytes read, 1073741824 bytes written")
    return 0


if __name__ == "__main__":
    sys.exit(main())
