"""Generate solver-labelled samples and train the shift-proposal network.

    python -m tailor_amd.train_shift_model --samples 2000 --steps 4000

1. Draw (alpha, Re, Ma) ~ U([0.5,3] x [500,3000] x [0.3,0.9]) from seed 2
   and solve each instance on the 100x50 grid with the plain fixed-shift
   method (omega_s = 0.02, 8 pairs, true residual <= 1e-10); the label is
   the 8 omegas ordered by growth rate.  The sample file is saved as it
   grows and resumed by the next run.
2. Train the network on the GPU with the library's kernels (full-batch Adam
   on the normalised data) and save weights and normalisation.

Samples go to tailor_amd/data/shift_samples_100x50.npz and the model to
tailor_amd/data/shift_model.npz; an existing sample file is reused unless
--regenerate is given.
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

DATA = Path(__file__).resolve().parent / "data"
SAMPLES = DATA / "shift_samples_100x50.npz"
MODEL = DATA / "shift_model.npz"


def generate(count: int, seed: int = 2, ny: int = 100, nz: int = 50, out=SAMPLES,
             log=print, save_every: int = 100, limit_s: float | None = None) -> dict:
    """Solve the training instances and save their labels to `out`.

    The file is rewritten every `save_every` samples and an existing file
    with the same draws is resumed (samples marked done are skipped), so the
    work can be spread over several runs; `limit_s` stops a run early."""
    from . import sweep
    insts = sweep.training_instances(count, seed, ny, nz)
    alpha = np.array([i.alpha for i in insts])
    re = np.array([i.reynolds for i in insts])
    ma = np.array([i.mach for i in insts])
    omegas = np.zeros((count, 8), dtype=np.complex128)
    resid = np.full((count, 8), np.inf)
    done = np.zeros(count, dtype=bool)
    out = Path(out)
    if out.exists():
        old = load_samples(out)
        if len(old["alpha"]) == count and np.array_equal(old["alpha"], alpha):
            omegas, resid = old["omegas"], old["residuals"]
            done = old["done"] if "done" in old else np.ones(count, dtype=bool)
    # group by Mach number so consecutive instances differ little (the base
    # flow depends on Ma only)
    order = [k for k in sorted(range(count), key=lambda k: (insts[k].mach, insts[k].alpha))
             if not done[k]]
    sw = sweep.Sweeper(ncv=64)

    def save():
        ok = done & np.all(resid <= 1e-10, axis=1)
        np.savez(out, alpha=alpha, reynolds=re, mach=ma, omegas=omegas, residuals=resid,
                 done=done, ok=ok)

    t0 = time.perf_counter()
    for n, k in enumerate(order, 1):
        case = insts[k].build(sw.omega_s)
        rec = sw.run([(insts[k], case)], form="plain").records[0]
        m = min(8, len(rec.omegas))
        if m == 8:
            omegas[k] = sweep.label_omegas(rec)
        resid[k, :m] = rec.residuals[:m]
        done[k] = True
        if n % 25 == 0:
            log(f"samples {int(done.sum())}/{count}  {time.perf_counter() - t0:.0f} s  last: alpha "
                f"{insts[k].alpha:.3f} Re {insts[k].reynolds:.0f} Ma {insts[k].mach:.3f} "
                f"good {rec.good}", flush=True)
        if n % save_every == 0:
            save()
        if limit_s is not None and time.perf_counter() - t0 > limit_s:
            break
    sw.close()
    save()
    return load_samples(out)


def load_samples(path=SAMPLES) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def train(samples: dict, steps: int = 4000, lr: float = 3e-3, seed: int = 2, out=MODEL,
          log=print):
    from .shift_model import ShiftModel
    ok = samples["ok"] & samples["done"] if "done" in samples else samples["ok"]
    m = ShiftModel(0, seed)
    hist = m.fit(samples["alpha"][ok], samples["reynolds"][ok], samples["mach"][ok],
                 samples["omegas"][ok], steps=steps, lr=lr)
    log(f"trained on {int(ok.sum())} samples: loss {hist[0]:.3e} -> {hist[-1]:.3e}", flush=True)
    if out is not None:
        m.save(out)
    return m, hist


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=2000)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--regenerate", action="store_true")
    ap.add_argument("--samples-file", default=str(SAMPLES))
    ap.add_argument("--model-file", default=str(MODEL))
    ap.add_argument("--limit-seconds", type=float, default=None,
                    help="stop generating after this long (resume in a later run)")
    a = ap.parse_args(argv)
    path = Path(a.samples_file)
    if a.regenerate and path.exists():
        path.unlink()
    samples = load_samples(path) if path.exists() else None
    if samples is not None and "done" not in samples:
        samples["done"] = np.ones(len(samples["alpha"]), dtype=bool)
    if samples is None or not samples["done"].all():
        samples = generate(a.samples, out=path, limit_s=a.limit_seconds)
    if not samples["done"].all():
        print(f"{int(samples['done'].sum())} of {a.samples} samples done; run again to continue")
        return 0
    train(samples, a.steps, a.lr, out=Path(a.model_file))
    return 0


if __name__ == "__main__":
    sys.exit(main())
