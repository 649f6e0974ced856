"""Wavenumber sweeps: many (grid, base flow, alpha) instances through the
assemble -> factor -> Krylov-Schur path of one GPU.

Forms of a sweep:

* ``plain``   complex FP64 fronts, the fixed search centre omega_s for every
              instance, cold start (the reference's method, instance by
              instance);
* ``mixed``   complex FP32 fronts with FP64 refinement of every solve to
              1e-12, fixed omega_s;
* ``learned`` shift centres proposed by a ShiftModel (centroid of the
              predicted 8 leading omega), instances ordered by wavenumber so
              each Krylov-Schur run starts from the previous one's Ritz
              vectors, and the plain fixed-shift solve as the fallback of any
              instance whose pairs miss the residual tolerance;
* ``fixed``   the learned form with the network switched off: the fixed
              search centre omega_s for every instance, with the same warm
              starts, early pass ends and fallback (what the network's
              shifts gain is the learned form's rate over this one's).

One context per grid size is kept for the whole sweep, so the symbolic
analysis and device buffers are built once per size; each instance uploads
its grid and base flow and reassembles.  With `contexts=C > 1` every form
runs C contexts per size as a pipeline: while one context runs the
current instance's Krylov-Schur, worker threads assemble and factor the
next C - 1 instances on the others (each context warm-starts from the
instance it solved before, C steps back in wavenumber).  The HIP streams
are non-blocking, so the GEMV-bound Krylov phase fills the GPU during the
factorization's latency-bound panel phases.  Each context drives two
streams, so the process needs at least 2 C hardware queues
(GPU_MAX_HW_QUEUES, read when HIP initialises; HIP's default is 4), or
unrelated streams wait for each other.

With `batch=B > 0` every form runs through joined factorizations instead:
the instances go in groups of B, each group on B contexts that are
assembled and then factored together by one launch sequence over their
joined elimination levels (a tailor_batch, assembly and factorization
replayed from one HIP graph), so every level launch covers the fronts of
the whole group.  Two groups of contexts alternate: a
worker thread assembles and factors group j + 1 while group j's Krylov-Schur
runs, one thread per context.  Groups may mix grid
sizes; each context keeps one size, so warm starts still follow a context's
previous instance.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import Batch, Context, PRECISION_FP64, PRECISION_MIXED
from .case import Physics, SplitMix64, synthetic_case


@dataclass
class Instance:
    ny: int
    nz: int
    seed: int
    alpha: float
    reynolds: float = 1000.0
    mach: float = 0.5

    def build(self, omega_s: complex):
        phys = Physics(reynolds=self.reynolds, mach=self.mach)
        return synthetic_case(self.ny, self.nz, seed=self.seed, alpha=self.alpha,
                              omega_shift=omega_s, physics=phys)


@dataclass
class Record:
    instance: Instance
    shift: complex
    converged: int
    good: int                 # pairs with true residual <= tol among the first nev
    omegas: np.ndarray
    residuals: np.ndarray     # max(residual, relative error) per pair
    restarts: int
    solves: int
    refinement_steps: int
    factorizations: int
    fallback: bool
    seconds: float


@dataclass
class SweepResult:
    form: str
    records: list = field(default_factory=list)
    seconds: float = 0.0

    @property
    def good_pairs(self) -> int:
        return sum(r.good for r in self.records)

    @property
    def eigenpairs_per_s(self) -> float:
        return self.good_pairs / self.seconds if self.seconds > 0 else 0.0

    def totals(self) -> dict:
        rs = self.records
        return {"instances": len(rs), "good_pairs": self.good_pairs,
                "seconds": round(self.seconds, 4),
                "eigenpairs_per_s": round(self.eigenpairs_per_s, 4),
                "factorizations": sum(r.factorizations for r in rs),
                "restarts": sum(r.restarts for r in rs),
                "solves": sum(r.solves for r in rs),
                "refinement_steps": sum(r.refinement_steps for r in rs),
                "fallbacks": sum(int(r.fallback) for r in rs)}


def sweep_instances(count: int, seed: int, sizes, alpha_range=(0.5, 3.0), reynolds=1000.0,
                    mach=0.5, per_instance_seeds: bool = True) -> list:
    """`count` instances, alpha ~ U(alpha_range) from a SplitMix64 stream;
    sizes cycle over `sizes` ((ny, nz) pairs); base-flow seeds 1000 + k."""
    rng = SplitMix64(seed)
    out = []
    for k in range(count):
        a = rng.uniform(*alpha_range)
        ny, nz = sizes[k % len(sizes)]
        out.append(Instance(ny, nz, 1000 + k if per_instance_seeds else 0, a, reynolds, mach))
    return out


def training_instances(count: int, seed: int = 2, ny: int = 100, nz: int = 50) -> list:
    """(alpha, Re, Ma) ~ U([0.5,3] x [500,3000] x [0.3,0.9]) on one grid."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(count):
        a = rng.uniform(0.5, 3.0)
        re = rng.uniform(500.0, 3000.0)
        ma = rng.uniform(0.3, 0.9)
        out.append(Instance(ny, nz, 0, a, re, ma))
    return out


# forms that warm-start from the previous instance, test convergence early
# and fall back to the plain solve
_WARM_FORMS = ("learned", "fixed")


class Sweeper:
    def __init__(self, device: int = 0, nev: int = 8, tol: float = 1e-10, ncv: int = 64,
                 omega_s: complex = 0.02 + 0j, maxit: int = 1000, contexts: int = 1,
                 check_from: int = 0, graph: bool = False, batch: int = 0,
                 threads: bool = True, block_pairs: bool | None = None):
        """`check_from` > 0: the learned form's Krylov-Schur runs also test
        convergence after every second step from that step on and end a pass
        early (the plain form and the fallback test only at pass ends).
        `graph`: every factorization is replayed from the context's HIP graph
        (tailor_factor_graph, the same results as the eager launches).  At
        200x100 a replay takes 169 ms against 177 ms for the eager launches
        on one context, but the eight-context pipeline measured slower with
        replays (33.6 against 45.7 eigenpairs/s), so eager is the default."""
        if not 1 <= contexts <= 8:
            raise ValueError("contexts must be 1 to 8")
        if not 0 <= batch <= 8:
            raise ValueError("batch must be 0 to 8")
        self.batch = batch
        self.batches = {}  # joined factorizations, by the contexts they join
        # joined groups: False runs each group's assembly, factorization and
        # Krylov-Schur runs one after another on the calling thread (for
        # profilers that cannot follow launches from several threads)
        self.threads = threads
        # two-level blocking of every context's factorization (K = 256
        # trailing updates; tailor_set_block_pairs); None: the library's
        # default (on)
        self.block_pairs = block_pairs
        if check_from < 0:
            raise ValueError("check_from must be >= 0")
        self.check_from = check_from
        self.graph = graph
        self.device = device
        self.nev, self.tol, self.ncv, self.maxit = nev, tol, ncv, maxit
        self.omega_s = complex(omega_s)
        self.contexts = contexts
        self.ctx = {}
        self.tuned = {}  # per grid size: the autotuner's report

    def close(self):
        for b in self.batches.values():
            b.close()
        self.batches = {}
        for c in self.ctx.values():
            c.close()
        self.ctx = {}

    def context(self, ny: int, nz: int, slot: int = 0) -> Context:
        key = (ny, nz, slot)
        if key not in self.ctx:
            self.ctx[key] = Context(self.device)
        return self.ctx[key]

    def _tune(self, ctx, case, precision):
        """First use of a context: build its plan and tune its GEMM tiles
        (not part of a sweep's time)."""
        ctx.load_case(case)
        ctx.assemble(diagnostics=False)
        ctx.set_precision(precision)
        if self.block_pairs is not None:
            ctx.set_block_pairs(self.block_pairs)
        return ctx.autotune(complex(self.omega_s.imag, -self.omega_s.real))

    def _factor(self, ctx, omega, precision):
        """LU of A_bc - sigma B_bc at the target of `omega`."""
        ctx.set_precision(precision)
        sigma = complex(omega.imag, -omega.real)
        if self.graph:
            ctx.factor_graph(sigma)
        else:
            ctx.factor(sigma)

    def _solve(self, ctx, omega, precision, warm, check_from=0, factored=False):
        """Krylov-Schur at the target of `omega`; `factored`: the context
        already holds the factorization at that shift (a pipeline stage made
        it), otherwise it is made here (the solve keeps a factorization at
        its own shift)."""
        if self.graph and not factored:
            self._factor(ctx, omega, precision)
        ctx.set_precision(precision)
        mixed = precision == PRECISION_MIXED
        d = ctx.solve(nev=self.nev, omega=omega, tol=self.tol, maxit=self.maxit, ncv=self.ncv,
                      refine_max=20 if mixed else 0, refine_tol=1e-12, raise_on_failure=False,
                      warm_start=warm, check_from=check_from)
        pairs = ctx.eigenpairs()[:self.nev]
        # a pair counts when both its residual and its relative error meet tol
        res = np.array([max(p["residual"], p["relative_error"]) for p in pairs])
        om = np.array([p["omega"] for p in pairs])
        good = int(np.sum(res <= self.tol))
        return d, om, res, good

    def run(self, cases, form: str = "plain", shifts=None, precision=None, order=None,
            on_record=None) -> SweepResult:
        """Run the sweep over prepared (instance, case) pairs.

        `shifts` (learned form) gives one omega shift centre per instance;
        the other forms use omega_s; `order` the processing order (defaults
        to the given order)."""
        if form not in ("plain", "mixed", "learned", "fixed"):
            raise ValueError(f"unknown sweep form {form!r}")
        if form == "learned" and shifts is None:
            raise ValueError("the learned form needs shifts")
        if precision is None:
            precision = PRECISION_MIXED if form == "mixed" else PRECISION_FP64
        if form != "learned":
            shifts = [self.omega_s] * len(cases)
        idx = list(range(len(cases))) if order is None else list(order)
        if self.batch > 0:
            return self._run_batched(cases, form, shifts, precision, idx, on_record)
        if self.contexts > 1 and len(idx) > 1:
            return self._run_pipelined(cases, form, shifts, precision, idx, on_record)
        result = SweepResult(form)
        last_size = None
        t_all = time.perf_counter()
        t_tune = 0.0  # one-time autotuning, not part of the sweep's time
        for k in idx:
            inst, case = cases[k]
            ctx = self.context(inst.ny, inst.nz)
            t0 = time.perf_counter()
            ctx.load_case(case)
            ctx.assemble(diagnostics=False)
            if (inst.ny, inst.nz) not in self.tuned:
                # first instance of a size: pick the frontal GEMM tile order
                tt = time.perf_counter()
                ctx.set_precision(precision)
                if self.block_pairs is not None:
                    ctx.set_block_pairs(self.block_pairs)
                self.tuned[(inst.ny, inst.nz)] = ctx.autotune(
                    complex(self.omega_s.imag, -self.omega_s.real))
                t_tune += time.perf_counter() - tt
                t0 += time.perf_counter() - tt
            warm = form in _WARM_FORMS and last_size == (inst.ny, inst.nz)
            omega = complex(shifts[k])
            d, om, res, good = self._solve(ctx, omega, precision, warm,
                                           self.check_from if form in _WARM_FORMS else 0)
            facts = 1
            fallback = False
            restarts, solves, refine = d["iterations"], d["solves"], d["refinement_steps"]
            if form in _WARM_FORMS and good < self.nev:
                fallback = True
                d, om, res, good = self._solve(ctx, self.omega_s, precision, False)
                facts += 1
                restarts += d["iterations"]
                solves += d["solves"]
                refine += d["refinement_steps"]
                omega = self.omega_s
            el = time.perf_counter() - t0
            last_size = (inst.ny, inst.nz)
            rec = Record(inst, omega, d["converged"], good, om, res, restarts, solves, refine,
                         facts, fallback, el)
            result.records.append(rec)
            if on_record:
                on_record(rec)
        result.seconds = time.perf_counter() - t_all - t_tune
        return result


    def _run_pipelined(self, cases, form, shifts, precision, idx, on_record) -> SweepResult:
        """A form over `self.contexts` contexts per grid size (see the
        module docstring): instance i runs its Krylov-Schur on context
        i mod C while worker threads assemble and factor the next C - 1
        instances on the others.  Instance k's record covers its Krylov stage
        and any fallback; the sweep's time is the wall time of the whole
        pipeline."""
        C = self.contexts
        result = SweepResult(form)
        warm_form = form in _WARM_FORMS
        # plans and tuning of every context of every size first (untimed)
        for k in idx:
            inst, case = cases[k]
            size = (inst.ny, inst.nz)
            for slot in range(C):
                if (inst.ny, inst.nz, slot) not in self.ctx:
                    rep = self._tune(self.context(inst.ny, inst.nz, slot), case, precision)
                    self.tuned.setdefault(size, rep)
        solved = set()  # contexts that hold Ritz vectors of an earlier instance
        errors = []

        def stage(slot, k):
            try:
                inst, case = cases[k]
                ctx = self.context(inst.ny, inst.nz, slot)
                ctx.load_case(case)
                ctx.assemble(diagnostics=False)
                self._factor(ctx, complex(shifts[k]), precision)
            except Exception as e:  # re-raised by the main thread
                errors.append(e)

        def start(i):
            th = threading.Thread(target=stage, args=(i % C, idx[i]))
            th.start()
            return th

        t_all = time.perf_counter()
        workers = {i: start(i) for i in range(min(C - 1, len(idx)))}
        for i, k in enumerate(idx):
            if i + C - 1 < len(idx):
                workers[i + C - 1] = start(i + C - 1)  # its context finished instance i - 1
            workers.pop(i).join()
            if errors:
                for th in workers.values():
                    th.join()
                raise errors[0]
            slot = i % C
            inst, _ = cases[k]
            ctx = self.context(inst.ny, inst.nz, slot)
            t0 = time.perf_counter()
            omega = complex(shifts[k])
            key = (inst.ny, inst.nz, slot)
            d, om, res, good = self._solve(ctx, omega, precision, warm_form and key in solved,
                                           self.check_from if warm_form else 0, factored=True)
            solved.add(key)
            facts, fallback = 1, False
            restarts, solves, refine = d["iterations"], d["solves"], d["refinement_steps"]
            if warm_form and good < self.nev:
                fallback = True
                d, om, res, good = self._solve(ctx, self.omega_s, precision, False)
                facts += 1
                restarts += d["iterations"]
                solves += d["solves"]
                refine += d["refinement_steps"]
                omega = self.omega_s
            el = time.perf_counter() - t0
            rec = Record(inst, omega, d["converged"], good, om, res, restarts, solves, refine,
                         facts, fallback, el)
            result.records.append(rec)
            if on_record:
                on_record(rec)
        result.seconds = time.perf_counter() - t_all
        return result

    def _krylov_record(self, ctx, inst, omega, form, precision, warm, t0):
        """Krylov-Schur of one instance on a context that holds its
        factorization, with the fixed-shift fallback of the warm forms."""
        warm_form = form in _WARM_FORMS
        d, om, res, good = self._solve(ctx, omega, precision, warm,
                                       self.check_from if warm_form else 0, factored=True)
        facts, fallback = 1, False
        restarts, solves, refine = d["iterations"], d["solves"], d["refinement_steps"]
        if warm_form and good < self.nev:
            fallback = True
            d, om, res, good = self._solve(ctx, self.omega_s, precision, False)
            facts += 1
            restarts += d["iterations"]
            solves += d["solves"]
            refine += d["refinement_steps"]
            omega = self.omega_s
        return Record(inst, omega, d["converged"], good, om, res, restarts, solves, refine,
                      facts, fallback, time.perf_counter() - t0)

    def _run_batched(self, cases, form, shifts, precision, idx, on_record) -> SweepResult:
        """A form through joined factorizations (see the module docstring).
        Instance k's record covers its Krylov stage and any fallback; the
        sweep's time is the wall time of the whole run."""
        B = self.batch
        result = SweepResult(form)
        warm_form = form in _WARM_FORMS
        groups = [idx[j:j + B] for j in range(0, len(idx), B)]

        def key(j, pos, k):
            inst = cases[k][0]
            return (inst.ny, inst.nz, (j % 2) * B + pos)

        def group_contexts(j):
            return [self.ctx[key(j, pos, k)] for pos, k in enumerate(groups[j])]

        # plans, tuning and the joined factorizations' graphs first (untimed)
        for j, g in enumerate(groups):
            for pos, k in enumerate(g):
                kk = key(j, pos, k)
                if kk not in self.ctx:
                    rep = self._tune(self.context(*kk), cases[k][1], precision)
                    self.tuned.setdefault(kk[:2], rep)
        for j in range(len(groups)):
            ctxs = group_contexts(j)
            bk = tuple(id(c) for c in ctxs)
            if bk not in self.batches:
                self.batches[bk] = Batch(ctxs)
                for ctx, k in zip(ctxs, groups[j]):
                    ctx.load_case(cases[k][1])
                    ctx.set_precision(precision)
                self.batches[bk].assemble_factor(
                    [complex(self.omega_s.imag, -self.omega_s.real)] * len(ctxs))
        solved = set()  # contexts that hold Ritz vectors of an earlier instance
        errors = []

        def stage(j):
            try:
                ctxs = group_contexts(j)
                for ctx, k in zip(ctxs, groups[j]):
                    ctx.load_case(cases[k][1])
                    ctx.set_precision(precision)
                sig = [complex(complex(shifts[k]).imag, -complex(shifts[k]).real) for k in groups[j]]
                # the group's assembly and joined factorization: one graph replay
                self.batches[tuple(id(c) for c in ctxs)].assemble_factor(sig, graph=True)
            except Exception as e:  # re-raised by the main thread
                errors.append(e)

        class _Done:
            def join(self):
                pass

        def start(j):
            if not self.threads:
                stage(j)
                return _Done()
            th = threading.Thread(target=stage, args=(j,))
            th.start()
            return th

        t_all = time.perf_counter()
        worker = start(0)
        for j, g in enumerate(groups):
            worker.join()
            if errors:
                raise errors[0]
            # group j + 1 takes the contexts group j - 1 has finished with
            worker = start(j + 1) if j + 1 < len(groups) else None
            # the group's Krylov-Schur runs, one thread per context, so their
            # latency-bound steps overlap each other and the next group's
            # factorization
            recs, errs = [None] * len(g), []

            def krylov(pos, k, warm):
                try:
                    recs[pos] = self._krylov_record(self.ctx[key(j, pos, k)], cases[k][0],
                                                    complex(shifts[k]), form, precision, warm,
                                                    time.perf_counter())
                except Exception as e:  # re-raised below
                    errs.append(e)

            jobs = [(pos, k, warm_form and key(j, pos, k) in solved) for pos, k in enumerate(g)]
            if self.threads:
                ths = [threading.Thread(target=krylov, args=job) for job in jobs]
                for th in ths:
                    th.start()
                for th in ths:
                    th.join()
            else:
                for job in jobs:
                    krylov(*job)
            if errs:
                if worker:
                    worker.join()
                raise errs[0]
            for pos, k in enumerate(g):
                solved.add(key(j, pos, k))
                result.records.append(recs[pos])
                if on_record:
                    on_record(recs[pos])
        result.seconds = time.perf_counter() - t_all
        return result


def prepare(instances, omega_s: complex = 0.02 + 0j) -> list:
    """Host-side case construction (preprocessing, outside any timed region)."""
    return [(inst, inst.build(omega_s)) for inst in instances]


def label_omegas(record: Record) -> np.ndarray:
    """Training label of an instance: its 8 omegas by growth rate."""
    from .shift_model import leading_order
    return leading_order(record.omegas)


def alpha_order(instances) -> list:
    """Indices ordered by grid size, then wavenumber (the learned form's
    processing order: consecutive instances are close in alpha)."""
    return sorted(range(len(instances)),
                  key=lambda k: (instances[k].ny * instances[k].nz, instances[k].alpha))


def isclose_alpha(a: float, b: float) -> bool:
    return math.isclose(a, b, rel_tol=0.0, abs_tol=1e-12)
