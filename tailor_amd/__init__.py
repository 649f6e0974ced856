"""tailor_amd -- MI355X-native BiGlobal assemble-and-solve path of Tailor.

The package is a thin Python host over ``libtailor.so`` (HIP kernels for
gfx950 plus C++20 host code behind the C ABI declared in include/tailor.h).
Python only builds inputs, marshals them across the C ABI and formats the
diagnostics; every arithmetic step of the hot path runs in the library.

The stage objects mirror the reference's stage classes
(Metrics / BaseFlowDerivatives / GlobalOperator / BoundaryConditions /
EigenSolver, src/CPlusPlus/*.hpp): ``Context.assemble`` is Metrics::compute
+ BaseFlowDerivatives::compute + GlobalOperator::assemble +
BoundaryConditions::apply (Process.cpp:53-56).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .case import Case, Physics, synthetic_case, uniform_case  # noqa: F401

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libtailor.so"

OP_L, OP_M, OP_A, OP_B = 0, 1, 2, 3

STATUS_NAMES = {
    0: "OK", 1: "FILE_UNEXPECTED", 2: "FP", 3: "ARG_OUTOFRANGE", 4: "SUP", 5: "PLIB",
    6: "NOT_CONVERGED", 7: "USER_INPUT", 8: "ARG_WRONGSTATE", 9: "MEM", 10: "HIP",
}


class TailorError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{STATUS_NAMES.get(status, status)}] {message}")
        self.status = status
        self.message = message


class _Rule(C.Structure):
    _fields_ = [("node_count", C.c_int32), ("q", C.c_int32), ("periodic", C.c_int32),
                ("period", C.c_double), ("nodes", C.POINTER(C.c_double)),
                ("stencil_indices", C.POINTER(C.c_int64)),
                ("stencil_offsets", C.POINTER(C.c_int64)), ("weights", C.POINTER(C.c_double))]


class _Case(C.Structure):
    _fields_ = [("ny", C.c_int32), ("nz", C.c_int32), ("xi", _Rule), ("eta", _Rule),
                ("grid", C.POINTER(C.c_double)), ("baseflow", C.POINTER(C.c_double)),
                ("spanwise_period", C.c_double), ("reynolds", C.c_double), ("mach", C.c_double),
                ("prandtl", C.c_double), ("gamma", C.c_double), ("gas_constant", C.c_double),
                ("reference_mu", C.c_double), ("reference_temperature", C.c_double),
                ("sutherland_constant", C.c_double), ("alpha_re", C.c_double),
                ("alpha_im", C.c_double)]


class AssemblyDiag(C.Structure):
    _fields_ = [("min_jacobian", C.c_double), ("max_jacobian", C.c_double),
                ("max_grid_imag", C.c_double), ("metric_norms", C.c_double * 10),
                ("max_baseflow_imag", C.c_double), ("derivative_norms", C.c_double * 25),
                ("rows", C.c_int64), ("block_size", C.c_int64), ("mass_blocks", C.c_int64),
                ("spatial_blocks", C.c_int64), ("mass_fro", C.c_double),
                ("spatial_fro", C.c_double), ("wall_nodes", C.c_int32),
                ("farfield_nodes", C.c_int32), ("incoming_modes", C.c_int32),
                ("neutral_modes", C.c_int32), ("min_incoming_per_node", C.c_int32),
                ("max_incoming_per_node", C.c_int32), ("max_characteristic_imag", C.c_double),
                ("max_eigenvector_condition", C.c_double), ("a_fro", C.c_double),
                ("b_fro", C.c_double)]

    def as_dict(self) -> dict:
        out = {}
        for name, _t in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out


class FactorDiag(C.Structure):
    _fields_ = [("fronts", C.c_int32), ("levels", C.c_int32), ("max_front", C.c_int32),
                ("perturbed_pivots", C.c_int32), ("arena_bytes", C.c_int64),
                ("factor_entries", C.c_int64), ("flops", C.c_double),
                ("symbolic_ms", C.c_double), ("factor_ms", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("nev", C.c_int32), ("ncv", C.c_int32), ("maxit", C.c_int32),
                ("tol", C.c_double), ("omega_re", C.c_double), ("omega_im", C.c_double),
                ("refine_max", C.c_int32), ("refine_tol", C.c_double),
                ("warm_start", C.c_int32), ("check_from", C.c_int32)]


class TuneReport(C.Structure):
    _fields_ = [("levels", C.c_int32), ("order", C.c_int32 * 64), ("gemm_ms", C.c_double * 8),
                ("tuned_gemm_ms", C.c_double)]


class SolveDiag(C.Structure):
    _fields_ = [("requested", C.c_int32), ("converged", C.c_int32), ("iterations", C.c_int32),
                ("solves", C.c_int32), ("refinement_steps", C.c_int32),
                ("perturbed_pivots", C.c_int32), ("lambda_target_re", C.c_double),
                ("lambda_target_im", C.c_double), ("max_residual", C.c_double),
                ("max_relative_error", C.c_double), ("symbolic_ms", C.c_double),
                ("factor_ms", C.c_double), ("solve_ms", C.c_double), ("fronts", C.c_int64),
                ("factor_bytes", C.c_int64), ("factor_flops", C.c_double)]


def _sdict(st) -> dict:
    return {name: getattr(st, name) for name, _t in st._fields_}


_lib = None


def library() -> C.CDLL:
    """Load libtailor.so; raises if the HIP library has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m tailor_amd.build` "
                           "(the HIP library is required; there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    P = C.POINTER
    lib.tailor_create.argtypes = [P(C.c_void_p), C.c_int]
    lib.tailor_destroy.argtypes = [C.c_void_p]
    lib.tailor_destroy.restype = None
    lib.tailor_last_error.argtypes = [C.c_void_p]
    lib.tailor_last_error.restype = C.c_char_p
    lib.tailor_load_case.argtypes = [C.c_void_p, P(_Case)]
    lib.tailor_set_alpha.argtypes = [C.c_void_p, C.c_double, C.c_double]
    lib.tailor_assemble.argtypes = [C.c_void_p, P(AssemblyDiag)]
    lib.tailor_assemble_unconstrained.argtypes = [C.c_void_p]
    lib.tailor_bsr_shape.argtypes = [C.c_void_p, P(C.c_int64), P(C.c_int64)]
    lib.tailor_export_bsr.argtypes = [C.c_void_p, C.c_int, P(C.c_int32), P(C.c_int32),
                                      P(C.c_double)]
    lib.tailor_export_fields.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double)]
    lib.tailor_time_assembly.argtypes = [C.c_void_p, C.c_int, P(C.c_double), P(C.c_double)]
    lib.tailor_factor.argtypes = [C.c_void_p, C.c_double, C.c_double, P(FactorDiag)]
    lib.tailor_factor_graph.argtypes = [C.c_void_p, C.c_double, C.c_double, P(FactorDiag)]
    lib.tailor_factor_batch.argtypes = [P(C.c_void_p), C.c_int32, P(C.c_double), P(C.c_double)]
    lib.tailor_linear_solve.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double), C.c_int32,
                                        C.c_double, P(C.c_double), P(C.c_int32)]
    lib.tailor_solve.argtypes = [C.c_void_p, P(SolveOpts), P(SolveDiag)]
    lib.tailor_set_precision.argtypes = [C.c_void_p, C.c_int32]
    lib.tailor_autotune.argtypes = [C.c_void_p, C.c_double, C.c_double, P(TuneReport)]
    lib.tailor_set_gemm_config.argtypes = [C.c_void_p, C.c_int32]
    lib.tailor_set_panel_kernel.argtypes = [C.c_void_p, C.c_int32]
    lib.tailor_set_block_pairs.argtypes = [C.c_void_p, C.c_int32]
    lib.tailor_profile_factor.argtypes = [C.c_void_p, C.c_double, C.c_double, P(C.c_double),
                                          P(C.c_double), P(C.c_double), P(C.c_int32)]
    lib.tailor_profile_gemm.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int32,
                                        P(C.c_int32), P(C.c_double), P(C.c_double),
                                        P(C.c_double), P(C.c_double)]
    lib.tailor_time_linear_solve.argtypes = [C.c_void_p, C.c_int32, P(C.c_double), P(C.c_double)]
    lib.tailor_time_operator.argtypes = [C.c_void_p, C.c_int32, P(C.c_double), P(C.c_double)]
    lib.tailor_linear_solve_pair.argtypes = [C.c_void_p, P(C.c_double), P(C.c_double),
                                             P(C.c_double), P(C.c_double)]
    # device pointers (caller-owned device memory), no device allocation
    lib.tailor_linear_solve_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
    lib.tailor_batch_create.argtypes = [P(C.c_void_p), P(C.c_void_p), C.c_int32]
    lib.tailor_batch_destroy.argtypes = [C.c_void_p]
    lib.tailor_batch_destroy.restype = None
    lib.tailor_batch_last_error.argtypes = [C.c_void_p]
    lib.tailor_batch_last_error.restype = C.c_char_p
    lib.tailor_batch_factor.argtypes = [C.c_void_p, P(C.c_double), C.c_int32, P(C.c_double)]
    lib.tailor_batch_assemble_factor.argtypes = [C.c_void_p, P(C.c_double), C.c_int32,
                                                 P(C.c_double)]
    lib.tailor_batch_autotune.argtypes = [C.c_void_p, C.c_double, C.c_double, P(TuneReport)]
    lib.tailor_batch_set_gemm_config.argtypes = [C.c_void_p, C.c_int32]
    lib.tailor_measure_peaks.argtypes = [C.c_int, P(C.c_double), P(C.c_double)]
    lib.tailor_eigenpair_count.argtypes = [C.c_void_p]
    lib.tailor_get_eigenpair.argtypes = [C.c_void_p, C.c_int32, P(C.c_double), P(C.c_double),
                                         P(C.c_double), P(C.c_double)]
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _rule_struct(rule, keep: list) -> _Rule:
    nodes = np.ascontiguousarray(rule.nodes, dtype=np.float64)
    idx = np.ascontiguousarray(rule.stencil_indices, dtype=np.int64)
    off = np.ascontiguousarray(rule.stencil_offsets, dtype=np.int64)
    w = np.ascontiguousarray(rule.weights, dtype=np.float64)
    keep += [nodes, idx, off, w]
    return _Rule(rule.node_count, rule.q, 1 if rule.topology == "periodic" else 0,
                 float(rule.period), _dptr(nodes), idx.ctypes.data_as(C.POINTER(C.c_int64)),
                 off.ctypes.data_as(C.POINTER(C.c_int64)), _dptr(w))


class Context:
    """One GPU context: device buffers of one case and its operators."""

    def __init__(self, device: int = 0):
        self.lib = library()
        self.ptr = C.c_void_p()
        st = self.lib.tailor_create(C.byref(self.ptr), device)
        if st:
            raise TailorError(st, "tailor_create failed")
        self.case: Case | None = None
        self._precision = PRECISION_FP64

    def close(self):
        if self.ptr:
            self.lib.tailor_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int):
        if st:
            raise TailorError(st, self.lib.tailor_last_error(self.ptr).decode())

    # -- case --------------------------------------------------------------
    def load_case(self, case: Case) -> None:
        keep: list = []
        grid = np.ascontiguousarray(case.grid, dtype=np.complex128).view(np.float64)
        base = np.ascontiguousarray(case.baseflow, dtype=np.complex128).view(np.float64)
        ph = case.physics
        cs = _Case(case.ny, case.nz, _rule_struct(case.xi_rule, keep),
                   _rule_struct(case.eta_rule, keep), _dptr(grid), _dptr(base),
                   float(case.spanwise_period), ph.reynolds, ph.mach, ph.prandtl, ph.gamma,
                   ph.gas_constant, ph.reference_mu, ph.reference_temperature,
                   ph.sutherland_constant, case.alpha.real, case.alpha.imag)
        self._check(self.lib.tailor_load_case(self.ptr, C.byref(cs)))
        self.case = case

    def set_alpha(self, alpha: complex) -> None:
        alpha = complex(alpha)
        self._check(self.lib.tailor_set_alpha(self.ptr, alpha.real, alpha.imag))

    # -- assembly ----------------------------------------------------------
    def assemble(self, diagnostics: bool = True) -> dict | None:
        if diagnostics:
            d = AssemblyDiag()
            self._check(self.lib.tailor_assemble(self.ptr, C.byref(d)))
            return d.as_dict()
        self._check(self.lib.tailor_assemble(self.ptr, None))
        return None

    def assemble_unconstrained(self) -> None:
        self._check(self.lib.tailor_assemble_unconstrained(self.ptr))

    def shape(self) -> tuple[int, int]:
        rows, nnzb = C.c_int64(), C.c_int64()
        self._check(self.lib.tailor_bsr_shape(self.ptr, C.byref(rows), C.byref(nnzb)))
        return rows.value, nnzb.value

    def export_bsr(self, which: int):
        rows, nnzb = self.shape()
        if which in (OP_L, OP_A):
            rowptr = np.zeros(rows + 1, dtype=np.int32)
            colidx = np.zeros(nnzb, dtype=np.int32)
            vals = np.zeros(nnzb * 50)
        else:
            rowptr = np.zeros(rows + 1, dtype=np.int32)
            colidx = np.zeros(rows, dtype=np.int32)
            vals = np.zeros(rows * 50)
        self._check(self.lib.tailor_export_bsr(
            self.ptr, which, rowptr.ctypes.data_as(C.POINTER(C.c_int32)),
            colidx.ctypes.data_as(C.POINTER(C.c_int32)), _dptr(vals)))
        return rowptr, colidx, vals.view(np.complex128).reshape(-1, 5, 5)

    def export_fields(self):
        rows, _ = self.shape()
        m = np.zeros((rows, 10))
        dd = np.zeros((rows, 25))
        self._check(self.lib.tailor_export_fields(self.ptr, _dptr(m), _dptr(dd)))
        return m, dd

    def time_assembly(self, reps: int = 10) -> tuple[float, float]:
        w, f = C.c_double(), C.c_double()
        self._check(self.lib.tailor_time_assembly(self.ptr, reps, C.byref(w), C.byref(f)))
        return w.value, f.value


PRECISION_FP64, PRECISION_MIXED = 0, 1
PANEL_AUTO, PANEL_STREAMING = 0, 1


def _set_precision(self, precision: int) -> None:
    """PRECISION_FP64 (plain form) or PRECISION_MIXED (complex FP32 fronts,
    FP64 refinement of every solve)."""
    self._check(self.lib.tailor_set_precision(self.ptr, int(precision)))
    self._precision = int(precision)


def _profile_factor(self, sigma: complex) -> dict:
    """Factor with HIP events around every frontal GEMM launch."""
    sigma = complex(sigma)
    f, g, fl, n = C.c_double(), C.c_double(), C.c_double(), C.c_int32()
    self._check(self.lib.tailor_profile_factor(self.ptr, sigma.real, sigma.imag, C.byref(f),
                                               C.byref(g), C.byref(fl), C.byref(n)))
    return {"factor_ms": f.value, "gemm_ms": g.value, "gemm_flops": fl.value,
            "gemm_launches": n.value}


# indices 1 and 3 are unused (the block rows of the trailing columns and of
# Z are finished by triangular sweeps)
GEMM_REGIONS = ("inner", "unused1", "trail", "unused3", "trailhead", "trailrest", "zblk",
                "pairrows")


def _profile_gemm(self, sigma: complex, max_levels: int = 64) -> dict:
    """Factor with HIP events around every frontal GEMM launch and return
    the GEMM time and flops (8 per complex multiply-add) per elimination
    level and per update region."""
    sigma = complex(sigma)
    n = C.c_int32()
    lms, lfl = (C.c_double * max_levels)(), (C.c_double * max_levels)()
    rms, rfl = (C.c_double * 8)(), (C.c_double * 8)()
    self._check(self.lib.tailor_profile_gemm(self.ptr, sigma.real, sigma.imag, int(max_levels),
                                             C.byref(n), lms, lfl, rms, rfl))
    nl = min(n.value, max_levels)
    return {"level_ms": list(lms)[:nl], "level_flops": list(lfl)[:nl],
            "region_ms": dict(zip(GEMM_REGIONS, list(rms))),
            "region_flops": dict(zip(GEMM_REGIONS, list(rfl)))}


def _autotune(self, sigma: complex) -> dict:
    """Choose the frontal GEMM tile shape and order per elimination level by
    timing (configuration c = 2 * shape + order)."""
    sigma = complex(sigma)
    r = TuneReport()
    self._check(self.lib.tailor_autotune(self.ptr, sigma.real, sigma.imag, C.byref(r)))
    return {"levels": r.levels, "config": list(r.order)[:r.levels],
            "gemm_ms": [round(x, 3) for x in list(r.gemm_ms)],
            "tuned_gemm_ms": r.tuned_gemm_ms}


def _factor(self, sigma: complex) -> dict:
    d = FactorDiag()
    sigma = complex(sigma)
    self._check(self.lib.tailor_factor(self.ptr, sigma.real, sigma.imag, C.byref(d)))
    return _sdict(d)


def _factor_graph(self, sigma: complex) -> dict:
    """tailor_factor replayed from the context's captured HIP graph."""
    d = FactorDiag()
    sigma = complex(sigma)
    self._check(self.lib.tailor_factor_graph(self.ptr, sigma.real, sigma.imag, C.byref(d)))
    return _sdict(d)


def factor_batch(contexts, sigmas) -> float:
    """Factor several contexts (one device) at their shifts: each context's
    factorization graph on its own stream, launched from this thread, so
    they run concurrently; returns the batch's time in ms."""
    contexts = list(contexts)
    if len(contexts) != len(sigmas) or not contexts:
        raise ValueError("one shift per context")
    ptrs = (C.c_void_p * len(contexts))(*[c.ptr.value for c in contexts])
    sg = np.array([[complex(s).real, complex(s).imag] for s in sigmas], dtype=np.float64).ravel()
    ms = C.c_double()
    lib = contexts[0].lib
    st = lib.tailor_factor_batch(ptrs, len(contexts), _dptr(sg), C.byref(ms))
    if st:
        msgs = [lib.tailor_last_error(c.ptr).decode() for c in contexts]
        raise TailorError(st, "; ".join(m for m in msgs if m) or "tailor_factor_batch failed")
    return ms.value


class Batch:
    """The factorizations of several contexts (one device, any grid sizes) as
    one launch sequence over their joined elimination levels
    (tailor_batch_*); each context then holds its own factors, as after
    Context.factor, for its solves and Krylov-Schur."""

    def __init__(self, contexts):
        self.contexts = list(contexts)
        if not self.contexts:
            raise ValueError("a batch needs at least one context")
        self.lib = self.contexts[0].lib
        self.ptr = C.c_void_p()
        ptrs = (C.c_void_p * len(self.contexts))(*[c.ptr.value for c in self.contexts])
        st = self.lib.tailor_batch_create(C.byref(self.ptr), ptrs, len(self.contexts))
        if st:
            raise TailorError(st, "tailor_batch_create failed (contexts must be distinct, "
                                  "on one device, each with a case loaded)")

    def close(self):
        if self.ptr:
            self.lib.tailor_batch_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st: int):
        if st:
            raise TailorError(st, self.lib.tailor_batch_last_error(self.ptr).decode())

    def _run(self, fn, sigmas, graph: bool) -> float:
        sigmas = list(sigmas)
        if len(sigmas) != len(self.contexts):
            raise ValueError("one shift per context")
        sg = np.array([[complex(x).real, complex(x).imag] for x in sigmas],
                      dtype=np.float64).ravel()
        ms = C.c_double()
        self._check(fn(self.ptr, _dptr(sg), 1 if graph else 0, C.byref(ms)))
        return ms.value

    def factor(self, sigmas, graph: bool = True) -> float:
        """Factor context k at sigmas[k] (each assembled); returns the
        batch's time in ms."""
        return self._run(self.lib.tailor_batch_factor, sigmas, graph)

    def assemble_factor(self, sigmas, graph: bool = True) -> float:
        """Assemble every context's A_bc, B_bc (each with a case loaded)
        and factor context k at sigmas[k], as one launch sequence; returns
        the batch's time in ms."""
        return self._run(self.lib.tailor_batch_assemble_factor, sigmas, graph)

    def autotune(self, sigma: complex) -> dict:
        sigma = complex(sigma)
        r = TuneReport()
        self._check(self.lib.tailor_batch_autotune(self.ptr, sigma.real, sigma.imag, C.byref(r)))
        return {"levels": r.levels, "config": list(r.order)[:r.levels],
                "gemm_ms": [round(x, 3) for x in list(r.gemm_ms)],
                "tuned_gemm_ms": r.tuned_gemm_ms}

    def set_gemm_config(self, config: int) -> None:
        self._check(self.lib.tailor_batch_set_gemm_config(self.ptr, int(config)))


def measure_peaks(device: int = 0) -> dict:
    """The device's sustained dense FP64 MFMA rate (TFLOP/s) and streaming
    HBM read rate (GB/s), each from a kernel that does nothing else."""
    lib = library()
    tf, gbs = C.c_double(), C.c_double()
    st = lib.tailor_measure_peaks(int(device), C.byref(tf), C.byref(gbs))
    if st:
        raise TailorError(st, "tailor_measure_peaks failed")
    return {"fp64_mfma_tflops": tf.value, "read_gbs": gbs.value}


def _linear_solve(self, b: np.ndarray, refine_max: int = 0, refine_tol: float = 1e-14):
    """x = (A_bc - sigma B_bc)^-1 b for a natural-order complex vector."""
    b = np.ascontiguousarray(b, dtype=np.complex128)
    x = np.zeros_like(b)
    rel = C.c_double(-1.0)
    steps = C.c_int32(0)
    self._check(self.lib.tailor_linear_solve(self.ptr, _dptr(b.view(np.float64)),
                                             _dptr(x.view(np.float64)), refine_max, refine_tol,
                                             C.byref(rel), C.byref(steps)))
    return x, rel.value, steps.value


def _solve(self, nev: int = 8, omega: complex = 0j, tol: float = 1e-10, maxit: int = 1000,
           ncv: int = 0, refine_max: int | None = None, refine_tol: float = 1e-12,
           raise_on_failure: bool = True, warm_start: bool = False, check_from: int = 0) -> dict:
    """Shift-invert Krylov-Schur (EigenSolver::solve).  With warm_start the
    iteration starts from the previous solve's Ritz vectors; with
    check_from = m > 0 convergence is also tested after every second step
    from step m on, and a pass ends as soon as the wanted pairs' Ritz
    estimates meet tol.  refine_max defaults to 0 on FP64 factors (the plain
    form: one pass over the factors per solve, as PCLU + KSPPREONLY) and to
    20 on complex FP32 factors, where every solve is refined in FP64 to a
    relative residual of refine_tol."""
    omega = complex(omega)
    if check_from < 0:
        raise ValueError("check_from must be >= 0")
    if refine_max is None:
        refine_max = 20 if getattr(self, "_precision", PRECISION_FP64) == PRECISION_MIXED else 0
    o = SolveOpts(nev, ncv, maxit, tol, omega.real, omega.imag, refine_max, refine_tol,
                  1 if warm_start else 0, check_from)
    d = SolveDiag()
    st = self.lib.tailor_solve(self.ptr, C.byref(o), C.byref(d))
    out = _sdict(d)
    out["status"] = st
    if st and raise_on_failure:
        self._check(st)
    return out


def _eigenpairs(self, vectors: bool = False) -> list:
    cnt = self.lib.tailor_eigenpair_count(self.ptr)
    rows, _ = self.shape()
    out = []
    for k in range(cnt):
        lam = np.zeros(2)
        res, rel = C.c_double(), C.c_double()
        q = np.zeros(2 * 5 * rows) if vectors else None
        self._check(self.lib.tailor_get_eigenpair(self.ptr, k, _dptr(lam),
                                                  _dptr(q) if vectors else None,
                                                  C.byref(res), C.byref(rel)))
        out.append({"lambda": complex(lam[0], lam[1]), "omega": 1j * complex(lam[0], lam[1]),
                    "residual": res.value, "relative_error": rel.value,
                    "mode": q.view(np.complex128) if vectors else None})
    return out


def _linear_solve_pair(self, b1: np.ndarray, b2: np.ndarray):
    """x1, x2 = K^-1 b1, K^-1 b2 in one pass over the factors (no refinement)."""
    b1 = np.ascontiguousarray(b1, dtype=np.complex128)
    b2 = np.ascontiguousarray(b2, dtype=np.complex128)
    if b1.shape != b2.shape:
        raise ValueError("right-hand sides differ in shape")
    x1, x2 = np.zeros_like(b1), np.zeros_like(b2)
    self._check(self.lib.tailor_linear_solve_pair(
        self.ptr, _dptr(b1.view(np.float64)), _dptr(b2.view(np.float64)),
        _dptr(x1.view(np.float64)), _dptr(x2.view(np.float64))))
    return x1, x2


Context.linear_solve_pair = _linear_solve_pair


def _time_linear_solve(self, reps: int = 20):
    ms, nbytes = C.c_double(), C.c_double()
    self._check(self.lib.tailor_time_linear_solve(self.ptr, reps, C.byref(ms), C.byref(nbytes)))
    return ms.value, nbytes.value


Context.time_linear_solve = _time_linear_solve


def _time_operator(self, reps: int = 20):
    """Mean ms and algorithmic bytes of one y = (A_bc - sigma B_bc) x."""
    ms, nbytes = C.c_double(), C.c_double()
    self._check(self.lib.tailor_time_operator(self.ptr, reps, C.byref(ms), C.byref(nbytes)))
    return ms.value, nbytes.value


Context.time_operator = _time_operator
Context.factor = _factor
Context.factor_graph = _factor_graph
Context.set_precision = _set_precision
Context.profile_factor = _profile_factor
Context.profile_gemm = _profile_gemm
Context.autotune = _autotune
Context.set_gemm_config = lambda self, c: self._check(self.lib.tailor_set_gemm_config(self.ptr, int(c)))
Context.set_panel_kernel = lambda self, k: self._check(self.lib.tailor_set_panel_kernel(self.ptr, int(k)))
Context.set_block_pairs = lambda self, on: self._check(self.lib.tailor_set_block_pairs(self.ptr, 1 if on else 0))
Context.linear_solve = _linear_solve
Context.solve = _solve
Context.eigenpairs = _eigenpairs


def format_assembly_report(case: Case, d: dict) -> str:
    """The stdout lines of Process.cpp:62-152 (regex-parsed by tests)."""
    grid_norm = float(np.linalg.norm(case.grid))
    base_norm = float(np.linalg.norm(case.baseflow))
    names = ["xi_y", "xi_z", "eta_y", "eta_z", "xi_yy", "xi_zz", "xi_yz", "eta_yy", "eta_zz",
             "eta_yz"]
    dn = ["dy", "dz", "dyy", "dzz", "dyz"]
    width = max(int(np.abs(case.xi_rule.stencil_offsets).max()),
                int(np.abs(case.eta_rule.stencil_offsets).max()))
    lines = [
        f"Prepared case '{case.title}'",
        f"  Ny x Nz: {case.ny} x {case.nz}",
        f"  alpha: {case.alpha.real:.16g} {case.alpha.imag:+.16g}i",
        f"  FD-q degrees: y={case.xi_rule.q}, z={case.eta_rule.q}; DMDA stencil width: {width}",
        f"  xi rule: {case.ny} nodes, {case.xi_rule.stencil_size}-point stencil",
        f"  eta rule: {case.nz} nodes, {case.eta_rule.stencil_size}-point stencil",
        f"  ||baseflow||2: {base_norm:.16g}, ||grid||2: {grid_norm:.16g}",
        f"  Jacobian range: [{d['min_jacobian']:.16g}, {d['max_jacobian']:.16g}]",
        f"  max |Im(grid y/z)|: {d['max_grid_imag']:.16g}",
    ]
    for k, nm in enumerate(names):
        lines.append(f"  metric norm {nm}: {d['metric_norms'][k]:.16g}")
    lines.append(f"  base-flow derivative components: 25; max |Im(baseflow)|: "
                 f"{d['max_baseflow_imag']:.16g}")
    for k, nm in enumerate(dn):
        vals = " ".join(f"{v:.16g}" for v in d["derivative_norms"][5 * k:5 * k + 5])
        lines.append(f"  base-flow derivative norm {nm}: {vals}")
    lines += [
        f"  global matrices: {d['rows']} x {d['rows']}, block size {d['block_size']}",
        f"  M_Gamma blocks used/allocated: {d['mass_blocks']} / {d['mass_blocks']}; "
        f"||M_Gamma||F: {d['mass_fro']:.16g}",
        f"  L blocks used/allocated: {d['spatial_blocks']} / {d['spatial_blocks']}; "
        f"||L||F: {d['spatial_fro']:.16g}",
        "  eigenvalue convention: (-L) q = lambda M_Gamma q, lambda = -i omega",
        f"  boundaries: wall nodes={d['wall_nodes']}, far-field nodes={d['farfield_nodes']}",
        f"  far-field incoming modes={d['incoming_modes']} (per node "
        f"{d['min_incoming_per_node']}..{d['max_incoming_per_node']}), "
        f"neutral modes={d['neutral_modes']}",
        f"  max characteristic |imag(c)|: {d['max_characteristic_imag']:.16g}; "
        f"max cond(R): {d['max_eigenvector_condition']:.16g}",
        f"  ||A_bc||F: {d['a_fro']:.16g}; ||B_bc||F: {d['b_fro']:.16g}",
    ]
    return "\n".join(lines)
