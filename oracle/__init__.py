"""CPU FP64 oracle of the BiGlobal assemble path (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this package, and only as the checker.  The product path (tailor_amd) never
imports it.

``tailor_oracle.c`` restates the reference's Metrics, BaseFlowDerivatives,
LNSCoefficients, StreamwiseFourier, CurvilinearTransform, GlobalOperator and
BoundaryConditions stages in plain C (file:line citations are in the C file).
``eigen.py`` restates the shift-invert Krylov-Schur solve.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SOURCE = HERE / "tailor_oracle.c"
LIB = HERE / "liboracle.so"

_lib = None


def build(force: bool = False) -> Path:
    """Compile the oracle with gcc (no FP contraction, IEEE order kept)."""
    if force or not LIB.exists() or LIB.stat().st_mtime < SOURCE.stat().st_mtime:
        tmp = LIB.with_suffix(".so.tmp")
        subprocess.run(["gcc", "-O2", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
                        "-fPIC", "-shared", "-Wall", "-Wextra", "-o", str(tmp), str(SOURCE), "-lm"],
                       check=True)
        os.replace(tmp, LIB)
    return LIB


class _Rule(C.Structure):
    _fields_ = [("n", C.c_int), ("ss", C.c_int),
                ("idx", C.POINTER(C.c_long)), ("off", C.POINTER(C.c_long)),
                ("w0", C.POINTER(C.c_double)), ("w1", C.POINTER(C.c_double)),
                ("w2", C.POINTER(C.c_double))]


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class _RuleHolder:
    def __init__(self, rule):
        self.idx = np.ascontiguousarray(rule.stencil_indices, dtype=np.int64)
        self.off = np.ascontiguousarray(rule.stencil_offsets, dtype=np.int64)
        self.w = [np.ascontiguousarray(rule.weights[m], dtype=np.float64) for m in range(3)]
        self.s = _Rule(rule.node_count, rule.stencil_size,
                       self.idx.ctypes.data_as(C.POINTER(C.c_long)),
                       self.off.ctypes.data_as(C.POINTER(C.c_long)),
                       _dp(self.w[0]), _dp(self.w[1]), _dp(self.w[2]))


class OracleError(RuntimeError):
    def __init__(self, stage: str, code: int):
        super().__init__(f"oracle {stage} failed with status {code}")
        self.stage = stage
        self.code = code


def lns_physical(recipe: np.ndarray, point: np.ndarray) -> np.ndarray:
    """11 real 5x5 blocks from a (10, 5) point (value, x, y, z, xx, yy, zz, xy, xz, yz)."""
    out = np.zeros(275)
    pt = np.ascontiguousarray(point, dtype=np.float64).reshape(50)
    st = lib().tor_lns_physical(_dp(np.ascontiguousarray(recipe, dtype=np.float64)), _dp(pt), _dp(out))
    if st:
        raise OracleError("lns", st)
    return out.reshape(11, 5, 5)


def fourier(physical: np.ndarray, alpha: complex) -> np.ndarray:
    out = np.zeros(350)
    ph = np.ascontiguousarray(physical, dtype=np.float64).reshape(275)
    st = lib().tor_fourier(_dp(ph), C.c_double(alpha.real), C.c_double(alpha.imag), _dp(out))
    if st:
        raise OracleError("fourier", st)
    return out.view(np.complex128).reshape(7, 5, 5)


def curvilinear(four: np.ndarray, metrics: np.ndarray) -> np.ndarray:
    out = np.zeros(350)
    f = np.ascontiguousarray(four, dtype=np.complex128).reshape(175).view(np.float64)
    m = np.ascontiguousarray(metrics, dtype=np.float64)
    st = lib().tor_curvilinear(_dp(f), _dp(m), _dp(out))
    if st:
        raise OracleError("curvilinear", st)
    return out.view(np.complex128).reshape(7, 5, 5)


def characteristics(recipe: np.ndarray, value: np.ndarray, metrics: np.ndarray):
    R = np.zeros(50); Ri = np.zeros(50); P = np.zeros(50); sp = np.zeros(5)
    fl = np.zeros(5, dtype=np.int32); ex = np.zeros(2)
    st = lib().tor_characteristics(_dp(np.ascontiguousarray(recipe, dtype=np.float64)),
                                   _dp(np.ascontiguousarray(value, dtype=np.float64)),
                                   _dp(np.ascontiguousarray(metrics, dtype=np.float64)),
                                   _dp(R), _dp(Ri), _dp(P), _dp(sp), _ip(fl), _dp(ex))
    if st:
        raise OracleError("characteristics", st)
    c = lambda a: a.view(np.complex128).reshape(5, 5)
    return dict(R=c(R), Rinv=c(Ri), P=c(P), speeds=sp, incoming=(fl & 1) == 1,
                neutral=(fl & 2) == 2, max_imag=ex[0], cond=ex[1])


@dataclass
class Assembly:
    metrics: np.ndarray      # (N, 10)
    metric_diag: np.ndarray  # min J, max J, max |Im grid|
    deriv: np.ndarray        # (N, 25)
    rowptr: np.ndarray       # (N+1,) int32
    colidx: np.ndarray       # (nnzb,) int32
    L: np.ndarray            # (nnzb, 5, 5) complex128
    M: np.ndarray            # (N, 5, 5) complex128
    A: np.ndarray            # (nnzb, 5, 5) complex128  (A_bc)
    B: np.ndarray            # (N, 5, 5) complex128     (B_bc, block diagonal)
    bc_diag: np.ndarray      # 8 values, see tor_apply_bc


def assemble(case, with_bc: bool = True) -> Assembly:
    """Metrics -> derivatives -> L, M -> A_bc, B_bc, all on the CPU."""
    L_ = lib()
    xr, er = _RuleHolder(case.xi_rule), _RuleHolder(case.eta_rule)
    ny, nz = case.ny, case.nz
    n = ny * nz
    grid = np.ascontiguousarray(case.grid, dtype=np.complex128).view(np.float64).reshape(-1)
    base = np.ascontiguousarray(case.baseflow, dtype=np.complex128).view(np.float64).reshape(-1)
    metrics = np.zeros((n, 10))
    mdiag = np.zeros(3)
    st = L_.tor_metrics(ny, nz, C.byref(xr.s), C.byref(er.s), C.c_double(case.spanwise_period),
                        _dp(grid), _dp(metrics), _dp(mdiag))
    if st:
        raise OracleError("metrics", st)
    deriv = np.zeros((n, 25))
    ddiag = np.zeros(1)
    st = L_.tor_derivatives(ny, nz, C.byref(xr.s), C.byref(er.s), _dp(base), _dp(metrics),
                            _dp(deriv), _dp(ddiag))
    if st:
        raise OracleError("derivatives", st)
    per = case.xi_rule.stencil_size * case.eta_rule.stencil_size
    nnzb = n * per
    rowptr = np.zeros(n + 1, dtype=np.int32)
    colidx = np.zeros(nnzb, dtype=np.int32)
    lv = np.zeros(nnzb * 50)
    mv = np.zeros(n * 50)
    recipe = case.physics.as_array()
    st = L_.tor_assemble(ny, nz, C.byref(xr.s), C.byref(er.s), _dp(recipe),
                         C.c_double(case.alpha.real), C.c_double(case.alpha.imag),
                         _dp(base), _dp(deriv), _dp(metrics), _ip(rowptr), _ip(colidx),
                         _dp(lv), _dp(mv))
    if st:
        raise OracleError("assemble", st)
    av = np.zeros_like(lv) if with_bc else None
    bv = np.zeros_like(mv) if with_bc else None
    bcd = np.zeros(8)
    if with_bc:
        st = L_.tor_apply_bc(ny, nz, C.byref(xr.s), C.byref(er.s), _dp(recipe), _dp(base),
                             _dp(deriv), _dp(metrics), _ip(colidx), _dp(lv), _dp(mv),
                             _dp(av), _dp(bv), _dp(bcd))
        if st:
            raise OracleError("boundary", st)
    cv = lambda a, k: None if a is None else a.view(np.complex128).reshape(k, 5, 5)
    return Assembly(metrics, mdiag, deriv, rowptr, colidx, cv(lv, nnzb), cv(mv, n),
                    cv(av, nnzb), cv(bv, n), bcd)


def to_scipy(rowptr, colidx, vals, n_block_rows):
    """BSR arrays -> scipy.sparse.bsr_matrix (5x5 blocks)."""
    import scipy.sparse as sp
    return sp.bsr_matrix((vals, colidx, rowptr), shape=(5 * n_block_rows, 5 * n_block_rows))


def block_diag_to_scipy(vals):
    import scipy.sparse as sp
    n = vals.shape[0]
    return sp.bsr_matrix((vals, np.arange(n, dtype=np.int32), np.arange(n + 1, dtype=np.int32)),
                         shape=(5 * n, 5 * n))


def lns_inviscid(recipe: np.ndarray, value: np.ndarray) -> np.ndarray:
    """Gamma, Bc, Cc (3, 5, 5) at a base-flow state [rho, U, V, W, T]."""
    out = np.zeros(75)
    st = lib().tor_lns_inviscid(_dp(np.ascontiguousarray(recipe, dtype=np.float64)),
                                _dp(np.ascontiguousarray(value, dtype=np.float64)), _dp(out))
    if st:
        raise OracleError("inviscid", st)
    return out.reshape(3, 5, 5)
