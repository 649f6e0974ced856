"""FD-q one-dimensional differentiation rules (host side).

A rule gives, for every node of a 1-D grid, the q+1 stencil columns, their
unwrapped offsets and the d0/d1/d2 weights, stored row-major over
``(node, slot)``.  This mirrors the rule data that the reference loads from
its FD-q cache (ProblemData.hpp:26-96, Prepare.cpp:588-668) and generates in
src/Python/fdq_nodes.py:

* bounded stencils start at ``clip(i - q//2, 0, n - q - 1)``
  (fdq_nodes.py:65-69), so the first and last q/2 rows are one-sided;
* periodic rules are uniform, half-open, centred, with offsets -q/2..q/2 and
  indices wrapped modulo n (fdq_nodes.py:307-353);
* weights come from Fornberg's recursion (B. Fornberg, "Generation of finite
  difference formulas on arbitrarily spaced grids", Math. Comp. 51 (1988)),
  the published method fdq_nodes.py:226-277 also follows;
* bounded nodes are the equal-error FD-q nodes: the node set on [-1, 1] for
  which the local interpolation error amplitude of every stencil is the same
  (fdq_nodes.py:152-194).  ``solve_equal_error_nodes`` solves that system with
  its own continuation/Newton scheme.

Everything here is host preprocessing; no array from this module is computed
on the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np

DATA_DIR = Path(__file__).resolve().parent / "data"


@dataclass(frozen=True)
class FDQRule:
    """One-dimensional rule.  ``weights[m, i, s]`` is the order-m weight."""

    nodes: np.ndarray            # (n,) float64, computational coordinate
    q: int
    stencil_indices: np.ndarray  # (n, q+1) int64
    stencil_offsets: np.ndarray  # (n, q+1) int64
    weights: np.ndarray          # (3, n, q+1) float64
    topology: str                # "bounded" or "periodic"
    period: float = 2.0

    @property
    def node_count(self) -> int:
        return int(self.nodes.size)

    @property
    def stencil_size(self) -> int:
        return self.q + 1


def stencil_starts(node_count: int, q: int) -> np.ndarray:
    """Centred, boundary-clamped starts of the q+1 point stencils."""
    rows = np.arange(node_count, dtype=np.int64)
    return np.clip(rows - q // 2, 0, node_count - q - 1)


def fornberg_weights(x0: float, points: np.ndarray, max_order: int = 2) -> np.ndarray:
    """Weights of derivative orders 0..max_order at ``x0`` (Fornberg 1988).

    Returns an array of shape (max_order + 1, len(points)).
    """
    x = np.asarray(points, dtype=np.float64)
    n = x.size
    if n == 0 or max_order < 0 or max_order >= n:
        raise ValueError("need 0 <= max_order < number of points")
    if np.unique(x).size != n:
        raise ValueError("stencil points must be distinct")
    c = np.zeros((max_order + 1, n), dtype=np.float64)
    c[0, 0] = 1.0
    prod_prev = 1.0
    dist_prev = x[0] - x0
    for i in range(1, n):
        top = min(i, max_order)
        prod = 1.0
        dist_last = dist_prev
        dist_prev = x[i] - x0
        for j in range(i):
            gap = x[i] - x[j]
            prod *= gap
            if j == i - 1:
                for k in range(top, 0, -1):
                    c[k, i] = prod_prev * (k * c[k - 1, i - 1] - dist_last * c[k, i - 1]) / prod
                c[0, i] = -prod_prev * dist_last * c[0, i - 1] / prod
            for k in range(top, 0, -1):
                c[k, j] = (dist_prev * c[k, j] - k * c[k - 1, j]) / gap
            c[0, j] = dist_prev * c[0, j] / gap
        prod_prev = prod
    return c


def bounded_rule(nodes: np.ndarray, q: int) -> FDQRule:
    """Bounded rule on the given strictly increasing nodes."""
    nodes = np.asarray(nodes, dtype=np.float64)
    n = nodes.size
    if not (n >= 3 and 2 <= q < n):
        raise ValueError("bounded rule needs node_count > q >= 2")
    starts = stencil_starts(n, q)
    idx = starts[:, None] + np.arange(q + 1, dtype=np.int64)
    off = idx - np.arange(n, dtype=np.int64)[:, None]
    w = np.empty((3, n, q + 1), dtype=np.float64)
    for row in range(n):
        w[:, row, :] = fornberg_weights(nodes[row], nodes[idx[row]], 2)
    return FDQRule(nodes, q, idx, off, w, "bounded", 2.0)


def periodic_rule(node_count: int, q: int, period: float = 2.0) -> FDQRule:
    """Uniform centred rule on the half-open periodic interval [-p/2, p/2)."""
    if node_count < 3 or q < 2 or q >= node_count or q % 2:
        raise ValueError("periodic rule needs even q with 2 <= q < node_count")
    h = float(period) / node_count
    nodes = -0.5 * float(period) + h * np.arange(node_count, dtype=np.float64)
    offs = np.arange(-(q // 2), q // 2 + 1, dtype=np.int64)
    off = np.broadcast_to(offs, (node_count, q + 1)).copy()
    idx = (np.arange(node_count, dtype=np.int64)[:, None] + off) % node_count
    w = np.empty((3, node_count, q + 1), dtype=np.float64)
    for row in range(node_count):
        w[:, row, :] = fornberg_weights(nodes[row], nodes[row] + h * offs, 2)
    return FDQRule(nodes, q, idx, off, w, "periodic", float(period))


# --------------------------------------------------------------------------
# Equal-error node solver
# --------------------------------------------------------------------------

def _interval_stencils(n_aux: int, order: int) -> np.ndarray:
    """Stencil (index rows) of ``order`` auxiliary nodes for each interval."""
    r = np.arange(n_aux - 1, dtype=np.int64)
    st = np.clip(r - (order - 1) // 2, 0, n_aux - order)
    return st[:, None] + np.arange(order, dtype=np.int64)


def _extrema(aux: np.ndarray, stencils: np.ndarray) -> np.ndarray:
    """Root of sum 1/(x - s) inside each (aux[r], aux[r+1]); safeguarded Newton."""
    s = aux[stencils]                      # (m, order)
    lo = aux[:-1].copy()
    hi = aux[1:].copy()
    x = 0.5 * (lo + hi)
    for _ in range(200):
        d = x[:, None] - s
        f = np.sum(1.0 / d, axis=1)
        fp = -np.sum(1.0 / (d * d), axis=1)
        # f decreases monotonically from +inf to -inf across the interval
        pos = f > 0.0
        lo = np.where(pos, x, lo)
        hi = np.where(pos, hi, x)
        step = f / fp
        xn = x - step
        bad = ~((xn > lo) & (xn < hi))
        xn = np.where(bad, 0.5 * (lo + hi), xn)
        done = np.abs(xn - x) <= 4e-16 * np.maximum(1.0, np.abs(x))
        x = xn
        if np.all(done):
            break
    return x


def _equal_error_residual(aux: np.ndarray, order: int):
    """Configuration nodes and adjacent log-error differences."""
    st = _interval_stencils(aux.size, order)
    ext = _extrema(aux, st)
    s = aux[st]
    e_mid = np.sum(np.log(np.abs(ext[:, None] - s)), axis=1)
    e_first = np.sum(np.log(np.abs(-1.0 - s[0])))
    e_last = np.sum(np.log(np.abs(1.0 - s[-1])))
    log_err = np.concatenate(([e_first], e_mid, [e_last]))
    nodes = np.concatenate(([-1.0], ext, [1.0]))
    return nodes, log_err, np.diff(log_err)


def _to_params(aux: np.ndarray) -> np.ndarray:
    gaps = np.diff(np.concatenate(([-1.0], aux, [1.0])))
    return np.log(gaps[:-1] / gaps[-1])


def _from_params(p: np.ndarray) -> np.ndarray:
    lg = np.concatenate((p, [0.0]))
    lg = lg - lg.max()
    g = np.exp(lg)
    g /= g.sum()
    return -1.0 + 2.0 * np.cumsum(g[:-1])


def solve_equal_error_nodes(intervals: int, q: int, verbose: bool = False) -> np.ndarray:
    """Equal-error FD-q nodes: ``intervals + 1`` nodes on [-1, 1].

    Continuation over the local degree 2..q; at each degree a damped Newton
    iteration (central-difference Jacobian in log-gap coordinates) drives the
    adjacent log-error differences to zero.
    """
    N = int(intervals)
    if N < 2 or q < 2 or q > N:
        raise ValueError("need 2 <= q <= intervals")
    aux = -1.0 + (2.0 * np.arange(N, dtype=np.float64) + 1.0) / N
    p = _to_params(aux)
    for order in range(2, q + 1):
        def res(pp):
            return _equal_error_residual(_from_params(pp), order)[2]
        r = res(p)
        for it in range(60):
            nr = np.linalg.norm(r, np.inf)
            if nr < 1e-13:
                break
            h = 1e-7
            J = np.empty((r.size, p.size))
            for k in range(p.size):
                e = np.zeros_like(p)
                e[k] = h
                J[:, k] = (res(p + e) - res(p - e)) / (2 * h)
            dp = np.linalg.lstsq(J, -r, rcond=None)[0]
            t = 1.0
            while True:
                pn = p + t * dp
                rn = res(pn)
                if np.linalg.norm(rn, np.inf) < nr or t < 1e-4:
                    break
                t *= 0.5
            p, r = pn, rn
            if verbose:
                print(f"order {order} it {it} |r|={np.linalg.norm(r, np.inf):.3e}")
        if np.linalg.norm(r, np.inf) > 1e-9:
            raise RuntimeError(f"equal-error solve failed at order {order}")
    nodes, log_err, _ = _equal_error_residual(_from_params(p), q)
    nodes = 0.5 * (nodes - nodes[::-1])
    nodes[0] = -1.0
    nodes[-1] = 1.0
    return nodes


def node_file(intervals: int, q: int) -> Path:
    return DATA_DIR / f"fdq_nodes_N{intervals}_q{q}.txt"


def load_bounded_nodes(intervals: int, q: int) -> np.ndarray:
    """Equal-error nodes from the packaged cache (generated by this module)."""
    path = node_file(intervals, q)
    if path.exists():
        return np.loadtxt(path, dtype=np.float64)
    nodes = solve_equal_error_nodes(intervals, q)
    return nodes


def write_bounded_nodes(intervals: int, q: int) -> Path:
    nodes = solve_equal_error_nodes(intervals, q)
    path = node_file(intervals, q)
    np.savetxt(path, nodes, fmt="%.17e")
    return path


def validate_rule(rule: FDQRule) -> None:
    """Moment validation of Prepare.cpp:328-437 (tolerance 5e-9)."""
    n, ss = rule.node_count, rule.stencil_size
    if not (n >= 3 and rule.q >= 2 and rule.q < n):
        raise ValueError("FD-q metadata must satisfy node_count > q >= 2")
    if np.any(np.diff(rule.nodes) <= 0):
        raise ValueError("FD-q nodes must be strictly increasing")
    for row in range(n):
        cols = rule.stencil_indices[row]
        offs = rule.stencil_offsets[row]
        if np.any(cols < 0) or np.any(cols >= n):
            raise ValueError(f"FD-q row {row} contains out-of-range index")
        if rule.topology == "periodic":
            if np.any((row + offs) % n != cols):
                raise ValueError(f"periodic row {row} has inconsistent stencil index/offset")
            xs = rule.nodes[row] + offs * rule.period / n
        else:
            if np.any(row + offs != cols):
                raise ValueError(f"bounded row {row} has inconsistent stencil index/offset")
            xs = rule.nodes[cols]
        if row not in cols:
            raise ValueError(f"FD-q stencil row {row} does not contain its evaluation node")
        if not np.all(np.isfinite(rule.weights[:, row, :])):
            raise ValueError(f"FD-q weight is not finite at row {row}")
        x0 = rule.nodes[row]
        # d^m/dx^m of x^pw at x0, for m, pw in 0..2
        exact_table = ((1.0, x0, x0 * x0), (0.0, 1.0, 2.0 * x0), (0.0, 0.0, 2.0))
        for m in range(3):
            for pw in range(3):
                approx = np.sum(rule.weights[m, row, :ss] * xs ** pw)
                exact = exact_table[m][pw]
                if abs(approx - exact) > 5e-9 * max(1.0, abs(exact)):
                    raise ValueError(
                        f"FD-q order-{m} weights fail degree-{pw} polynomial check at row {row}")
