"""FD-q preprocessing: a structured source flow onto the FD-q grid.

Restates prepare_fdq_case (interpolate.py:501-605), the step ``Tailor.py``
runs before the solver.  The source file holds ``/grid`` (Nz, Ny, 3, 2) and
``/baseflow`` (Nz, Ny, 5, 2) on uniform logical nodes (ξ = linspace(-1, 1,
Ny), half-open periodic η = -1 + 2k/Nz) with root attributes Ny, Nz,
spanwise_periodic = 1 and spanwise_period.  Each real component is
interpolated by quintic B-splines, first along ξ (not-a-knot end
conditions) and then along η with periodic closure, after the spanwise
translation of the grid's z component (z grows by spanwise_period per
period) has been taken out; the published algorithm is SciPy's
``make_interp_spline`` (de Boor's B-spline interpolation), which the
reference calls (interpolate.py:82-157).  The order ξ then η and the
per-component affine translation are needed for the same values.

The result is checked like interpolate.py:274-324 (finite, ρ > 0, T > 0,
corners preserved within 2e-11, FD-q grid Jacobian > 0) and written as a
schema-v2 cache.  This is host-side input preparation, run once per case.
"""
from __future__ import annotations

import hashlib
import os
from pathlib import Path

import numpy as np
from scipy.interpolate import make_interp_spline

from . import fdq, h5

INTERPOLATION_METHOD = "bounded_y_periodic_z_quintic_bspline"
INTERPOLATION_VERSION = 3


class PrepareError(ValueError):
    pass


def sha256_file(path: Path) -> str:
    h = hashlib.sha256()
    with Path(path).open("rb") as f:
        for block in iter(lambda: f.read(1 << 20), b""):
            h.update(block)
    return h.hexdigest()


def rule_digest(rule: fdq.FDQRule) -> str:
    """SHA-256 over the rule's arrays (nodes, indices, offsets, d0-d2)."""
    h = hashlib.sha256()
    for a in (rule.nodes, rule.stencil_indices, rule.stencil_offsets, rule.weights):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def read_source(path: Path):
    """The structured source case (interpolate.py:170-217 checks)."""
    root = h5.read(path)
    if "grid" not in root or "baseflow" not in root:
        raise PrepareError("sample HDF5 must contain /grid and /baseflow")
    if "Ny" not in root.attrs or "Nz" not in root.attrs:
        raise PrepareError("sample HDF5 must define root attributes Ny and Nz")
    if int(root.attrs.get("spanwise_periodic", 0)) != 1:
        raise PrepareError("sample HDF5 must declare spanwise_periodic=1")
    if "spanwise_period" not in root.attrs:
        raise PrepareError("sample HDF5 must define spanwise_period")
    ny, nz = int(root.attrs["Ny"]), int(root.attrs["Nz"])
    period = float(root.attrs["spanwise_period"])
    grid = np.asarray(root["grid"].data, dtype=np.float64)
    base = np.asarray(root["baseflow"].data, dtype=np.float64)
    if ny < 6 or nz < 6:
        raise PrepareError("sample Ny and Nz must both be at least 6")
    if not (np.isfinite(period) and period > 0):
        raise PrepareError("sample spanwise_period must be finite and positive")
    if grid.shape != (nz, ny, 3, 2):
        raise PrepareError(f"/grid must have shape {(nz, ny, 3, 2)}; got {grid.shape}")
    if base.shape != (nz, ny, 5, 2):
        raise PrepareError(f"/baseflow must have shape {(nz, ny, 5, 2)}; got {base.shape}")
    if not (np.all(np.isfinite(grid)) and np.all(np.isfinite(base))):
        raise PrepareError("sample grid and baseflow must contain only finite values")
    grid = grid[..., 0] + 1j * grid[..., 1]
    base = base[..., 0] + 1j * base[..., 1]
    if np.any(base[..., 0].real <= 0):
        raise PrepareError("sample density must be strictly positive")
    if np.any(base[..., 4].real <= 0):
        raise PrepareError("sample temperature must be strictly positive")
    return grid, base, ny, nz, period


def spline_fields(values: np.ndarray, xi_new: np.ndarray, eta_new: np.ndarray,
                  translation: np.ndarray) -> np.ndarray:
    """Quintic tensor B-spline of ``values`` (nz, ny, C) complex onto the
    nodes (eta_new x xi_new); component c is periodic up to translation[c]
    per period of η (period 2 in logical coordinates)."""
    nz, ny, ncomp = values.shape
    xi_old = np.linspace(-1.0, 1.0, ny)
    eta_old = -1.0 + 2.0 * np.arange(nz, dtype=np.float64) / nz
    shift = np.asarray(translation, dtype=np.complex128)
    src_phase = (eta_old - eta_old[0]) / 2.0
    dst_phase = (np.asarray(eta_new) - eta_old[0]) / 2.0
    periodic = values - src_phase[:, None, None] * shift[None, None, :]
    out = np.empty((len(eta_new), len(xi_new), ncomp), dtype=np.complex128)
    eta_closed = np.append(eta_old, eta_old[0] + 2.0)
    for c in range(ncomp):
        parts = []
        for part in (periodic[..., c].real, periodic[..., c].imag):
            along_xi = make_interp_spline(xi_old, np.ascontiguousarray(part), k=5, axis=1)(xi_new)
            closed = np.concatenate([along_xi, along_xi[:1]], axis=0)
            parts.append(make_interp_spline(eta_closed, closed, k=5, bc_type="periodic",
                                            axis=0)(eta_new))
        out[..., c] = parts[0] + 1j * parts[1]
    return out + dst_phase[:, None, None] * shift[None, None, :]


def _first_derivative(values: np.ndarray, rule: fdq.FDQRule, axis: int,
                      translation: float = 0.0) -> np.ndarray:
    w = rule.weights[1]
    if axis == 1:
        return np.einsum("zis,is->zi", values[:, rule.stencil_indices], w)
    taps = values[rule.stencil_indices, :]                      # (nz, s, ny)
    if translation:
        rows = np.arange(rule.node_count)[:, None]
        wraps = (rows + rule.stencil_offsets - rule.stencil_indices) // rule.node_count
        taps = taps + wraps[:, :, None] * translation
    return np.einsum("jsi,js->ji", taps, w)


def check_fields(src_grid, src_base, grid, base, xi, eta, period) -> float:
    """interpolate.py:274-324: physical checks; returns the minimum Jacobian."""
    if not (np.all(np.isfinite(grid)) and np.all(np.isfinite(base))):
        raise PrepareError("interpolated grid and baseflow must be finite")
    if np.any(base[..., 0].real <= 0):
        raise PrepareError("interpolated density is not strictly positive")
    if np.any(base[..., 4].real <= 0):
        raise PrepareError("interpolated temperature is not strictly positive")
    for src, dst, name in ((src_grid, grid, "grid"), (src_base, base, "baseflow")):
        a = np.array([src[0, 0], src[0, -1], src[-1, 0], src[-1, -1]])
        b = np.array([dst[0, 0], dst[0, -1], dst[-1, 0], dst[-1, -1]])
        err = float(np.max(np.abs(b - a)))
        if err > 2e-11 * max(1.0, float(np.max(np.abs(a)))):
            raise PrepareError(f"{name} corner preservation failed: error={err:.3e}")
    y, z = grid[..., 1].real, grid[..., 2].real
    jac = (_first_derivative(y, xi, 1) * _first_derivative(z, eta, 0, period)
           - _first_derivative(y, eta, 0) * _first_derivative(z, xi, 1))
    if not np.all(np.isfinite(jac)):
        raise PrepareError("interpolated grid Jacobian contains non-finite values")
    jmin = float(np.min(jac))
    if jmin <= 0:
        raise PrepareError(f"interpolated grid Jacobian must be strictly positive; "
                           f"minimum={jmin:.6e}")
    return jmin


def expected_attributes(source: Path, ny: int, nz: int, q_y: int, q_z: int, period: float,
                        xi: fdq.FDQRule, eta: fdq.FDQRule) -> dict:
    return {"schema_version": 2, "interpolation_method": INTERPOLATION_METHOD,
            "interpolation_version": INTERPOLATION_VERSION, "source_file": source.name,
            "source_sha256": sha256_file(source), "y_rule_sha256": rule_digest(xi),
            "z_rule_sha256": rule_digest(eta), "ordering": "k_j_dof", "Ny": ny, "Nz": nz,
            "q_y": q_y, "q_z": q_z, "spanwise_periodic": 1, "spanwise_period": period}


def prepare_case(source: Path, output: Path, q_y: int, q_z: int, log=print) -> Path:
    """Build the schema-v2 cache ``output`` from ``source`` (interpolate.py:501-605)."""
    source, output = Path(source), Path(output)
    if not source.is_file():
        raise FileNotFoundError(f"structured sample file does not exist: {source}")
    grid, base, ny, nz, period = read_source(source)
    if not 2 <= q_y <= ny - 1:
        raise PrepareError(f"q_y must satisfy 2 <= q_y <= {ny - 1}")
    if not 2 <= q_z <= nz - 1:
        raise PrepareError(f"q_z must satisfy 2 <= q_z <= {nz - 1}")
    if q_z % 2:
        raise PrepareError("periodic q_z must be even")
    xi = fdq.bounded_rule(fdq.load_bounded_nodes(ny - 1, q_y), q_y)
    eta = fdq.periodic_rule(nz, q_z, 2.0)
    g = spline_fields(grid, xi.nodes, eta.nodes, np.array([0, 0, period], dtype=np.complex128))
    b = spline_fields(base, xi.nodes, eta.nodes, np.zeros(5, dtype=np.complex128))
    jmin = check_fields(grid, base, g, b, xi, eta, period)

    from .case import Case
    from .cli import write_cache
    case = Case(ny, nz, xi, eta, g.reshape(-1, 3), b.reshape(-1, 5), period)
    attrs = expected_attributes(source, ny, nz, q_y, q_z, period, xi, eta)
    tmp = output.with_name(f".{output.stem}.tmp-{os.getpid()}{output.suffix}")
    try:
        write_cache(case, tmp, extra_attrs=attrs)
        os.replace(tmp, output)
    finally:
        if tmp.exists():
            tmp.unlink()
    log(f"FD-q cache generated: {output}\n"
        f"  Ny={ny}, q_y={q_y}; Nz={nz}, q_z={q_z}\n"
        f"  minimum Jacobian={jmin:.16g}\n"
        f"  rho range=[{b[..., 0].real.min():.16g}, {b[..., 0].real.max():.16g}]\n"
        f"  T range=[{b[..., 4].real.min():.16g}, {b[..., 4].real.max():.16g}]")
    return output
