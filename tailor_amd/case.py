"""Synthetic BiGlobal cases and the in-memory case descriptor.

A case holds what the reference's FD-q cache holds (Prepare.cpp:443-668,
interpolate.py:399-436): the two 1-D rules, ``/grid`` (N,3) and ``/baseflow``
(N,5) complex128 arrays in natural ``k_j_dof`` order (node index
``j*Ny + i``, i = xi index, j = eta index), the spanwise period, plus the
YAML physics / eigen-solver settings (Parser.cpp:41-108).

The synthetic generator builds the workloads the benchmark is quoted on:
an erf boundary layer over a tanh-clustered, wall-bumped curvilinear grid.
Grid and base flow are evaluated analytically at the FD-q nodes, so no
interpolation step is involved.  A SplitMix64 stream (53-bit uniforms) makes
the same seed give the same case on every machine.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import fdq

MASK64 = (1 << 64) - 1


class SplitMix64:
    """Portable 64-bit generator (Steele, Lea, Flood 2014)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * 2.0 ** -53)


@dataclass
class Physics:
    """Physics block of the YAML (Parser.cpp:84-108, Recipe.hpp:17-61)."""

    reynolds: float = 1000.0
    mach: float = 0.5
    prandtl: float = 0.72
    gamma: float = 1.4
    gas_constant: float = 287.05          # required and validated, unused
    reference_mu: float = 1.716e-5        # required and validated, unused
    reference_temperature: float = 288.15
    sutherland_constant: float = 110.4

    def as_array(self) -> np.ndarray:
        """[Re, Ma, Pr, gamma, Tref, S] as passed across the C ABI."""
        return np.array([self.reynolds, self.mach, self.prandtl, self.gamma,
                         self.reference_temperature, self.sutherland_constant],
                        dtype=np.float64)


@dataclass
class Case:
    ny: int
    nz: int
    xi_rule: fdq.FDQRule
    eta_rule: fdq.FDQRule
    grid: np.ndarray          # (N, 3) complex128, natural order
    baseflow: np.ndarray      # (N, 5) complex128, natural order
    spanwise_period: float
    physics: Physics = field(default_factory=Physics)
    alpha: complex = 2.43 + 0.0j
    omega_shift: complex = 0.0 + 0.0j
    nev: int = 8
    tol: float = 1e-10
    maxit: int = 1000
    title: str = "synthetic"

    @property
    def nodes(self) -> int:
        return self.ny * self.nz

    @property
    def unknowns(self) -> int:
        return 5 * self.nodes

    @property
    def lambda_shift(self) -> complex:
        """Target lambda_s = Im(omega_s) - i Re(omega_s) (EigenSolver.cpp:46-48)."""
        return complex(self.omega_shift.imag, -self.omega_shift.real)

    def with_alpha(self, alpha: complex) -> "Case":
        return replace(self, alpha=complex(alpha))


def synthetic_case(ny: int = 100, nz: int = 50, q_y: int = 10, q_z: int = 6,
                   seed: int = 0, alpha: complex = 2.43 + 0j,
                   omega_shift: complex = 0j, physics: Physics | None = None,
                   wall_temperature: float = 1.5, height: float = 40.0,
                   stretch: float = 2.5, bump: float = 0.05,
                   span: float = 2.0 * math.pi, nodes: np.ndarray | None = None) -> Case:
    """Boundary-layer case on an Ny x Nz FD-q grid.

    Mapping: yhat = (xi+1)/2, y = H [1 + tanh(s (yhat-1)) / tanh s], then the
    wall bump y <- y + b (1 - y/H) cos z; z = Lz (eta+1)/2.
    Base flow: U = erf(y / delta(z)), delta = 1 + 0.2 cos z
    + sum_{k=2..4} a_k cos(k z + phi_k), a_k ~ U(0, 0.03), phi_k ~ U(0, 2 pi)
    drawn in the order a2, phi2, a3, phi3, a4, phi4; V = W = 0;
    Crocco-Busemann T = Tw + (1 + r - Tw) U - r U^2 with r = (gamma-1) Ma^2 / 2;
    rho = 1/T.
    """
    phys = physics or Physics()
    if nodes is None:
        xi_nodes = fdq.load_bounded_nodes(ny - 1, q_y)
    else:
        xi_nodes = np.asarray(nodes, dtype=np.float64)
    xi_rule = fdq.bounded_rule(xi_nodes, q_y)
    eta_rule = fdq.periodic_rule(nz, q_z, 2.0)

    rng = SplitMix64(seed)
    amps = []
    for _k in range(2, 5):
        a = rng.uniform(0.0, 0.03)
        ph = rng.uniform(0.0, 2.0 * math.pi)
        amps.append((a, ph))

    n = ny * nz
    grid = np.zeros((n, 3), dtype=np.complex128)
    base = np.zeros((n, 5), dtype=np.complex128)
    r = 0.5 * (phys.gamma - 1.0) * phys.mach ** 2
    th = math.tanh(stretch)
    for j in range(nz):
        z = span * (eta_rule.nodes[j] + 1.0) / 2.0
        delta = 1.0 + 0.2 * math.cos(z)
        for k, (a, ph) in zip(range(2, 5), amps):
            delta += a * math.cos(k * z + ph)
        for i in range(ny):
            yh = 0.5 * (xi_rule.nodes[i] + 1.0)
            y = height * (1.0 + math.tanh(stretch * (yh - 1.0)) / th)
            y = y + bump * (1.0 - y / height) * math.cos(z)
            u = math.erf(y / delta)
            t = wall_temperature + (1.0 + r - wall_temperature) * u - r * u * u
            idx = j * ny + i
            grid[idx] = (0.0, y, z)
            base[idx] = (1.0 / t, u, 0.0, 0.0, t)
    return Case(ny, nz, xi_rule, eta_rule, grid, base, span, phys,
                complex(alpha), complex(omega_shift), title=f"synthetic {ny}x{nz} seed {seed}")


def uniform_case(ny: int = 24, nz: int = 8, q_y: int = 4, q_z: int = 2,
                 physics: Physics | None = None) -> Case:
    """Uniform base flow [rho,U,V,W,T] = [1,1,0,0,1] on a Cartesian grid.

    The far-field characteristic split of this flow has exactly one incoming
    and three neutral modes per node (test_boundary_uniform.py:32-61).
    """
    nodes = np.cos(np.pi * np.arange(ny)[::-1] / (ny - 1))
    nodes[0], nodes[-1] = -1.0, 1.0
    xi_rule = fdq.bounded_rule(nodes, q_y)
    eta_rule = fdq.periodic_rule(nz, q_z, 2.0)
    n = ny * nz
    grid = np.zeros((n, 3), dtype=np.complex128)
    base = np.zeros((n, 5), dtype=np.complex128)
    for j in range(nz):
        for i in range(ny):
            idx = j * ny + i
            grid[idx] = (0.0, 10.0 * (xi_rule.nodes[i] + 1.0), math.pi * (eta_rule.nodes[j] + 1.0))
            base[idx] = (1.0, 1.0, 0.0, 0.0, 1.0)
    return Case(ny, nz, xi_rule, eta_rule, grid, base, 2.0 * math.pi,
                physics or Physics(), 1.0 + 0j, 0j, title="uniform")
