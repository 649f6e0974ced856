"""Shared fixtures.  GPU tests are marked ``gpu`` and skipped without one."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def gpu_available() -> bool:
    try:
        import torch
        return bool(torch.cuda.is_available())
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle


def small_case(ny=24, nz=12, alpha=2.43 + 0.0j, seed=0, q_y=10, q_z=6):
    """Synthetic case on Chebyshev-like nodes small enough for dense checks."""
    from tailor_amd import case as cs
    nodes = -np.cos(np.pi * np.arange(ny) / (ny - 1))
    nodes[0], nodes[-1] = -1.0, 1.0
    return cs.synthetic_case(ny, nz, q_y=q_y, q_z=q_z, seed=seed, alpha=alpha, nodes=nodes)


# Skewed affine grid and separable base flow with closed-form derivatives.
SKEW_A, SKEW_B, SKEW_C = 5.0, 6.0, 0.4


def skewed_product_case(ny=24, nz=12):
    """y = a xi + b, z = Lz (eta + 1) / 2 + c xi, every base-flow field
    f = P_f(y) H_f(eta) with P_f a cubic and H_f periodic in eta.

    Returns the case and the exact fields (N, 25) at d*5 + f with d in
    [Dy, Dz, Dyy, Dzz, Dyz] (BaseFlowDerivatives.hpp:60-65), formed from the
    physical function by the product and chain rules:
      eta(y, z) = 2 (z - c (y - b) / a) / Lz - 1, so e_y = -2c / (a Lz), e_z = 2 / Lz;
      f_y = P' H + P H' e_y,  f_z = P H' e_z,
      f_yy = P'' H + 2 P' H' e_y + P H'' e_y^2,  f_zz = P H'' e_z^2,
      f_yz = P' H' e_z + P H'' e_y e_z.
    P', P'' are exact (a cubic in xi is differentiated exactly by the
    11-point xi rule); H', H'' are the periodic rule's own 1-D derivatives of
    the samples of H (the eta rule is not exact for trigonometric H)."""
    c = small_case(ny, nz)
    a, b, sk = SKEW_A, SKEW_B, SKEW_C
    lz = c.spanwise_period
    xi = np.tile(c.xi_rule.nodes, nz)
    eta = np.repeat(c.eta_rule.nodes, ny)
    y = a * xi + b
    grid = np.zeros((ny * nz, 3), dtype=np.complex128)
    grid[:, 1] = y
    grid[:, 2] = lz * (eta + 1.0) / 2.0 + sk * xi
    # P_f(y) = p0 + p1 s + p2 s^2 + p3 s^3 with s = y / 10; H_f = 1 + h1 cos(pi(eta+1)) + h2 sin(2 pi(eta+1))
    pc = np.array([[1.0, 0.10, -0.05, 0.02], [0.3, 0.40, 0.10, -0.03], [0.05, -0.02, 0.01, 0.004],
                   [-0.04, 0.03, 0.02, -0.01], [1.2, 0.15, -0.06, 0.01]])
    hc = np.array([[0.05, 0.02], [0.10, -0.04], [0.20, 0.10], [-0.15, 0.05], [0.04, -0.03]])
    s = y / 10.0
    P = pc[:, 0] + pc[:, 1] * s[:, None] + pc[:, 2] * s[:, None] ** 2 + pc[:, 3] * s[:, None] ** 3
    dP = (pc[:, 1] + 2 * pc[:, 2] * s[:, None] + 3 * pc[:, 3] * s[:, None] ** 2) / 10.0
    d2P = (2 * pc[:, 2] + 6 * pc[:, 3] * s[:, None]) / 100.0
    ph = np.pi * (c.eta_rule.nodes + 1.0)
    H1 = 1.0 + hc[:, 0] * np.cos(ph)[:, None] + hc[:, 1] * np.sin(2 * ph)[:, None]  # (nz, 5)

    def eta_derivative(m):
        out = np.zeros_like(H1)
        for j in range(nz):
            for t in range(c.eta_rule.stencil_size):
                out[j] += c.eta_rule.weights[m, j, t] * H1[c.eta_rule.stencil_indices[j, t]]
        return out

    H, dH, d2H = (np.repeat(v, ny, axis=0) for v in (H1, eta_derivative(1), eta_derivative(2)))
    c.grid = grid
    c.baseflow = (P * H).astype(np.complex128)
    ey, ez = -2.0 * sk / (a * lz), 2.0 / lz
    fy = dP * H + P * dH * ey
    fz = P * dH * ez
    fyy = d2P * H + 2 * dP * dH * ey + P * d2H * ey ** 2
    fzz = P * d2H * ez ** 2
    fyz = dP * dH * ez + P * d2H * ey * ez
    exact = np.concatenate([fy, fz, fyy, fzz, fyz], axis=1)
    metrics = np.zeros(10)
    metrics[0], metrics[1], metrics[2], metrics[3] = 1.0 / a, 0.0, ey, ez
    return c, exact, metrics
