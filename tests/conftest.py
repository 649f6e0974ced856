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
