"""FD-q rules: generator vs the reference's shipped rules and invariants.

Pins: tailor_amd.fdq against fdqNodes/N100_q10.h5, N200_q10.h5, N50_q6p.h5
(copied as .npz fixtures), the NOTEBOOK_N100_Q10 vector of
tests/Python/test_fdq_nodes.py:24-27 (atol 2e-9), the Chebyshev-Gauss-Lobatto
limit q = N (test_fdq_nodes.py:62-68) and polynomial exactness of Fornberg
weights (test_fdq_nodes.py:117-134).
"""
import numpy as np
import pytest

from conftest import GOLDEN
from tailor_amd import fdq


def test_fornberg_polynomial_exactness():
    rng = np.random.default_rng(3)
    pts = np.sort(rng.uniform(-1, 1, 7))
    x0 = 0.13
    w = fdq.fornberg_weights(x0, pts, 2)
    for p in range(7):
        f = pts ** p
        exact = [x0 ** p, p * x0 ** (p - 1) if p >= 1 else 0.0,
                 p * (p - 1) * x0 ** (p - 2) if p >= 2 else 0.0]
        for m in range(3):
            assert abs(w[m] @ f - exact[m]) < 1e-11


def test_fornberg_matches_shipped_weights_n200():
    g = np.load(GOLDEN / "fdq_N200_q10.npz")
    rule = fdq.bounded_rule(g["nodes"], 10)
    assert np.array_equal(rule.stencil_indices, g["stencil_indices"])
    assert np.array_equal(rule.stencil_offsets, g["stencil_offsets"])
    for m, key in enumerate(("d0", "d1", "d2")):
        ref = g[key]
        assert np.max(np.abs(rule.weights[m] - ref) / np.maximum(1, np.abs(ref))) < 1e-12


def test_periodic_rule_matches_shipped_n50():
    g = np.load(GOLDEN / "fdq_N50_q6p.npz")
    rule = fdq.periodic_rule(50, 6, 2.0)
    assert np.allclose(rule.nodes, g["nodes"], atol=0, rtol=0)
    assert np.array_equal(rule.stencil_indices, g["stencil_indices"])
    assert np.array_equal(rule.stencil_offsets, g["stencil_offsets"])
    for m, key in enumerate(("d0", "d1", "d2")):
        assert np.max(np.abs(rule.weights[m] - g[key])) < 1e-12
    fdq.validate_rule(rule)


def test_periodic_wrap_indices():
    rule = fdq.periodic_rule(10, 4)
    assert list(rule.stencil_indices[0]) == [8, 9, 0, 1, 2]
    assert list(rule.stencil_indices[9]) == [7, 8, 9, 0, 1]


def test_stencil_starts_one_sided_rows():
    st = fdq.stencil_starts(200, 10)
    assert list(st[:7]) == [0, 0, 0, 0, 0, 0, 1]
    assert st[-1] == 189 and st[-6] == 189 and st[-7] == 188


def test_cgl_limit():
    nodes = fdq.solve_equal_error_nodes(8, 8)
    expected = -np.cos(np.arange(9) * np.pi / 8)
    assert np.max(np.abs(nodes - expected)) < 2e-11


NOTEBOOK_N100_Q10_HEAD = [-1.0, -0.9963295146396931, -0.9856865078911036, -0.9691449371148457,
                          -0.9483991819705881, -0.9256722077896464, -0.9045316749727464]


def test_generator_matches_shipped_n100_nodes():
    nodes = fdq.solve_equal_error_nodes(100, 10)
    g = np.load(GOLDEN / "fdq_N100_q10.npz")
    assert np.max(np.abs(nodes - g["nodes"])) < 2e-9
    assert np.max(np.abs(nodes[:7] - NOTEBOOK_N100_Q10_HEAD)) < 2e-9


@pytest.mark.parametrize("intervals", [99, 199])
def test_packaged_nodes_are_equal_error(intervals):
    nodes = fdq.load_bounded_nodes(intervals, 10)
    assert nodes.size == intervals + 1
    assert nodes[0] == -1.0 and nodes[-1] == 1.0
    assert np.all(np.diff(nodes) > 0)
    assert np.max(np.abs(nodes + nodes[::-1])) < 5e-11
    rule = fdq.bounded_rule(nodes, 10)
    fdq.validate_rule(rule)


def test_packaged_n199_close_to_shipped_n200_shape():
    # the N=199 and N=200 node sets are different grids but must share the
    # equal-error wall clustering: first gaps within a few percent
    a = fdq.load_bounded_nodes(199, 10)
    b = np.load(GOLDEN / "fdq_N200_q10.npz")["nodes"]
    ga, gb = np.diff(a)[:5], np.diff(b)[:5]
    assert np.all(np.abs(ga / gb - 1) < 0.05)
