"""Pin the oracle's quadrature tables and Tet10 basis.

Reference unit tests restated: src/fem/quadrature.rs:117-160,
src/fem/basis.rs:387-455, benchmarks/shape_functions/mod.rs:100-247,
tests/curved_element.rs:6-48.  Exactness of every rule on every monomial up
to its degree is checked against the closed form
int L1^a L2^b L3^c dV = a! b! c! / (a+b+c+3)!  over the reference tet.
"""
from math import factorial

import numpy as np
import pytest

RULES = {1: 1, 4: 2, 5: 3, 11: 4}


def unit_tet10():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    return tet10_from_vertices(v)


def tet10_from_vertices(v):
    edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    return np.vstack([v] + [(v[a] + v[b]) / 2 for a, b in edges])


def monomial_exact(a, b, c):
    return factorial(a) * factorial(b) * factorial(c) / factorial(a + b + c + 3)


def rule_integral(pts, w, a, b, c):
    return float(np.sum(w * pts[:, 1] ** a * pts[:, 2] ** b * pts[:, 3] ** c))


@pytest.mark.parametrize("npts", sorted(RULES))
def test_weights_sum_and_barycentric_sum(oracle, npts):
    pts, w = oracle.quad_rule(npts)
    assert abs(w.sum() - 1.0 / 6.0) < 1e-14
    assert np.all(np.abs(pts.sum(axis=1) - 1.0) < 1e-14)


@pytest.mark.parametrize("npts", sorted(RULES))
def test_rule_exact_on_monomials(oracle, npts):
    pts, w = oracle.quad_rule(npts)
    deg = RULES[npts]
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            for c in range(deg + 1 - a - b):
                got = rule_integral(pts, w, a, b, c)
                assert abs(got - monomial_exact(a, b, c)) < 2e-15, (npts, a, b, c)


def test_rule_degree_is_tight(oracle):
    # the 4-point rule is not exact for every cubic (guards against a rule
    # that silently over-integrates because of a wrong table)
    pts, w = oracle.quad_rule(4)
    assert abs(rule_integral(pts, w, 3, 0, 0) - monomial_exact(3, 0, 0)) > 1e-6


def test_quadratic_moment_1_over_60(oracle):
    pts, w = oracle.quad_rule(4)
    assert abs(rule_integral(pts, w, 2, 0, 0) - 1.0 / 60.0) < 1e-14


def test_wrong_4pt_constant_would_fail():
    # The code comment at src/fem/quadrature.rs:25 reads "(5 + sqrt5)/20";
    # the literal is (5 + 3 sqrt5)/20.  The exactness check rejects the former.
    a = (5 + 5 ** 0.5) / 20
    b = (1 - a) / 3
    pts = np.array([[a, b, b, b], [b, a, b, b], [b, b, a, b], [b, b, b, a]])
    w = np.full(4, 1 / 24)
    assert abs(rule_integral(pts, w, 2, 0, 0) - 1 / 60) > 1e-4
    a = (5 + 3 * 5 ** 0.5) / 20
    assert abs(a - 0.5854101966249685) < 1e-16


POINTS = [[0.25, 0.25, 0.25, 0.25], [1, 0, 0, 0], [0, 1, 0, 0], [0.5, 0.5, 0, 0],
          [0.1, 0.2, 0.3, 0.4]]


@pytest.mark.parametrize("L", POINTS)
def test_partition_of_unity(oracle, L):
    assert abs(oracle.shape_functions(L).sum() - 1.0) < 1e-14


def test_kronecker_delta_vertices_and_edges(oracle):
    for i in range(4):
        L = np.zeros(4)
        L[i] = 1
        N = oracle.shape_functions(L)
        assert abs(N[i] - 1) < 1e-14
        assert np.all(np.abs(np.delete(N, i)) < 1e-14)
    edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    for k, (a, b) in enumerate(edges):
        L = np.zeros(4)
        L[a] = L[b] = 0.5
        N = oracle.shape_functions(L)
        assert abs(N[4 + k] - 1) < 1e-14
        assert np.all(np.abs(np.delete(N, 4 + k)) < 1e-14)


def test_bary_derivatives_match_finite_differences(oracle):
    L = np.array([0.1, 0.2, 0.3, 0.4])
    dN = oracle.shape_derivs_bary(L)
    h = 1e-6
    for m in range(4):
        e = np.zeros(4)
        e[m] = h
        fd = (oracle.shape_functions(L + e) - oracle.shape_functions(L - e)) / (2 * h)
        assert np.allclose(dN[:, m], fd, atol=1e-8)


def test_reference_volume(oracle):
    assert abs(oracle.element_volume(unit_tet10()) - 1 / 6) < 1e-14


@pytest.mark.parametrize("scale", [1.0, 2.0, 1000.0 / 6])
def test_scaled_volume(oracle, scale):
    v = scale * np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    assert abs(oracle.element_volume(tet10_from_vertices(v)) / (scale ** 3 / 6) - 1) < 1e-14


def test_jacobian_equals_vertex_edge_matrix_for_affine_tets(oracle):
    rng = np.random.default_rng(3)
    for _ in range(5):
        v = rng.normal(size=(4, 3))
        nodes = tet10_from_vertices(v)
        Jlin = np.column_stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]])
        for L in POINTS:
            assert np.allclose(oracle.jacobian(L, nodes), Jlin, atol=1e-13)


def test_derivative_sums_vanish(oracle):
    nodes = unit_tet10()
    for L in ([0.25] * 4, [0.1, 0.2, 0.3, 0.4]):
        d = oracle.shape_derivs_cart(L, nodes)
        assert np.all(np.abs(d.sum(axis=0)) < 1e-12)


def test_linear_field_reproduction(oracle):
    # benchmarks/shape_functions/mod.rs:100-157 (barycentric coordinates of
    # a point of the reference tet are (1-x-y-z, x, y, z))
    nodes = unit_tet10()
    field = lambda p: 2 * p[0] + 3 * p[1] + 4 * p[2] + 5
    vals = np.array([field(p) for p in nodes])
    for p in ([0.25, 0.25, 0.25], [0.1, 0.2, 0.3], [0.4, 0.3, 0.2], [0.6, 0.2, 0.1]):
        L = [1 - sum(p), p[0], p[1], p[2]]
        assert abs(oracle.shape_functions(L) @ vals - field(p)) < 1e-12


def test_gradient_of_linear_field_is_exact(oracle):
    rng = np.random.default_rng(5)
    v = rng.normal(size=(4, 3)) + np.eye(4, 3) * 2
    nodes = tet10_from_vertices(v)
    g = np.array([1.5, -2.0, 0.25])
    vals = nodes @ g + 3.0
    for L in POINTS[:1] + POINTS[4:]:
        d = oracle.shape_derivs_cart(L, nodes)
        assert np.allclose(vals @ d, g, atol=1e-12)


def test_curved_element_volume(oracle):
    # tests/curved_element.rs:6-48
    nodes = unit_tet10()
    assert abs(oracle.element_volume(nodes) - 1 / 6) < 1e-10
    nodes[4, 1] -= 0.1
    assert oracle.element_volume(nodes) > 1 / 6


def test_singular_jacobian_reported(oracle):
    v = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 0, 1]], float)
    with pytest.raises(oracle.OracleError) as ei:
        oracle.shape_derivs_cart([0.25] * 4, tet10_from_vertices(v))
    assert ei.value.code == oracle.OR_ERR_SINGULAR_JACOBIAN
