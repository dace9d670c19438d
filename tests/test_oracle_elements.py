"""Pin the oracle's element kernels.

Golden vectors: exact rational element matrices/loads from
tests/golden/make_element_golden.py (sympy) for the reference tet and the six
tets of the 6-tet hex subdivision at two spacings.  Invariants restated from
the reference's own tests: src/fem/element.rs:166-225,
src/mechanics/element.rs:321-389, src/mechanics/body_force.rs:119-211,
src/mechanics/constitutive.rs:403-469, src/mechanics/strain.rs:137-236.
"""
import numpy as np
import pytest

from test_oracle_basis import tet10_from_vertices, unit_tet10


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_golden_thermal_stiffness(oracle, golden):
    for name, c in golden.items():
        K = oracle.thermal_stiffness(np.array(c["nodes"]), 1.0)
        assert rel(K, np.array(c["thermal_K_k1"])) < 2e-15, name


def test_golden_thermal_load(oracle, golden):
    for name, c in golden.items():
        f = oracle.thermal_load(np.array(c["nodes"]), 1.0)
        assert rel(f, np.array(c["int_N"])) < 2e-15, name


def test_golden_elastic_stiffness(oracle, golden):
    for name, c in golden.items():
        K = oracle.elastic_stiffness(np.array(c["nodes"]), 1.0, 0.25)
        assert rel(K, np.array(c["elastic_K_E1_nu025"])) < 2e-15, name


def test_golden_gravity_load(oracle, golden):
    g = np.array([0.3, -1.2, -9.81])
    for name, c in golden.items():
        f = oracle.gravity_load(np.array(c["nodes"]), 3300.0, g)
        expect = 3300.0 * np.outer(np.array(c["int_N"]), g).ravel()
        assert rel(f, expect) < 4e-15, name


def test_golden_volume(oracle, golden):
    for name, c in golden.items():
        assert abs(oracle.element_volume(np.array(c["nodes"])) / c["volume"] - 1) < 1e-14


def test_planted_transpose_mistake_is_caught(golden):
    # A plausible bug -- using J instead of J^-T for the gradients -- must
    # fail the golden comparison.  Restated here with numpy on a Kuhn tet.
    c = golden["kuhn_thermal_6x6x6_2"]
    nodes = np.array(c["nodes"])
    v = nodes[:4]
    J = np.column_stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]])
    # reference-coordinate gradients of the P2 basis at the centroid
    from oracle import oracle as O
    K_wrong = np.zeros((10, 10))
    pts, w = O.quad_rule(4)
    for L, wq in zip(pts, w):
        dN = O.shape_derivs_bary(L)
        dref = dN[:, 1:] - dN[:, [0]]
        G = dref @ J  # wrong: should be dref @ inv(J)
        K_wrong += wq * abs(np.linalg.det(J)) * G @ G.T
    assert rel(K_wrong, np.array(c["thermal_K_k1"])) > 1e-3


@pytest.mark.parametrize("npts", [5, 11])
def test_thermal_rule_invariance_on_affine_tets(oracle, golden, npts):
    # integrand degree 2 on straight tets: every rule of degree >= 2 agrees
    for name, c in golden.items():
        nodes = np.array(c["nodes"])
        assert rel(oracle.thermal_stiffness(nodes, 1.0, npts=npts),
                   oracle.thermal_stiffness(nodes, 1.0)) < 1e-13, name


def test_thermal_stiffness_invariants(oracle):
    K = oracle.thermal_stiffness(unit_tet10(), 1.0)
    assert np.all(np.abs(K - K.T) < 1e-14)
    assert np.all(np.diag(K) > 0)
    assert np.all(np.abs(K.sum(axis=1)) < 1e-12)
    assert np.allclose(oracle.thermal_stiffness(unit_tet10(), 3.0), 3.0 * K, rtol=1e-15)


def test_thermal_load_sum(oracle):
    nodes = unit_tet10()
    f = oracle.thermal_load(nodes, 6.0)
    assert abs(f.sum() - 6.0 * oracle.element_volume(nodes)) < 1e-12


def test_elastic_stiffness_invariants(oracle):
    K = oracle.elastic_stiffness(unit_tet10(), 100e9, 0.25)
    assert np.all(np.abs(K - K.T) <= 1e-12 * np.abs(K) + 1e-6)
    ev = np.sort(np.linalg.eigvalsh(0.5 * (K + K.T)))
    assert np.all(np.abs(ev[:6]) < 1e3)
    assert np.all(ev[6:] > 1e6)
    K2 = oracle.elastic_stiffness(unit_tet10(), 200e9, 0.25)
    assert np.all(np.abs(K2 - 2 * K) < 1e-6)


def test_elastic_rigid_modes_in_nullspace(oracle):
    rng = np.random.default_rng(11)
    v = rng.normal(size=(4, 3)) + 2 * np.eye(4, 3)
    nodes = tet10_from_vertices(v)
    K = oracle.elastic_stiffness(nodes, 1.0, 0.3)
    modes = []
    for c in range(3):
        t = np.zeros((10, 3))
        t[:, c] = 1
        modes.append(t.ravel())
    for axis in range(3):
        w = np.zeros(3)
        w[axis] = 1
        modes.append(np.cross(w, nodes).ravel())
    for m in modes:
        assert np.max(np.abs(K @ m)) < 1e-12 * np.max(np.abs(K)) * np.max(np.abs(m))


def test_gravity_load_properties(oracle):
    nodes = unit_tet10()
    V = oracle.element_volume(nodes)
    f = oracle.gravity_load(nodes, 3000.0, [0, 0, -9.81])
    assert abs(f[2::3].sum() - (-3000.0 * V * 9.81)) < 1e-6
    assert np.all(np.abs(f[2::3]) > 1e-10)
    fx = oracle.gravity_load(nodes, 3000.0, [9.81, 0, 0])
    assert abs(fx[0::3].sum() - 3000.0 * V * 9.81) < 1e-6
    assert abs(fx[1::3].sum()) < 1e-10 and abs(fx[2::3].sum()) < 1e-10
    assert np.all(oracle.gravity_load(nodes, 3000.0, [0, 0, 0]) == 0)


def test_constitutive(oracle):
    D = oracle.constitutive(100e9, 0.25)
    assert np.all(np.abs(D - D.T) < 1e-10)
    D0 = oracle.constitutive(100e9, 0.0)
    assert abs(D0[0, 1]) < 1e-6 and abs(D0[0, 2]) < 1e-6 and abs(D0[1, 2]) < 1e-6
    E, nu = 100e9, 0.25
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    assert abs(D[0, 1] - lam) < 1e-3 and abs(D[3, 3] - mu) < 1e-3
    assert abs(D[0, 0] - (lam + 2 * mu)) < 1e-3
    for bad in [(-100e9, 0.25), (100e9, 0.6), (100e9, -1.0), (0.0, 0.2)]:
        with pytest.raises(oracle.OracleError) as ei:
            oracle.constitutive(*bad)
        assert ei.value.code == oracle.OR_ERR_INVALID_MATERIAL


def test_b_matrix_structure(oracle):
    d = np.zeros((10, 3))
    d[0] = [1, 2, 3]
    B = oracle.b_matrix(d)
    assert B[0, 0] == 1 and B[1, 1] == 2 and B[2, 2] == 3
    assert B[3, 0] == 2 and B[3, 1] == 1
    assert B[4, 1] == 3 and B[4, 2] == 2
    assert B[5, 0] == 3 and B[5, 2] == 1
    assert np.all(B[:, 3:] == 0)


def test_b_matrix_rigid_translation(oracle):
    B = oracle.b_matrix(oracle.shape_derivs_cart([0.25] * 4, unit_tet10()))
    u = np.zeros(30)
    u[0::3] = 1
    assert np.all(np.abs(B @ u) < 1e-10)
