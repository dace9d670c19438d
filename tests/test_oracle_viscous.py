"""Pins of the oracle's viscosity and Maxwell restatements
(src/mechanics/constitutive.rs:156-325, src/mechanics/element.rs:87-118,
257-295, src/mechanics/stress_update.rs:39-147, src/fem/assembly.rs:646-694,
959-1020).

Each check is fixed by the mathematics, not by retyping the oracle's
formulas: quadratic forms of linear velocity fields (uniform strain rate, so
u^T K u = V eps . D eps exactly under the 4-point rule), rigid modes, the
divergence theorem for the stress-history force of a uniform stress, the
closed-form Maxwell update of a uniform strain increment, and the
reference's own unit tests.
"""
import numpy as np
import pytest
import scipy.sparse as sp

UNIT_TET = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1],
                     [.5, 0, 0], [.5, .5, 0], [0, .5, 0], [0, 0, .5], [.5, 0, .5], [0, .5, .5]],
                    dtype=float)


def kuhn_tets(oracle, h=(1.0, 1.5, 0.75)):
    xyz, conn = oracle.generate_cube(1, 1, 1, *h)
    return [xyz[c] for c in conn]


def linear_field(nodes, A, b=(0.0, 0.0, 0.0)):
    """Nodal values (30) of u(x) = A x + b."""
    return (nodes @ np.asarray(A).T + np.asarray(b)).ravel()


def voigt_strain(A):
    """Engineering-shear Voigt strain [xx, yy, zz, xy, yz, zx] of u = A x."""
    A = np.asarray(A)
    return np.array([A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[1, 2] + A[2, 1],
                     A[2, 0] + A[0, 2]])


def rigid_modes(nodes):
    out = []
    for c in range(3):
        t = np.zeros(3)
        t[c] = 1.0
        out.append(linear_field(nodes, np.zeros((3, 3)), t))
    for W in ([[0, -1, 0], [1, 0, 0], [0, 0, 0]], [[0, 0, 0], [0, 0, -1], [0, 1, 0]],
              [[0, 0, 1], [0, 0, 0], [-1, 0, 0]]):
        out.append(linear_field(nodes, W))
    return out


def test_viscous_d_reference_pattern(oracle):
    mu, zeta = 3.0, 5.0
    D = oracle.viscous_constitutive(mu, zeta)
    # tau = 2 mu eps_dev + zeta tr(eps) I with engineering shear on 2mu (the
    # reference's choice, constitutive.rs:214-219): checked through its action
    eps = np.array([0.3, -0.2, 0.7, 0.1, -0.4, 0.25])
    tr = eps[:3].sum()
    tau = D @ eps
    assert np.allclose(tau[:3], 2 * mu * (eps[:3] - tr / 3) + zeta * tr, rtol=1e-14)
    assert np.allclose(tau[3:], 2 * mu * eps[3:], rtol=1e-14)
    with pytest.raises(Exception):
        oracle.viscous_constitutive(0.0)


@pytest.mark.parametrize("zeta", [0.0, 7.0])
def test_viscosity_element_quadratic_forms(oracle, zeta):
    mu = 1000.0
    tets = [UNIT_TET] + kuhn_tets(oracle)
    rng = np.random.default_rng(5)
    for nodes in tets:
        K = oracle.viscosity_stiffness(nodes, mu, zeta)
        V = oracle.element_volume(nodes)
        assert np.array_equal(K, K.T) or np.abs(K - K.T).max() <= 1e-12 * np.abs(K).max()
        for r in rigid_modes(nodes):
            assert np.abs(K @ r).max() <= 1e-12 * np.abs(K).max() * np.abs(r).max()
        # uniform dilatation: deviatoric part zero, u^T K u = 9 zeta V
        u = linear_field(nodes, np.eye(3))
        assert abs(u @ K @ u - 9.0 * zeta * V) <= 1e-12 * mu * V * 10
        # simple shear u = (y, 0, 0): gamma_xy = 1 -> 2 mu V
        u = linear_field(nodes, [[0, 1, 0], [0, 0, 0], [0, 0, 0]])
        assert u @ K @ u == pytest.approx(2 * mu * V, rel=1e-13)
        # isochoric extension (x, -y, 0): eps = (1, -1, 0) -> 4 mu V
        u = linear_field(nodes, np.diag([1.0, -1.0, 0.0]))
        assert u @ K @ u == pytest.approx(4 * mu * V, rel=1e-13)
        # random uniform strain rate: V eps . (2 mu eps_dev + zeta tr eps) with
        # engineering shear
        A = rng.standard_normal((3, 3))
        u = linear_field(nodes, A)
        e = voigt_strain(A)
        tr = e[:3].sum()
        energy = V * (e[:3] @ (2 * mu * (e[:3] - tr / 3) + zeta * tr) + 2 * mu * e[3:] @ e[3:])
        assert u @ K @ u == pytest.approx(energy, rel=1e-12)


def test_viscosity_reference_unit_tests(oracle):
    """src/mechanics/element.rs:413-483: symmetry, non-negative spectrum,
    linear in the viscosity."""
    K1 = oracle.viscosity_stiffness(UNIT_TET, 1000.0)
    K2 = oracle.viscosity_stiffness(UNIT_TET, 2000.0)
    assert np.allclose(K1, K1.T, atol=1e-10)
    assert np.linalg.eigvalsh(0.5 * (K1 + K1.T)).min() >= -1e-6
    assert np.array_equal(K2, 2.0 * K1)  # a power-of-two scale is exact


def test_maxwell_element(oracle):
    E, nu = 100e9, 0.25
    rng = np.random.default_rng(9)
    for nodes in [UNIT_TET] + kuhn_tets(oracle):
        V = oracle.element_volume(nodes)
        # K is the elastic stiffness, whatever the stress
        sig = rng.standard_normal(6) * 1e6
        K, f = oracle.maxwell_element(nodes, E, nu, sig)
        assert np.array_equal(K, oracle.elastic_stiffness(nodes, E, nu))
        # zero history -> zero force (element.rs:541-556)
        _, f0 = oracle.maxwell_element(nodes, E, nu, np.zeros(6))
        assert not f0.any()
        # uniform stress: the force components sum to zero (sum_i grad N_i = 0)
        assert np.abs(f.reshape(10, 3).sum(axis=0)).max() <= 1e-9 * np.abs(f).max()
        # virtual work of a linear field: u . f = -V eps(u) . sigma (Voigt,
        # engineering shear)
        for _ in range(3):
            A = rng.standard_normal((3, 3))
            u = linear_field(nodes, A)
            assert u @ f == pytest.approx(-V * voigt_strain(A) @ sig, rel=1e-11)
        # linear in sigma
        _, f2 = oracle.maxwell_element(nodes, E, nu, 2.0 * sig)
        assert np.array_equal(f2, 2.0 * f)


def maxwell_closed_form(E, nu, visc, dt, eps, sig):
    """sigma_{n+1} for a uniform Voigt strain increment eps: elastic
    predictor through the Lame form, volumetric part kept, deviatoric part
    relaxed by 1 / (1 + dt / tau), tau = visc / G."""
    G = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    trial = np.array(sig, dtype=float)
    trial[:3] += lam * eps[:3].sum() + 2 * G * eps[:3]
    trial[3:] += G * eps[3:]
    p = trial[:3].mean()
    dev = trial.copy()
    dev[:3] -= p
    out = dev / (1 + dt * G / visc)
    out[:3] += p
    return out


def test_maxwell_stress_update(oracle):
    E, nu, visc = 100e9, 0.25, 1e19
    G = E / (2 * (1 + nu))
    tau = visc / G
    rng = np.random.default_rng(4)
    for nodes in [UNIT_TET] + kuhn_tets(oracle):
        for _ in range(3):
            A = rng.standard_normal((3, 3)) * 1e-4
            sig = rng.standard_normal(6) * 1e6
            dt = tau * rng.uniform(0.1, 3.0)
            u0 = linear_field(nodes, rng.standard_normal((3, 3)) * 1e-4, rng.standard_normal(3))
            du = linear_field(nodes, A)
            got = oracle.maxwell_stress(nodes, E, nu, visc, dt, du, sig)
            want = maxwell_closed_form(E, nu, visc, dt, voigt_strain(A), sig)
            assert np.allclose(got, want, rtol=1e-9, atol=1e-9 * np.abs(want).max())
            # the rigid part of an increment does not change the stress
            r = rigid_modes(nodes)[4] * 1e-3
            assert np.allclose(oracle.maxwell_stress(nodes, E, nu, visc, dt, du + r, sig), got,
                               rtol=1e-9, atol=1e-6)
            del u0


def test_maxwell_stress_reference_tests(oracle):
    """src/mechanics/stress_update.rs:154-248 on the 1x1x1 cube."""
    xyz, conn = oracle.generate_cube(1, 1, 1, 1.0, 1.0, 1.0)
    E, nu, visc = 100e9, 0.25, 1e19
    n = 3 * len(xyz)
    zeros = np.zeros((len(conn), 6))
    out = oracle.update_stresses_maxwell(xyz, conn, E, nu, visc, 1.0, np.zeros(n), np.zeros(n),
                                         zeros)
    assert not out.any()
    out = oracle.update_stresses_maxwell(xyz, conn, E, nu, visc, 1.0, np.zeros(n),
                                         np.full(n, 0.01), zeros)
    assert out.shape == (len(conn), 6)
    sig = zeros.copy()
    sig[0, 0] = 1e6
    tau = visc / (E / (2 * (1 + nu)))
    out = oracle.update_stresses_maxwell(xyz, conn, E, nu, visc, tau, np.full(n, 0.01),
                                         np.full(n, 0.01), sig)
    assert abs(out[0, 0] - 2e6 / 3) / (2e6 / 3) < 1e-6


def test_global_viscosity_and_maxwell_assembly(oracle):
    """Global assembly = scipy's COO->CSR duplicate summation of the element
    matrices with DOF 3 node + c; per-element viscosity; the Maxwell history
    force of a uniform stress vanishes at interior nodes and sums to zero."""
    xyz, conn = oracle.generate_cube(2, 2, 2, 1.0, 1.2, 0.8)
    ne = len(conn)
    rng = np.random.default_rng(2)
    mu_e = rng.uniform(0.5, 2.0, ne)
    zeta_e = rng.uniform(0.0, 3.0, ne)
    K = oracle.assemble_viscosity(xyz, conn, 1.0, 0.0, mu_elem=mu_e, zeta_elem=zeta_e)
    rows, cols, vals = [], [], []
    for e in range(ne):
        Ke = oracle.viscosity_stiffness(xyz[conn[e]], mu_e[e], zeta_e[e])
        d = (3 * conn[e][:, None] + np.arange(3)).ravel()
        rows.append(np.repeat(d, 30))
        cols.append(np.tile(d, 30))
        vals.append(Ke.ravel())
    ref = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                        shape=K.shape).tocsr()
    ref.sort_indices()
    assert np.array_equal(K.indptr, ref.indptr) and np.array_equal(K.indices, ref.indices)
    assert np.allclose(K.data, ref.data, rtol=1e-14, atol=1e-14 * np.abs(ref.data).max())
    sig = np.tile(rng.standard_normal(6), (ne, 1))
    Km, f = oracle.assemble_maxwell(xyz, conn, 7.0, 0.3, 1e3, 1.0, sig)
    Ke_ = oracle.assemble_elastic(xyz, conn, 7.0, 0.3)
    assert np.array_equal(Km.data, Ke_.data) and np.array_equal(Km.indices, Ke_.indices)
    f = f.reshape(-1, 3)
    inner = np.nonzero((xyz[:, 0] > 1e-9) & (xyz[:, 0] < 1 - 1e-9) & (xyz[:, 1] > 1e-9)
                       & (xyz[:, 1] < 1.2 - 1e-9) & (xyz[:, 2] > 1e-9)
                       & (xyz[:, 2] < 0.8 - 1e-9))[0]
    assert len(inner) > 0
    assert np.abs(f[inner]).max() <= 1e-12 * np.abs(f).max()
    assert np.abs(f.sum(axis=0)).max() <= 1e-12 * np.abs(f).max()
