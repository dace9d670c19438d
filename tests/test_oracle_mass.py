"""Pin the oracle's thermal mass matrix (src/fem/element.rs:74-105, 5-point
rule) and its global assembly.

Reference test restated: symmetry (src/fem/element.rs:200-212).  Beyond it:
the 5-point rule is exact to degree 3 and the shape functions sum to one,
so each row of M sums to rho_cp * int N_i exactly (-V/20 at vertices, V/5
at edge nodes, affine tets); the entries equal the rule's weighted outer
products of the (separately pinned) shape functions; the assembled M times
ones equals the load vector with Q = rho_cp; M is linear in rho_cp."""
import numpy as np
import pytest

from test_oracle_basis import tet10_from_vertices, unit_tet10


def test_symmetry_reference(oracle):
    M = oracle.thermal_mass(unit_tet10(), 1.0, 1.0)
    assert np.allclose(M, M.T, rtol=0, atol=1e-14)


@pytest.mark.parametrize("verts", [
    [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]],
    [[1, 2, 3], [4, 2.5, 3], [1.5, 5, 3.2], [1.2, 2.2, 7]],
])
def test_row_sums_and_total(oracle, verts):
    nodes = tet10_from_vertices(np.array(verts, float))
    rho, cp = 3300.0, 1250.0
    M = oracle.thermal_mass(nodes, rho, cp)
    V = oracle.element_volume(nodes)
    want = rho * cp * V * np.array([-1 / 20] * 4 + [1 / 5] * 6)
    assert np.allclose(M.sum(axis=1), want, rtol=1e-13, atol=1e-13 * abs(want).max())
    assert abs(M.sum() - rho * cp * V) <= 1e-13 * rho * cp * V


def test_entries_are_the_five_point_sum(oracle):
    pts, w = oracle.quad_rule(5)
    ref = sum(wq * np.outer(oracle.shape_functions(L), oracle.shape_functions(L))
              for L, wq in zip(pts, w))
    M = oracle.thermal_mass(unit_tet10(), 2.0, 0.5)
    assert np.allclose(M, ref, rtol=0, atol=1e-16)
    # and differs from the 4-point rule's (N_i N_j is quartic; neither rule is exact)
    p4, w4 = oracle.quad_rule(4)
    ref4 = sum(wq * np.outer(oracle.shape_functions(L), oracle.shape_functions(L))
               for L, wq in zip(p4, w4))
    assert np.abs(M - ref4).max() > 1e-4


def test_global_mass(oracle):
    xyz, conn = oracle.generate_cube(3, 2, 2, 1000.0, 800.0, 600.0)
    rho, cp = 3300.0, 1250.0
    M = oracle.assemble_thermal_mass(xyz, conn, rho, cp)
    f = oracle.assemble_thermal_load(xyz, conn, rho * cp)
    assert np.allclose(M @ np.ones(M.shape[0]), f, rtol=1e-12, atol=1e-12 * abs(f).max())
    d = abs(M - M.T).max()
    assert d <= 1e-14 * abs(M).max()
    rce = np.full(len(conn), 2.0 * rho * cp)
    M2 = oracle.assemble_thermal_mass(xyz, conn, 0.0, 0.0, rho_cp_elem=rce)
    assert np.array_equal(M2.indptr, M.indptr) and np.array_equal(M2.indices, M.indices)
    assert np.allclose(M2.data, 2.0 * M.data, rtol=1e-15, atol=0)
