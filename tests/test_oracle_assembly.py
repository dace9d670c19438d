"""Pin the oracle's mesh generator, global assembly and Dirichlet elimination.

Reference tests restated: src/mesh_generator_improved.rs:177-227,
src/fem/assembly.rs:1029-1090 (structural), src/fem/dof.rs:188-245.
Independent checks: scipy COO->CSR duplicate summation of the element
matrices, K 1 = 0 for thermal, total loads, closed-form nodal solutions.
"""
import numpy as np
import pytest
import scipy.sparse as sp


def test_generator_counts(oracle):
    xyz, conn = oracle.generate_cube(1, 1, 1, 1.0, 1.0, 1.0)
    assert conn.shape == (6, 10)
    xyz, conn = oracle.generate_cube(2, 2, 2, 1.0, 1.0, 1.0)
    assert conn.shape == (48, 10)
    xyz, conn = oracle.generate_cube(6, 6, 6, 1000.0, 1000.0, 600.0)
    assert xyz.shape == (2197, 3) and conn.shape == (1296, 10)


def test_generator_no_duplicate_nodes_and_conforming(oracle):
    xyz, conn = oracle.generate_cube(3, 2, 2, 1.0, 1.0, 1.0)
    keys = np.round(xyz * 1000).astype(np.int64)
    assert len({tuple(k) for k in keys}) == len(xyz)
    # every node is used, midpoints are edge averages, orientation positive
    assert set(np.unique(conn)) == set(range(len(xyz)))
    edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    for e in conn:
        v = xyz[e[:4]]
        for k, (a, b) in enumerate(edges):
            assert np.allclose(xyz[e[4 + k]], (v[a] + v[b]) / 2)
        assert np.linalg.det(np.column_stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]])) > 0
    # interior faces shared by exactly two tets, boundary faces by one
    faces = {}
    for i, e in enumerate(conn):
        for f in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
            faces.setdefault(tuple(sorted(e[list(f)])), []).append(i)
    assert all(len(v) in (1, 2) for v in faces.values())


def test_generator_elements_span_two_z_levels(oracle):
    xyz, conn = oracle.generate_cube(1, 1, 1, 1.0, 1.0, 1.0)
    for e in conn:
        assert len(np.unique(np.round(xyz[e[:4], 2], 10))) >= 2


def test_generator_vertex_numbering(oracle):
    nx, ny, nz = 3, 2, 2
    xyz, _ = oracle.generate_cube(nx, ny, nz, 3.0, 2.0, 2.0)
    n = 0
    for iz in range(nz + 1):
        for iy in range(ny + 1):
            for ix in range(nx + 1):
                assert np.all(xyz[n] == [ix, iy, iz])
                n += 1


def elem_triplets(oracle, xyz, conn, kind, k=1.0, E=1.0, nu=0.25):
    I, J, V = [], [], []
    for e in conn:
        nodes = xyz[e]
        if kind == "thermal":
            K = oracle.thermal_stiffness(nodes, k)
            dofs = e
        else:
            K = oracle.elastic_stiffness(nodes, E, nu)
            dofs = (3 * e[:, None] + np.arange(3)).ravel()
        I.append(np.repeat(dofs, len(dofs)))
        J.append(np.tile(dofs, len(dofs)))
        V.append(K.ravel())
    return np.concatenate(I), np.concatenate(J), np.concatenate(V)


@pytest.mark.parametrize("kind", ["thermal", "elastic"])
def test_assembly_matches_scipy_coo(oracle, kind):
    xyz, conn = oracle.generate_cube(2, 2, 1, 2.0, 2.0, 1.5)
    n = len(xyz) * (1 if kind == "thermal" else 3)
    I, J, V = elem_triplets(oracle, xyz, conn, kind)
    ref = sp.coo_matrix((V, (I, J)), shape=(n, n)).tocsr()
    ref.sort_indices()
    K = (oracle.assemble_thermal(xyz, conn, 1.0) if kind == "thermal"
         else oracle.assemble_elastic(xyz, conn, 1.0, 0.25))
    assert np.array_equal(K.indptr, ref.indptr)
    assert np.array_equal(K.indices, ref.indices)
    assert np.max(np.abs(K.data - ref.data)) <= 1e-14 * np.max(np.abs(ref.data))


def test_dense_brute_force_1x1x1(oracle):
    xyz, conn = oracle.generate_cube(1, 1, 1, 1.0, 1.0, 1.0)
    n = len(xyz)
    Kd = np.zeros((n, n))
    for e in conn:
        Kd[np.ix_(e, e)] += oracle.thermal_stiffness(xyz[e], 2.0)
    K = oracle.assemble_thermal(xyz, conn, 2.0).toarray()
    assert np.max(np.abs(K - Kd)) < 1e-14 * np.max(np.abs(Kd))


def test_assembly_thread_count_is_bitwise_irrelevant(oracle):
    xyz, conn = oracle.generate_cube(3, 3, 2, 3.0, 3.0, 2.0)
    a = oracle.assemble_thermal(xyz, conn, 3.0, nthreads=1)
    b = oracle.assemble_thermal(xyz, conn, 3.0, nthreads=8)
    assert np.array_equal(a.data, b.data) and np.array_equal(a.indices, b.indices)


def test_config1_structure(oracle):
    xyz, conn = oracle.generate_cube(6, 6, 6, 1000.0, 1000.0, 600.0)
    K = oracle.assemble_thermal(xyz, conn, 3.0)
    assert K.nnz == 54793
    assert np.max(np.abs(K @ np.ones(K.shape[0]))) < 1e-12 * np.max(np.abs(K.data))
    assert abs(K - K.T).max() == 0.0
    mask, vals = oracle.thermal_dirichlet(xyz, 600.0)
    assert mask.sum() == 338
    f = oracle.assemble_thermal_load(xyz, conn, 1e-6)
    assert abs(f.sum() - 1e-6 * 1000 * 1000 * 600) < 1e-9 * 600
    Kbc, fbc = oracle.apply_dirichlet(K, f, mask, vals)
    assert Kbc.nnz == 44513


def test_gravity_total(oracle):
    xyz, conn = oracle.generate_cube(2, 2, 2, 1e6, 1e6, 6e5)
    f = oracle.assemble_gravity_load(xyz, conn, 3300.0, [0, 0, -9.81])
    V = 1e6 * 1e6 * 6e5
    assert abs(f[2::3].sum() / (-3300.0 * 9.81 * V) - 1) < 1e-12
    assert np.all(f[0::3] == 0) and np.all(f[1::3] == 0)


def test_dirichlet_structure(oracle):
    xyz, conn = oracle.generate_cube(2, 2, 2, 1000.0, 1000.0, 600.0)
    K = oracle.assemble_thermal(xyz, conn, 3.0)
    f = oracle.assemble_thermal_load(xyz, conn, 1e-6)
    mask, vals = oracle.thermal_dirichlet(xyz, 600.0)
    Kbc, fbc = oracle.apply_dirichlet(K, f, mask, vals)
    D = mask.astype(bool)
    F = ~D
    c = max(1.0, np.max(np.abs(K.diagonal())))
    Kd, Kbd = K.toarray(), Kbc.toarray()
    assert np.array_equal(Kbd[np.ix_(F, F)], Kd[np.ix_(F, F)])
    assert np.all(Kbd[np.ix_(F, D)] == 0) and np.all(Kbd[np.ix_(D, F)] == 0)
    assert np.array_equal(Kbd[np.ix_(D, D)], c * np.eye(D.sum()))
    assert np.allclose(fbc[D], c * vals[D], rtol=0, atol=0)
    lift = f[F] - Kd[np.ix_(F, D)] @ vals[D]
    assert np.allclose(fbc[F], lift, rtol=1e-13, atol=1e-9)
    assert abs(Kbc - Kbc.T).max() == 0.0


def test_dirichlet_scaling_floor_is_one(oracle):
    # c = max(1, max |K_ii|): with tiny entries the constrained diagonal is 1
    A = sp.csr_matrix(np.array([[1e-3, 1e-4], [1e-4, 2e-3]]))
    Kbc, fbc = oracle.apply_dirichlet(A, np.array([1.0, 2.0]), np.array([1, 0], np.uint8),
                                      np.array([5.0, 0.0]))
    assert Kbc.toarray()[0, 0] == 1.0 and fbc[0] == 5.0
    assert fbc[1] == 2.0 - 1e-4 * 5.0


@pytest.mark.parametrize("dims", [(1, 1, 1), (3, 2, 2), (4, 3, 5)])
def test_generator_numbering_first_touch(oracle, dims):
    """The numbering callers index results by (src/mesh_generator_improved.rs:
    29-56,139-146): vertex (ix, iy, iz) is node ix + (nx+1)(iy + (ny+1) iz) at
    (ix dx, iy dy, iz dz); edge midpoints are numbered in the order they are
    first requested -- elements in order, edges e01, e12, e20, e03, e13, e23 --
    so the first occurrences of midpoint ids, scanned that way, count up from
    the number of vertices without gaps; a repeated edge reuses its id."""
    nx, ny, nz = dims
    lx, ly, lz = 3.0, 2.0, 1.5
    xyz, conn = oracle.generate_cube(nx, ny, nz, lx, ly, lz)
    nv = (nx + 1) * (ny + 1) * (nz + 1)
    ix, iy, iz = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1),
                             indexing="ij")
    ids = (ix + (nx + 1) * (iy + (ny + 1) * iz)).ravel()
    expect = np.stack([ix.ravel() * lx / nx, iy.ravel() * ly / ny, iz.ravel() * lz / nz], 1)
    assert np.allclose(xyz[ids], expect, rtol=0, atol=1e-12)
    assert conn[:, :4].max() < nv and conn[:, 4:].min() >= nv
    seen, nxt = {}, nv
    for e in conn:
        for k in range(6):
            a, b = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)][k]
            key = tuple(sorted((e[a], e[b])))
            mid = e[4 + k]
            if key in seen:
                assert seen[key] == mid
            else:
                assert mid == nxt
                seen[key] = mid
                nxt += 1
    assert nxt == len(xyz)
