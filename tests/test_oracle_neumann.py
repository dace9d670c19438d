"""Pin the oracle's Neumann face restatement (src/fem/boundary.rs:60-253,
src/fem/assembly.rs:248-309).  The reference has no test of it; the checks
are the face-node table against the Tet10 edge numbering, the closed form of
a constant flux on a flat quadratic triangle (vertex shares 0, edge shares
A/3 each), the area, and global balance on a box surface."""
import numpy as np
import pytest

EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]  # Tet10 nodes 4..9


def test_face_nodes_follow_the_edge_numbering(oracle):
    for lf in range(4):
        fn = oracle.face_nodes(lf)
        v = list(fn[:3])
        assert sorted(v) == sorted(set(range(4)) - {lf})
        for k, (a, b) in enumerate([(0, 1), (1, 2), (2, 0)]):
            e = fn[3 + k]
            assert set(EDGES[e - 4]) == {v[a], v[b]}
    with pytest.raises(Exception):
        oracle.face_nodes(4)


@pytest.mark.parametrize("v", [
    [[0, 0, 0], [2, 0, 0], [0, 1, 0]],
    [[1, 2, 3], [4, 2.5, 3.5], [1.5, 5, 2.0]],
])
def test_flat_face_closed_form(oracle, v):
    v = np.array(v, float)
    c = np.vstack([v, (v[0] + v[1]) / 2, (v[1] + v[2]) / 2, (v[2] + v[0]) / 2])
    A = 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
    assert abs(oracle.face_area(c) - A) <= 1e-14 * A
    f = oracle.face_flux(c, 7.0)
    assert np.allclose(f[:3], 0.0, atol=1e-14 * 7.0 * A)
    assert np.allclose(f[3:], 7.0 * A / 3, rtol=1e-14)


def test_curved_face_sums_to_flux_times_area(oracle):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    c = np.vstack([v, [0.5, 0.0, 0.1], [0.5, 0.5, 0.15], [0.0, 0.5, 0.05]])
    a = oracle.face_area(c)
    assert a > 0.5
    assert abs(oracle.face_flux(c, 2.0).sum() - 2.0 * a) <= 1e-14 * a


def test_box_surface_balance(oracle):
    xyz, conn = oracle.generate_cube(3, 2, 2, 1000.0, 800.0, 600.0)
    faces = [(e, lf) for e in range(len(conn)) for lf in range(4)
             if np.all(np.abs(xyz[conn[e][oracle.face_nodes(lf)], 2] - 600.0) < 1.0)]
    assert len(faces) == 3 * 2 * 2
    el = [e for e, _ in faces]
    lf = [l for _, l in faces]
    f = oracle.apply_neumann(xyz, conn, el, lf, np.full(len(faces), 0.065), np.zeros(len(xyz)))
    assert abs(f.sum() - 0.065 * 1000.0 * 800.0) <= 1e-12 * 0.065 * 1000.0 * 800.0
    top = np.abs(xyz[:, 2] - 600.0) < 1.0
    assert np.all(f[~top] == 0.0)
    with pytest.raises(Exception):
        oracle.apply_neumann(xyz, conn, [0], [5], [1.0], np.zeros(len(xyz)))
