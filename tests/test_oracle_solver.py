"""Pin the oracle's Jacobi-PCG and the end-to-end path.

Reference KATs: src/linalg/iterative.rs:491-504, src/linalg/preconditioner.rs:
213-241, src/linalg/solver.rs:185-211.  End to end: the exact solutions of
the benchmark problems are quadratic in z, hence in the P2 space, and the
4-point rule integrates their loads exactly, so the Galerkin solution equals
them at the nodes (benchmarks/thermal_1d_improved_mesh.rs:9-13,
benchmarks/elasticity_halfspace_gravity.rs:23-27).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def test_cg_2x2_kat(oracle):
    A = sp.csr_matrix(np.array([[2.0, 1.0], [1.0, 2.0]]))
    x, st = oracle.cg_solve(A, [3.0, 3.0])
    assert np.allclose(x, [1, 1], atol=1e-6) and st.converged


def test_jacobi_kat(oracle):
    A = sp.csr_matrix(np.diag([2.0, 4.0, 8.0]))
    d = oracle.jacobi_diag_inv(A)
    assert np.allclose(d * np.array([2.0, 4.0, 8.0]), 1.0, atol=1e-14)
    # missing / tiny diagonal -> 1
    B = sp.csr_matrix(np.array([[0.0, 1.0], [1.0, 1e-15]]))
    B.eliminate_zeros()
    assert np.array_equal(oracle.jacobi_diag_inv(B), [1.0, 1.0])


def test_spmv_and_norm(oracle):
    A = sp.csr_matrix(np.array([[2.0, 1.0], [1.0, 2.0]]))
    assert np.allclose(oracle.spmv(A, [1.0, 1.0]), [3.0, 3.0], atol=0)
    assert abs(np.sqrt(3.0 ** 2 + 4.0 ** 2) - 5.0) < 1e-14


def test_cg_zero_rhs(oracle):
    A = sp.csr_matrix(np.eye(3))
    x, st = oracle.cg_solve(A, np.zeros(3))
    assert st.converged and st.iterations == 0 and np.all(x == 0)
    assert st.residual_norm == 0.0 and st.relative_residual == 0.0


def test_cg_iteration_semantics(oracle):
    # the converging pass is not counted; max_iter caps completed passes
    A = sp.csr_matrix(np.diag([1.0, 2.0, 3.0, 4.0]))
    x, st = oracle.cg_solve(A, [1.0, 1.0, 1.0, 1.0], use_precond=True)
    assert st.converged and st.iterations == 0  # Jacobi makes it one step
    x, st = oracle.cg_solve(A, [1.0, 1.0, 1.0, 1.0], use_precond=False, max_iter=1)
    assert not st.converged and st.iterations == 1
    x, st = oracle.cg_solve(A, [1.0, 1.0, 1.0, 1.0], use_precond=False, max_iter=0)
    assert not st.converged and st.iterations == 0 and st.relative_residual == 1.0


def thermal_system(oracle, n, k=3.0, Q=1e-6, lz=600.0):
    xyz, conn = oracle.generate_cube(n, n, n, 1000.0, 1000.0, lz)
    K = oracle.assemble_thermal(xyz, conn, k)
    f = oracle.assemble_thermal_load(xyz, conn, Q)
    mask, vals = oracle.thermal_dirichlet(xyz, lz)
    Kbc, fbc = oracle.apply_dirichlet(K, f, mask, vals)
    return xyz, conn, Kbc, fbc, mask


def test_config1_matches_spsolve_and_closed_form(oracle):
    xyz, conn, Kbc, fbc, mask = thermal_system(oracle, 6)
    x, st = oracle.cg_solve(Kbc, fbc, max_iter=5000, tol=1e-13, abs_tol=0.0)
    assert st.converged
    xs = sla.spsolve(Kbc.tocsc(), fbc)
    ex = oracle.thermal_analytic(xyz[:, 2], 600.0, 1280.0, 0.0, 3.0, 1e-6)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-12
    assert np.linalg.norm(x - ex) / np.linalg.norm(ex) < 1e-12
    assert np.linalg.norm(xs - ex) / np.linalg.norm(ex) < 1e-13


def test_config1_reference_tolerance(oracle):
    # the benchmark's own solver settings (tol 1e-8, max 1000)
    xyz, conn, Kbc, fbc, mask = thermal_system(oracle, 6)
    x, st = oracle.cg_solve(Kbc, fbc, max_iter=1000, tol=1e-8)
    assert st.converged and st.relative_residual < 1e-8
    r = fbc - Kbc @ x
    assert np.linalg.norm(r) / np.linalg.norm(fbc) < 2e-8
    ex = oracle.thermal_analytic(xyz[:, 2], 600.0, 1280.0, 0.0, 3.0, 1e-6)
    err = np.sqrt(np.mean((x - ex)[mask == 0] ** 2)) / 1280.0 * 100
    assert err < 1e-4  # README.md:107-111 reports 0.0000 %


def test_elastic_confined_column_closed_form(oracle):
    lx = ly = 1e6
    lz = 6e5
    xyz, conn = oracle.generate_cube(2, 2, 3, lx, ly, lz)
    K = oracle.assemble_elastic(xyz, conn, 1e11, 0.25)
    f = oracle.assemble_gravity_load(xyz, conn, 3300.0, [0, 0, -9.81])
    mask, vals = oracle.elastic_dirichlet(xyz, lx, ly)
    Kbc, fbc = oracle.apply_dirichlet(K, f, mask, vals)
    x, st = oracle.cg_solve(Kbc, fbc, max_iter=20000, tol=1e-14, abs_tol=0.0)
    assert st.converged
    uz = oracle.elastic_analytic_uz(xyz[:, 2], lz, 3300.0, 9.81, 1e11, 0.25)
    assert np.linalg.norm(x[2::3] - uz) / np.linalg.norm(uz) < 1e-11
    assert np.max(np.abs(x[0::3])) < 1e-9 * np.max(np.abs(uz))
    assert np.max(np.abs(x[1::3])) < 1e-9 * np.max(np.abs(uz))


def test_elastic_patch_test_linear_field(oracle):
    # benchmarks/elasticity_patch_test.rs: a linear displacement imposed on the
    # whole boundary is reproduced exactly inside (no body force).
    xyz, conn = oracle.generate_cube(2, 2, 2, 1.0, 1.0, 1.0)
    n = len(xyz)
    G = np.array([[1e-3, 2e-4, 0.0], [-1e-4, 5e-4, 3e-4], [2e-4, 0.0, -4e-4]])
    u_exact = xyz @ G.T
    on_b = np.any((np.abs(xyz) < 1e-12) | (np.abs(xyz - 1.0) < 1e-12), axis=1)
    mask = np.zeros(3 * n, np.uint8)
    vals = np.zeros(3 * n)
    for c in range(3):
        mask[3 * np.nonzero(on_b)[0] + c] = 1
        vals[3 * np.arange(n) + c] = u_exact[:, c]
    K = oracle.assemble_elastic(xyz, conn, 1e11, 0.25)
    Kbc, fbc = oracle.apply_dirichlet(K, np.zeros(3 * n), mask, vals)
    x, st = oracle.cg_solve(Kbc, fbc, max_iter=10000, tol=1e-14, abs_tol=0.0)
    assert np.max(np.abs(x.reshape(n, 3) - u_exact)) < 1e-12


def test_parallel_cg_timing_column_matches_serial(oracle):
    """The OpenMP CG (a timing column of the CPU baseline, not a restatement)
    runs the serial loop's arithmetic up to the order of its reductions: the
    same iteration count within one and the same solution to rounding."""
    xyz, conn = oracle.generate_cube(10, 10, 6, 1000.0, 1000.0, 600.0)
    mask, vals = oracle.thermal_dirichlet(xyz, 600.0)
    K = oracle.assemble_thermal(xyz, conn, 3.0)
    f = oracle.assemble_thermal_load(xyz, conn, 1e-6)
    Kb, fb = oracle.apply_dirichlet(K, f, mask, vals)
    x1, s1 = oracle.cg_solve(Kb, fb, max_iter=5000, tol=1e-10)
    for threads in (1, 4):
        x2, s2 = oracle.cg_solve_omp(Kb, fb, max_iter=5000, tol=1e-10, nthreads=threads)
        assert s2.converged and abs(s2.iterations - s1.iterations) <= 1
        assert np.linalg.norm(x2 - x1) <= 1e-10 * np.linalg.norm(x1)
