"""Pin the oracle's BiCGSTAB restatement (src/linalg/iterative.rs:177-291).

The reference has no BiCGSTAB test, so the checks come from the algorithm
(van der Vorst, SIAM J. Sci. Stat. Comput. 13, 1992) and from an
independent implementation: a first pass worked by hand on a 2x2
nonsymmetric system, the exact one-step solve of the identity, the zero
right-hand side, the true residual against the recursive one, the solution
of random nonsymmetric systems against a direct solve, and iteration counts
against scipy's BiCGSTAB with the same Jacobi preconditioner.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla


def nonsym(n, seed, dens=0.01, skew=0.5):
    """Diagonally dominant nonsymmetric sparse matrix (a convection-like
    skew part on top of an SPD one)."""
    rng = np.random.default_rng(seed)
    k = int(dens * n * n)
    r, c = rng.integers(0, n, k), rng.integers(0, n, k)
    B = sp.coo_matrix((rng.uniform(-1, 1, k), (r, c)), shape=(n, n)).tocsr()
    S = B + B.T
    K = B - B.T
    A = S + skew * K
    A = A + sp.diags(np.asarray(abs(A).sum(axis=1)).ravel() + 1.0)
    A = sp.csr_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    return A


def test_first_pass_by_hand(oracle):
    # A = [[4,1],[2,3]], b = [1,2], Jacobi: rho = 5, p = b, p^ = (1/4, 2/3),
    # v = (5/3, 5/2), r^.v = 20/3, alpha = 3/4, s = (-1/4, 1/8),
    # s^ = (-1/16, 1/24), t = (-5/24, 0), t.t = 25/576, t.s = 5/96,
    # omega = 6/5, x = (9/80, 11/20), r = s - omega t = (0, 1/8)
    A = sp.csr_matrix(np.array([[4.0, 1.0], [2.0, 3.0]]))
    x, st = oracle.bicgstab_solve(A, [1.0, 2.0], max_iter=1, tol=1e-12)
    assert np.allclose(x, [9 / 80, 11 / 20], rtol=1e-14, atol=0)
    assert st.iterations == 1 and not st.converged
    assert abs(st.residual_norm - 0.125) < 1e-15
    # the next pass ends on s (the iteration is not counted) at the solution
    x, st = oracle.bicgstab_solve(A, [1.0, 2.0], tol=1e-12)
    assert st.converged and st.iterations == 1
    assert np.allclose(x, np.linalg.solve(A.toarray(), [1.0, 2.0]), rtol=1e-14, atol=1e-15)


def test_identity_and_zero_rhs(oracle):
    I = sp.identity(5, format="csr")
    b = np.arange(1.0, 6.0)
    x, st = oracle.bicgstab_solve(I, b)
    assert st.converged and st.iterations == 0 and np.array_equal(x, b)
    x, st = oracle.bicgstab_solve(I, np.zeros(5))
    assert st.converged and st.iterations == 0 and np.all(x == 0.0)
    assert st.residual_norm == 0.0


def test_solution_and_true_residual(oracle):
    for seed, precond in ((1, True), (2, False), (3, True)):
        A = nonsym(400, seed)
        b = np.random.default_rng(seed + 10).standard_normal(400)
        x, st = oracle.bicgstab_solve(A, b, tol=1e-11, abs_tol=0.0, use_precond=precond)
        assert st.converged
        xr = sla.spsolve(A.tocsc(), b)
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-9
        true = np.linalg.norm(b - A @ x)
        assert abs(true - st.residual_norm) <= 1e-3 * st.residual_norm + 1e-13


def test_iterations_match_scipy(oracle):
    """scipy's BiCGSTAB is the same recurrence with the preconditioner
    applied as p^ = M p, s^ = M s; with its tolerance on ||r|| the pass
    counts agree within one (scipy may stop on s one pass earlier)."""
    for seed in (4, 5, 6):
        A = nonsym(600, seed, skew=1.0)
        b = np.random.default_rng(seed).standard_normal(600)
        d = 1.0 / A.diagonal()
        M = sla.LinearOperator(A.shape, matvec=lambda v, d=d: d * v)
        count = [0]
        xs, info = sla.bicgstab(A, b, rtol=1e-10, atol=0.0, M=M, maxiter=1000,
                                callback=lambda xk: count.__setitem__(0, count[0] + 1))
        assert info == 0
        x, st = oracle.bicgstab_solve(A, b, tol=1e-10, abs_tol=0.0)
        assert st.converged
        assert abs(st.iterations - count[0]) <= 1
        assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8


def test_max_iterations_and_semantics(oracle):
    A = nonsym(300, 7, skew=1.0)
    b = np.ones(300)
    x, st = oracle.bicgstab_solve(A, b, max_iter=3, tol=1e-14, abs_tol=0.0)
    assert st.iterations == 3 and not st.converged
    assert abs(np.linalg.norm(b - A @ x) - st.residual_norm) <= 1e-8 * np.linalg.norm(b)
    _, st0 = oracle.bicgstab_solve(A, b, max_iter=0)
    assert st0.iterations == 0 and not st0.converged
