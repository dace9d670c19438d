"""Pin the oracle's GMRES restatement (src/linalg/iterative.rs:313-486).

The reference has no GMRES test, so the checks come from the method
(Saad & Schultz, SIAM J. Sci. Stat. Comput. 7, 1986): one right-
preconditioned step is the minimal-residual step along M r (closed form),
GMRES(m) with m >= n ends at the exact solution within n steps, the Givens
residual estimate |g_k| equals the true residual ||b - A x||, and restarted
solves of random nonsymmetric systems reach the direct solution.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

from test_oracle_bicgstab import nonsym


def test_one_step_is_minimal_residual(oracle):
    A = sp.csr_matrix(np.array([[4.0, 1.0], [2.0, 3.0]]))
    b = np.array([1.0, 2.0])
    for precond in (True, False):
        d = 1.0 / A.diagonal() if precond else np.ones(2)
        z = d * b
        Az = A @ z
        alpha = (b @ Az) / (Az @ Az)
        x, st = oracle.gmres_solve(A, b, max_iter=1, restart=5, tol=1e-14,
                                   use_precond=precond)
        assert st.iterations == 1 and not st.converged
        assert np.allclose(x, alpha * z, rtol=1e-14, atol=0)
        assert abs(st.residual_norm - np.linalg.norm(b - alpha * Az)) < 1e-15


def test_exact_within_n_steps(oracle):
    rng = np.random.default_rng(3)
    D = rng.standard_normal((6, 6)) + 6 * np.eye(6)
    A = sp.csr_matrix(D)
    b = rng.standard_normal(6)
    x, st = oracle.gmres_solve(A, b, restart=10, tol=1e-13, abs_tol=0.0)
    assert st.converged and st.iterations <= 6
    assert np.allclose(x, np.linalg.solve(D, b), rtol=1e-12, atol=1e-14)


def test_zero_rhs_and_max_iter_zero(oracle):
    A = sp.identity(4, format="csr") * 2.0
    x, st = oracle.gmres_solve(A, np.zeros(4))
    assert st.converged and st.iterations == 0 and np.all(x == 0.0)
    x, st = oracle.gmres_solve(A, np.ones(4), max_iter=0)
    assert not st.converged and st.iterations == 0 and st.residual_norm == 2.0


def test_restarted_solves_and_residual_estimate(oracle):
    for seed, m, precond in ((1, 10, True), (2, 30, True), (3, 10, False)):
        A = nonsym(400, seed, skew=1.0)
        b = np.random.default_rng(seed).standard_normal(400)
        x, st = oracle.gmres_solve(A, b, max_iter=5000, restart=m, tol=1e-11, abs_tol=0.0,
                                   use_precond=precond)
        assert st.converged
        xr = sla.spsolve(A.tocsc(), b)
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-9
        true = np.linalg.norm(b - A @ x)
        assert abs(true - st.residual_norm) <= 1e-3 * true + 1e-14
        if m == 10:
            # cycles restart from the true residual: a cap of whole cycles
            # stops exactly there, unconverged, with the counted inner steps
            _, s2 = oracle.gmres_solve(A, b, max_iter=2 * m, restart=m, tol=1e-14,
                                       abs_tol=0.0, use_precond=precond)
            assert s2.iterations == 2 * m and not s2.converged


def test_matches_scipy_solution(oracle):
    A = nonsym(500, 8, skew=1.0)
    b = np.random.default_rng(8).standard_normal(500)
    xs, info = sla.gmres(A, b, rtol=1e-11, atol=0.0, restart=20, maxiter=500)
    assert info == 0
    x, st = oracle.gmres_solve(A, b, max_iter=5000, restart=20, tol=1e-11, abs_tol=0.0,
                               use_precond=False)
    assert st.converged
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
