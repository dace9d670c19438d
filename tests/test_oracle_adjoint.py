"""Pin the oracle's adjoint element backward with central finite differences.

L(theta) = 1/2 || d - P u(theta) ||^2, theta_e = log k_e (thermal) or
log E_e (elastic); dL/dtheta_e = -lambda_e^T K_e u_e with K_bc lambda = dL/du
(zero on Dirichlet DOFs).  The FD side re-assembles and re-solves the full
problem, so it checks the adjoint derivation, the Dirichlet-lift term and the
element kernel together.
"""
import numpy as np


def solve_tight(oracle, K, f):
    x, st = oracle.cg_solve(K, f, max_iter=100000, tol=1e-15, abs_tol=0.0)
    return x


def thermal_loss(oracle, xyz, conn, logk, mask, vals, obs_idx, d):
    K = oracle.assemble_thermal(xyz, conn, 0.0, k_elem=np.exp(logk))
    f = oracle.assemble_thermal_load(xyz, conn, 1e-6)
    Kbc, fbc = oracle.apply_dirichlet(K, f, mask, vals)
    T = solve_tight(oracle, Kbc, fbc)
    r = T[obs_idx] - d
    return 0.5 * r @ r, T, Kbc


def test_thermal_adjoint_vs_fd(oracle):
    rng = np.random.default_rng(7)
    xyz, conn = oracle.generate_cube(3, 3, 2, 1000.0, 1000.0, 600.0)
    mask, vals = oracle.thermal_dirichlet(xyz, 600.0)
    logk = np.log(3.0) + 0.3 * rng.standard_normal(len(conn))
    free = np.nonzero(mask == 0)[0]
    obs = rng.choice(free, 12, replace=False)
    d = rng.uniform(300, 900, size=len(obs))
    L0, T, Kbc = thermal_loss(oracle, xyz, conn, logk, mask, vals, obs, d)
    g = np.zeros(len(xyz))
    g[obs] = T[obs] - d
    g[mask == 1] = 0
    lam = solve_tight(oracle, Kbc, g)
    grad = oracle.thermal_element_backward(xyz, conn, np.exp(logk), T, lam)
    h = 1e-6
    for e in rng.choice(len(conn), 8, replace=False):
        lp, lm = logk.copy(), logk.copy()
        lp[e] += h
        lm[e] -= h
        fd = (thermal_loss(oracle, xyz, conn, lp, mask, vals, obs, d)[0] -
              thermal_loss(oracle, xyz, conn, lm, mask, vals, obs, d)[0]) / (2 * h)
        assert abs(fd - grad[e]) <= 1e-5 * max(abs(fd), 1e-3 * np.max(np.abs(grad))), (e, fd, grad[e])


def test_elastic_adjoint_vs_fd(oracle):
    rng = np.random.default_rng(9)
    lx = ly = 1e6
    lz = 6e5
    xyz, conn = oracle.generate_cube(2, 2, 2, lx, ly, lz)
    mask, vals = oracle.elastic_dirichlet(xyz, lx, ly)
    logE = np.log(1e11) + 0.5 * rng.standard_normal(len(conn))
    top = np.nonzero(np.abs(xyz[:, 2] - lz) < 1.0)[0]
    obs = 3 * top + 2
    d = rng.normal(-4e4, 5e3, size=len(obs))

    def loss(le):
        K = oracle.assemble_elastic(xyz, conn, 0.0, 0.25, E_elem=np.exp(le))
        f = oracle.assemble_gravity_load(xyz, conn, 3300.0, [0, 0, -9.81])
        Kbc, fbc = oracle.apply_dirichlet(K, f, mask, vals)
        u = solve_tight(oracle, Kbc, fbc)
        r = u[obs] - d
        return 0.5 * r @ r, u, Kbc

    L0, u, Kbc = loss(logE)
    g = np.zeros(3 * len(xyz))
    g[obs] = u[obs] - d
    g[mask == 1] = 0
    lam = solve_tight(oracle, Kbc, g)
    grad = oracle.elastic_element_backward(xyz, conn, np.exp(logE), 0.25, u, lam)
    h = 1e-6
    for e in rng.choice(len(conn), 8, replace=False):
        lp, lm = logE.copy(), logE.copy()
        lp[e] += h
        lm[e] -= h
        fd = (loss(lp)[0] - loss(lm)[0]) / (2 * h)
        assert abs(fd - grad[e]) <= 1e-5 * max(abs(fd), 1e-3 * np.max(np.abs(grad))), (e, fd, grad[e])
