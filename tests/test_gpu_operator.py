"""GPU tests of ConjugateGradient.solve_with_operator (src/linalg/solver.rs:
112-120, src/linalg/iterative.rs:133-140): the reference's CG loop (:46-119)
on a caller's operator and preconditioner, called back on device tensors.

Pinned against the oracle's CG (the same loop) on the thermal system and on
a matrix-free tridiagonal operator, the reference's 2x2 test, a direct
solve, and the loop's edge semantics (zero right-hand side, no iterations,
an exact preconditioner's uncounted converging pass, failing callbacks).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

import geo_simulator_amd as G
from geo_simulator_amd import workloads as W

pytestmark = pytest.mark.gpu


def laplacian_1d(n):
    """y_i = 2 v_i - v_{i-1} - v_{i+1} (zero outside), on device tensors."""
    def apply(v):
        y = 2.0 * v
        y[1:] -= v[:-1]
        y[:-1] -= v[1:]
        return y
    A = sp.diags([-np.ones(n - 1), 2.0 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()
    return G.LinearOperator(apply, n), A


def test_operator_cg_on_the_assembled_system_matches_oracle(oracle):
    mesh, dof, _ = W.thermal_box(G, 6, 6, 6)
    K, f = G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    Ks = K.to_scipy()
    fh = f.cpu().numpy()
    cg = G.ConjugateGradient().with_max_iterations(1000).with_tolerance(1e-8)
    x, st = cg.solve_with_operator(K, f, G.JacobiPreconditioner(K))
    xo, so = oracle.cg_solve(Ks, fh, 1000, 1e-8)
    xs, ss = cg.solve(K, f)
    assert st.converged and so.converged
    assert abs(st.iterations - so.iterations) <= 2
    assert abs(st.iterations - ss.iterations) <= 2
    x = x.cpu().numpy()
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
    assert np.linalg.norm(fh - Ks @ x) / np.linalg.norm(fh) < 1e-8
    assert st.spmv_count == st.iterations + 1  # the converging pass is not counted


def test_operator_cg_reference_2x2():
    Am = torch.tensor([[2.0, 1.0], [1.0, 2.0]], dtype=torch.float64, device="cuda")
    op = G.LinearOperator(lambda v: Am @ v, 2)
    x, st = G.ConjugateGradient().solve_with_operator(op, np.array([3.0, 3.0]),
                                                      G.IdentityPreconditioner())
    assert isinstance(x, np.ndarray)
    assert np.allclose(x, [1.0, 1.0], atol=1e-6) and st.converged


def test_operator_cg_matrix_free_tridiagonal(oracle):
    n = 400
    op, A = laplacian_1d(n)
    b = np.random.default_rng(5).standard_normal(n)
    cg = (G.ConjugateGradient().with_max_iterations(5000).with_tolerance(1e-12)
          .with_abs_tolerance(0.0))
    x, st = cg.solve_with_operator(op, b)
    xo, so = oracle.cg_solve(A, b, 5000, 1e-12, 0.0, use_precond=False)
    xr = spla.spsolve(A.tocsc(), b)
    assert st.converged and so.converged
    assert abs(st.iterations - so.iterations) <= 2
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-9
    # a Python preconditioner (the exact diagonal of A: Jacobi by hand)
    prec = G.LinearOperator(lambda r: r / 2.0, n)
    x2, st2 = cg.solve_with_operator(op, b, prec)
    xo2, so2 = oracle.cg_solve(A, b, 5000, 1e-12, 0.0, use_precond=True)
    assert st2.converged and abs(st2.iterations - so2.iterations) <= 2
    assert np.linalg.norm(x2 - xr) / np.linalg.norm(xr) < 1e-9


def test_operator_cg_edge_semantics():
    n = 64
    d = torch.linspace(1.0, 4.0, n, dtype=torch.float64, device="cuda")
    op = G.LinearOperator(lambda v: d * v, n)
    exact = G.LinearOperator(lambda r: r / d, n)
    b = torch.ones(n, dtype=torch.float64, device="cuda")
    cg = G.ConjugateGradient()
    # an exact preconditioner: solved on the first pass, which is not counted
    x, st = cg.solve_with_operator(op, b, exact)
    assert st.converged and st.iterations == 0
    assert torch.allclose(x, 1.0 / d, rtol=1e-14, atol=0.0)
    # zero right-hand side: converged at once, x = 0
    x, st = cg.solve_with_operator(op, torch.zeros_like(b))
    assert st.converged and st.iterations == 0 and torch.count_nonzero(x) == 0
    assert st.residual_norm == 0.0
    # no iterations allowed: x = 0, residual = ||b||
    x, st = G.ConjugateGradient().with_max_iterations(0).solve_with_operator(op, b)
    assert not st.converged and st.iterations == 0 and torch.count_nonzero(x) == 0
    assert abs(st.relative_residual - 1.0) < 1e-15


def test_operator_cg_callback_errors():
    n = 32
    b = np.ones(n)

    def boom(v):
        raise RuntimeError("operator failed on purpose")
    with pytest.raises(RuntimeError, match="on purpose"):
        G.ConjugateGradient().solve_with_operator(G.LinearOperator(boom, n), b)
    wrong = G.LinearOperator(lambda v: v[:-1].clone(), n)
    with pytest.raises(ValueError, match="shape"):
        G.ConjugateGradient().solve_with_operator(wrong, b)
    op = G.LinearOperator(lambda v: 2.0 * v, n)
    with pytest.raises(ValueError, match="FP64"):
        G.ConjugateGradient().with_precision("mixed").solve_with_operator(op, b)
    # the context still works after a failed callback
    x, st = G.ConjugateGradient().solve_with_operator(op, b)
    assert st.converged and np.allclose(x, 0.5)


def test_apply_result_complete_on_return():
    """Entry points return after their work is complete: a device result is
    read at once on torch's stream (the library's own stream is another)."""
    mesh, dof, _ = W.thermal_box(G, 8, 8, 6)
    K, _ = G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    Ks = K.to_scipy()
    x = np.random.default_rng(9).standard_normal(K.shape[0])
    ref = Ks @ x
    xd = torch.as_tensor(x, device="cuda")
    copies = [K.apply(xd).clone() for _ in range(20)]
    bound = 1e-13 * (abs(Ks) @ np.abs(x))
    for y in copies:
        assert np.all(np.abs(y.cpu().numpy() - ref) <= bound + 1e-300)
