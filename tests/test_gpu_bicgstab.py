"""GPU BiCGSTAB (src/linalg/iterative.rs:149-291) against the oracle's
restatement: same pass counts (within the rounding of a different dot-product
order), same solutions, same stopping semantics (a pass that ends on the
intermediate residual s is not counted; absolute breakdown thresholds)."""
import numpy as np
import pytest
import scipy.sparse as sp

import geo_simulator_amd as G
from geo_simulator_amd import workloads as W

pytestmark = pytest.mark.gpu


def nonsym(n, seed, dens, skew=1.0, wide=False):
    rng = np.random.default_rng(seed)
    k = int(dens * n * n)
    r = rng.integers(0, n, k)
    c = rng.integers(0, n, k) if wide else np.clip(r + rng.integers(-150, 151, k), 0, n - 1)
    B = sp.coo_matrix((rng.uniform(-1, 1, k), (r, c)), shape=(n, n)).tocsr()
    A = (B + B.T) + skew * (B - B.T)
    A = A + sp.diags(np.asarray(abs(A).sum(axis=1)).ravel() + 1.0)
    A = sp.csr_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    return A


@pytest.mark.parametrize("precond", [True, False])
@pytest.mark.parametrize("wide", [False, True])
def test_bicgstab_matches_oracle(oracle, precond, wide):
    n = 4000 if not wide else 70000
    A = nonsym(n, 3, 0.003 if not wide else 6e-5, wide=wide)
    b = np.random.default_rng(9).standard_normal(n)
    xo, so = oracle.bicgstab_solve(A, b, 5000, 1e-10, 0.0, precond)
    x, st = (G.BiCGSTAB().with_tolerance(1e-10).with_abs_tolerance(0.0)
             .with_max_iterations(5000).with_preconditioner(precond).solve(A, b))
    assert so.converged and st.converged
    assert abs(st.iterations - so.iterations) <= 2
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
    true = np.linalg.norm(b - A @ x)
    assert true < 2e-10 * np.linalg.norm(b)
    assert abs(true - st.residual_norm) <= 1e-2 * true + 1e-14
    assert st.spmv_count >= 2 * st.iterations


def test_bicgstab_first_passes_and_edges(oracle):
    A = nonsym(3000, 5, 0.004)
    b = np.ones(3000)
    for k in (1, 2, 3):  # after k passes the iterates agree to rounding
        xo, so = oracle.bicgstab_solve(A, b, k, 1e-14, 0.0)
        x, st = (G.BiCGSTAB().with_tolerance(1e-14).with_abs_tolerance(0.0)
                 .with_max_iterations(k).solve(A, b))
        assert st.iterations == so.iterations == k and not st.converged
        assert np.linalg.norm(x - xo) <= 1e-12 * np.linalg.norm(xo)
        assert abs(st.residual_norm - so.residual_norm) <= 1e-10 * so.residual_norm
    # b = 0: converged at once, x = 0
    x, st = G.BiCGSTAB().solve(A, np.zeros(3000))
    assert st.converged and st.iterations == 0 and np.all(x == 0.0)
    # identity: the first pass ends on s with x = b, uncounted
    I = sp.identity(500, format="csr")
    bb = np.arange(500.0) + 1.0
    x, st = G.BiCGSTAB().solve(I, bb)
    assert st.converged and st.iterations == 0 and np.array_equal(x, bb)
    # max_iterations 0: nothing runs
    x, st = G.BiCGSTAB().with_max_iterations(0).solve(A, b)
    assert st.iterations == 0 and not st.converged and np.all(x == 0.0)


def test_bicgstab_on_thermal_system_and_errors():
    mesh, dof, exact = W.thermal_box(G, 8, 8, 6)
    K, f = G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    x, st = G.BiCGSTAB().with_tolerance(1e-12).with_abs_tolerance(0.0).solve(K, f)
    xc, sc = (G.ConjugateGradient().with_tolerance(1e-12).with_abs_tolerance(0.0)
              .with_max_iterations(5000).solve(K, f))
    assert st.converged and sc.converged
    xh, xch = x.cpu().numpy(), xc.cpu().numpy()
    assert np.linalg.norm(xh - xch) / np.linalg.norm(xch) < 1e-9
    assert np.linalg.norm(xh - exact) / np.linalg.norm(exact) < 1e-8
    M = G.CsrMatrix.from_scipy(nonsym(500, 2, 0.01))
    M.set_spmv_shape(0, 0, G._lib.GS_SPMV_CSR)
    with pytest.raises(G.GeoFemError):
        G.BiCGSTAB().solve(M, np.ones(500))
    R = G.CsrMatrix.from_scipy(sp.random(30, 40, 0.2, format="csr", random_state=1))
    with pytest.raises(G.GeoFemError):
        G.BiCGSTAB().solve(R, np.ones(30))
