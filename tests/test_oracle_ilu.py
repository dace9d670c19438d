"""Oracle ILU(0) (ILUPreconditioner, src/linalg/preconditioner.rs:78-204)
pinned independently of its own formulas:
  - the reference's unit tests (:241-292): the hand-worked 3x3 factors and
    solve, and the zero-diagonal failure;
  - the defining property of ILU(0): (L U)_ij = A_ij on A's pattern;
  - on a full pattern, the LU factorization without pivoting (scipy's
    partial-pivoting LU of a column diagonally dominant matrix does not
    pivot);
  - on a pattern closed under elimination (tridiagonal), the exact solve;
  - apply(r) solves (L U) z = r.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp


def lu_parts(F):
    D = F.toarray()
    L = np.tril(D, -1) + np.eye(D.shape[0])
    U = np.triu(D)
    return L, U


def random_sparse(n, density, seed, dominant=True):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=rng, format="csr")
    A.data = rng.standard_normal(A.nnz)
    d = np.abs(A).sum(axis=1).A1 + np.abs(A).sum(axis=0).A1 + 1.0
    A = (A + sp.diags(d if dominant else rng.standard_normal(n))).tocsr()
    A.sort_indices()
    return A


def test_reference_kat(oracle):
    A = sp.csr_matrix(np.array([[2.0, -1, 0], [-1, 2, -1], [0, -1, 2]]))
    F, diag = oracle.ilu0_factor(A)
    # i=1: -1/2 = -0.5, 2 - (-0.5)(-1) = 1.5; i=2: -1/1.5, 2 - (-1/1.5)(-1) = 4/3
    assert F[1, 0] == -0.5 and F[1, 1] == 1.5 and F[0, 0] == 2.0 and F[0, 1] == -1.0
    assert abs(F[2, 1] + 2.0 / 3.0) < 1e-16 and abs(F[2, 2] - 4.0 / 3.0) < 3e-16
    assert list(diag) == [0, 3, 6]
    z = oracle.ilu0_apply(F, diag, [0.0, 0.0, 4.0])
    assert np.allclose(z, [1.0, 2.0, 3.0], rtol=0, atol=1e-15)


def test_reference_fail_fast(oracle):
    A = sp.csr_matrix((np.array([0.0, 1.0]), np.array([0, 1]), np.array([0, 1, 2])),
                      shape=(2, 2))  # explicit stored zero on the diagonal
    with pytest.raises(oracle.IluFactorError, match="Zero diag after update at 0"):
        oracle.ilu0_factor(A)
    B = sp.csr_matrix(np.array([[0.0, 1.0], [1.0, 0.0]]))
    with pytest.raises(oracle.IluFactorError, match="Row 0 has no diag"):
        oracle.ilu0_factor(B)
    C = sp.csr_matrix(np.array([[1.0, 1.0], [1.0, 1.0]]))  # pivot 1 - 1*1 = 0
    with pytest.raises(oracle.IluFactorError, match="Zero diag after update at 1"):
        oracle.ilu0_factor(C)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_product_matches_on_pattern(oracle, seed):
    A = random_sparse(80, 0.06, seed)
    F, diag = oracle.ilu0_factor(A)
    L, U = lu_parts(F)
    P = (L @ U)
    r, c = A.nonzero()
    Ad = A.toarray()
    scale = np.abs(Ad).max()
    assert np.abs(P[r, c] - Ad[r, c]).max() <= 1e-13 * scale
    # zero fill: the factors live on A's pattern, and the product differs
    # from A off it (the dropped fill), so this is not a full LU
    assert F.nnz == A.nnz and np.array_equal(F.indices, A.indices)
    mask = np.ones_like(Ad, dtype=bool)
    mask[r, c] = False
    assert np.abs(P[mask]).max() > 1e-6


def test_full_pattern_is_lu(oracle):
    rng = np.random.default_rng(5)
    n = 12
    D = rng.standard_normal((n, n))
    D += np.diag(np.abs(D).sum(axis=0) + 1.0)  # column dominant: no pivoting
    F, diag = oracle.ilu0_factor(sp.csr_matrix(D))
    L, U = lu_parts(F)
    P, Ls, Us = sla.lu(D)
    assert np.array_equal(P, np.eye(n))
    assert np.abs(L - Ls).max() < 1e-13 and np.abs(U - Us).max() < 1e-12 * np.abs(Us).max()


def test_tridiagonal_exact_solve(oracle):
    rng = np.random.default_rng(6)
    n = 50
    A = sp.diags([rng.uniform(-1, 1, n - 1), rng.uniform(3, 4, n), rng.uniform(-1, 1, n - 1)],
                 [-1, 0, 1]).tocsr()
    F, diag = oracle.ilu0_factor(A)
    b = rng.standard_normal(n)
    z = oracle.ilu0_apply(F, diag, b)
    x = np.linalg.solve(A.toarray(), b)
    assert np.abs(z - x).max() <= 1e-13 * np.abs(x).max()


def test_apply_solves_lu(oracle):
    A = random_sparse(120, 0.04, 9)
    F, diag = oracle.ilu0_factor(A)
    L, U = lu_parts(F)
    r = np.random.default_rng(9).standard_normal(120)
    z = oracle.ilu0_apply(F, diag, r)
    assert np.linalg.norm(L @ (U @ z) - r) <= 1e-12 * np.linalg.norm(r)
