"""Oracle AMG (AMGPreconditioner, src/linalg/amg.rs) pinned independently of
its own loops:
  - the reference's unit test (:499-525): the 1D Laplacian hierarchy with
    coarse_size 2 and a finite V-cycle of the right length;
  - literal transcriptions of the reference's loops, in pure Python on small
    matrices (the O(n^2) scans of coarsen and strength_matrix, the dense
    column loop of matrix_multiply, the recursive v_cycle): the oracle's
    efficient loops must give the same coarse points and the same bits;
  - mathematics: every fine row of P sums to 1 and coarse rows are unit
    vectors, A_c = P^T A P against scipy, the Galerkin operator of a
    symmetric matrix is symmetric, a planted mistake (Jacobi weight 2/3
    instead of 0.67) changes the V-cycle's bits, and AMG-preconditioned CG
    needs fewer iterations than Jacobi-PCG on a 2D Poisson matrix.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp


# ---- literal transcriptions of src/linalg/amg.rs (pure Python, small n) ------

def rows_of(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return [list(zip(A.indices[A.indptr[i]:A.indptr[i + 1]].tolist(),
                     A.data[A.indptr[i]:A.indptr[i + 1]].tolist()))
            for i in range(A.shape[0])]


def get(rows, i, j):
    for c, v in rows[i]:
        if c == j:
            return v
    return None


def ref_strength(rows, theta):  # :359-390
    n = len(rows)
    S = [[] for _ in range(n)]
    for i in range(n):
        max_off = 0.0
        for j in range(n):
            if i != j:
                v = get(rows, i, j)
                if v is not None:
                    max_off = max(max_off, abs(v))
        threshold = theta * max_off
        for j in range(n):
            if i != j:
                v = get(rows, i, j)
                if v is not None and abs(v) >= threshold:
                    S[i].append((j, v))
    return S


def ref_coarsen(rows, theta):  # :286-355
    n = len(rows)
    S = ref_strength(rows, theta)
    lam = [sum(1 for j, _ in S[i] if j != i) for i in range(n)]
    cf = [False] * n
    und = [True] * n
    nc = 0
    while any(und):
        max_lambda, max_node = 0, 0
        for i in range(n):
            if und[i] and lam[i] > max_lambda:
                max_lambda, max_node = lam[i], i
        if max_lambda == 0:
            for i in range(n):
                if und[i]:
                    cf[i], und[i] = True, False
                    nc += 1
            break
        cf[max_node], und[max_node] = True, False
        nc += 1
        for j, _ in S[max_node]:
            if j != max_node and und[j]:
                und[j] = False
                for k, _ in S[j]:
                    if k != j and und[k]:
                        lam[k] += 1
    return cf, nc


def ref_prolongation(rows, cf, nc, theta):  # :396-487
    n = len(rows)
    S = ref_strength(rows, theta)
    cidx, c = [0] * n, 0
    for i in range(n):
        if cf[i]:
            cidx[i] = c
            c += 1
    P = [[] for _ in range(n)]
    for i in range(n):
        if cf[i]:
            P[i].append((cidx[i], 1.0))
            continue
        nb, s = [], 0.0
        for j, v in S[i]:
            if cf[j]:
                nb.append((j, abs(v)))
                s += abs(v)
        if s > 1e-14:
            P[i] += [(cidx[j], w / s) for j, w in nb]
        else:
            nb, s = [], 0.0
            for j, v in rows[i]:
                if cf[j]:
                    nb.append((j, abs(v)))
                    s += abs(v)
            if s > 1e-14:
                P[i] += [(cidx[j], w / s) for j, w in nb]
            elif nc > 0:
                P[i].append((0, 1.0))
    return P


def ref_matmul(arows, brows, ncols):  # :195-226, the dense column loop
    out = []
    for i, arow in enumerate(arows):
        row = []
        for j in range(ncols):
            s = 0.0
            for k, a in arow:
                b = get(brows, k, j)
                if b is not None:
                    s += a * b
            if abs(s) > 1e-14:
                row.append((j, s))
        out.append(row)
    return out


def ref_transpose(rows, ncols):
    T = [[] for _ in range(ncols)]
    for i, row in enumerate(rows):
        for j, v in row:
            T[j].append((i, v))
    return T


def ref_diag_inv(rows):  # :266-280
    d = []
    for i in range(len(rows)):
        v = get(rows, i, i)
        d.append(1.0 / v if v is not None and abs(v) > 1e-14 else 1.0)
    return d


def ref_build(rows, max_levels, coarse_size, theta):  # :44-129
    levels = []
    cur = rows
    for _ in range(max_levels):
        n = len(cur)
        if n <= coarse_size:
            levels.append((cur, None, ref_diag_inv(cur)))
            break
        dinv = ref_diag_inv(cur)
        cf, nc = ref_coarsen(cur, theta)
        if nc == 0 or nc >= n:
            levels.append((cur, None, dinv))
            break
        P = ref_prolongation(cur, cf, nc, theta)
        PT = ref_transpose(P, nc)
        AP = ref_matmul(cur, P, nc)
        Ac = ref_matmul(PT, AP, nc)
        levels.append((cur, P, dinv))
        cur = Ac
    return levels


def ref_vcycle(levels, r, level=0, omega=0.67):  # :132-185
    A, P, dinv = levels[level]
    n = len(r)
    x = [0.0] * n
    if level == len(levels) - 1:
        for _ in range(10):
            for i in range(n):
                ax = 0.0
                for j, v in A[i]:
                    ax += v * x[j]
                x[i] += (r[i] - ax) * dinv[i]
        return x

    def jacobi(x):
        xn = list(x)
        for i in range(n):
            ax = 0.0
            for j, v in A[i]:
                ax += v * x[j]
            xn[i] = x[i] + omega * (r[i] - ax) * dinv[i]
        return xn

    def matvec(M, v):
        out = []
        for row in M:
            s = 0.0
            for j, a in row:
                s += a * v[j]
            out.append(s)
        return out

    for _ in range(2):
        x = jacobi(x)
    ax = matvec(A, x)
    res = [r[i] - ax[i] for i in range(n)]
    nc = max((j for row in P for j, _ in row), default=-1) + 1
    PT = ref_transpose(P, nc)
    ec = ref_vcycle(levels, matvec(PT, res), level + 1, omega)
    corr = matvec(P, ec)
    x = [x[i] + corr[i] for i in range(n)]
    for _ in range(2):
        x = jacobi(x)
    return x


def to_rows_csr(rows, ncols):
    indptr, idx, val = [0], [], []
    for row in rows:
        for j, v in row:
            idx.append(j)
            val.append(v)
        indptr.append(len(idx))
    return sp.csr_matrix((np.array(val, dtype=float), np.array(idx, dtype=np.int64),
                          np.array(indptr, dtype=np.int64)), shape=(len(rows), ncols))


# ---- matrices ----------------------------------------------------------------

def lap1d(n):
    return sp.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1]).tocsr()


def lap2d(m):
    T = lap1d(m)
    I = sp.identity(m)
    return (sp.kron(T, I) + sp.kron(I, T)).tocsr()


def random_spd(n, density, seed):
    rng = np.random.default_rng(seed)
    B = sp.random(n, n, density=density, random_state=rng, format="csr")
    B.data = -np.abs(rng.standard_normal(B.nnz))
    B = B + B.T
    B.setdiag(0.0)
    B.eliminate_zeros()
    d = np.abs(B).sum(axis=1).A1 + rng.uniform(0.1, 1.0, n)
    A = (B + sp.diags(d)).tocsr()
    A.sort_indices()
    return A


def mesh_matrix(oracle, n=3):
    xyz, conn = oracle.generate_cube(n, n, n, 1000.0, 1000.0, 600.0)
    return oracle.assemble_thermal(xyz, conn, 3.0)


# ---- tests -------------------------------------------------------------------

def test_reference_unit_test(oracle):
    A = lap1d(10)
    M = oracle.Amg(A, 10, 2, 0.25)
    x = M.apply(np.ones(10))
    assert x.shape == (10,) and np.all(np.isfinite(x))
    # hand-checked coarsening: every odd node (node 1 has the first largest
    # lambda; each choice makes its neighbours fine and raises the next one)
    cf, nc = oracle.amg_coarsen(A, 0.25)
    assert nc == 5 and cf.tolist() == [False, True] * 5
    assert [a.shape[0] for a in M.A] == [10, 5, 2]


@pytest.mark.parametrize("theta", [0.0, 0.25, 0.5, 0.9])
@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_strength_and_coarsen_match_reference_scan(oracle, seed, theta):
    A = random_spd(60, 0.08, seed)
    rows = rows_of(A)
    S = oracle.amg_strength(A, theta)
    want = to_rows_csr(ref_strength(rows, theta), A.shape[1])
    assert np.array_equal(S.indptr, want.indptr) and np.array_equal(S.indices, want.indices)
    assert np.array_equal(S.data, want.data)
    cf, nc = oracle.amg_coarsen(A, theta)
    rcf, rnc = ref_coarsen(rows, theta)
    assert nc == rnc and cf.tolist() == rcf


def test_coarsen_ties_and_meshes(oracle):
    # uniform rows (ties everywhere) and a Tet10 mesh matrix
    for A in (lap2d(9), mesh_matrix(oracle, 2)):
        rows = rows_of(A)
        cf, nc = oracle.amg_coarsen(A, 0.25)
        rcf, rnc = ref_coarsen(rows, 0.25)
        assert nc == rnc and cf.tolist() == rcf


@pytest.mark.parametrize("seed", [5, 6])
def test_prolongation_matches_reference(oracle, seed):
    A = random_spd(50, 0.1, seed)
    rows = rows_of(A)
    cf, nc = oracle.amg_coarsen(A, 0.25)
    P = oracle.amg_prolongation(A, cf, nc, 0.25)
    want = to_rows_csr(ref_prolongation(rows, cf.tolist(), nc, 0.25), nc)
    assert np.array_equal(P.indptr, want.indptr) and np.array_equal(P.indices, want.indices)
    assert np.array_equal(P.data, want.data)
    # coarse rows are unit vectors in coarse order; fine rows sum to 1
    c = np.cumsum(cf) - 1
    for i in range(A.shape[0]):
        row = P.getrow(i)
        if cf[i]:
            assert row.indices.tolist() == [c[i]] and row.data.tolist() == [1.0]
        else:
            assert abs(row.data.sum() - 1.0) < 1e-14


def test_prolongation_fallbacks(oracle):
    # node 3 is fine and its only strong neighbours are fine (2 and 4 pull
    # hard, the coarse ones weakly): the all-coarse fallback over A's row;
    # node 5 connects to nothing coarse at all: the island, first coarse point
    dense = np.zeros((6, 6))
    for i in range(6):
        dense[i, i] = 10.0
    for i, j, v in [(0, 1, -1.0), (1, 2, -0.01), (2, 3, -5.0), (3, 4, -5.0), (3, 1, -0.01),
                    (3, 0, -0.01), (4, 0, -1.0), (5, 5, 10.0)]:
        if i != j:
            dense[i, j] = dense[j, i] = v
    A = sp.csr_matrix(dense)
    cf = np.array([True, True, False, False, False, False])
    P = oracle.amg_prolongation(A, cf, 2, 0.25)
    want = to_rows_csr(ref_prolongation(rows_of(A), cf.tolist(), 2, 0.25), 2)
    assert np.array_equal(P.indices, want.indices) and np.array_equal(P.data, want.data)
    assert P.getrow(3).data.tolist() == [0.5, 0.5]      # |a_30|, |a_31| equal
    assert P.getrow(5).indices.tolist() == [0] and P.getrow(5).data.tolist() == [1.0]


@pytest.mark.parametrize("seed", [7, 8])
def test_matmul_matches_reference_and_scipy(oracle, seed):
    A = random_spd(40, 0.1, seed)
    cf, nc = oracle.amg_coarsen(A, 0.25)
    P = oracle.amg_prolongation(A, cf, nc, 0.25)
    AP = oracle.csr_matmul(A, P)
    want = to_rows_csr(ref_matmul(rows_of(A), rows_of(P), nc), nc)
    assert np.array_equal(AP.indptr, want.indptr) and np.array_equal(AP.indices, want.indices)
    assert np.array_equal(AP.data, want.data)
    Ac = oracle.csr_matmul(P.T.tocsr(), AP)
    ref = (P.T @ A @ P).toarray()
    ref[np.abs(ref) <= 1e-14] = 0.0
    assert np.allclose(Ac.toarray(), ref, rtol=1e-13, atol=1e-14)
    D = Ac.toarray()
    assert np.allclose(D, D.T, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("case", ["lap1d", "lap2d", "random", "mesh"])
def test_vcycle_bits_match_reference(oracle, case):
    A = {"lap1d": lambda: lap1d(40), "lap2d": lambda: lap2d(8),
         "random": lambda: random_spd(70, 0.06, 11),
         "mesh": lambda: mesh_matrix(oracle, 2)}[case]()
    n = A.shape[0]
    coarse = 6
    M = oracle.Amg(A, 10, coarse, 0.25)
    levels = ref_build(rows_of(A), 10, coarse, 0.25)
    assert M.nlevels == len(levels) >= 2
    for l, (Ar, Pr, dr) in enumerate(levels):
        want = to_rows_csr(Ar, len(Ar))
        assert np.array_equal(M.A[l].indices, want.indices)
        assert np.array_equal(M.A[l].data, want.data)
        assert M.dinv[l].tolist() == dr
    r = np.random.default_rng(3).standard_normal(n)
    z = M.apply(r)
    zr = ref_vcycle(levels, r.tolist())
    assert z.tolist() == zr
    assert z.tolist() != ref_vcycle(levels, r.tolist(), omega=2.0 / 3.0)


def test_max_levels_cut(oracle):
    # all levels coarsened: the last kept level has its P, the last Galerkin
    # operator is dropped, and the V-cycle ends on that level's own operator
    A = lap1d(40)
    M = oracle.Amg(A, 2, 2, 0.25)
    assert M.nlevels == 2 and M.P[1] is not None and M.A[1].shape == (20, 20)
    levels = ref_build(rows_of(A), 2, 2, 0.25)
    r = np.linspace(-1.0, 1.0, 40)
    assert M.apply(r).tolist() == ref_vcycle(levels, r.tolist())
    with pytest.raises(oracle.OracleError):
        oracle.Amg(A, 0, 2, 0.25)


def _pcg(A, b, M, tol=1e-10, maxit=500):
    x = np.zeros_like(b)
    r = b.copy()
    z = M(r)
    p = z.copy()
    rz = r @ z
    bn = np.linalg.norm(b)
    for it in range(maxit):
        q = A @ p
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) < tol * bn:
            return it + 1
        z = M(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    return maxit


def test_amg_preconditions_poisson(oracle):
    A = lap2d(24)
    b = np.random.default_rng(0).standard_normal(A.shape[0])
    M = oracle.Amg(A, 10, 20, 0.25)
    d = 1.0 / A.diagonal()
    it_amg = _pcg(A, b, M.apply)
    it_jac = _pcg(A, b, lambda r: d * r)
    assert it_amg < 0.6 * it_jac, (it_amg, it_jac)
    assert math.isfinite(it_amg)
