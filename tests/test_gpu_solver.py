"""GPU tests of the SELL-64 operator, the system-plan cache and the
mixed-precision solver (beyond the oracle parity suite).

SpMV results are compared with scipy in FP64 (summation order differs, so
within 1e-13 of the row-wise absolute sums); CG solutions with a direct
solve or the FP64 path.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import geo_simulator_amd as G
from geo_simulator_amd import workloads as W

pytestmark = pytest.mark.gpu


def spd_random(n, density, seed, wide=False):
    """SPD matrix with row lengths from 1 to well above 64: diagonally
    dominant sum of a random sparse matrix and its transpose, plus a few
    dense rows; `wide` places columns far from the diagonal."""
    rng = np.random.default_rng(seed)
    k = int(density * n * n)
    rows = rng.integers(0, n, k)
    cols = rng.integers(0, n, k)
    if not wide:
        cols = np.clip(rows + rng.integers(-200, 201, k), 0, n - 1)
    dense_rows = rng.choice(n, 3, replace=False)
    rows = np.concatenate([rows, np.repeat(dense_rows, 150)])
    cols = np.concatenate([cols] + [rng.choice(n, 150, replace=False) for _ in dense_rows])
    B = sp.coo_matrix((rng.uniform(0.1, 1.0, len(rows)), (rows, cols)), shape=(n, n)).tocsr()
    A = (B + B.T).tocsr()
    A = A + sp.diags(np.asarray(abs(A).sum(axis=1)).ravel() + 1.0)
    A = sp.csr_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    return A


def check_spmv(M, A):
    x = np.random.default_rng(3).standard_normal(A.shape[1])
    y = M.apply(x)
    ref = A @ x
    bound = 1e-13 * (abs(A) @ np.abs(x))
    assert np.all(np.abs(y - ref) <= bound + 1e-300)


@pytest.mark.parametrize("wide", [False, True])
def test_sell_spmv_and_cg_vs_scipy(wide):
    """Banded (16-bit SELL offsets) and scattered columns over 70,000 rows
    (offsets beyond 16 bits: int32 SELL columns)."""
    n, dens = (5000, 0.002) if not wide else (70000, 4e-5)
    A = spd_random(n, dens, 11, wide=wide)
    lens = np.diff(A.indptr)
    assert lens.max() > 64 and lens.min() >= 1
    if wide:
        r = np.repeat(np.arange(n), lens)
        assert np.abs(A.indices - r).max() > 40000
    M = G.CsrMatrix.from_scipy(A)
    check_spmv(M, A)
    b = np.random.default_rng(5).standard_normal(A.shape[0])
    sol = {}
    for prec in ("fp64", "mixed"):
        x, st = (G.ConjugateGradient().with_precision(prec).with_tolerance(1e-12)
                 .with_abs_tolerance(0.0).with_max_iterations(20000).solve(M, b))
        assert st.converged, prec
        assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 2e-12, prec
        sol[prec] = x
    if not wide:
        xr = spla.spsolve(A.tocsc(), b)
        assert np.linalg.norm(sol["fp64"] - xr) / np.linalg.norm(xr) < 1e-10
    assert np.linalg.norm(sol["mixed"] - sol["fp64"]) / np.linalg.norm(sol["fp64"]) < 1e-10


def test_spmv_engines_and_index_widths():
    A = spd_random(3000, 0.003, 21)
    M = G.CsrMatrix.from_scipy(A)
    for flags in (0, G._lib.GS_SPMV_INDEX32, G._lib.GS_SPMV_CSR,
                  G._lib.GS_SPMV_CSR | G._lib.GS_SPMV_VECTOR,
                  G._lib.GS_SPMV_CSR | G._lib.GS_SPMV_INDEX32):
        M.set_spmv_shape(0, 0, flags)
        check_spmv(M, A)
    M.set_spmv_shape(0, 96, 0)  # persistent SELL grid
    check_spmv(M, A)
    with pytest.raises(G.GeoFemError):
        M.set_spmv_shape(4, 0, 0)  # lanes belong to the CSR engines
    # rectangular matrices run on the CSR engine
    rng = np.random.default_rng(4)
    R = sp.coo_matrix((rng.standard_normal(7700), (rng.integers(0, 700, 7700),
                                                  rng.integers(0, 1100, 7700))),
                      shape=(700, 1100)).tocsr()
    check_spmv(G.CsrMatrix.from_scipy(R), R)


def test_time_spmv_reports_positive():
    A = spd_random(2000, 0.004, 31)
    M = G.CsrMatrix.from_scipy(A)
    assert M.time_spmv("fp64", 3) > 0.0 and M.time_spmv("mixed", 3) > 0.0


def test_time_spmv_mixed_then_solve_negative_diagonal():
    """The mixed timing hook on an operator with a negative Jacobi entry
    times the three-launch recursion's SpMV; a mixed solve afterwards on the
    same matrix still reaches x = b / d (the operator is diagonal)."""
    n = 2500
    rng = np.random.default_rng(29)
    d = rng.uniform(1.0, 4.0, n)
    d[5] = -2.0
    M = G.CsrMatrix.from_scipy(sp.diags(d).tocsr())
    assert M.time_spmv("mixed", 3) > 0.0
    b = rng.standard_normal(n)
    x, st = (G.ConjugateGradient().with_precision("mixed").with_tolerance(1e-12)
             .with_abs_tolerance(0.0).solve(M, b))
    assert st.converged
    assert np.linalg.norm(x - b / d) / np.linalg.norm(b / d) < 1e-12


def test_mixed_pass_cap():
    """max_iterations caps the FP32 passes of a mixed solve; an unconverged
    solve reports its true residual."""
    mesh, dof, _ = W.thermal_box(G, 8, 8, 6)
    K, f = G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    x, st = (G.ConjugateGradient().with_precision("mixed").with_tolerance(1e-12)
             .with_abs_tolerance(0.0).with_max_iterations(7).solve(K, f))
    assert not st.converged and st.iterations <= 7
    Ks = K.to_scipy()
    fh = f.cpu().numpy()
    true = np.linalg.norm(fh - Ks @ x.cpu().numpy()) / np.linalg.norm(fh)
    assert abs(true - st.relative_residual) <= 1e-6 * true
    assert 1e-12 < true < 1.0


def test_system_plans_per_dofmap():
    """Two DOF maps on one mesh: each fused system equals the separate
    assemble / load / eliminate path bit for bit, also when alternated."""
    mesh = G.ImprovedMeshGenerator.generate_cube(5, 4, 3, 1000.0, 1000.0, 600.0)
    z = mesh.nodes[:, 2]
    d1 = G.DofManager(mesh.num_nodes(), 1)
    d1.set_dirichlet_many(np.nonzero(np.abs(z) < 1.0)[0], 1280.0)
    d2 = G.DofManager(mesh.num_nodes(), 1)
    d2.set_dirichlet_many(np.nonzero(np.abs(z - 600.0) < 1.0)[0], 5.0)
    d2.set_dirichlet_many(np.array([3, 17]), -2.0)
    for dof in (d1, d2, d1, d2):
        K = G.Assembler.assemble_thermal_stiffness_parallel(mesh, dof, 2.5)
        f = G.Assembler.assemble_thermal_load(mesh, dof, 1e-6)
        Kb, fb = G.Assembler.apply_dirichlet_bcs(K, f, dof)
        Ks, fs = G.Assembler.assemble_thermal_system(mesh, dof, 2.5, 1e-6)
        A, B = Kb.to_scipy(), Ks.to_scipy()
        assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
        assert np.array_equal(A.data, B.data)
        assert np.array_equal(np.asarray(fb), fs.cpu().numpy())


def test_mixed_heterogeneous_elastic_matches_fp64():
    """Log-normal Young's modulus (sigma 1): the FP32 operator's rounding
    is far from negligible; the mixed path must still converge to the FP64
    solution."""
    mesh, dof, _ = W.elastic_box(G, 6, 6, 4)
    rng = np.random.default_rng(7)
    E = np.exp(np.log(1e11) + 1.0 * rng.standard_normal(mesh.num_elements()))
    import torch
    Et = torch.tensor(E, device="cuda")
    K, f = G.Assembler.assemble_elasticity_system(mesh, dof, G.IsotropicElasticity(1.0, 0.25),
                                                  3300.0, [0, 0, -9.81], E_per_elem=Et)
    sol = {}
    for prec in ("fp64", "mixed"):
        x, st = (G.ConjugateGradient().with_precision(prec).with_tolerance(1e-12)
                 .with_abs_tolerance(0.0).with_max_iterations(100000).solve(K, f))
        assert st.converged, prec
        assert st.relative_residual < 1e-12
        sol[prec] = x.cpu().numpy()
    err = np.linalg.norm(sol["mixed"] - sol["fp64"]) / np.linalg.norm(sol["fp64"])
    assert err < 1e-9


def test_mixed_heterogeneous_thermal_matches_fp64():
    """Log-normal conductivity (sigma 1.5) on a scalar operator, where the
    mixed solve runs its two-launch recursion on the Jacobi-scaled FP32
    operator: it must reach the FP64 solution at 1e-12."""
    mesh, dof, _ = W.thermal_box(G, 10, 10, 8)
    rng = np.random.default_rng(17)
    k = np.exp(1.5 * rng.standard_normal(mesh.num_elements()))
    import torch
    kt = torch.tensor(k, device="cuda")
    K, f = G.Assembler.assemble_thermal_system(mesh, dof, 1.0, 1e-6, k_per_elem=kt)
    sol = {}
    for prec in ("fp64", "mixed"):
        x, st = (G.ConjugateGradient().with_precision(prec).with_tolerance(1e-12)
                 .with_abs_tolerance(0.0).with_max_iterations(100000).solve(K, f))
        assert st.converged, prec
        assert st.relative_residual < 1e-12
        sol[prec] = x.cpu().numpy()
    Ks = K.to_scipy()
    fh = f.cpu().numpy()
    assert np.linalg.norm(fh - Ks @ sol["mixed"]) / np.linalg.norm(fh) < 1e-12
    err = np.linalg.norm(sol["mixed"] - sol["fp64"]) / np.linalg.norm(sol["fp64"])
    assert err < 1e-9


@pytest.mark.parametrize("signs", ["positive", "mixed"])
def test_mixed_diagonal_operator(signs):
    """Diagonal operators: Jacobi makes the recursion exact in one pass.  A
    negative diagonal entry has no square root, so the solve keeps D^-1 in
    its own z instead of scaling the operator; both reach x = b / d."""
    n = 3000
    rng = np.random.default_rng(23)
    d = rng.uniform(1.0, 10.0, n)
    if signs == "mixed":
        d[::7] *= -1.0
    A = sp.diags(d).tocsr()
    M = G.CsrMatrix.from_scipy(A)
    b = rng.standard_normal(n)
    x, st = (G.ConjugateGradient().with_precision("mixed").with_tolerance(1e-12)
             .with_abs_tolerance(0.0).solve(M, b))
    assert st.converged
    assert np.linalg.norm(x - b / d) / np.linalg.norm(b / d) < 1e-12


def test_mixed_scaled_entry_out_of_fixed_point_range():
    """The scaled operator S A S is held in int32 fixed point for entries in
    (-2, 2), which holds for every symmetric positive definite matrix.  A
    symmetric indefinite matrix with a positive diagonal and |a_ij| = 3
    sqrt(a_ii a_jj) does not fit, and the solve keeps the unscaled FP32
    operator.  Spectrum {4, -2, 1} (2x2 blocks [[1, 3], [3, 1]] and unit
    rows): FP64 CG ends in three steps; the mixed solve's outer line search
    assumes a positive definite matrix, so it need not converge, but what it
    reports must be the true residual of what it returns."""
    nb, n1 = 200, 600
    blk = sp.csr_matrix(np.array([[1.0, 3.0], [3.0, 1.0]]))
    A = sp.block_diag([blk] * nb + [sp.identity(n1)]).tocsr()
    M = G.CsrMatrix.from_scipy(A)
    b = np.random.default_rng(29).standard_normal(A.shape[0])
    xr = spla.spsolve(A.tocsc(), b)
    cg = lambda p: (G.ConjugateGradient().with_precision(p).with_tolerance(1e-12)
                    .with_abs_tolerance(0.0).with_max_iterations(2000))
    x, st = cg("fp64").solve(M, b)
    assert st.converged and st.iterations <= 4
    assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-12
    x, st = cg("mixed").solve(M, b)
    true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert np.all(np.isfinite(x))
    assert abs(true - st.relative_residual) <= 1e-3 * true + 1e-15
    assert st.converged == (true <= 1e-12 * (1 + 1e-3))
    assert true < 1e-6


def test_mixed_reports_true_residual_and_counts():
    mesh, dof, exact = W.thermal_box(G, 8, 8, 6)
    K, f = G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    x, st = (G.ConjugateGradient().with_precision("mixed").with_tolerance(1e-10)
             .with_abs_tolerance(0.0).solve(K, f))
    Ks = K.to_scipy()
    fh = f.cpu().numpy()
    xh = x.cpu().numpy()
    true = np.linalg.norm(fh - Ks @ xh) / np.linalg.norm(fh)
    assert st.converged and true < 1e-10
    assert abs(true - st.relative_residual) <= 1e-3 * true + 1e-16
    assert st.outer_iterations >= 1 and st.spmv_count >= st.iterations
    assert np.linalg.norm(xh - exact) / np.linalg.norm(exact) < 1e-8


def test_mixed_inner_cap_and_outer_cap():
    """max_inner_iterations bounds the FP32 passes between FP64 steps (more
    outer steps, same answer); max_outer_iterations stops the solve
    unconverged with the true residual reported."""
    mesh, dof, exact = W.thermal_box(G, 8, 8, 6)
    K, f = G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    base = (G.ConjugateGradient().with_precision("mixed").with_tolerance(1e-10)
            .with_abs_tolerance(0.0))
    x0, s0 = base.solve(K, f)
    capped = (G.ConjugateGradient().with_precision("mixed").with_tolerance(1e-10)
              .with_abs_tolerance(0.0).with_max_inner_iterations(5)
              .with_max_outer_iterations(1000))
    x1, s1 = capped.solve(K, f)
    assert s0.converged and s1.converged
    assert s1.outer_iterations > s0.outer_iterations
    xa, xb = x0.cpu().numpy(), x1.cpu().numpy()
    assert np.linalg.norm(xa - xb) / np.linalg.norm(xa) < 1e-8
    one = (G.ConjugateGradient().with_precision("mixed").with_tolerance(1e-14)
           .with_abs_tolerance(0.0).with_max_inner_iterations(3)
           .with_max_outer_iterations(1))
    _, s2 = one.solve(K, f)
    assert not s2.converged and s2.outer_iterations == 1 and s2.relative_residual > 1e-14


@pytest.mark.parametrize("case", ["thermal", "elastic"])
@pytest.mark.parametrize("cap", [2, 4, 9])
def test_mixed_many_device_decided_cycles(case, cap):
    """Short recursion cycles (max_inner_iterations) make most outer steps
    ones the recursion predicts will not converge, which the device decides
    (restart flag, check, end) without a host read: the solve still
    converges, reports the true residual of what it returns, counts at least
    one SpMV per pass and per outer step, and matches a direct solve."""
    if case == "thermal":
        mesh, dof, _ = W.thermal_box(G, 6, 6, 4)
        K, f = G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    else:
        mesh, dof, _ = W.elastic_box(G, 4, 4, 3)
        mat = G.IsotropicElasticity(1e11, 0.25)
        K, f = G.Assembler.assemble_elasticity_system(mesh, dof, mat, 3300.0, [0, 0, -9.81])
    x, st = (G.ConjugateGradient().with_precision("mixed").with_tolerance(1e-11)
             .with_abs_tolerance(0.0).with_max_inner_iterations(cap)
             .with_max_outer_iterations(5000).with_max_iterations(100000).solve(K, f))
    Ks = K.to_scipy()
    fh, xh = f.cpu().numpy(), x.cpu().numpy()
    true = np.linalg.norm(fh - Ks @ xh) / np.linalg.norm(fh)
    assert st.converged and true < 1e-11 * (1 + 1e-3)
    assert abs(true - st.relative_residual) <= 1e-3 * true + 1e-16
    assert st.outer_iterations >= 2
    assert st.spmv_count >= st.iterations + st.outer_iterations
    xr = spla.spsolve(Ks.tocsc(), fh)
    assert np.linalg.norm(xh - xr) / np.linalg.norm(xr) < 1e-8


def block_random(nb, seed, wide=False, holes=0.3):
    """Square matrix of order 3 nb with 3x3-block structure, some entries of
    the stored blocks missing (holes), block rows of 1 to about 40 blocks."""
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for I in range(nb):
        k = int(rng.integers(1, 40))
        if wide:
            J = rng.choice(nb, k, replace=False)
        else:
            J = np.unique(np.clip(I + rng.integers(-300, 301, k), 0, nb - 1))
        J = np.union1d(J, [I])
        for j in J:
            for a in range(3):
                for b in range(3):
                    if j != I and rng.random() < holes:
                        continue
                    rows.append(3 * I + a)
                    cols.append(3 * j + b)
    vals = rng.uniform(-1.0, 1.0, len(rows))
    B = sp.coo_matrix((vals, (rows, cols)), shape=(3 * nb, 3 * nb)).tocsr()
    A = B + B.T
    A = A + sp.diags(np.asarray(abs(A).sum(axis=1)).ravel() + 1.0)
    A = sp.csr_matrix(A)
    A.sum_duplicates()
    A.sort_indices()
    return A


@pytest.mark.parametrize("wide", [False, True])
def test_block_sell_spmv_and_cg(wide):
    """3x3-block SELL view (asked for with GS_SPMV_BLOCK3) against scipy and
    against the scalar view: the same products per row in the same order."""
    nb = 3000 if not wide else 24000
    A = block_random(nb, 17, wide=wide)
    M = G.CsrMatrix.from_scipy(A)
    M.set_spmv_shape(0, 0, G._lib.GS_SPMV_BLOCK3)
    check_spmv(M, A)
    x = np.random.default_rng(2).standard_normal(A.shape[0])
    yb = M.apply(x)
    M.set_spmv_shape(0, 0, G._lib.GS_SPMV_SCALAR)
    ys = M.apply(x)
    assert np.max(np.abs(yb - ys)) <= 1e-15 * np.max(np.abs(ys))
    b = np.random.default_rng(5).standard_normal(A.shape[0])
    sol = {}
    for flags in (G._lib.GS_SPMV_BLOCK3, G._lib.GS_SPMV_SCALAR):
        M.set_spmv_shape(0, 0, flags)
        for prec in ("fp64", "mixed"):
            xs, st = (G.ConjugateGradient().with_precision(prec).with_tolerance(1e-12)
                      .with_abs_tolerance(0.0).with_max_iterations(20000).solve(M, b))
            assert st.converged, (flags, prec)
            assert np.linalg.norm(b - A @ xs) / np.linalg.norm(b) < 2e-12
            sol[(flags, prec)] = (xs, st.iterations)
    xb, ib = sol[(G._lib.GS_SPMV_BLOCK3, "fp64")]
    xs_, is_ = sol[(G._lib.GS_SPMV_SCALAR, "fp64")]
    assert abs(ib - is_) <= 2
    assert np.linalg.norm(xb - xs_) / np.linalg.norm(xs_) < 1e-10
    with pytest.raises(G.GeoFemError):
        M.set_spmv_shape(0, 0, G._lib.GS_SPMV_BLOCK3 | G._lib.GS_SPMV_SCALAR)


def test_elastic_system_block_view_matches_scalar():
    """The elastic system uses the block view by default; SpMV, FP64 CG and
    the mixed solve agree with the scalar view."""
    import torch
    mesh, dof, uz = W.elastic_box(G, 6, 5, 4)
    mat = G.IsotropicElasticity(1e11, 0.25)
    K, f = G.Assembler.assemble_elasticity_system(mesh, dof, mat, 3300.0, [0, 0, -9.81])
    Ks = K.to_scipy()
    x = torch.tensor(np.random.default_rng(4).standard_normal(Ks.shape[0]), device="cuda")
    yb = K.apply(x).cpu().numpy()
    ref = Ks @ x.cpu().numpy()
    assert np.all(np.abs(yb - ref) <= 1e-13 * (abs(Ks) @ np.abs(x.cpu().numpy())) + 1e-300)
    res = {}
    for flags in (0, G._lib.GS_SPMV_SCALAR):
        K.set_spmv_shape(0, 0, flags)
        for prec in ("fp64", "mixed"):
            u, st = (G.ConjugateGradient().with_precision(prec).with_tolerance(1e-12)
                     .with_abs_tolerance(0.0).with_max_iterations(20000).solve(K, f))
            assert st.converged
            res[(flags, prec)] = (u.cpu().numpy(), st.iterations)
    u0, i0 = res[(0, "fp64")]
    u1, i1 = res[(G._lib.GS_SPMV_SCALAR, "fp64")]
    assert abs(i0 - i1) <= 2
    assert np.linalg.norm(u0 - u1) / np.linalg.norm(u1) < 1e-10
    assert np.linalg.norm(u0[2::3] - uz) / np.linalg.norm(uz) < 1e-8
    um, _ = res[(0, "mixed")]
    assert np.linalg.norm(um - u1) / np.linalg.norm(u1) < 1e-9


@pytest.mark.parametrize("elastic", [False, True])
def test_fused_system_csr_values_on_demand(elastic):
    """A fused system keeps only its SELL values; its CSR values are produced
    when a caller asks for the CSR form.  Each consumer, reached first on a
    fresh matrix, sees the same matrix: the Jacobi diagonal, the CSR SpMV
    engine, a SELL view rebuilt with 32-bit indices (then a mixed solve), CSR
    timing, and the export."""
    from geo_simulator_amd import _lib
    if elastic:
        mesh, dof, _ = W.elastic_box(G, 4, 4, 3)

        def build():
            return G.Assembler.assemble_elasticity_system(
                mesh, dof, G.IsotropicElasticity(1e11, 0.25), 3300.0, [0.0, 0.0, -9.81])
    else:
        mesh, dof, _ = W.thermal_box(G, 6, 6, 4)

        def build():
            return G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    Kref, f = build()
    A = Kref.to_scipy()
    d = A.diagonal()
    x = np.random.default_rng(41).standard_normal(A.shape[0])
    bound = 1e-13 * (abs(A) @ np.abs(x))

    K1, _ = build()
    dinv = K1.jacobi_diag_inv().cpu().numpy()
    assert np.array_equal(dinv, np.where(np.abs(d) > 1e-14, 1.0 / d, 1.0))

    K2, _ = build()
    K2.set_spmv_shape(flags=_lib.GS_SPMV_CSR)
    assert np.all(np.abs(K2.apply(x) - A @ x) <= bound + 1e-300)
    assert K2.time_spmv("fp64", 2) > 0.0

    K3, f3 = build()
    K3.set_spmv_shape(flags=_lib.GS_SPMV_INDEX32)
    assert np.all(np.abs(K3.apply(x) - A @ x) <= bound + 1e-300)
    u, st = (G.ConjugateGradient().with_precision("mixed").with_tolerance(1e-10)
             .with_abs_tolerance(0.0).solve(K3, f3))
    assert st.converged
    B = K3.to_scipy()
    assert np.array_equal(B.indptr, A.indptr) and np.array_equal(B.indices, A.indices)
    assert np.array_equal(B.data, A.data)
