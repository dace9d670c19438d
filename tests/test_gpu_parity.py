"""GPU parity: libgeofem (HIP, gfx950) against the FP64 oracle.

Bar: element matrices, assembled CSR values, loads and Dirichlet elimination
are bit-identical to the oracle (same operation order, no FP contraction, the
same ascending-element summation per CSR slot).  SpMV and CG are floating
point reductions in a different order: SpMV within 1e-14 relative, CG
solutions within 1e-10 relative L2 of the oracle's at tight tolerance, and
iteration counts within a few passes of the oracle's.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import geo_simulator_amd as G

pytestmark = pytest.mark.gpu


def mesh_pair(oracle, nx, ny, nz, lx, ly, lz, perturb=0.0, seed=0):
    xyz, conn = oracle.generate_cube(nx, ny, nz, lx, ly, lz)
    if perturb:
        rng = np.random.default_rng(seed)
        nv = (nx + 1) * (ny + 1) * (nz + 1)
        h = min(lx / nx, ly / ny, lz / nz)
        xyz = xyz.copy()
        xyz[nv:] += perturb * h * rng.standard_normal((len(xyz) - nv, 3))
    return xyz, conn, G.Mesh(xyz, conn)


def same_csr(A, B):
    A = sp.csr_matrix(A)
    B = sp.csr_matrix(B)
    return (A.shape == B.shape and np.array_equal(A.indptr, B.indptr)
            and np.array_equal(A.indices, B.indices) and np.array_equal(A.data, B.data))


# ---- element kernels + assembly: bit-exact ------------------------------------
@pytest.mark.parametrize("perturb", [0.0, 0.08])
def test_thermal_assembly_bitexact(oracle, perturb):
    xyz, conn, mesh = mesh_pair(oracle, 6, 6, 6, 1000.0, 1000.0, 600.0, perturb, 1)
    dof = G.DofManager(mesh.num_nodes(), 1)
    K = G.Assembler.assemble_thermal_stiffness_parallel(mesh, dof, 3.0)
    assert same_csr(K.to_scipy(), oracle.assemble_thermal(xyz, conn, 3.0))
    f = G.Assembler.assemble_thermal_load(mesh, dof, 1e-6)
    assert np.array_equal(f, oracle.assemble_thermal_load(xyz, conn, 1e-6))


def test_thermal_assembly_per_element_k_bitexact(oracle):
    xyz, conn, mesh = mesh_pair(oracle, 4, 3, 5, 1000.0, 800.0, 600.0, 0.05, 2)
    k = np.exp(np.random.default_rng(4).normal(1.0, 0.4, len(conn)))
    dof = G.DofManager(mesh.num_nodes(), 1)
    K = G.Assembler.assemble_thermal_stiffness_parallel(mesh, dof, 0.0, k_per_elem=k)
    assert same_csr(K.to_scipy(), oracle.assemble_thermal(xyz, conn, 0.0, k_elem=k))


@pytest.mark.parametrize("perturb", [0.0, 0.08])
def test_thermal_mass_bitexact(oracle, perturb):
    """5-point thermal mass (src/fem/element.rs:74-105), all 100 entries per
    element, uniform and per-element rho c_p, straight and curved elements."""
    xyz, conn, mesh = mesh_pair(oracle, 5, 4, 3, 1000.0, 800.0, 600.0, perturb, 6)
    dof = G.DofManager(mesh.num_nodes(), 1)
    M = G.Assembler.assemble_thermal_mass(mesh, dof, 3300.0, 1250.0)
    assert same_csr(M.to_scipy(), oracle.assemble_thermal_mass(xyz, conn, 3300.0, 1250.0))
    rc = np.exp(np.random.default_rng(5).normal(15.0, 0.3, len(conn)))
    M2 = G.Assembler.assemble_thermal_mass(mesh, dof, 0.0, 0.0, rho_cp_per_elem=rc)
    assert same_csr(M2.to_scipy(), oracle.assemble_thermal_mass(xyz, conn, 0.0, 0.0,
                                                                rho_cp_elem=rc))


@pytest.mark.parametrize("perturb", [0.0, 0.08])
def test_neumann_faces_bitexact(oracle, perturb):
    """Neumann fluxes on the top and one side (src/fem/assembly.rs:248-309):
    the GPU face integrals and the per-node sums in face order equal the
    oracle's sequential loop bit for bit, on flat and curved faces."""
    import torch
    xyz, conn, mesh = mesh_pair(oracle, 4, 3, 3, 1000.0, 800.0, 600.0, perturb, 7)
    # faces found on the flat mesh; the perturbed one curves them
    flat = G.Mesh(*oracle.generate_cube(4, 3, 3, 1000.0, 800.0, 600.0))
    top = G.Assembler.find_boundary_faces(flat, lambda p: np.abs(p[:, 2] - 600.0) < 1.0)
    side = G.Assembler.find_boundary_faces(flat, lambda p: abs(p[0]) < 1.0)  # per point
    assert len(top) == 2 * 4 * 3 and len(side) == 2 * 3 * 3
    bcs = G.BoundaryConditions()
    rng = np.random.default_rng(3)
    for e, lf in top + side:
        bcs.add_neumann(G.BoundaryFace(e, lf), float(rng.uniform(0.01, 0.1)))
    faces = list(bcs.neumann_faces())
    el = [fc.element_id for fc, _ in faces]
    lfs = [fc.local_face_id for fc, _ in faces]
    fl = [bc.flux for _, bc in faces]
    f0 = rng.standard_normal(mesh.num_nodes())
    want = oracle.apply_neumann(xyz, conn, el, lfs, fl, f0)
    got = G.Assembler.apply_neumann_bcs(mesh, f0.copy(), bcs)
    assert np.array_equal(got, want)
    fd = torch.tensor(f0, device="cuda")
    G.Assembler.apply_neumann_bcs(mesh, fd, bcs)  # in place on the device
    assert np.array_equal(fd.cpu().numpy(), want)
    bad = G.BoundaryConditions()
    bad.add_neumann(G.BoundaryFace(0, 4), 1.0)
    with pytest.raises(G.GeoFemError):
        G.Assembler.apply_neumann_bcs(mesh, f0.copy(), bad)
    with pytest.raises(ValueError):
        G.Tet10Face.from_element_face(conn[0], 4)
    assert np.array_equal(G.Tet10Face.from_element_face(conn[0], 0).nodes,
                          conn[0][oracle.face_nodes(0)])


@pytest.mark.parametrize("perturb", [0.0, 0.06])
def test_elastic_assembly_bitexact(oracle, perturb):
    xyz, conn, mesh = mesh_pair(oracle, 2, 3, 2, 1e6, 1e6, 6e5, perturb, 3)
    dof = G.DofManager(mesh.num_nodes(), 3)
    mat = G.IsotropicElasticity(1e11, 0.25)
    K = G.Assembler.assemble_elasticity_stiffness_parallel(mesh, dof, mat)
    assert same_csr(K.to_scipy(), oracle.assemble_elastic(xyz, conn, 1e11, 0.25))
    f = G.Assembler.assemble_gravity_load(mesh, dof, 3300.0, [0.0, 0.0, -9.81])
    assert np.array_equal(f, oracle.assemble_gravity_load(xyz, conn, 3300.0, [0, 0, -9.81]))


def test_elastic_per_element_E_bitexact(oracle):
    xyz, conn, mesh = mesh_pair(oracle, 2, 2, 2, 1e6, 1e6, 6e5, 0.04, 5)
    E = 1e11 * np.exp(0.5 * np.random.default_rng(6).standard_normal(len(conn)))
    dof = G.DofManager(mesh.num_nodes(), 3)
    K = G.Assembler.assemble_elasticity_stiffness_parallel(
        mesh, dof, G.IsotropicElasticity(1.0, 0.3), E_per_elem=E)
    assert same_csr(K.to_scipy(), oracle.assemble_elastic(xyz, conn, 0.0, 0.3, E_elem=E))


def test_dirichlet_bitexact_and_fused(oracle):
    xyz, conn, mesh = mesh_pair(oracle, 6, 6, 6, 1000.0, 1000.0, 600.0)
    mask, vals = oracle.thermal_dirichlet(xyz, 600.0)
    dof = G.DofManager(mesh.num_nodes(), 1)
    dof.set_dirichlet_many(np.nonzero(mask)[0], vals[mask == 1])
    K = G.Assembler.assemble_thermal_stiffness_parallel(mesh, dof, 3.0)
    f = G.Assembler.assemble_thermal_load(mesh, dof, 1e-6)
    Kbc, fbc = G.Assembler.apply_dirichlet_bcs(K, f, dof)
    Ko = oracle.assemble_thermal(xyz, conn, 3.0)
    fo = oracle.assemble_thermal_load(xyz, conn, 1e-6)
    Kbo, fbo = oracle.apply_dirichlet(Ko, fo, mask, vals)
    assert Kbc.nnz() == 44513
    assert same_csr(Kbc.to_scipy(), Kbo)
    assert np.array_equal(fbc, fbo)
    Kf, ff = G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    assert same_csr(Kf.to_scipy(), Kbo)
    assert np.array_equal(ff.cpu().numpy(), fbo)


def test_elastic_dirichlet_bitexact_and_fused(oracle):
    lx = ly = 1e6
    lz = 6e5
    xyz, conn, mesh = mesh_pair(oracle, 2, 2, 3, lx, ly, lz)
    mask, vals = oracle.elastic_dirichlet(xyz, lx, ly)
    dof = G.DofManager(mesh.num_nodes(), 3)
    dof.set_dirichlet_many(np.nonzero(mask)[0], 0.0)
    mat = G.IsotropicElasticity(1e11, 0.25)
    Ko = oracle.assemble_elastic(xyz, conn, 1e11, 0.25)
    fo = oracle.assemble_gravity_load(xyz, conn, 3300.0, [0, 0, -9.81])
    Kbo, fbo = oracle.apply_dirichlet(Ko, fo, mask, vals)
    K = G.Assembler.assemble_elasticity_stiffness_parallel(mesh, dof, mat)
    f = G.Assembler.assemble_gravity_load(mesh, dof, 3300.0, [0, 0, -9.81])
    Kbc, fbc = G.Assembler.apply_dirichlet_bcs(K, f, dof)
    assert same_csr(Kbc.to_scipy(), Kbo) and np.array_equal(fbc, fbo)
    Kf, ff = G.Assembler.assemble_elasticity_system(mesh, dof, mat, 3300.0, [0, 0, -9.81])
    assert same_csr(Kf.to_scipy(), Kbo) and np.array_equal(ff.cpu().numpy(), fbo)


def test_dirichlet_on_uploaded_matrix_bitexact(oracle):
    # identity-numbered path: apply_dirichlet_bcs on a host CsMat
    xyz, conn, mesh = mesh_pair(oracle, 3, 3, 2, 1000.0, 1000.0, 600.0, 0.05, 9)
    mask, vals = oracle.thermal_dirichlet(xyz, 600.0)
    dof = G.DofManager(mesh.num_nodes(), 1)
    dof.set_dirichlet_many(np.nonzero(mask)[0], vals[mask == 1])
    Ko = oracle.assemble_thermal(xyz, conn, 3.0)
    fo = oracle.assemble_thermal_load(xyz, conn, 1e-6)
    Kbo, fbo = oracle.apply_dirichlet(Ko, fo, mask, vals)
    Kbc, fbc = G.Assembler.apply_dirichlet_bcs(Ko, fo, dof)
    assert same_csr(Kbc.to_scipy(), Kbo) and np.array_equal(fbc, fbo)


def test_jacobi_and_elastic_spmv_caller_numbering(oracle):
    xyz, conn, mesh = mesh_pair(oracle, 2, 3, 2, 1e6, 1e6, 6e5, 0.03, 10)
    dof = G.DofManager(mesh.num_nodes(), 3)
    K = G.Assembler.assemble_elasticity_stiffness_parallel(
        mesh, dof, G.IsotropicElasticity(1e11, 0.25))
    Ko = oracle.assemble_elastic(xyz, conn, 1e11, 0.25)
    d = G.JacobiPreconditioner(K).diag_inv.cpu().numpy()
    assert np.array_equal(d, oracle.jacobi_diag_inv(Ko))
    x = np.random.default_rng(2).standard_normal(K.shape[0])
    y = K.apply(x)
    assert np.max(np.abs(y - Ko @ x)) <= 1e-13 * np.max(np.abs(Ko) @ np.abs(x))


def test_assembly_deterministic():
    mesh = G.ImprovedMeshGenerator.generate_cube(8, 8, 8, 1000.0, 1000.0, 600.0)
    dof = G.DofManager(mesh.num_nodes(), 1)
    a = G.Assembler.assemble_thermal_stiffness_parallel(mesh, dof, 3.0).to_scipy()
    b = G.Assembler.assemble_thermal_stiffness_parallel(mesh, dof, 3.0).to_scipy()
    assert same_csr(a, b)


def test_empty_mesh():
    mesh = G.Mesh(np.zeros((4, 3)), np.zeros((0, 10), dtype=np.int64))
    dof = G.DofManager(4, 1)
    K = G.Assembler.assemble_thermal_stiffness_parallel(mesh, dof, 1.0)
    assert K.nnz() == 0 and K.shape == (4, 4)
    assert np.all(G.Assembler.assemble_thermal_load(mesh, dof, 1.0) == 0)


# ---- operator + solver ----------------------------------------------------------
def test_spmv_matches_scipy(oracle):
    xyz, conn, mesh = mesh_pair(oracle, 5, 5, 4, 1000.0, 1000.0, 600.0, 0.05, 8)
    dof = G.DofManager(mesh.num_nodes(), 1)
    K = G.Assembler.assemble_thermal_stiffness_parallel(mesh, dof, 3.0)
    Ks = K.to_scipy()
    x = np.random.default_rng(1).standard_normal(K.shape[0])
    y = K.apply(x)
    ref = Ks @ x
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(Ks) @ np.abs(x))


def test_cg_kat_2x2():
    A = sp.csr_matrix(np.array([[2.0, 1.0], [1.0, 2.0]]))
    x, st = G.ConjugateGradient().solve(A, np.array([3.0, 3.0]))
    assert np.allclose(x, [1.0, 1.0], atol=1e-6) and st.converged


def test_cg_zero_rhs_and_max_iter(oracle):
    A = sp.csr_matrix(np.diag([1.0, 2.0, 3.0, 4.0]))
    x, st = G.ConjugateGradient().solve(A, np.zeros(4))
    assert st.converged and st.iterations == 0 and np.all(x == 0)
    assert st.residual_norm == 0.0
    b = np.ones(4)
    _, st = G.ConjugateGradient().with_preconditioner(False).with_max_iterations(1).solve(A, b)
    _, so = oracle.cg_solve(A, b, max_iter=1, use_precond=False)
    assert (st.iterations, st.converged) == (so.iterations, so.converged) == (1, False)
    _, st = G.ConjugateGradient().with_max_iterations(0).solve(A, b)
    assert st.iterations == 0 and not st.converged and abs(st.relative_residual - 1) < 1e-15
    _, st = G.ConjugateGradient().solve(A, b)  # Jacobi: exact after one pass
    assert st.converged and st.iterations == 0


def thermal_problem(oracle, n=6):
    xyz, conn, mesh = mesh_pair(oracle, n, n, n, 1000.0, 1000.0, 600.0)
    mask, vals = oracle.thermal_dirichlet(xyz, 600.0)
    dof = G.DofManager(mesh.num_nodes(), 1)
    dof.set_dirichlet_many(np.nonzero(mask)[0], vals[mask == 1])
    return xyz, conn, mesh, dof, mask, vals


def test_cg_fp64_matches_oracle(oracle):
    xyz, conn, mesh, dof, mask, vals = thermal_problem(oracle)
    K, f = G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    Ko, fo = oracle.apply_dirichlet(oracle.assemble_thermal(xyz, conn, 3.0),
                                    oracle.assemble_thermal_load(xyz, conn, 1e-6), mask, vals)
    # the benchmark's settings: tol 1e-8, max 1000 (thermal_1d_improved_mesh.rs:62-64)
    x, st = G.ConjugateGradient().with_max_iterations(1000).with_tolerance(1e-8).solve(K, f)
    xo, so = oracle.cg_solve(Ko, fo, 1000, 1e-8)
    assert st.converged and so.converged
    assert abs(st.iterations - so.iterations) <= 2
    x = x.cpu().numpy()
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
    # tight: the solutions agree to 1e-10 and equal the closed form
    x, st = G.ConjugateGradient().with_max_iterations(5000).with_tolerance(1e-13) \
        .with_abs_tolerance(0.0).solve(K, f)
    xo, so = oracle.cg_solve(Ko, fo, 5000, 1e-13, 0.0)
    x = x.cpu().numpy()
    assert abs(st.iterations - so.iterations) <= 3
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-10
    ex = oracle.thermal_analytic(xyz[:, 2], 600.0, 1280.0, 0.0, 3.0, 1e-6)
    assert np.linalg.norm(x - ex) / np.linalg.norm(ex) < 1e-11


def test_cg_mixed_matches_oracle(oracle):
    xyz, conn, mesh, dof, mask, vals = thermal_problem(oracle)
    K, f = G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)
    cg = G.ConjugateGradient().with_precision("mixed").with_max_iterations(20000) \
        .with_tolerance(1e-13).with_abs_tolerance(0.0)
    x, st = cg.solve(K, f)
    assert st.converged and st.relative_residual < 1e-13 and st.outer_iterations >= 2
    x = x.cpu().numpy()
    ex = oracle.thermal_analytic(xyz[:, 2], 600.0, 1280.0, 0.0, 3.0, 1e-6)
    assert np.linalg.norm(x - ex) / np.linalg.norm(ex) < 1e-11
    # the reported residual is the true FP64 residual
    Ks = K.to_scipy()
    fh = f.cpu().numpy()
    assert abs(np.linalg.norm(fh - Ks @ x) / np.linalg.norm(fh) - st.relative_residual) \
        < 1e-3 * st.relative_residual + 1e-16


def test_elastic_solve_closed_form(oracle):
    lx = ly = 1e6
    lz = 6e5
    xyz, conn, mesh = mesh_pair(oracle, 3, 3, 4, lx, ly, lz)
    mask, vals = oracle.elastic_dirichlet(xyz, lx, ly)
    dof = G.DofManager(mesh.num_nodes(), 3)
    dof.set_dirichlet_many(np.nonzero(mask)[0], 0.0)
    mat = G.IsotropicElasticity(1e11, 0.25)
    K, f = G.Assembler.assemble_elasticity_system(mesh, dof, mat, 3300.0, [0, 0, -9.81])
    uz = oracle.elastic_analytic_uz(xyz[:, 2], lz, 3300.0, 9.81, 1e11, 0.25)
    for prec in ("fp64", "mixed"):
        cg = G.ConjugateGradient().with_precision(prec).with_max_iterations(50000) \
            .with_tolerance(1e-13).with_abs_tolerance(0.0)
        u, st = cg.solve(K, f)
        u = u.cpu().numpy()
        assert st.converged, prec
        assert np.linalg.norm(u[2::3] - uz) / np.linalg.norm(uz) < 1e-10, prec
        assert np.max(np.abs(u[0::3])) < 1e-8 * np.max(np.abs(uz)), prec


# ---- adjoint ---------------------------------------------------------------------
def test_element_backward_matches_oracle(oracle):
    xyz, conn, mesh = mesh_pair(oracle, 3, 3, 2, 1000.0, 1000.0, 600.0, 0.05, 12)
    rng = np.random.default_rng(13)
    k = np.exp(rng.normal(1.0, 0.3, len(conn)))
    u = rng.standard_normal(len(xyz))
    lam = rng.standard_normal(len(xyz))
    g = G.thermal_element_backward(mesh, k, u, lam).cpu().numpy()
    go = oracle.thermal_element_backward(xyz, conn, k, u, lam)
    assert np.max(np.abs(g - go)) <= 1e-12 * np.max(np.abs(go))
    E = 1e11 * np.exp(rng.normal(0, 0.5, len(conn)))
    u3 = rng.standard_normal(3 * len(xyz))
    l3 = rng.standard_normal(3 * len(xyz))
    g = G.elastic_element_backward(mesh, E, 0.25, u3, l3).cpu().numpy()
    go = oracle.elastic_element_backward(xyz, conn, E, 0.25, u3, l3)
    assert np.max(np.abs(g - go)) <= 1e-12 * np.max(np.abs(go))


def test_thermal_autograd_vs_fd_and_oracle(oracle):
    import torch
    from geo_simulator_amd.adjoint import ThermalProblem, thermal_solve
    xyz, conn, mesh, dof, mask, vals = thermal_problem(oracle, 3)
    rng = np.random.default_rng(21)
    logk = torch.tensor(np.log(3.0) + 0.3 * rng.standard_normal(len(conn)),
                        device="cuda", requires_grad=True)
    free = np.nonzero(mask == 0)[0]
    obs = torch.tensor(rng.choice(free, 10, replace=False), device="cuda")
    d = torch.tensor(rng.uniform(300, 900, 10), device="cuda")
    prob = ThermalProblem(mesh, dof, 1e-6)

    def loss(lk):
        T = thermal_solve(lk, prob)
        return 0.5 * ((T[obs] - d) ** 2).sum()

    L = loss(logk)
    L.backward()
    grad = logk.grad.cpu().numpy()
    h = 1e-6
    with torch.no_grad():
        for e in rng.choice(len(conn), 6, replace=False):
            lp = logk.detach().clone()
            lm = logk.detach().clone()
            lp[e] += h
            lm[e] -= h
            fd = (loss(lp).item() - loss(lm).item()) / (2 * h)
            assert abs(fd - grad[e]) <= 1e-5 * max(abs(fd), 1e-3 * np.abs(grad).max())


def test_reverse_solve_independent_of_loss_units(oracle):
    """A loss scaled by 1e-40 gives the gradient scaled by 1e-40: the reverse
    solve runs on the unit-norm right-hand side, so the FP64 path's absolute
    breakdown threshold (|p.Ap| < 1e-30) does not stop it early."""
    import torch
    from geo_simulator_amd.adjoint import ThermalProblem, thermal_solve
    xyz, conn, mesh, dof, mask, vals = thermal_problem(oracle, 3)
    rng = np.random.default_rng(23)
    free = np.nonzero(mask == 0)[0]
    obs = torch.tensor(rng.choice(free, 10, replace=False), device="cuda")
    d = torch.tensor(rng.uniform(300, 900, 10), device="cuda")
    prob = ThermalProblem(mesh, dof, 1e-6)
    grads = []
    for scale in (1.0, 1e-40):
        lk = torch.full((len(conn),), np.log(3.0), dtype=torch.float64, device="cuda",
                        requires_grad=True)
        T = thermal_solve(lk, prob)
        (scale * 0.5 * ((T[obs] - d) ** 2).sum()).backward()
        assert prob.backward_stats.converged
        grads.append(lk.grad.cpu().numpy() / scale)
    assert np.max(np.abs(grads[1] - grads[0])) <= 1e-12 * np.max(np.abs(grads[0]))


def test_elastic_autograd_vs_oracle_adjoint(oracle):
    import torch
    from geo_simulator_amd.adjoint import ElasticProblem, elastic_solve
    lx = ly = 1e6
    lz = 6e5
    xyz, conn, mesh = mesh_pair(oracle, 2, 2, 2, lx, ly, lz)
    mask, vals = oracle.elastic_dirichlet(xyz, lx, ly)
    dof = G.DofManager(mesh.num_nodes(), 3)
    dof.set_dirichlet_many(np.nonzero(mask)[0], 0.0)
    rng = np.random.default_rng(31)
    logE = np.log(1e11) + 0.5 * rng.standard_normal(len(conn))
    top = np.nonzero(np.abs(xyz[:, 2] - lz) < 1.0)[0]
    obs = 3 * top + 2
    d = rng.normal(-4e4, 5e3, size=len(obs))
    prob = ElasticProblem(mesh, dof, 0.25, 3300.0, [0, 0, -9.81])
    lE = torch.tensor(logE, device="cuda", requires_grad=True)
    u = elastic_solve(lE, prob)
    L = 0.5 * ((u[torch.tensor(obs, device="cuda")] - torch.tensor(d, device="cuda")) ** 2).sum()
    L.backward()
    # oracle adjoint
    Ko = oracle.assemble_elastic(xyz, conn, 0.0, 0.25, E_elem=np.exp(logE))
    fo = oracle.assemble_gravity_load(xyz, conn, 3300.0, [0, 0, -9.81])
    Kb, fb = oracle.apply_dirichlet(Ko, fo, mask, vals)
    uo, _ = oracle.cg_solve(Kb, fb, 100000, 1e-15, 0.0)
    g = np.zeros(3 * len(xyz))
    g[obs] = uo[obs] - d
    lam, _ = oracle.cg_solve(Kb, g, 100000, 1e-15, 0.0)
    go = oracle.elastic_element_backward(xyz, conn, np.exp(logE), 0.25, uo, lam)
    gr = lE.grad.cpu().numpy()
    assert np.max(np.abs(gr - go)) <= 1e-8 * np.max(np.abs(go))


# ---- failure handling (reference panics -> statuses) --------------------------------
def test_singular_element_reported():
    xyz = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [0, 0, 1]], float)
    edges = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
    nodes = np.vstack([xyz] + [(xyz[a] + xyz[b]) / 2 for a, b in edges])
    mesh = G.Mesh(nodes, np.arange(10)[None, :])
    dof = G.DofManager(10, 1)
    with pytest.raises(G.GeoFemError) as ei:
        G.Assembler.assemble_thermal_stiffness_parallel(mesh, dof, 1.0)
    assert ei.value.code == G._lib.GS_ERR_SINGULAR_JACOBIAN


def test_bad_dofs_and_material():
    mesh = G.ImprovedMeshGenerator.generate_cube(1, 1, 1, 1.0, 1.0, 1.0)
    with pytest.raises(G.GeoFemError) as ei:
        G.Assembler.assemble_elasticity_stiffness_parallel(
            mesh, G.DofManager(mesh.num_nodes(), 1), G.IsotropicElasticity(1e11, 0.25))
    assert ei.value.code == G._lib.GS_ERR_BAD_DOFS
    with pytest.raises(G.GeoFemError) as ei:
        G.Assembler.assemble_elasticity_stiffness_parallel(
            mesh, G.DofManager(mesh.num_nodes(), 3), G.IsotropicElasticity(1.0, 0.25),
            E_per_elem=np.array([1.0, 1.0, -1.0, 1.0, 1.0, 1.0]))
    assert ei.value.code == G._lib.GS_ERR_INVALID_MATERIAL


@pytest.mark.parametrize("case", ["thermal", "elastic"])
def test_fused_system_max_diagonal_at_constrained_node(oracle, case):
    """c = max(1, max_i |K_ii|) runs over every diagonal of K, also those of
    constrained rows (src/fem/assembly.rs:184-193): the node holding the
    largest diagonal is constrained (all three DOFs in the elastic case, so
    no entry of its diagonal block is kept in K_bc)."""
    xyz, conn, mesh = mesh_pair(oracle, 2, 3, 2, 1e6, 1e6, 6e5)
    p = np.exp(np.random.default_rng(8).normal(0.0, 0.5, len(conn)))
    if case == "elastic":
        mask, vals = oracle.elastic_dirichlet(xyz, 1e6, 1e6)
        dof = G.DofManager(mesh.num_nodes(), 3)
        Ko = oracle.assemble_elastic(xyz, conn, 1.0, 0.25, E_elem=p * 1e11)
        fo = oracle.assemble_gravity_load(xyz, conn, 3300.0, [0, 0, -9.81])
        node = int(np.argmax(np.abs(Ko.diagonal()))) // 3
        mask[3 * node:3 * node + 3] = 1
        vals[3 * node:3 * node + 3] = [1e-3, -2e-3, 3e-3]
    else:
        mask, vals = oracle.thermal_dirichlet(xyz, 6e5)
        dof = G.DofManager(mesh.num_nodes(), 1)
        Ko = oracle.assemble_thermal(xyz, conn, 1.0, k_elem=p * 3.0)
        fo = oracle.assemble_thermal_load(xyz, conn, 1e-6)
        node = int(np.argmax(np.abs(Ko.diagonal())))
        mask[node] = 1
        vals[node] = 640.0
    assert xyz[node, 2] > 1.0 and xyz[node, 2] < 6e5 - 1.0  # an interior node
    dof.set_dirichlet_many(np.nonzero(mask)[0], vals[mask == 1])
    Kbo, fbo = oracle.apply_dirichlet(Ko, fo, mask, vals)
    if case == "elastic":
        Kf, ff = G.Assembler.assemble_elasticity_system(
            mesh, dof, G.IsotropicElasticity(1.0, 0.25), 3300.0, [0, 0, -9.81],
            E_per_elem=p * 1e11)
    else:
        Kf, ff = G.Assembler.assemble_thermal_system(mesh, dof, 1.0, 1e-6, k_per_elem=p * 3.0)
    assert same_csr(Kf.to_scipy(), Kbo)
    assert np.array_equal(ff.cpu().numpy(), fbo)


def star_mesh(n_tets):
    """n_tets Tet10 elements sharing only vertex 0 (rotated copies of one
    tet), so node 0 couples to 9 n_tets + 1 nodes: rows far longer than a
    structured mesh's 65."""
    base = np.array([[0, 0, 0], [1.0, 0.1, 0.0], [0.2, 1.1, 0.1], [0.1, 0.2, 0.9]])
    xyz, conn = [np.zeros(3)], []
    for k in range(n_tets):
        th = 2 * np.pi * k / n_tets
        ph = 0.3 * np.sin(3 * th)
        rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
        rx = np.array([[1, 0, 0], [0, np.cos(ph), -np.sin(ph)], [0, np.sin(ph), np.cos(ph)]])
        v = base @ (rz @ rx).T * (1.0 + 0.01 * k)
        ids = [0]
        for p in v[1:]:
            ids.append(len(xyz))
            xyz.append(p)
        for a, b in ((0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)):
            ids.append(len(xyz))
            xyz.append(0.5 * (v[a] + v[b]))
        conn.append(ids)
    return np.array(xyz), np.array(conn, dtype=np.int64)


@pytest.mark.parametrize("n_tets", [20, 30])
def test_thermal_system_long_rows(oracle, n_tets):
    """Rows longer than one accumulator pass of the row gather (20 tets: 181
    entries, 1-byte codes) and longer than 255 (30 tets: 271 entries,
    2-byte codes), with lift entries in the long row: the fused system
    equals the oracle's assembly + elimination bit for bit."""
    xyz, conn = star_mesh(n_tets)
    mesh = G.Mesh(xyz, conn)
    k = np.exp(np.random.default_rng(n_tets).normal(0.0, 0.5, len(conn)))
    mask = np.zeros(len(xyz), dtype=np.uint8)
    mask[conn[::3, 1]] = 1   # a vertex of every third tet
    mask[conn[1::4, 7]] = 1  # an edge node of every fourth
    vals = np.where(mask == 1, 1.0 + xyz[:, 2], 0.0)
    dof = G.DofManager(mesh.num_nodes(), 1)
    dof.set_dirichlet_many(np.nonzero(mask)[0], vals[mask == 1])
    Ko = oracle.assemble_thermal(xyz, conn, 1.0, k_elem=k)
    assert np.diff(Ko.indptr).max() == 9 * n_tets + 1
    fo = oracle.assemble_thermal_load(xyz, conn, 2.5)
    Kbo, fbo = oracle.apply_dirichlet(Ko, fo, mask, vals)
    Kf, ff = G.Assembler.assemble_thermal_system(mesh, dof, 1.0, 2.5, k_per_elem=k)
    assert same_csr(Kf.to_scipy(), Kbo)
    assert np.array_equal(ff.cpu().numpy(), fbo)
    # the row gather's SELL values also feed the solver: one CG solve
    x, st = G.ConjugateGradient().with_tolerance(1e-12).with_max_iterations(2000).solve(Kf, ff)
    assert st.converged
    x = x.cpu().numpy() if hasattr(x, "cpu") else np.asarray(x)
    xo = np.linalg.solve(Kbo.toarray(), fbo)
    assert np.linalg.norm(x - xo) <= 1e-9 * np.linalg.norm(xo)


def test_thermal_system_subnormal_values(oracle):
    """A conductivity of 1e-312 leaves K's entries subnormal (about 1e-310):
    the element kernel, the row gather's LDS sums and the elimination keep
    them bit for bit."""
    xyz, conn, mesh = mesh_pair(oracle, 3, 3, 2, 1000.0, 1000.0, 600.0, 0.05, 3)
    mask, vals = oracle.thermal_dirichlet(xyz, 600.0)
    dof = G.DofManager(mesh.num_nodes(), 1)
    dof.set_dirichlet_many(np.nonzero(mask)[0], vals[mask == 1])
    Ko = oracle.assemble_thermal(xyz, conn, 1e-312)
    assert 0 < np.abs(Ko.data[Ko.data != 0]).max() < 2.2e-308  # all subnormal
    fo = oracle.assemble_thermal_load(xyz, conn, 1e-315)
    Kbo, fbo = oracle.apply_dirichlet(Ko, fo, mask, vals)
    Kf, ff = G.Assembler.assemble_thermal_system(mesh, dof, 1e-312, 1e-315)
    assert same_csr(Kf.to_scipy(), Kbo)
    assert np.array_equal(ff.cpu().numpy(), fbo)


@pytest.mark.parametrize("n_points", [1, 4, 5, 11])
def test_thermal_element_matrices_every_rule(oracle, n_points):
    """ElementMatrix::thermal_stiffness (src/fem/element.rs:23-61) with each
    GaussQuadrature rule on curved elements and per-element k: all 100 entries
    of every element bit-identical to the oracle's element routine."""
    xyz, conn, mesh = mesh_pair(oracle, 3, 2, 2, 1000.0, 800.0, 600.0, 0.08, 11)
    k = np.exp(np.random.default_rng(n_points).normal(1.0, 0.3, len(conn)))
    Ke = G.ElementMatrix.thermal_stiffness(mesh, 0.0, n_points, k_per_elem=k).cpu().numpy()
    for e in range(len(conn)):
        assert np.array_equal(Ke[e], oracle.thermal_stiffness(xyz[conn[e]], k[e], n_points)), e


def test_thermal_element_rule_invariance_and_4pt_assembly(oracle):
    """On affine elements the integrand is quadratic: the 4-, 5- and 11-point
    rules agree to rounding (the 1-point rule does not); the 4-point matrices
    are the assembly's (their scatter equals the assembled K)."""
    xyz, conn, mesh = mesh_pair(oracle, 2, 2, 2, 1000.0, 1000.0, 600.0)
    K = {n: G.ElementMatrix.thermal_stiffness(mesh, 3.0, G.GaussQuadrature(n)).cpu().numpy()
         for n in (1, 4, 5, 11)}
    scale = np.abs(K[4]).max()
    assert np.abs(K[5] - K[4]).max() < 1e-13 * scale
    assert np.abs(K[11] - K[4]).max() < 1e-13 * scale
    assert np.abs(K[1] - K[4]).max() > 1e-3 * scale
    rows = np.repeat(conn, 10, axis=1).ravel()
    cols = np.tile(conn, (1, 10)).ravel()
    S = sp.csr_matrix((K[4].ravel(), (rows, cols)), shape=(len(xyz), len(xyz)))
    dof = G.DofManager(mesh.num_nodes(), 1)
    A = G.Assembler.assemble_thermal_stiffness_parallel(mesh, dof, 3.0).to_scipy()
    assert np.abs((S - A).toarray()).max() <= 1e-12 * scale


def test_thermal_element_matrices_singular_and_bad_rule(oracle):
    xyz, conn = star_mesh(2)
    xyz = xyz.copy()
    xyz[conn[1, 1:]] = xyz[conn[1, 0]]  # element 1 collapsed to a point: J = 0
    mesh = G.Mesh(xyz, conn)
    with pytest.raises(G.GeoFemError):
        G.ElementMatrix.thermal_stiffness(mesh, 1.0, 4)
    with pytest.raises(G.GeoFemError):
        G.ElementMatrix.thermal_stiffness(mesh, 1.0, 3)
