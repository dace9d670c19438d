"""Parity at full size: the 64x64x32 thermal box (1,081,665 DOFs, 29,420,281
nonzeros after the Dirichlet elimination), the workload the benchmark times.

The fused GPU system (assembly, load and the scaled elimination of
src/fem/assembly.rs:69-230) must equal the oracle's bit for bit, and the
FP64 Jacobi-PCG (src/linalg/iterative.rs:46-131) must take the oracle CG's
iteration count (within the rounding of parallel dot products) to the
reference's tolerance and reach the same solution; both solutions, and the
mixed-precision one, are compared with the closed form
(benchmarks/thermal_1d_improved_mesh.rs:9-13), which the quadratic elements
reproduce at the nodes, so their distance from it is the solver's error.
"""
import numpy as np
import pytest

import geo_simulator_amd as G

pytestmark = pytest.mark.gpu

NX, NY, NZ = 64, 64, 32
LX, LY, LZ = 1000.0, 1000.0, 600.0
K_COND, Q = 3.0, 1e-6


@pytest.fixture(scope="module")
def c2(oracle):
    xyz, conn = oracle.generate_cube(NX, NY, NZ, LX, LY, LZ)
    mask, vals = oracle.thermal_dirichlet(xyz, LZ)
    mesh = G.Mesh(xyz, conn)
    dof = G.DofManager(mesh.num_nodes(), 1)
    dof.set_dirichlet_many(np.nonzero(mask)[0], vals[mask == 1])
    Ko, fo = oracle.apply_dirichlet(oracle.assemble_thermal(xyz, conn, K_COND),
                                    oracle.assemble_thermal_load(xyz, conn, Q), mask, vals)
    K, f = G.Assembler.assemble_thermal_system(mesh, dof, K_COND, Q)
    exact = oracle.thermal_analytic(xyz[:, 2], LZ, 1280.0, 0.0, K_COND, Q)
    return Ko, fo, K, f, exact


def test_fullsize_thermal_system_bitexact(c2):
    Ko, fo, K, f, _ = c2
    assert Ko.shape == (1081665, 1081665) and Ko.nnz == 29420281
    A = K.to_scipy()
    assert np.array_equal(A.indptr, Ko.indptr)
    assert np.array_equal(A.indices, Ko.indices)
    assert np.array_equal(A.data, Ko.data)
    assert np.array_equal(f.cpu().numpy(), np.asarray(fo))


def test_fullsize_thermal_cg_matches_oracle(oracle, c2):
    Ko, fo, K, f, exact = c2
    # the benchmark's settings: tol 1e-8 (thermal_1d_improved_mesh.rs:62-64)
    xo, so = oracle.cg_solve(Ko, fo, 5000, 1e-8)
    x, st = G.ConjugateGradient().with_max_iterations(5000).with_tolerance(1e-8).solve(K, f)
    assert st.converged and so.converged
    assert abs(st.iterations - so.iterations) <= 2, (st.iterations, so.iterations)
    x = x.cpu().numpy()
    err_o = np.linalg.norm(xo - exact) / np.linalg.norm(exact)
    err = np.linalg.norm(x - exact) / np.linalg.norm(exact)
    # both stop at rel-res 1e-8 on the same recursion: the same iterate when
    # the counts agree; otherwise within one pass's error decay (about 7%)
    assert err < 2e-6 and err_o < 2e-6
    if st.iterations == so.iterations:
        assert np.linalg.norm(x - xo) / np.linalg.norm(xo) < 1e-8
    else:
        assert abs(err - err_o) <= 0.25 * err_o
    # mixed precision reaches the same accuracy to the same true residual
    xm, sm = (G.ConjugateGradient().with_precision("mixed").with_max_iterations(5000)
              .with_tolerance(1e-8).solve(K, f))
    assert sm.converged and sm.relative_residual < 1e-8
    xm = xm.cpu().numpy()
    err_m = np.linalg.norm(xm - exact) / np.linalg.norm(exact)
    assert err_m < 2e-6


def test_fullsize_thermal_fp64_bar(oracle, c2):
    """The plain FP64 path's accuracy bar at full size: the GPU FP64 CG and
    the oracle's serial CG (src/linalg/iterative.rs:46-131), both to rel-res
    1e-13 with abs_tol 0, give solutions within 1e-10 relative L2."""
    Ko, fo, K, f, exact = c2
    xo, so = oracle.cg_solve(Ko, fo, 20000, 1e-13, 0.0)
    x, st = (G.ConjugateGradient().with_max_iterations(20000).with_tolerance(1e-13)
             .with_abs_tolerance(0.0).solve(K, f))
    assert st.converged and so.converged
    x = x.cpu().numpy()
    rel = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    print(f"[fullsize] FP64 CG vs oracle CG at 1e-13: rel L2 {rel:.2e} "
          f"({st.iterations} / {so.iterations} iterations)", flush=True)
    assert rel <= 1e-10, rel
    assert abs(st.iterations - so.iterations) <= 5


def test_fullsize_elastic_system_bitexact_and_closed_form(oracle):
    """The 48x48x24 elastic column in metres (1,383,123 DOFs, 111,247,603
    nonzeros after the elimination; src/fem/assembly.rs:389-437, 906-937,
    168-230): the fused system equals the oracle's bit for bit, and both
    precisions reach the closed-form u_z of
    benchmarks/elasticity_halfspace_gravity.rs:23-27 (exact at the nodes)
    with u_x = u_y = 0.  The oracle's serial CG is not run at this size (about
    two minutes); its parity with the GPU CG is pinned on the thermal box."""
    lx, ly, lz = 1e6, 1e6, 6e5
    E, nu, rho, g = 1e11, 0.25, 3300.0, 9.81
    xyz, conn = oracle.generate_cube(48, 48, 24, lx, ly, lz)
    mask, vals = oracle.elastic_dirichlet(xyz, lx, ly)
    Ko, fo = oracle.apply_dirichlet(oracle.assemble_elastic(xyz, conn, E, nu),
                                    oracle.assemble_gravity_load(xyz, conn, rho, [0.0, 0.0, -g]),
                                    mask, vals)
    mesh = G.Mesh(xyz, conn)
    dof = G.DofManager(mesh.num_nodes(), 3)
    dof.set_dirichlet_many(np.nonzero(mask)[0], vals[mask == 1])
    K, f = G.Assembler.assemble_elasticity_system(mesh, dof, G.IsotropicElasticity(E, nu),
                                                  rho, [0.0, 0.0, -g])
    A = K.to_scipy()
    assert A.shape == (1383123, 1383123) and A.nnz == 111247603
    assert np.array_equal(A.indptr, Ko.indptr)
    assert np.array_equal(A.indices, Ko.indices)
    assert np.array_equal(A.data, Ko.data)
    assert np.array_equal(f.cpu().numpy(), np.asarray(fo))
    del A, Ko, fo
    uz = oracle.elastic_analytic_uz(xyz[:, 2], lz, rho, g, E, nu)
    for prec in ("fp64", "mixed"):
        u, st = (G.ConjugateGradient().with_precision(prec).with_max_iterations(20000)
                 .with_tolerance(1e-8).solve(K, f))
        assert st.converged, prec
        u = u.cpu().numpy().reshape(-1, 3)
        assert np.linalg.norm(u[:, 2] - uz) / np.linalg.norm(uz) < 1e-8, prec
        assert np.abs(u[:, :2]).max() < 1e-8 * np.abs(uz).max(), prec
