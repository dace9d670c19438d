"""gs_cg_solve_noalloc: the FP64 CG on the library's fixed device scratch.

It must equal gs_cg_solve bit for bit (the same kernels in the same order),
on two calls in a row with the second system smaller than the first, and
refuse what it cannot do without allocating."""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import geo_simulator_amd as G
from geo_simulator_amd import _lib
from geo_simulator_amd._lib import lib
from geo_simulator_amd.runtime import dptr

pytestmark = pytest.mark.gpu


def thermal_system(n):
    mesh = G.ImprovedMeshGenerator.generate_cube(n, n, n, 1000.0, 1000.0, 600.0)
    z = mesh.nodes[:, 2]
    dof = G.DofManager(mesh.num_nodes(), 1)
    for v, at in ((1280.0, 0.0), (0.0, 600.0)):
        ids = np.nonzero(np.abs(z - at) < 1.0)[0]
        dof.set_dirichlet_many(ids, np.full(len(ids), v))
    return G.Assembler.assemble_thermal_system(mesh, dof, 3.0, 1e-6)


def run(entry, K, f, tol=1e-10):
    ctx = K.ctx
    o = G.ConjugateGradient().with_tolerance(tol).with_max_iterations(5000).opts()
    x = ctx.empty(K.rows())
    st = _lib.SolverStatsC()
    ctx.sync_torch()
    status = getattr(lib(), entry)(ctx.handle, K.handle, dptr(f), C.byref(o), dptr(x),
                                   C.byref(st))
    return status, x.cpu().numpy(), st


@pytest.mark.parametrize("sizes", [(10, 4)])
def test_noalloc_equals_cg_solve(sizes):
    for n in sizes:  # the second system is smaller than the first
        print(f"[no_alloc] thermal {n}^3", flush=True)
        K, f = thermal_system(n)
        for _ in range(2):  # two calls in a row on the shared scratch
            s0, x0, st0 = run("gs_cg_solve", K, f)
            s1, x1, st1 = run("gs_cg_solve_noalloc", K, f)
            assert s0 == 0 and s1 == 0, K.ctx.check(s1, "gs_cg_solve_noalloc")
            assert st1.converged and st1.iterations == st0.iterations > 0
            assert st1.residual_norm == st0.residual_norm
            assert st1.spmv_count == st0.spmv_count
            assert np.array_equal(x1, x0)


def test_noalloc_refusals():
    K, f = thermal_system(3)
    # mixed precision needs the library's allocations
    ctx = K.ctx
    o = G.ConjugateGradient().with_precision("mixed").opts()
    x = ctx.empty(K.rows())
    st = _lib.SolverStatsC()
    assert lib().gs_cg_solve_noalloc(ctx.handle, K.handle, dptr(f), C.byref(o), dptr(x),
                                     C.byref(st)) == _lib.GS_ERR_INVALID_ARG
    # a matrix whose solver arrays do not exist yet
    A = sp.diags(np.arange(1.0, 101.0)).tocsr()
    M = G.CsrMatrix.from_scipy(A)
    b = ctx.to_device(np.ones(100))
    s, _, _ = run("gs_cg_solve_noalloc", M, b)
    assert s == _lib.GS_ERR_INVALID_ARG
    assert b"gs_cg_solve" in lib().gs_last_error(ctx.handle)
    # once solved by gs_cg_solve, it is prepared
    s0, x0, _ = run("gs_cg_solve", M, b)
    s1, x1, _ = run("gs_cg_solve_noalloc", M, b)
    assert s0 == 0 and s1 == 0 and np.array_equal(x0, x1)
    assert np.allclose(x1, 1.0 / np.arange(1.0, 101.0), rtol=1e-12)
    # beyond the scratch's 2^21 rows
    n = (1 << 21) + 1
    big = G.CsrMatrix.from_scipy(sp.identity(n, format="csr"))
    bb = ctx.to_device(np.ones(n))
    s, _, _ = run("gs_cg_solve_noalloc", big, bb)
    assert s == _lib.GS_ERR_TOO_LARGE
