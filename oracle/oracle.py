"""ctypes front-end of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker.  The product package
(geo_simulator_amd) never imports this module.

Each wrapper names the reference function it stands for; the C file cites the
reference file:line of every statement block.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

OR_OK = 0
OR_ERR_INVALID_ARG = 1
OR_ERR_SINGULAR_JACOBIAN = 3
OR_ERR_BAD_DOFS = 4
OR_ERR_INVALID_MATERIAL = 5
OR_ERR_ZERO_PIVOT = 9


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: oracle status {code}")
        self.code = code


class _Csr(C.Structure):
    _fields_ = [("rows", C.c_int64), ("cols", C.c_int64), ("nnz", C.c_int64),
                ("indptr", C.POINTER(C.c_int64)),
                ("indices", C.POINTER(C.c_int64)),
                ("data", C.POINTER(C.c_double))]


class _Amg(C.Structure):
    _fields_ = [("nlevels", C.c_int64), ("A", C.POINTER(_Csr)), ("P", C.POINTER(_Csr)),
                ("PT", C.POINTER(_Csr)), ("dinv", C.POINTER(C.POINTER(C.c_double)))]


class _Stats(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("residual_norm", C.c_double),
                ("relative_residual", C.c_double), ("converged", C.c_int32),
                ("pad_", C.c_int32), ("solve_time", C.c_double)]


@dataclass
class SolverStats:
    iterations: int
    residual_norm: float
    relative_residual: float
    converged: bool
    solve_time: float


def build() -> None:
    """Compile oracle/liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)


def _declare(L):
    L.or_quad_rule.argtypes = [C.c_int, C.POINTER(_dp), C.POINTER(_dp)]
    L.or_shape_functions.argtypes = [_dp, _dp]
    L.or_shape_derivs_bary.argtypes = [_dp, _dp]
    L.or_jacobian.argtypes = [_dp, _dp, _dp]
    L.or_det3.argtypes = [_dp]
    L.or_det3.restype = C.c_double
    L.or_inv3.argtypes = [_dp, _dp]
    L.or_shape_derivs_cart.argtypes = [_dp, _dp, _dp]
    L.or_element_volume.argtypes = [_dp]
    L.or_element_volume.restype = C.c_double
    L.or_thermal_stiffness.argtypes = [_dp, C.c_double, _dp]
    L.or_thermal_stiffness_rule.argtypes = [_dp, C.c_double, C.c_int, _dp]
    L.or_thermal_load.argtypes = [_dp, C.c_double, _dp]
    L.or_constitutive.argtypes = [C.c_double, C.c_double, _dp]
    L.or_b_matrix.argtypes = [_dp, _dp]
    L.or_elastic_stiffness.argtypes = [_dp, C.c_double, C.c_double, _dp]
    L.or_gravity_load.argtypes = [_dp, C.c_double, _dp, _dp]
    L.or_cube_sizes.argtypes = [C.c_int64] * 3 + [_i64p, _i64p]
    L.or_generate_cube.argtypes = [C.c_int64] * 3 + [C.c_double] * 3 + [_dp, _i64p]
    L.or_csr_free.argtypes = [C.POINTER(_Csr)]
    L.or_csr_from_triplets.argtypes = [C.c_int64, C.c_int64, C.c_int64, _i64p,
                                       _i64p, _dp, C.POINTER(_Csr)]
    L.or_assemble_thermal.argtypes = [C.c_int64, _dp, C.c_int64, _i64p,
                                      C.c_double, _dp, C.c_int, C.POINTER(_Csr)]
    L.or_assemble_elastic.argtypes = [C.c_int64, _dp, C.c_int64, _i64p,
                                      C.c_double, C.c_double, _dp, C.c_int,
                                      C.POINTER(_Csr)]
    L.or_viscous_constitutive.argtypes = [C.c_double, C.c_double, _dp]
    L.or_viscosity_stiffness.argtypes = [_dp, C.c_double, C.c_double, _dp]
    L.or_maxwell_element.argtypes = [_dp, C.c_double, C.c_double, _dp, _dp, _dp]
    L.or_maxwell_stress.argtypes = [_dp, C.c_double, C.c_double, C.c_double, C.c_double,
                                    _dp, _dp, _dp]
    L.or_assemble_viscosity.argtypes = [C.c_int64, _dp, C.c_int64, _i64p, C.c_double,
                                        C.c_double, _dp, _dp, C.c_int, C.POINTER(_Csr)]
    L.or_assemble_maxwell.argtypes = [C.c_int64, _dp, C.c_int64, _i64p, C.c_double,
                                      C.c_double, C.c_double, C.c_double, _dp, C.c_int,
                                      C.POINTER(_Csr), _dp]
    L.or_update_stresses_maxwell.argtypes = [C.c_int64, _dp, C.c_int64, _i64p, C.c_double,
                                             C.c_double, C.c_double, C.c_double, _dp, _dp,
                                             _dp, _dp]
    L.or_assemble_thermal_load.argtypes = [C.c_int64, _dp, C.c_int64, _i64p,
                                           C.c_double, _dp]
    L.or_face_nodes.argtypes = [C.c_int, C.POINTER(C.c_int)]
    L.or_face_area.argtypes = [_dp]
    L.or_face_area.restype = C.c_double
    L.or_face_flux.argtypes = [_dp, C.c_double, _dp]
    L.or_face_flux.restype = None
    L.or_apply_neumann.argtypes = [C.c_int64, _dp, C.c_int64, _i64p, C.c_int64, _i64p,
                                   C.POINTER(C.c_int32), _dp, _dp]
    L.or_thermal_mass.argtypes = [_dp, C.c_double, C.c_double, _dp]
    L.or_thermal_mass.restype = None
    L.or_assemble_thermal_mass.argtypes = [C.c_int64, _dp, C.c_int64, _i64p, C.c_double,
                                           C.c_double, _dp, C.c_int, C.POINTER(_Csr)]
    L.or_assemble_gravity_load.argtypes = [C.c_int64, _dp, C.c_int64, _i64p,
                                           C.c_double, _dp, _dp]
    L.or_apply_dirichlet.argtypes = [C.POINTER(_Csr), _dp, _u8p, _dp,
                                     C.POINTER(_Csr), _dp]
    L.or_spmv.argtypes = [C.POINTER(_Csr), _dp, _dp]
    L.or_jacobi_diag_inv.argtypes = [C.POINTER(_Csr), _dp]
    L.or_cg_solve.argtypes = [C.POINTER(_Csr), _dp, C.c_int64, C.c_double,
                              C.c_double, C.c_int, _dp, C.POINTER(_Stats)]
    L.or_cg_solve_omp.argtypes = [C.POINTER(_Csr), _dp, C.c_int64, C.c_double,
                                  C.c_double, C.c_int, _dp, C.POINTER(_Stats)]
    L.or_bicgstab_solve.argtypes = [C.POINTER(_Csr), _dp, C.c_int64, C.c_double,
                                    C.c_double, C.c_int, _dp, C.POINTER(_Stats)]
    L.or_gmres_solve.argtypes = [C.POINTER(_Csr), _dp, C.c_int64, C.c_int64, C.c_double,
                                 C.c_double, C.c_int, _dp, C.POINTER(_Stats)]
    L.or_ilu0_factor.argtypes = [C.POINTER(_Csr), _dp, _i64p, _i64p, C.POINTER(C.c_int32)]
    L.or_ilu0_apply.argtypes = [C.POINTER(_Csr), _dp, _i64p, _dp, _dp]
    L.or_amg_strength.argtypes = [C.POINTER(_Csr), C.c_double, C.POINTER(_Csr)]
    L.or_amg_coarsen.argtypes = [C.POINTER(_Csr), C.c_double, _u8p, C.POINTER(C.c_int64)]
    L.or_amg_prolongation.argtypes = [C.POINTER(_Csr), _u8p, C.c_int64, C.c_double,
                                      C.POINTER(_Csr)]
    L.or_csr_matmul.argtypes = [C.POINTER(_Csr), C.POINTER(_Csr), C.POINTER(_Csr)]
    L.or_csr_transpose.argtypes = [C.POINTER(_Csr), C.POINTER(_Csr)]
    L.or_amg_diag_inv.argtypes = [C.POINTER(_Csr), _dp]
    L.or_amg_create.argtypes = [C.POINTER(_Csr), C.c_int64, C.c_int64, C.c_double,
                                C.POINTER(_Amg)]
    L.or_amg_apply.argtypes = [C.POINTER(_Amg), _dp, _dp]
    L.or_amg_free.argtypes = [C.POINTER(_Amg)]
    L.or_thermal_element_backward.argtypes = [C.c_int64, _dp, C.c_int64, _i64p,
                                              _dp, _dp, _dp, _dp]
    L.or_elastic_element_backward.argtypes = [C.c_int64, _dp, C.c_int64, _i64p,
                                              _dp, C.c_double, _dp, _dp, _dp]


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_i64p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _check(code, what):
    if code != OR_OK:
        raise OracleError(code, what)


# ---- quadrature / basis -------------------------------------------------

def quad_rule(npts: int):
    """GaussQuadrature::tet_{1,4,5,11}point (src/fem/quadrature.rs:13-104)."""
    p, w = _dp(), _dp()
    _check(lib().or_quad_rule(npts, C.byref(p), C.byref(w)), "quad_rule")
    pts = np.ctypeslib.as_array(p, shape=(npts, 4)).copy()
    wts = np.ctypeslib.as_array(w, shape=(npts,)).copy()
    return pts, wts


def shape_functions(L):
    out = np.zeros(10)
    lib().or_shape_functions(_d(_f64(L)), _d(out))
    return out


def shape_derivs_bary(L):
    out = np.zeros((10, 4))
    lib().or_shape_derivs_bary(_d(_f64(L)), _d(out))
    return out


def jacobian(L, nodes):
    out = np.zeros((3, 3))
    lib().or_jacobian(_d(_f64(L)), _d(_f64(nodes, (10, 3))), _d(out))
    return out


def shape_derivs_cart(L, nodes):
    out = np.zeros((10, 3))
    _check(lib().or_shape_derivs_cart(_d(_f64(L)), _d(_f64(nodes, (10, 3))),
                                      _d(out)), "shape_derivs_cart")
    return out


def element_volume(nodes):
    return lib().or_element_volume(_d(_f64(nodes, (10, 3))))


# ---- element kernels ----------------------------------------------------

def thermal_stiffness(nodes, k, npts=4):
    out = np.zeros((10, 10))
    _check(lib().or_thermal_stiffness_rule(_d(_f64(nodes, (10, 3))), k, npts,
                                           _d(out)), "thermal_stiffness")
    return out


def thermal_load(nodes, Q):
    out = np.zeros(10)
    _check(lib().or_thermal_load(_d(_f64(nodes, (10, 3))), Q, _d(out)),
           "thermal_load")
    return out


def thermal_mass(nodes, density, specific_heat):
    """ElementMatrix::thermal_mass (src/fem/element.rs:74-105)."""
    out = np.zeros((10, 10))
    lib().or_thermal_mass(_d(_f64(nodes, (10, 3))), density, specific_heat, _d(out))
    return out


def face_nodes(local_face):
    """Tet10Face::from_element_face (src/fem/boundary.rs:72-115), element-local."""
    out = (C.c_int * 6)()
    _check(lib().or_face_nodes(local_face, out), "face_nodes")
    return np.array(list(out))


def face_area(coords):
    return lib().or_face_area(_d(_f64(coords, (6, 3))))


def face_flux(coords, flux):
    out = np.zeros(6)
    lib().or_face_flux(_d(_f64(coords, (6, 3))), flux, _d(out))
    return out


def apply_neumann(xyz, conn, elem_ids, local_faces, fluxes, f):
    """Assembler::apply_neumann_bcs (src/fem/assembly.rs:248-273), faces in order."""
    xyz = _f64(xyz)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    f = np.array(f, dtype=np.float64)
    el = np.ascontiguousarray(elem_ids, dtype=np.int64)
    lf = np.ascontiguousarray(local_faces, dtype=np.int32)
    fl = _f64(fluxes)
    _check(lib().or_apply_neumann(len(xyz), _d(xyz), len(conn), _i(conn), len(el), _i(el),
                                  lf.ctypes.data_as(C.POINTER(C.c_int32)), _d(fl), _d(f)),
           "apply_neumann")
    return f


def constitutive(E, nu):
    out = np.zeros((6, 6))
    _check(lib().or_constitutive(E, nu, _d(out)), "constitutive")
    return out


def b_matrix(dNdx):
    out = np.zeros((6, 30))
    lib().or_b_matrix(_d(_f64(dNdx, (10, 3))), _d(out))
    return out


def elastic_stiffness(nodes, E, nu):
    out = np.zeros((30, 30))
    _check(lib().or_elastic_stiffness(_d(_f64(nodes, (10, 3))), E, nu, _d(out)),
           "elastic_stiffness")
    return out


def viscous_constitutive(mu, zeta=0.0):
    """NewtonianViscosity::constitutive_matrix (src/mechanics/constitutive.rs:227-255)."""
    out = np.zeros((6, 6))
    _check(lib().or_viscous_constitutive(mu, zeta, _d(out)), "viscous_constitutive")
    return out


def viscosity_stiffness(nodes, mu, zeta=0.0):
    """ElasticityElement::viscosity_matrix (src/mechanics/element.rs:87-118)."""
    out = np.zeros((30, 30))
    _check(lib().or_viscosity_stiffness(_d(_f64(nodes, (10, 3))), mu, zeta, _d(out)),
           "viscosity_stiffness")
    return out


def maxwell_element(nodes, E, nu, sigma):
    """ElasticityElement::maxwell_viscoelastic (src/mechanics/element.rs:257-295)
    -> (K_e, f_history_e)."""
    K = np.zeros((30, 30))
    f = np.zeros(30)
    _check(lib().or_maxwell_element(_d(_f64(nodes, (10, 3))), E, nu, _d(_f64(sigma)),
                                    _d(K), _d(f)), "maxwell_element")
    return K, f


def maxwell_stress(nodes, E, nu, visc, dt, du, sigma_n):
    """One element of update_stresses_maxwell (src/mechanics/stress_update.rs:53-145)."""
    out = np.zeros(6)
    _check(lib().or_maxwell_stress(_d(_f64(nodes, (10, 3))), E, nu, visc, dt, _d(_f64(du)),
                                   _d(_f64(sigma_n)), _d(out)), "maxwell_stress")
    return out


def gravity_load(nodes, rho, g):
    out = np.zeros(30)
    _check(lib().or_gravity_load(_d(_f64(nodes, (10, 3))), rho, _d(_f64(g)),
                                 _d(out)), "gravity_load")
    return out


# ---- mesh ---------------------------------------------------------------

def generate_cube(nx, ny, nz, lx, ly, lz):
    """ImprovedMeshGenerator::generate_cube -> (xyz (n,3) f64, conn (E,10) i64)."""
    nn, ne = C.c_int64(), C.c_int64()
    lib().or_cube_sizes(nx, ny, nz, C.byref(nn), C.byref(ne))
    xyz = np.zeros((nn.value, 3))
    conn = np.zeros((ne.value, 10), dtype=np.int64)
    _check(lib().or_generate_cube(nx, ny, nz, lx, ly, lz, _d(xyz), _i(conn)),
           "generate_cube")
    return xyz, conn


# ---- assembly -----------------------------------------------------------

def _take_csr(m: _Csr) -> sp.csr_matrix:
    rows, cols, nnz = m.rows, m.cols, m.nnz
    indptr = np.ctypeslib.as_array(m.indptr, shape=(rows + 1,)).copy()
    if nnz > 0:
        indices = np.ctypeslib.as_array(m.indices, shape=(nnz,)).copy()
        data = np.ctypeslib.as_array(m.data, shape=(nnz,)).copy()
    else:
        indices = np.zeros(0, dtype=np.int64)
        data = np.zeros(0)
    lib().or_csr_free(C.byref(m))
    return sp.csr_matrix((data, indices, indptr), shape=(rows, cols))


def _give_csr(A: sp.csr_matrix):
    A = sp.csr_matrix(A)
    indptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
    indices = np.ascontiguousarray(A.indices, dtype=np.int64)
    data = np.ascontiguousarray(A.data, dtype=np.float64)
    m = _Csr(A.shape[0], A.shape[1], A.nnz, _i(indptr), _i(indices), _d(data))
    return m, (indptr, indices, data)


def csr_from_triplets(rows, cols, ti, tj, tv):
    ti = np.ascontiguousarray(ti, dtype=np.int64)
    tj = np.ascontiguousarray(tj, dtype=np.int64)
    tv = _f64(tv)
    m = _Csr()
    _check(lib().or_csr_from_triplets(rows, cols, len(ti), _i(ti), _i(tj), _d(tv),
                                      C.byref(m)), "csr_from_triplets")
    return _take_csr(m)


def assemble_thermal(xyz, conn, k, k_elem=None, nthreads=0):
    """Assembler::assemble_thermal_stiffness_parallel (src/fem/assembly.rs:69-117)."""
    xyz = _f64(xyz)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    ke = None if k_elem is None else _f64(k_elem)
    m = _Csr()
    _check(lib().or_assemble_thermal(len(xyz), _d(xyz), len(conn), _i(conn), k,
                                     None if ke is None else _d(ke), nthreads,
                                     C.byref(m)), "assemble_thermal")
    return _take_csr(m)


def assemble_thermal_mass(xyz, conn, density, specific_heat, rho_cp_elem=None, nthreads=0):
    """Global thermal mass: element.rs:74-105 scattered like the stiffness."""
    xyz = _f64(xyz)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    re = None if rho_cp_elem is None else _f64(rho_cp_elem)
    m = _Csr()
    _check(lib().or_assemble_thermal_mass(len(xyz), _d(xyz), len(conn), _i(conn), density,
                                          specific_heat, None if re is None else _d(re),
                                          nthreads, C.byref(m)), "assemble_thermal_mass")
    return _take_csr(m)


def assemble_elastic(xyz, conn, E, nu, E_elem=None, nthreads=0):
    """Assembler::assemble_elasticity_stiffness_parallel (src/fem/assembly.rs:389-437)."""
    xyz = _f64(xyz)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    ee = None if E_elem is None else _f64(E_elem)
    m = _Csr()
    _check(lib().or_assemble_elastic(len(xyz), _d(xyz), len(conn), _i(conn), E,
                                     nu, None if ee is None else _d(ee), nthreads,
                                     C.byref(m)), "assemble_elastic")
    return _take_csr(m)


def assemble_viscosity(xyz, conn, mu, zeta=0.0, mu_elem=None, zeta_elem=None, nthreads=0):
    """Assembler::assemble_viscosity_stiffness_parallel (src/fem/assembly.rs:646-694)."""
    xyz = _f64(xyz)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    me = None if mu_elem is None else _f64(mu_elem)
    ze = None if zeta_elem is None else _f64(zeta_elem)
    m = _Csr()
    _check(lib().or_assemble_viscosity(len(xyz), _d(xyz), len(conn), _i(conn), mu, zeta,
                                       None if me is None else _d(me),
                                       None if ze is None else _d(ze), nthreads,
                                       C.byref(m)), "assemble_viscosity")
    return _take_csr(m)


def assemble_maxwell(xyz, conn, E, nu, visc, dt, stress, nthreads=0):
    """Assembler::assemble_maxwell_viscoelastic_parallel (src/fem/assembly.rs:959-1020)
    -> (K, f_history); stress (E, 6) per element."""
    xyz = _f64(xyz)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    sg = _f64(stress)
    m = _Csr()
    f = np.zeros(3 * len(xyz))
    _check(lib().or_assemble_maxwell(len(xyz), _d(xyz), len(conn), _i(conn), E, nu, visc, dt,
                                     _d(sg), nthreads, C.byref(m), _d(f)), "assemble_maxwell")
    return _take_csr(m), f


def update_stresses_maxwell(xyz, conn, E, nu, visc, dt, u_n, u_next, stress):
    """update_stresses_maxwell (src/mechanics/stress_update.rs:39-147) -> (E, 6)."""
    xyz = _f64(xyz)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    out = np.zeros((len(conn), 6))
    _check(lib().or_update_stresses_maxwell(len(xyz), _d(xyz), len(conn), _i(conn), E, nu,
                                            visc, dt, _d(_f64(u_n)), _d(_f64(u_next)),
                                            _d(_f64(stress)), _d(out)), "update_stresses_maxwell")
    return out


def assemble_thermal_load(xyz, conn, Q):
    xyz = _f64(xyz)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    f = np.zeros(len(xyz))
    _check(lib().or_assemble_thermal_load(len(xyz), _d(xyz), len(conn), _i(conn),
                                          Q, _d(f)), "assemble_thermal_load")
    return f


def assemble_gravity_load(xyz, conn, rho, g):
    xyz = _f64(xyz)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    f = np.zeros(3 * len(xyz))
    _check(lib().or_assemble_gravity_load(len(xyz), _d(xyz), len(conn), _i(conn),
                                          rho, _d(_f64(g)), _d(f)),
           "assemble_gravity_load")
    return f


def apply_dirichlet(K, f, is_dir, values):
    """Assembler::apply_dirichlet_bcs (src/fem/assembly.rs:168-230)."""
    m, keep = _give_csr(K)
    f = _f64(f)
    mask = np.ascontiguousarray(is_dir, dtype=np.uint8)
    vals = _f64(values)
    out = _Csr()
    fbc = np.zeros(len(f))
    _check(lib().or_apply_dirichlet(C.byref(m), _d(f), mask.ctypes.data_as(_u8p),
                                    _d(vals), C.byref(out), _d(fbc)),
           "apply_dirichlet")
    del keep
    return _take_csr(out), fbc


# ---- linear algebra -----------------------------------------------------

def spmv(A, x):
    m, keep = _give_csr(A)
    x = _f64(x)
    y = np.zeros(A.shape[0])
    lib().or_spmv(C.byref(m), _d(x), _d(y))
    return y


def jacobi_diag_inv(A):
    m, keep = _give_csr(A)
    d = np.zeros(A.shape[0])
    lib().or_jacobi_diag_inv(C.byref(m), _d(d))
    return d


def cg_solve(A, b, max_iter=1000, tol=1e-8, abs_tol=1e-12, use_precond=True):
    """ConjugateGradient::solve (src/linalg/iterative.rs:46-131)."""
    m, keep = _give_csr(A)
    b = _f64(b)
    x = np.zeros(A.shape[0])
    st = _Stats()
    _check(lib().or_cg_solve(C.byref(m), _d(b), max_iter, tol, abs_tol,
                             1 if use_precond else 0, _d(x), C.byref(st)),
           "cg_solve")
    return x, SolverStats(st.iterations, st.residual_norm, st.relative_residual,
                          bool(st.converged), st.solve_time)


def cg_solve_omp(A, b, max_iter=1000, tol=1e-8, abs_tol=1e-12, nthreads=0):
    """Timing column, not a restatement: cg_solve's Jacobi-PCG loop on
    nthreads OpenMP threads (0: OpenMP's default), dot products reduced in a
    thread-dependent order."""
    m, keep = _give_csr(A)
    b = _f64(b)
    x = np.zeros(A.shape[0])
    st = _Stats()
    _check(lib().or_cg_solve_omp(C.byref(m), _d(b), max_iter, tol, abs_tol, nthreads,
                                 _d(x), C.byref(st)), "cg_solve_omp")
    return x, SolverStats(st.iterations, st.residual_norm, st.relative_residual,
                          bool(st.converged), st.solve_time)


def bicgstab_solve(A, b, max_iter=1000, tol=1e-8, abs_tol=1e-12, use_precond=True):
    """BiCGSTAB::solve (src/linalg/iterative.rs:177-291)."""
    m, keep = _give_csr(A)
    b = _f64(b)
    x = np.zeros(A.shape[0])
    st = _Stats()
    _check(lib().or_bicgstab_solve(C.byref(m), _d(b), max_iter, tol, abs_tol,
                                   1 if use_precond else 0, _d(x), C.byref(st)),
           "bicgstab_solve")
    return x, SolverStats(st.iterations, st.residual_norm, st.relative_residual,
                          bool(st.converged), st.solve_time)


def gmres_solve(A, b, max_iter=1000, restart=10, tol=1e-8, abs_tol=1e-12, use_precond=True):
    """GMRES::solve (src/linalg/iterative.rs:313-486)."""
    m, keep = _give_csr(A)
    b = _f64(b)
    x = np.zeros(A.shape[0])
    st = _Stats()
    _check(lib().or_gmres_solve(C.byref(m), _d(b), max_iter, restart, tol, abs_tol,
                                1 if use_precond else 0, _d(x), C.byref(st)),
           "gmres_solve")
    return x, SolverStats(st.iterations, st.residual_norm, st.relative_residual,
                          bool(st.converged), st.solve_time)


class IluFactorError(ValueError):
    """ILUPreconditioner::new / update failed: kind 0 'Row {row} has no diag',
    1 'Zero diag at {row}', 2 'Zero diag after update at {row}'."""

    MESSAGES = ("Row {} has no diag", "Zero diag at {}", "Zero diag after update at {}")

    def __init__(self, row, kind):
        super().__init__(self.MESSAGES[kind].format(row))
        self.row, self.kind = row, kind


def ilu0_factor(A):
    """ILUPreconditioner::new (src/linalg/preconditioner.rs:86-166): returns
    (factors on A's pattern, diagonal positions); raises IluFactorError."""
    A = sp.csr_matrix(A)
    A.sort_indices()
    m, keep = _give_csr(A)
    F = np.empty(A.nnz)
    diag = np.zeros(A.shape[0], dtype=np.int64)
    bad = C.c_int64(-1)
    kind = C.c_int32(-1)
    code = lib().or_ilu0_factor(C.byref(m), _d(F), _i(diag), C.byref(bad), C.byref(kind))
    if code == OR_ERR_ZERO_PIVOT:
        raise IluFactorError(bad.value, kind.value)
    _check(code, "ilu0_factor")
    return sp.csr_matrix((F, A.indices.copy(), A.indptr.copy()), shape=A.shape), diag


def ilu0_apply(F, diag, r):
    """ILUPreconditioner::apply (src/linalg/preconditioner.rs:168-203)."""
    m, keep = _give_csr(F)
    r = _f64(r)
    z = np.empty_like(r)
    dg = np.ascontiguousarray(diag, dtype=np.int64)
    lib().or_ilu0_apply(C.byref(m), _d(keep[2]), _i(dg), _d(r), _d(z))
    return z


# ---- AMG (src/linalg/amg.rs) -------------------------------------------------

def _copy_csr(m: _Csr) -> sp.csr_matrix:
    """A copy of an oracle-owned matrix (not freed here)."""
    rows, cols, nnz = m.rows, m.cols, m.nnz
    indptr = np.ctypeslib.as_array(m.indptr, shape=(rows + 1,)).copy()
    if nnz > 0:
        indices = np.ctypeslib.as_array(m.indices, shape=(nnz,)).copy()
        data = np.ctypeslib.as_array(m.data, shape=(nnz,)).copy()
    else:
        indices = np.zeros(0, dtype=np.int64)
        data = np.zeros(0)
    return sp.csr_matrix((data, indices, indptr), shape=(rows, cols))


def _sorted(A):
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A


def amg_strength(A, theta):
    """strength_matrix (src/linalg/amg.rs:359-390)."""
    m, keep = _give_csr(_sorted(A))
    out = _Csr()
    _check(lib().or_amg_strength(C.byref(m), theta, C.byref(out)), "amg_strength")
    return _take_csr(out)


def amg_coarsen(A, theta):
    """coarsen (src/linalg/amg.rs:286-355): (coarse mask, number of coarse points)."""
    A = _sorted(A)
    m, keep = _give_csr(A)
    cf = np.zeros(A.shape[0], dtype=np.uint8)
    nc = C.c_int64(0)
    _check(lib().or_amg_coarsen(C.byref(m), theta, cf.ctypes.data_as(C.POINTER(C.c_uint8)),
                                C.byref(nc)), "amg_coarsen")
    return cf.astype(bool), nc.value


def amg_prolongation(A, cf, num_coarse, theta):
    """build_prolongation (src/linalg/amg.rs:396-487)."""
    m, keep = _give_csr(_sorted(A))
    cf8 = np.ascontiguousarray(cf, dtype=np.uint8)
    out = _Csr()
    _check(lib().or_amg_prolongation(C.byref(m), cf8.ctypes.data_as(C.POINTER(C.c_uint8)),
                                     num_coarse, theta, C.byref(out)), "amg_prolongation")
    return _take_csr(out)


def csr_matmul(A, B):
    """matrix_multiply (src/linalg/amg.rs:195-226)."""
    ma, ka = _give_csr(_sorted(A))
    mb, kb = _give_csr(_sorted(B))
    out = _Csr()
    _check(lib().or_csr_matmul(C.byref(ma), C.byref(mb), C.byref(out)), "csr_matmul")
    return _take_csr(out)


def amg_diag_inv(A):
    """extract_diag_inv (src/linalg/amg.rs:266-280)."""
    A = _sorted(A)
    m, keep = _give_csr(A)
    d = np.empty(A.shape[0])
    lib().or_amg_diag_inv(C.byref(m), _d(d))
    return d


class Amg:
    """AMGPreconditioner::new (src/linalg/amg.rs:44-129) and apply (:188-192)."""

    def __init__(self, A, max_levels=10, coarse_size=500, theta=0.25):
        self._h = None
        m, keep = _give_csr(_sorted(A))
        h = _Amg()
        _check(lib().or_amg_create(C.byref(m), max_levels, coarse_size, theta, C.byref(h)),
               "amg_create")
        self._h = h
        self.nlevels = h.nlevels
        self.A = [_copy_csr(h.A[l]) for l in range(h.nlevels)]
        self.P = [_copy_csr(h.P[l]) if h.P[l].rows > 0 else None for l in range(h.nlevels)]
        self.dinv = [np.ctypeslib.as_array(h.dinv[l], shape=(h.A[l].rows,)).copy()
                     if h.A[l].rows > 0 else np.zeros(0) for l in range(h.nlevels)]

    def apply(self, r):
        r = _f64(r)
        z = np.empty_like(r)
        _check(lib().or_amg_apply(C.byref(self._h), _d(r), _d(z)), "amg_apply")
        return z

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            lib().or_amg_free(C.byref(self._h))
            self._h = None


# ---- adjoint --------------------------------------------------------------

def thermal_element_backward(xyz, conn, k_elem, u, lam):
    xyz = _f64(xyz)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    g = np.zeros(len(conn))
    _check(lib().or_thermal_element_backward(len(xyz), _d(xyz), len(conn),
                                             _i(conn), _d(_f64(k_elem)),
                                             _d(_f64(u)), _d(_f64(lam)), _d(g)),
           "thermal_element_backward")
    return g


def elastic_element_backward(xyz, conn, E_elem, nu, u, lam):
    xyz = _f64(xyz)
    conn = np.ascontiguousarray(conn, dtype=np.int64)
    g = np.zeros(len(conn))
    _check(lib().or_elastic_element_backward(len(xyz), _d(xyz), len(conn),
                                             _i(conn), _d(_f64(E_elem)), nu,
                                             _d(_f64(u)), _d(_f64(lam)), _d(g)),
           "elastic_element_backward")
    return g


# ---- reference-benchmark problem set-ups -----------------------------------

def thermal_dirichlet(xyz, lz, t_bottom=1280.0, t_top=0.0):
    """BC selection of benchmarks/thermal_1d_improved_mesh.rs:46-52."""
    z = xyz[:, 2]
    mask = np.zeros(len(xyz), dtype=np.uint8)
    vals = np.zeros(len(xyz))
    bot = np.abs(z) < 1.0
    top = (~bot) & (np.abs(z - lz) < 1.0)
    mask[bot] = 1
    vals[bot] = t_bottom
    mask[top] = 1
    vals[top] = t_top
    return mask, vals


def thermal_analytic(z, lz, t_bottom, t_top, k, q):
    """benchmarks/thermal_1d_improved_mesh.rs:9-13."""
    return t_bottom + (t_top - t_bottom) * (z / lz) + (q / (2.0 * k)) * z * (lz - z)


def elastic_dirichlet(xyz, lx, ly):
    """Confined column: bottom fixed (all components), u_x = 0 on x faces,
    u_y = 0 on y faces (BC pattern of benchmarks/elasticity_halfspace_gravity.rs:75-103
    with the bottom fully fixed)."""
    n = len(xyz)
    mask = np.zeros(3 * n, dtype=np.uint8)
    x, y, z = xyz[:, 0], xyz[:, 1], xyz[:, 2]
    bot = np.abs(z) < 1.0
    for c in range(3):
        mask[3 * np.nonzero(bot)[0] + c] = 1
    xf = (np.abs(x) < 1.0) | (np.abs(x - lx) < 1.0)
    yf = (np.abs(y) < 1.0) | (np.abs(y - ly) < 1.0)
    mask[3 * np.nonzero(xf)[0] + 0] = 1
    mask[3 * np.nonzero(yf)[0] + 1] = 1
    return mask, np.zeros(3 * n)


def elastic_analytic_uz(z, H, rho, g, E, nu):
    """benchmarks/elasticity_halfspace_gravity.rs:23-27."""
    M = E * (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu))
    return -(rho * g / M) * (H * z - z * z / 2.0)
