"""ctypes binding of libgeofem.so (the C ABI declared in include/geofem.h).

The library is built in-tree by geo_simulator_amd/csrc/build.py.  Loading it
needs no GPU; every compute entry point needs one.  There is no fallback: if
the library is missing the import of the compute API fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libgeofem.so")

GS_OK = 0
GS_ERR_INVALID_ARG = 1
GS_ERR_HIP = 2
GS_ERR_SINGULAR_JACOBIAN = 3
GS_ERR_BAD_DOFS = 4
GS_ERR_INVALID_MATERIAL = 5
GS_ERR_NONFINITE = 6
GS_ERR_TOO_LARGE = 7
GS_ERR_OOM = 8
GS_ERR_ZERO_PIVOT = 9

GS_PREC_FP64 = 0
GS_PREC_MIXED = 1

GS_SPMV_VECTOR = 4
GS_SPMV_INDEX32 = 8
GS_SPMV_CSR = 16
GS_SPMV_SCALAR = 32
GS_SPMV_BLOCK3 = 64

_STATUS_NAMES = {
    GS_ERR_INVALID_ARG: "GS_ERR_INVALID_ARG", GS_ERR_HIP: "GS_ERR_HIP",
    GS_ERR_SINGULAR_JACOBIAN: "GS_ERR_SINGULAR_JACOBIAN",
    GS_ERR_BAD_DOFS: "GS_ERR_BAD_DOFS",
    GS_ERR_INVALID_MATERIAL: "GS_ERR_INVALID_MATERIAL",
    GS_ERR_NONFINITE: "GS_ERR_NONFINITE", GS_ERR_TOO_LARGE: "GS_ERR_TOO_LARGE",
    GS_ERR_OOM: "GS_ERR_OOM", GS_ERR_ZERO_PIVOT: "GS_ERR_ZERO_PIVOT",
}


class GeoFemError(RuntimeError):
    """A non-zero gs_status (the reference would have panicked)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"{_STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


class CgOpts(C.Structure):
    _fields_ = [("max_iterations", C.c_int64), ("tolerance", C.c_double),
                ("abs_tolerance", C.c_double), ("use_preconditioner", C.c_int32),
                ("precision", C.c_int32), ("inner_tolerance", C.c_double),
                ("max_inner_iterations", C.c_int64),
                ("max_outer_iterations", C.c_int64),
                ("relative_breakdown", C.c_int32), ("reserved_", C.c_int32)]


class SolverStatsC(C.Structure):
    _fields_ = [("iterations", C.c_int64), ("residual_norm", C.c_double),
                ("relative_residual", C.c_double), ("converged", C.c_int32),
                ("pad_", C.c_int32), ("solve_time", C.c_double),
                ("outer_iterations", C.c_int64), ("spmv_count", C.c_int64)]


_vp = C.c_void_p

# gs_operator_fn: int32 (*)(void *user)
OPERATOR_FN = C.CFUNCTYPE(C.c_int32, C.c_void_p)
_pp = C.POINTER(C.c_void_p)
_i64 = C.c_int64
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)
_u8p = C.POINTER(C.c_uint8)

# name -> (restype, argtypes); the exact set of symbols include/geofem.h declares
SIGNATURES = {
    "gs_ctx_create": (C.c_int, [C.c_int, _pp]),
    "gs_ctx_destroy": (None, [_vp]),
    "gs_last_error": (C.c_char_p, [_vp]),
    "gs_ctx_synchronize": (C.c_int, [_vp]),
    "gs_device_alloc": (C.c_int, [_vp, C.c_size_t, _pp]),
    "gs_device_free": (C.c_int, [_vp, _vp]),
    "gs_memcpy_h2d": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "gs_memcpy_d2h": (C.c_int, [_vp, _vp, _vp, C.c_size_t]),
    "gs_generate_cube_sizes": (C.c_int, [_i64, _i64, _i64, _i64p, _i64p]),
    "gs_generate_cube": (C.c_int, [_i64, _i64, _i64, C.c_double, C.c_double,
                                   C.c_double, _dp, _i64p]),
    "gs_mesh_create": (C.c_int, [_vp, _dp, _i64, _i64p, _i64, _pp]),
    "gs_mesh_destroy": (None, [_vp]),
    "gs_dofmap_create": (C.c_int, [_vp, _i64, C.c_int, _u8p, _dp, _pp]),
    "gs_dofmap_destroy": (None, [_vp]),
    "gs_assemble_thermal_stiffness": (C.c_int, [_vp, _vp, _vp, C.c_double, _vp, _pp]),
    "gs_assemble_thermal_load": (C.c_int, [_vp, _vp, _vp, C.c_double, _vp]),
    "gs_assemble_elasticity_stiffness": (C.c_int, [_vp, _vp, _vp, C.c_double,
                                                   C.c_double, _vp, _pp]),
    "gs_assemble_thermal_mass": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, _vp,
                                           _pp]),
    "gs_apply_neumann_bcs": (C.c_int, [_vp, _vp, _i64, _i64p, C.POINTER(C.c_int32), _dp,
                                       _vp]),
    "gs_assemble_gravity_load": (C.c_int, [_vp, _vp, _vp, C.c_double, _dp, _vp]),
    "gs_apply_dirichlet_bcs": (C.c_int, [_vp, _vp, _vp, _vp, _pp, _vp]),
    "gs_assemble_thermal_system": (C.c_int, [_vp, _vp, _vp, C.c_double, _vp,
                                             C.c_double, _pp, _vp]),
    "gs_assemble_elasticity_system": (C.c_int, [_vp, _vp, _vp, C.c_double,
                                                C.c_double, _vp, C.c_double, _dp,
                                                _pp, _vp]),
    "gs_assemble_viscosity_stiffness": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double,
                                                  _vp, _vp, _pp]),
    "gs_assemble_viscosity_system": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, _vp,
                                               _vp, _vp, _pp, _vp]),
    "gs_assemble_maxwell_viscoelastic": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double,
                                                   C.c_double, C.c_double, _vp, _pp, _vp]),
    "gs_assemble_maxwell_system": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double,
                                             C.c_double, C.c_double, _vp, _vp, _pp, _vp]),
    "gs_update_stresses_maxwell": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double,
                                             C.c_double, C.c_double, _vp, _vp, _vp, _vp]),
    "gs_csr_info": (C.c_int, [_vp, _i64p, _i64p, _i64p]),
    "gs_csr_export": (C.c_int, [_vp, _vp, _i64p, _i64p, _dp]),
    "gs_csr_from_host": (C.c_int, [_vp, _i64, _i64, _i64, _i64p, _i64p, _dp, _pp]),
    "gs_csr_destroy": (None, [_vp]),
    "gs_csr_spmv": (C.c_int, [_vp, _vp, _vp, _vp]),
    "gs_jacobi_diag_inv": (C.c_int, [_vp, _vp, _vp]),
    "gs_quadrature_rule": (C.c_int, [C.c_int32, _dp, _dp]),
    "gs_thermal_element_matrices": (C.c_int, [_vp, _vp, C.c_double, _vp, C.c_int32, _vp]),
    "gs_ilu_create": (C.c_int, [_vp, _vp, _pp]),
    "gs_ilu_update": (C.c_int, [_vp, _vp, _vp]),
    "gs_ilu_apply": (C.c_int, [_vp, _vp, _vp, _vp]),
    "gs_ilu_info": (C.c_int, [_vp, _i64p, _i64p, _i64p, _i64p]),
    "gs_ilu_export": (C.c_int, [_vp, _vp, _i64p, _i64p, _dp]),
    "gs_ilu_destroy": (None, [_vp]),
    "gs_amg_create": (C.c_int, [_vp, _vp, _i64, _i64, C.c_double, _pp]),
    "gs_amg_apply": (C.c_int, [_vp, _vp, _vp, _vp]),
    "gs_amg_levels": (C.c_int, [_vp, _i64p]),
    "gs_amg_level_info": (C.c_int, [_vp, _i64, _i64p, _i64p, _i64p, _i64p, _i64p]),
    "gs_amg_export": (C.c_int, [_vp, _vp, _i64, C.c_int32, _i64p, _i64p, _dp]),
    "gs_amg_destroy": (None, [_vp]),
    "gs_cg_opts_default": (C.c_int, [C.POINTER(CgOpts)]),
    "gs_cg_solve": (C.c_int, [_vp, _vp, _vp, C.POINTER(CgOpts), _vp,
                              C.POINTER(SolverStatsC)]),
    "gs_cg_solve_noalloc": (C.c_int, [_vp, _vp, _vp, C.POINTER(CgOpts), _vp,
                                      C.POINTER(SolverStatsC)]),
    "gs_bicgstab_solve": (C.c_int, [_vp, _vp, _vp, C.POINTER(CgOpts), _vp,
                                    C.POINTER(SolverStatsC)]),
    "gs_cg_solve_operator": (C.c_int, [_vp, _i64, OPERATOR_FN, _vp, OPERATOR_FN, _vp, _vp,
                                       _vp, _vp, C.POINTER(CgOpts), _vp,
                                       C.POINTER(SolverStatsC)]),
    "gs_gmres_solve": (C.c_int, [_vp, _vp, _vp, C.POINTER(CgOpts), C.c_int32, _vp,
                                 C.POINTER(SolverStatsC)]),
    "gs_thermal_element_backward": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "gs_elastic_element_backward": (C.c_int, [_vp, _vp, _vp, C.c_double, _vp,
                                              _vp, _vp]),
    "gs_csr_set_spmv_shape": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32]),
    "gs_time_stream": (C.c_int, [_vp, C.c_size_t, C.c_int32, C.c_int32, _dp]),
    "gs_time_spmv": (C.c_int, [_vp, _vp, C.c_int32, C.c_int32, _dp]),
    "gs_time_elastic_elements": (C.c_int, [_vp, _vp, C.c_double, C.c_double, _vp,
                                           C.c_int32, C.c_int32, _dp, _dp]),
}

_lib = None


def lib():
    """Load libgeofem.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python -m geo_simulator_amd.csrc.build` (there is no "
                "non-HIP fallback)")
        # torch's HIP runtime first: the library then binds to the same
        # runtime instance instead of loading a second one beside it
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, ctx_handle=None, what: str = ""):
    if status != GS_OK:
        msg = ""
        if ctx_handle is not None:
            raw = lib().gs_last_error(ctx_handle)
            msg = raw.decode() if raw else ""
        raise GeoFemError(status, f"{what}: {msg}" if what else msg)
