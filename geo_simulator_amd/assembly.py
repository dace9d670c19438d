"""Global assembly on the GPU, mirroring `Assembler` (src/fem/assembly.rs) and
`IsotropicElasticity` (src/mechanics/constitutive.rs:16-135).

Stiffness matrices come back as device-resident `CsrMatrix` handles (export
with `.to_scipy()`); load vectors come back as numpy arrays like the
reference's Vec<f64>, or as device tensors with device=True.  The reference's
panics surface as `GeoFemError` (or ValueError for invalid materials, as the
constructor asserts).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import lib
from .linalg import CsrMatrix, _is_torch
from .runtime import default_context, dptr


class IsotropicElasticity:
    """src/mechanics/constitutive.rs:16-135."""

    def __init__(self, youngs_modulus: float, poisson_ratio: float):
        if not youngs_modulus > 0.0:
            raise ValueError("Young's modulus must be positive")
        if not (-1.0 < poisson_ratio < 0.5):
            raise ValueError("Poisson's ratio must be in (-1, 0.5) for physical stability")
        self.youngs_modulus = float(youngs_modulus)
        self.poisson_ratio = float(poisson_ratio)

    @classmethod
    def new(cls, E, nu):
        return cls(E, nu)

    def constitutive_matrix(self) -> np.ndarray:
        E, nu = self.youngs_modulus, self.poisson_ratio
        factor = E / ((1.0 + nu) * (1.0 - 2.0 * nu))
        D = np.zeros((6, 6))
        D[:3, :3] = nu
        np.fill_diagonal(D[:3, :3], 1.0 - nu)
        D[3, 3] = D[4, 4] = D[5, 5] = (1.0 - 2.0 * nu) / 2.0
        return D * factor

    def lame_parameters(self):
        E, nu = self.youngs_modulus, self.poisson_ratio
        return (E * nu) / ((1.0 + nu) * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))

    def bulk_modulus(self) -> float:
        E, nu = self.youngs_modulus, self.poisson_ratio
        return E / (3.0 * (1.0 - 2.0 * nu))


def _elem_array(ctx, a, n_elems):
    if a is None:
        return None
    t = ctx.to_device(a)
    if t.numel() != n_elems:
        raise ValueError("per-element array length differs from the element count")
    return t


def _out(ctx, t, device):
    return t if device else t.cpu().numpy()


class Assembler:
    # ---- thermal -----------------------------------------------------------
    @staticmethod
    def assemble_thermal_stiffness_parallel(mesh, dof_mgr, conductivity,
                                            k_per_elem=None, ctx=None):
        """src/fem/assembly.rs:69-117 (per-element k: additive overload)."""
        ctx = ctx or default_context()
        ke = _elem_array(ctx, k_per_elem, mesh.num_elements())
        h = C.c_void_p()
        ctx.sync_torch()
        ctx.check(lib().gs_assemble_thermal_stiffness(
            ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx), float(conductivity),
            dptr(ke), C.byref(h)), "assemble_thermal_stiffness_parallel")
        return CsrMatrix(ctx, h)

    assemble_thermal_stiffness_serial = assemble_thermal_stiffness_parallel

    @staticmethod
    def assemble_thermal_load(mesh, dof_mgr, source, device=False, ctx=None):
        """src/fem/assembly.rs:130-154."""
        ctx = ctx or default_context()
        f = ctx.empty(dof_mgr.total_dofs())
        ctx.check(lib().gs_assemble_thermal_load(
            ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx), float(source),
            dptr(f)), "assemble_thermal_load")
        return _out(ctx, f, device)

    @staticmethod
    def assemble_thermal_mass(mesh, dof_mgr, density, specific_heat,
                              rho_cp_per_elem=None, ctx=None):
        """Global thermal mass from ElementMatrix::thermal_mass
        (src/fem/element.rs:74-105, 5-point rule), scattered like the
        stiffness; rho_cp_per_elem overrides density * specific_heat."""
        ctx = ctx or default_context()
        rc = _elem_array(ctx, rho_cp_per_elem, mesh.num_elements())
        h = C.c_void_p()
        ctx.sync_torch()
        ctx.check(lib().gs_assemble_thermal_mass(
            ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx), float(density),
            float(specific_heat), dptr(rc), C.byref(h)), "assemble_thermal_mass")
        return CsrMatrix(ctx, h)

    @staticmethod
    def apply_neumann_bcs(mesh, f, bcs, ctx=None):
        """src/fem/assembly.rs:248-273: f += the flux of every Neumann face of
        `bcs` (a BoundaryConditions), in place on a numpy array or a device
        tensor; the face integrals run on the GPU."""
        ctx = ctx or default_context()
        faces = list(bcs.neumann_faces())
        if not faces:
            return f
        el = np.array([fc.element_id for fc, _ in faces], dtype=np.int64)
        lf = np.array([fc.local_face_id for fc, _ in faces], dtype=np.int32)
        fl = np.array([bc.flux for _, bc in faces], dtype=np.float64)
        import torch
        on_device = isinstance(f, torch.Tensor)
        fd = f if on_device and f.is_cuda and f.dtype == torch.float64 and f.is_contiguous() \
            else ctx.to_device(f)
        ctx.sync_torch()
        ctx.check(lib().gs_apply_neumann_bcs(
            ctx.handle, mesh.handle(ctx), len(el), el.ctypes.data_as(C.POINTER(C.c_int64)),
            lf.ctypes.data_as(C.POINTER(C.c_int32)), fl.ctypes.data_as(C.POINTER(C.c_double)),
            dptr(fd)), "apply_neumann_bcs")
        if fd is not f:
            if on_device:
                f.copy_(fd)
            else:
                f[...] = fd.cpu().numpy()
        return f

    @staticmethod
    def find_boundary_faces(mesh, condition):
        """src/fem/assembly.rs:285-309: (element_id, local_face_id) of every
        element face whose six nodes all satisfy `condition` (called with an
        (n, 3) array of points, or with one point at a time when it does not
        accept arrays), in element then face order."""
        from .boundary import FACE_NODES
        pts = mesh.nodes
        try:
            on = np.asarray(condition(pts), dtype=bool).reshape(-1)
            if on.shape[0] != pts.shape[0]:
                raise ValueError
        except Exception:
            on = np.array([bool(condition(p)) for p in pts])
        fn = mesh.elements[:, FACE_NODES]          # (E, 4, 6)
        hit = on[fn].all(axis=2)                   # (E, 4)
        e, lf = np.nonzero(hit)
        return list(zip(e.tolist(), lf.tolist()))

    # ---- elasticity -----------------------------------------------------------
    @staticmethod
    def assemble_elasticity_stiffness_parallel(mesh, dof_mgr, material,
                                               E_per_elem=None, ctx=None):
        """src/fem/assembly.rs:389-437 (per-element E: additive overload)."""
        ctx = ctx or default_context()
        ee = _elem_array(ctx, E_per_elem, mesh.num_elements())
        h = C.c_void_p()
        ctx.sync_torch()
        ctx.check(lib().gs_assemble_elasticity_stiffness(
            ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx),
            material.youngs_modulus, material.poisson_ratio, dptr(ee),
            C.byref(h)), "assemble_elasticity_stiffness_parallel")
        return CsrMatrix(ctx, h)

    assemble_elasticity_stiffness_serial = assemble_elasticity_stiffness_parallel

    @staticmethod
    def assemble_gravity_load(mesh, dof_mgr, density, gravity, device=False,
                              ctx=None):
        """src/fem/assembly.rs:906-937."""
        ctx = ctx or default_context()
        g = np.ascontiguousarray(gravity, dtype=np.float64).reshape(3)
        f = ctx.empty(dof_mgr.total_dofs())
        ctx.check(lib().gs_assemble_gravity_load(
            ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx), float(density),
            g.ctypes.data_as(C.POINTER(C.c_double)), dptr(f)),
            "assemble_gravity_load")
        return _out(ctx, f, device)

    # ---- Dirichlet ---------------------------------------------------------
    @staticmethod
    def apply_dirichlet_bcs(K, f, dof_mgr, device=None):
        """src/fem/assembly.rs:168-230 -> (K_bc, f_bc)."""
        if not isinstance(K, CsrMatrix):
            K = CsrMatrix.from_scipy(K)
        ctx = K.ctx
        host = not _is_torch(f) if device is None else not device
        fd = ctx.to_device(f)
        if fd.numel() != K.shape[0]:
            raise ValueError("load vector length differs from the matrix rows")
        fbc = ctx.empty(K.shape[0])
        h = C.c_void_p()
        ctx.sync_torch()
        ctx.check(lib().gs_apply_dirichlet_bcs(ctx.handle, K.handle, dptr(fd),
                                               dof_mgr.handle(ctx), C.byref(h),
                                               dptr(fbc)), "apply_dirichlet_bcs")
        return CsrMatrix(ctx, h), (fbc.cpu().numpy() if host else fbc)

    # ---- fused device-resident pipelines --------------------------------------
    @staticmethod
    def assemble_thermal_system(mesh, dof_mgr, conductivity, source,
                                k_per_elem=None, ctx=None):
        """assemble_thermal_stiffness_parallel + assemble_thermal_load +
        apply_dirichlet_bcs on the device -> (K_bc, f_bc tensor)."""
        ctx = ctx or default_context()
        ke = _elem_array(ctx, k_per_elem, mesh.num_elements())
        fbc = ctx.empty(dof_mgr.total_dofs())
        h = C.c_void_p()
        ctx.sync_torch()
        ctx.check(lib().gs_assemble_thermal_system(
            ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx), float(conductivity),
            dptr(ke), float(source), C.byref(h), dptr(fbc)),
            "assemble_thermal_system")
        return CsrMatrix(ctx, h), fbc

    @staticmethod
    def assemble_elasticity_system(mesh, dof_mgr, material, density, gravity,
                                   E_per_elem=None, ctx=None):
        """assemble_elasticity_stiffness_parallel + assemble_gravity_load +
        apply_dirichlet_bcs on the device -> (K_bc, f_bc tensor)."""
        ctx = ctx or default_context()
        ee = _elem_array(ctx, E_per_elem, mesh.num_elements())
        g = np.ascontiguousarray(gravity, dtype=np.float64).reshape(3)
        fbc = ctx.empty(dof_mgr.total_dofs())
        h = C.c_void_p()
        ctx.sync_torch()
        ctx.check(lib().gs_assemble_elasticity_system(
            ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx),
            material.youngs_modulus, material.poisson_ratio, dptr(ee),
            float(density), g.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h),
            dptr(fbc)), "assemble_elasticity_system")
        return CsrMatrix(ctx, h), fbc


    # ---- viscous flow and Maxwell viscoelasticity ------------------------------
    @staticmethod
    def assemble_viscosity_stiffness_parallel(mesh, dof_mgr, material, mu_per_elem=None,
                                              zeta_per_elem=None, ctx=None):
        """src/fem/assembly.rs:646-694 (material: NewtonianViscosity; per-element
        viscosity / bulk viscosity: additive overload)."""
        ctx = ctx or default_context()
        me = _elem_array(ctx, mu_per_elem, mesh.num_elements())
        ze = _elem_array(ctx, zeta_per_elem, mesh.num_elements())
        h = C.c_void_p()
        ctx.sync_torch()
        ctx.check(lib().gs_assemble_viscosity_stiffness(
            ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx), material.viscosity,
            material.bulk_viscosity, dptr(me), dptr(ze), C.byref(h)),
            "assemble_viscosity_stiffness_parallel")
        return CsrMatrix(ctx, h)

    @staticmethod
    def assemble_viscosity_system(mesh, dof_mgr, material, f=None, mu_per_elem=None,
                                  zeta_per_elem=None, ctx=None):
        """assemble_viscosity_stiffness_parallel + apply_dirichlet_bcs on the
        device (src/linalg/picard.rs:264-268) -> (K_bc, f_bc tensor); f is
        the right-hand side before the elimination (None: zero)."""
        ctx = ctx or default_context()
        me = _elem_array(ctx, mu_per_elem, mesh.num_elements())
        ze = _elem_array(ctx, zeta_per_elem, mesh.num_elements())
        fd = None if f is None else ctx.to_device(f)
        if fd is not None and fd.numel() != dof_mgr.total_dofs():
            raise ValueError("load vector length differs from the DOF count")
        fbc = ctx.empty(dof_mgr.total_dofs())
        h = C.c_void_p()
        ctx.sync_torch()
        ctx.check(lib().gs_assemble_viscosity_system(
            ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx), material.viscosity,
            material.bulk_viscosity, dptr(me), dptr(ze), dptr(fd), C.byref(h), dptr(fbc)),
            "assemble_viscosity_system")
        return CsrMatrix(ctx, h), fbc

    @staticmethod
    def assemble_maxwell_viscoelastic_parallel(mesh, dof_mgr, material, dt, device=False,
                                               ctx=None):
        """src/fem/assembly.rs:959-1020 -> (K, f_history); the stress history
        is mesh.stress_history (a rheology.StressHistory)."""
        from .rheology import _stress_of
        ctx = ctx or default_context()
        sig = _stress_of(mesh, ctx)
        f = ctx.empty(dof_mgr.total_dofs())
        h = C.c_void_p()
        ctx.sync_torch()
        ctx.check(lib().gs_assemble_maxwell_viscoelastic(
            ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx), material.youngs_modulus,
            material.poisson_ratio, material.viscosity, float(dt), dptr(sig), C.byref(h),
            dptr(f)), "assemble_maxwell_viscoelastic_parallel")
        return CsrMatrix(ctx, h), _out(ctx, f, device)

    @staticmethod
    def assemble_maxwell_system(mesh, dof_mgr, material, dt, f_ext, ctx=None):
        """BackwardEuler::step's assembly (src/fem/time_integration.rs:99-110):
        Maxwell K and f_history, f = f_ext + f_history, apply_dirichlet_bcs,
        on the device -> (K_bc, f_bc tensor)."""
        from .rheology import _stress_of
        ctx = ctx or default_context()
        sig = _stress_of(mesh, ctx)
        fe = ctx.to_device(f_ext)
        if fe.numel() != dof_mgr.total_dofs():
            raise ValueError("load vector length differs from the DOF count")
        fbc = ctx.empty(dof_mgr.total_dofs())
        h = C.c_void_p()
        ctx.sync_torch()
        ctx.check(lib().gs_assemble_maxwell_system(
            ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx), material.youngs_modulus,
            material.poisson_ratio, material.viscosity, float(dt), dptr(sig), dptr(fe),
            C.byref(h), dptr(fbc)), "assemble_maxwell_system")
        return CsrMatrix(ctx, h), fbc


class GaussQuadrature:
    """GaussQuadrature (src/fem/quadrature.rs:1-110): `points` (barycentric
    [L0, L1, L2, L3]) and `weights` of the library's literal tables, the ones
    its kernels integrate with."""

    def __init__(self, n_points: int):
        pts = np.zeros((max(1, int(n_points)), 4))
        w = np.zeros(max(1, int(n_points)))
        st = lib().gs_quadrature_rule(int(n_points), pts.ctypes.data_as(C.POINTER(C.c_double)),
                                      w.ctypes.data_as(C.POINTER(C.c_double)))
        if st != 0:
            raise ValueError("tetrahedral rules have 1, 4, 5 or 11 points")
        self.points = [list(p) for p in pts]
        self.weights = list(w)

    @classmethod
    def tet_1point(cls):
        return cls(1)

    @classmethod
    def tet_4point(cls):
        return cls(4)

    @classmethod
    def tet_5point(cls):
        return cls(5)

    @classmethod
    def tet_11point(cls):
        return cls(11)

    def num_points(self) -> int:
        return len(self.weights)


class ElementMatrix:
    """Element matrices (src/fem/element.rs) for every element of a mesh at
    once, on the device."""

    @staticmethod
    def thermal_stiffness(mesh, conductivity: float, quadrature=None, k_per_elem=None,
                          ctx=None):
        """ElementMatrix::thermal_stiffness (src/fem/element.rs:23-61) of each
        element, integrated with `quadrature` (a GaussQuadrature or its point
        count; default the reference's 4-point rule): a device tensor
        (n_elems, 10, 10)."""
        ctx = ctx or default_context()
        n = quadrature.num_points() if isinstance(quadrature, GaussQuadrature) else (
            4 if quadrature is None else int(quadrature))
        E = mesh.num_elements()
        kd = _elem_array(ctx, k_per_elem, E)
        out = ctx.empty(E * 100)
        ctx.sync_torch()
        ctx.check(lib().gs_thermal_element_matrices(
            ctx.handle, mesh.handle(ctx), float(conductivity),
            dptr(kd) if kd is not None else None, n, dptr(out)),
            "ElementMatrix::thermal_stiffness")
        return out.view(E, 10, 10)


def thermal_element_backward(mesh, k_per_elem, u, lam, ctx=None):
    """grad_e = -lambda_e^T K_e(k_e) u_e = dL/dlog k_e (device tensors)."""
    ctx = ctx or default_context()
    k = ctx.to_device(k_per_elem)
    ud, ld = ctx.to_device(u), ctx.to_device(lam)
    g = ctx.empty(mesh.num_elements())
    ctx.sync_torch()
    ctx.check(lib().gs_thermal_element_backward(ctx.handle, mesh.handle(ctx),
                                                dptr(k), dptr(ud), dptr(ld),
                                                dptr(g)),
              "thermal_element_backward")
    return g


def elastic_element_backward(mesh, E_per_elem, poisson, u, lam, ctx=None):
    """grad_e = -lambda_e^T K_e(E_e) u_e = dL/dlog E_e (device tensors)."""
    ctx = ctx or default_context()
    E = ctx.to_device(E_per_elem)
    ud, ld = ctx.to_device(u), ctx.to_device(lam)
    g = ctx.empty(mesh.num_elements())
    ctx.sync_torch()
    ctx.check(lib().gs_elastic_element_backward(ctx.handle, mesh.handle(ctx),
                                                dptr(E), float(poisson),
                                                dptr(ud), dptr(ld), dptr(g)),
              "elastic_element_backward")
    return g


ELEMENT_VARIANTS = {"valu": 0, "mfma": 1, "valu_f32": 2, "mfma_f32": 3, "valu_rows": 4}


def time_elastic_elements(mesh, material, E_per_elem=None, variant="valu", reps=10,
                          ctx=None):
    """Average time (ms) of one launch of the elastic element kernel (K_e and
    gravity load of every element) and the largest K_e entry difference from
    the VALU kernel relative to the largest entry -> (ms, rel_diff).
    variant "valu" is the kernel assembly uses (FP64, one lane per 3x3 node
    block); "mfma" evaluates K_e = sum_q w_q B_q^T D B_q as one 30x24 by
    24x30 product on the matrix cores (v_mfma_f64_16x16x4f64); "valu_f32" and
    "mfma_f32" are the same two in FP32 (v_mfma_f32_16x16x4f32), FP32 K_e;
    "valu_rows" is the FP64 kernel with one lane per K_e row, staged through
    LDS."""
    ctx = ctx or default_context()
    if variant not in ELEMENT_VARIANTS:
        raise ValueError(f"unknown element kernel variant {variant!r}")
    ee = _elem_array(ctx, E_per_elem, mesh.num_elements())
    ms, rel = C.c_double(), C.c_double()
    ctx.sync_torch()
    ctx.check(lib().gs_time_elastic_elements(
        ctx.handle, mesh.handle(ctx), material.youngs_modulus, material.poisson_ratio,
        dptr(ee), ELEMENT_VARIANTS[variant], int(reps), C.byref(ms), C.byref(rel)),
        "time_elastic_elements")
    return ms.value, rel.value
