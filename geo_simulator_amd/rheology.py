"""Viscous and Maxwell viscoelastic materials and the Backward-Euler step.

Mirrors `NewtonianViscosity` and `MaxwellViscoelasticity`
(src/mechanics/constitutive.rs:156-325), `StressHistory`
(src/mesh/state.rs:12-75), `update_stresses_maxwell`
(src/mechanics/stress_update.rs:39-147) and `BackwardEuler`
(src/fem/time_integration.rs:38-160).  The element kernels, the values-only
re-assembly on the mesh's cached pattern and elimination plan, and the
stress update run in libgeofem on the GPU; the stress history lives on the
device as an (n_elems, 6) float64 tensor, Voigt [xx, yy, zz, xy, yz, xz].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib
from .runtime import default_context, dptr


class NewtonianViscosity:
    """src/mechanics/constitutive.rs:156-260."""

    def __init__(self, viscosity: float):
        if not viscosity > 0.0:
            raise ValueError(f"Viscosity must be positive, got {viscosity}")
        self.viscosity = float(viscosity)
        self.bulk_viscosity = 0.0

    @classmethod
    def new(cls, viscosity):
        return cls(viscosity)

    def with_penalty(self, penalty: float):
        self.bulk_viscosity = float(penalty)
        return self

    def with_incompressible(self):
        self.bulk_viscosity = self.viscosity * 1e7
        return self

    def constitutive_matrix(self) -> np.ndarray:
        mu, zeta = self.viscosity, self.bulk_viscosity
        lam = zeta - (2.0 * mu / 3.0)
        D = np.zeros((6, 6))
        D[:3, :3] = lam
        D[np.arange(3), np.arange(3)] += 2.0 * mu
        D[3, 3] = D[4, 4] = D[5, 5] = 2.0 * mu
        return D

    def dynamic_viscosity(self) -> float:
        return self.viscosity


class MaxwellViscoelasticity:
    """src/mechanics/constitutive.rs:262-380."""

    def __init__(self, youngs_modulus: float, poisson_ratio: float, viscosity: float):
        if not youngs_modulus > 0.0:
            raise ValueError(f"Young's modulus must be positive, got {youngs_modulus}")
        if not (-1.0 < poisson_ratio < 0.5):
            raise ValueError(f"Poisson's ratio must be in (-1, 0.5), got {poisson_ratio}")
        if not viscosity > 0.0:
            raise ValueError(f"Viscosity must be positive, got {viscosity}")
        self.youngs_modulus = float(youngs_modulus)
        self.poisson_ratio = float(poisson_ratio)
        self.viscosity = float(viscosity)

    @classmethod
    def new(cls, E, nu, mu):
        return cls(E, nu, mu)

    def shear_modulus(self) -> float:
        return self.youngs_modulus / (2.0 * (1.0 + self.poisson_ratio))

    def relaxation_time(self) -> float:
        return self.viscosity / self.shear_modulus()

    def elastic_matrix(self) -> np.ndarray:
        from .assembly import IsotropicElasticity
        return IsotropicElasticity(self.youngs_modulus, self.poisson_ratio).constitutive_matrix()

    def effective_constitutive_matrix(self, dt: float) -> np.ndarray:
        return self.elastic_matrix() * (1.0 / (1.0 + dt / self.relaxation_time()))


class StressHistory:
    """Per-element stress, a device tensor (n_elements, 6) (src/mesh/state.rs:12-75)."""

    def __init__(self, n_elements: int, ctx=None):
        import torch
        self.ctx = ctx or default_context()
        self.stress = torch.zeros((int(n_elements), 6), dtype=torch.float64,
                                  device=self.ctx.torch_device())

    @classmethod
    def new(cls, n_elements, ctx=None):
        return cls(n_elements, ctx)

    def get(self, elem_id: int):
        return self.stress[elem_id].cpu().numpy()

    def set(self, elem_id: int, stress):
        import torch
        self.stress[elem_id] = torch.as_tensor(np.asarray(stress, dtype=np.float64))

    def update_all(self, new_stresses):
        t = self.ctx.to_device(new_stresses).reshape(-1, 6)
        if t.shape != self.stress.shape:
            raise ValueError("New stress vector size must match existing size")
        self.stress = t.contiguous()

    def num_elements(self) -> int:
        return int(self.stress.shape[0])


def _stress_of(mesh, ctx):
    sh = getattr(mesh, "stress_history", None)
    if sh is None:
        raise ValueError("Mesh must have stress_history for viscoelasticity")
    if sh.num_elements() != mesh.num_elements():
        raise ValueError("stress history size differs from the element count")
    return ctx.to_device(sh.stress).reshape(-1, 6).contiguous()


def update_stresses_maxwell(mesh, dof_mgr, material: MaxwellViscoelasticity, u_n, u_next,
                            dt: float, ctx=None):
    """src/mechanics/stress_update.rs:39-147 -> sigma_{n+1}, a device tensor
    (n_elements, 6); mesh.stress_history holds sigma_n."""
    ctx = ctx or default_context()
    sig = _stress_of(mesh, ctx)
    un, ux = ctx.to_device(u_n), ctx.to_device(u_next)
    n = dof_mgr.total_dofs()
    if un.numel() != n or ux.numel() != n:
        raise ValueError("displacement length differs from the DOF count")
    out = ctx.empty(6 * mesh.num_elements()).reshape(-1, 6)
    ctx.sync_torch()
    ctx.check(lib().gs_update_stresses_maxwell(
        ctx.handle, mesh.handle(ctx), dof_mgr.handle(ctx), material.youngs_modulus,
        material.poisson_ratio, material.viscosity, float(dt), dptr(un), dptr(ux), dptr(sig),
        dptr(out)), "update_stresses_maxwell")
    return out


class TimeStepStats:
    """src/fem/time_integration.rs:10-20."""

    def __init__(self, time, dt, iterations, residual):
        self.time, self.dt, self.iterations, self.residual = time, dt, iterations, residual


class BackwardEuler:
    """src/fem/time_integration.rs:38-160: K_eff u_{n+1} = f_ext + f_history
    per step, assembled on the device against the cached pattern and
    elimination plan (only the element kernel and the numeric gathers run
    per step)."""

    def __init__(self, n_dofs: int, dt: float, ctx=None):
        if not dt > 0.0:
            raise ValueError(f"Time step must be positive, got {dt}")
        if not n_dofs > 0:
            raise ValueError(f"Number of DOFs must be positive, got {n_dofs}")
        self.ctx = ctx or default_context()
        self.time = 0.0
        self.dt = float(dt)
        self.displacement = self.ctx.empty(int(n_dofs)).zero_()

    @classmethod
    def new(cls, n_dofs, dt, ctx=None):
        return cls(n_dofs, dt, ctx)

    def step(self, mesh, dof_mgr, material: MaxwellViscoelasticity, f_ext, solver):
        """One step t_n -> t_{n+1} -> (u_{n+1} device tensor, TimeStepStats);
        the stress history is not updated here (as in the reference)."""
        from .assembly import Assembler
        K, f = Assembler.assemble_maxwell_system(mesh, dof_mgr, material, self.dt, f_ext,
                                                 ctx=self.ctx)
        u, st = solver.solve(K, f)
        self.displacement = u
        self.time += self.dt
        return u, TimeStepStats(self.time, self.dt, st.iterations, st.residual_norm)

    def current_time(self) -> float:
        return self.time

    def current_displacement(self):
        return self.displacement

    def reset(self):
        self.time = 0.0
        self.displacement.zero_()

    def set_timestep(self, new_dt: float):
        if not new_dt > 0.0:
            raise ValueError(f"Time step must be positive, got {new_dt}")
        self.dt = float(new_dt)
