"""geo_simulator_amd -- MI355X (gfx950) assemble-and-solve path of
geo_simulator: Tet10 element kernels, deterministic CSR assembly with
Dirichlet elimination, and Jacobi-preconditioned CG, in hand-written HIP
behind a C ABI (include/geofem.h, libgeofem.so).

The Python names mirror the reference's Rust API:
    ImprovedMeshGenerator.generate_cube, Mesh, DofManager,
    Assembler.assemble_thermal_stiffness_parallel / assemble_thermal_load /
    assemble_elasticity_stiffness_parallel / assemble_gravity_load /
    apply_dirichlet_bcs, IsotropicElasticity, ConjugateGradient.solve,
    SolverStats, JacobiPreconditioner, ILUPreconditioner, AMGPreconditioner.
Importing the package needs no GPU; compute calls need one and the built
library (there is no CPU fallback).
"""
from ._lib import GeoFemError, LIB_PATH
from .mesh import ImprovedMeshGenerator, Mesh, Tet10Element
from .dof import DofManager
from .linalg import (GMRES, AMGPreconditioner, BiCGSTAB, ConjugateGradient, CsrMatrix,
                     IdentityPreconditioner,
                     ILUPreconditioner, JacobiPreconditioner, LinearOperator, SolverStats,
                     norm)
from .assembly import (Assembler, ElementMatrix, GaussQuadrature, IsotropicElasticity,
                       elastic_element_backward,
                       time_elastic_elements,
                       thermal_element_backward)
from .rheology import (BackwardEuler, MaxwellViscoelasticity, NewtonianViscosity,
                       StressHistory, TimeStepStats, update_stresses_maxwell)
from .runtime import Context, default_context
from .fields import FieldData, ScalarField, VectorField
from .boundary import BoundaryConditions, BoundaryFace, NeumannBC, Tet10Face
from .vtk import VtkWriter

__all__ = [
    "GeoFemError", "LIB_PATH", "ImprovedMeshGenerator", "Mesh", "Tet10Element",
    "DofManager", "ConjugateGradient", "BiCGSTAB", "GMRES", "CsrMatrix", "JacobiPreconditioner",
    "IdentityPreconditioner", "ILUPreconditioner", "AMGPreconditioner", "LinearOperator",
    "SolverStats", "norm", "Assembler", "IsotropicElasticity", "ElementMatrix",
    "GaussQuadrature",
    "thermal_element_backward", "elastic_element_backward", "time_elastic_elements",
    "Context", "FieldData", "ScalarField", "VectorField", "VtkWriter",
    "BoundaryConditions", "BoundaryFace", "NeumannBC", "Tet10Face",
    "default_context", "NewtonianViscosity", "MaxwellViscoelasticity", "StressHistory",
    "update_stresses_maxwell", "BackwardEuler", "TimeStepStats",
]


def build(verbose: bool = False) -> str:
    from .csrc.build import build as _build
    return _build(verbose)
