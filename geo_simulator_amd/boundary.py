"""Neumann boundary data (src/fem/boundary.rs:4-115).

BoundaryConditions keeps faces in insertion order (the reference's hash map
iterates in a run-dependent order, which only changes the summation order of
the flux contributions)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# element-local node of face node k: vertices, then the triangle's edges
# (01, 12, 20) (src/fem/boundary.rs:72-115)
FACE_NODES = np.array([[1, 2, 3, 5, 9, 8],
                       [0, 2, 3, 6, 9, 7],
                       [0, 1, 3, 4, 8, 7],
                       [0, 1, 2, 4, 5, 6]], dtype=np.int64)


@dataclass(frozen=True)
class BoundaryFace:
    """BoundaryFace (src/fem/boundary.rs:7-12)."""
    element_id: int
    local_face_id: int


@dataclass
class NeumannBC:
    """NeumannBC (src/fem/boundary.rs:15-19): flux (positive into the body)."""
    flux: float


class BoundaryConditions:
    """BoundaryConditions (src/fem/boundary.rs:22-56)."""

    def __init__(self):
        self._neumann: dict[BoundaryFace, NeumannBC] = {}

    @classmethod
    def new(cls):
        return cls()

    def add_neumann(self, face: BoundaryFace, flux: float):
        self._neumann[face] = NeumannBC(float(flux))

    def get_neumann(self, face: BoundaryFace):
        return self._neumann.get(face)

    def neumann_faces(self):
        return iter(self._neumann.items())

    def num_neumann_bcs(self) -> int:
        return len(self._neumann)


class Tet10Face:
    """Tet10Face::from_element_face (src/fem/boundary.rs:60-115)."""

    def __init__(self, nodes):
        self.nodes = np.asarray(nodes, dtype=np.int64).reshape(6)

    @classmethod
    def from_element_face(cls, element_nodes, local_face: int):
        if not 0 <= int(local_face) <= 3:
            raise ValueError(f"Invalid face index for tet10: {local_face}")
        return cls(np.asarray(element_nodes, dtype=np.int64)[FACE_NODES[int(local_face)]])
