"""Nodal and cell fields attached to a mesh (src/mesh/fields.rs:6-137).

The reference keeps fields in hash maps, so the order in which its VTU writer
emits them varies from run to run; here they keep insertion order."""
from __future__ import annotations

import numpy as np


class ScalarField:
    """ScalarField::new (src/mesh/fields.rs:11-17)."""

    def __init__(self, name: str, data):
        self.name = str(name)
        self.data = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1))

    def len(self) -> int:
        return int(self.data.shape[0])

    def is_empty(self) -> bool:
        return self.len() == 0


class VectorField:
    """VectorField::new / from_dof_vector (src/mesh/fields.rs:35-67)."""

    def __init__(self, name: str, data):
        self.name = str(name)
        self.data = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1, 3))

    @classmethod
    def from_dof_vector(cls, name: str, dof_vector):
        v = np.asarray(dof_vector, dtype=np.float64).reshape(-1)
        if v.shape[0] % 3 != 0:
            raise ValueError("DOF vector must have 3*n entries for vector field")
        return cls(name, v.reshape(-1, 3))

    def len(self) -> int:
        return int(self.data.shape[0])

    def is_empty(self) -> bool:
        return self.len() == 0


class FieldData:
    """FieldData (src/mesh/fields.rs:80-137), insertion-ordered."""

    def __init__(self):
        self.scalar_fields: dict[str, ScalarField] = {}
        self.vector_fields: dict[str, VectorField] = {}

    def add_field(self, field: ScalarField):
        self.scalar_fields[field.name] = field

    def add_vector_field(self, field: VectorField):
        self.vector_fields[field.name] = field

    def get_field(self, name: str):
        return self.scalar_fields.get(name)

    def get_vector_field(self, name: str):
        return self.vector_fields.get(name)

    def field_names(self):
        return list(self.scalar_fields)

    def vector_field_names(self):
        return list(self.vector_fields)

    def num_fields(self) -> int:
        return len(self.scalar_fields)

    def num_vector_fields(self) -> int:
        return len(self.vector_fields)

    def is_empty(self) -> bool:
        return not self.scalar_fields and not self.vector_fields
