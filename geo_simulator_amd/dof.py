"""DOF numbering and Dirichlet bookkeeping.

Mirrors `DofManager` (src/fem/dof.rs:8-166) for a fixed number of DOFs per
node: global DOF = node * dofs_per_node + component, a Dirichlet mask and a
dense value array.  The device copy (gs_dofmap) is rebuilt whenever the
Dirichlet set changes.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import lib


class DofManager:
    def __init__(self, num_nodes: int, dofs_per_node: int):
        if num_nodes < 0 or dofs_per_node < 1:
            raise ValueError("invalid DOF manager size")
        self._num_nodes = int(num_nodes)
        self._dpn = int(dofs_per_node)
        n = self._num_nodes * self._dpn
        self.dirichlet_mask = np.zeros(n, dtype=np.uint8)
        self.dirichlet_values = np.zeros(n, dtype=np.float64)
        self._version = 0
        self._handles = {}

    # src/fem/dof.rs:87-106
    def velocity_dof(self, node_id: int, component: int) -> int:
        return node_id * self._dpn + component

    def global_dof(self, node_id: int, local_dof: int) -> int:
        if not (0 <= local_dof < self._dpn):
            raise IndexError("Node has no pressure DOF")
        return node_id * self._dpn + local_dof

    # src/fem/dof.rs:113-135
    def set_dirichlet(self, dof: int, value: float):
        self.dirichlet_mask[dof] = 1
        self.dirichlet_values[dof] = value
        self._version += 1

    def set_dirichlet_node(self, node_id: int, value: float):
        for c in range(self._dpn):
            self.set_dirichlet(self.global_dof(node_id, c), value)

    def set_dirichlet_many(self, dofs, values):
        """Vectorised set_dirichlet over arrays of DOFs and values."""
        dofs = np.asarray(dofs, dtype=np.int64)
        self.dirichlet_mask[dofs] = 1
        self.dirichlet_values[dofs] = np.broadcast_to(
            np.asarray(values, dtype=np.float64), dofs.shape)
        self._version += 1

    def is_dirichlet(self, dof: int) -> bool:
        return bool(self.dirichlet_mask[dof])

    def get_dirichlet_value(self, dof: int) -> float:
        return float(self.dirichlet_values[dof])

    # src/fem/dof.rs:143-165
    def total_dofs(self) -> int:
        return self._num_nodes * self._dpn

    def num_free_dofs(self) -> int:
        return self.total_dofs() - int(self.dirichlet_mask.sum())

    def num_constrained_dofs(self) -> int:
        return int(self.dirichlet_mask.sum())

    def dofs_per_node(self) -> int:
        return self._dpn

    def num_nodes(self) -> int:
        return self._num_nodes

    def handle(self, ctx):
        key = id(ctx)
        ent = self._handles.get(key)
        if ent is not None and ent[2] == self._version:
            return ent[1]
        if ent is not None:
            lib().gs_dofmap_destroy(ent[1])
        h = C.c_void_p()
        ctx.check(lib().gs_dofmap_create(
            ctx.handle, self._num_nodes, self._dpn,
            self.dirichlet_mask.ctypes.data_as(C.POINTER(C.c_uint8)),
            self.dirichlet_values.ctypes.data_as(C.POINTER(C.c_double)),
            C.byref(h)), "gs_dofmap_create")
        self._handles[key] = (ctx, h, self._version)
        return h

    def __del__(self):
        try:
            for _, (ctx, h, _v) in list(self._handles.items()):
                if ctx.handle:
                    lib().gs_dofmap_destroy(h)
            self._handles.clear()
        except Exception:
            pass
