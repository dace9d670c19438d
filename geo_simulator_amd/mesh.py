"""Mesh container and the structured 6-tet box generator.

Mirrors `Mesh` (src/mesh/geometry.rs:8-11,41-50, src/mesh/topology.rs:13-15,
55-56) and `ImprovedMeshGenerator::generate_cube`
(src/mesh_generator_improved.rs:19-169).  The generator runs in libgeofem's
host code (no GPU needed) and reproduces the reference's node numbering:
vertex grid ix fastest, then edge midpoints in first-touch order over hexes
(iz, iy, ix) x 6 tets x edges (01, 12, 20, 03, 13, 23).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .fields import FieldData

from ._lib import GeoFemError, check, lib


class Tet10Element:
    """`Tet10Element { nodes: [usize; 10] }` (src/mesh/topology.rs:13-15)."""

    __slots__ = ("nodes",)

    def __init__(self, nodes):
        self.nodes = tuple(int(n) for n in nodes)

    def vertices(self):
        return self.nodes[:4]


class Mesh:
    """Nodes as an (n, 3) float64 array (AoS, like Vec<Point3<f64>>) and
    Tet10 connectivity as an (E, 10) int64 array (node order v0..v3, e01,
    e12, e20, e03, e13, e23)."""

    def __init__(self, nodes, elements):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.float64).reshape(-1, 3)
        self.elements = np.ascontiguousarray(elements, dtype=np.int64).reshape(-1, 10)
        self.field_data = FieldData()  # nodal fields (src/mesh/geometry.rs:44)
        self.cell_data = FieldData()   # element fields (src/mesh/geometry.rs:45)
        # per-element stress for viscoelasticity (src/mesh/geometry.rs:47): a
        # rheology.StressHistory or None
        self.stress_history = None
        self._handles = {}

    # reference accessors
    def num_nodes(self) -> int:
        return int(self.nodes.shape[0])

    def num_elements(self) -> int:
        return int(self.elements.shape[0])

    @property
    def tet10_elements(self):
        return [Tet10Element(r) for r in self.elements]

    # device handle (gs_mesh), cached per context
    def handle(self, ctx):
        key = id(ctx)
        h = self._handles.get(key)
        if h is None:
            h = C.c_void_p()
            ctx.check(lib().gs_mesh_create(
                ctx.handle, self.nodes.ctypes.data_as(C.POINTER(C.c_double)),
                self.num_nodes(),
                self.elements.ctypes.data_as(C.POINTER(C.c_int64)),
                self.num_elements(), C.byref(h)), "gs_mesh_create")
            self._handles[key] = (ctx, h)
            return h
        return h[1]

    def __del__(self):
        try:
            for _, (ctx, h) in list(self._handles.items()):
                if ctx.handle:
                    lib().gs_mesh_destroy(h)
            self._handles.clear()
        except Exception:
            pass


class ImprovedMeshGenerator:
    @staticmethod
    def generate_cube(nx: int, ny: int, nz: int, lx: float, ly: float,
                      lz: float) -> Mesh:
        nn, ne = C.c_int64(), C.c_int64()
        st = lib().gs_generate_cube_sizes(nx, ny, nz, C.byref(nn), C.byref(ne))
        if st != 0:
            raise GeoFemError(st, "generate_cube: nx, ny, nz must be positive")
        xyz = np.empty((nn.value, 3))
        conn = np.empty((ne.value, 10), dtype=np.int64)
        check(lib().gs_generate_cube(nx, ny, nz, float(lx), float(ly), float(lz),
                                     xyz.ctypes.data_as(C.POINTER(C.c_double)),
                                     conn.ctypes.data_as(C.POINTER(C.c_int64))),
              None, "gs_generate_cube")
        return Mesh(xyz, conn)
