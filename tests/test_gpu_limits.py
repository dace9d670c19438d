"""Size limits of the C ABI.  Counts past the library's int32 indexing are
refused with GS_ERR_TOO_LARGE before any array is read or allocated, so the
arrays passed with them are deliberately tiny: the mesh (100 * n_elems slot
entries, 3 * n_nodes coordinates), the DOF map (n_nodes * dofs_per_node) and
a host CSR (rows, cols, nnz).  The same context then accepts a small object
of each kind."""
import ctypes as C

import numpy as np
import pytest

from geo_simulator_amd import _lib
from geo_simulator_amd._lib import lib
from geo_simulator_amd.runtime import default_context

pytestmark = pytest.mark.gpu

INT32_MAX = 2**31 - 1


def test_mesh_too_large():
    ctx = default_context()
    xyz = np.zeros(30, dtype=np.float64)
    conn = np.zeros(10, dtype=np.int64)
    out = C.c_void_p()
    dp = xyz.ctypes.data_as(C.POINTER(C.c_double))
    cp = conn.ctypes.data_as(C.POINTER(C.c_int64))
    # 100 * n_elems reaches 2^31: refused before conn is read
    n_elems = (INT32_MAX + 99) // 100
    st = lib().gs_mesh_create(ctx.handle, dp, 10, cp, n_elems, C.byref(out))
    assert st == _lib.GS_ERR_TOO_LARGE and not out.value
    st = lib().gs_mesh_create(ctx.handle, dp, INT32_MAX // 3, cp, 1, C.byref(out))
    assert st == _lib.GS_ERR_TOO_LARGE and not out.value
    # a real one-element mesh still goes through on the same context
    st = lib().gs_mesh_create(ctx.handle, dp, 10, cp, 1, C.byref(out))
    assert st == _lib.GS_OK and out.value
    lib().gs_mesh_destroy(out)


def test_dofmap_too_large():
    ctx = default_context()
    mask = np.zeros(3, dtype=np.uint8)
    vals = np.zeros(3, dtype=np.float64)
    out = C.c_void_p()
    mp = mask.ctypes.data_as(C.POINTER(C.c_uint8))
    vp = vals.ctypes.data_as(C.POINTER(C.c_double))
    st = lib().gs_dofmap_create(ctx.handle, INT32_MAX // 3 + 1, 3, mp, vp, C.byref(out))
    assert st == _lib.GS_ERR_TOO_LARGE and not out.value
    st = lib().gs_dofmap_create(ctx.handle, 1, 3, mp, vp, C.byref(out))
    assert st == _lib.GS_OK and out.value
    lib().gs_dofmap_destroy(out)


def test_csr_too_large():
    ctx = default_context()
    indptr = np.array([0, 1], dtype=np.int64)
    indices = np.array([0], dtype=np.int64)
    data = np.array([2.0])
    out = C.c_void_p()
    ip = indptr.ctypes.data_as(C.POINTER(C.c_int64))
    jp = indices.ctypes.data_as(C.POINTER(C.c_int64))
    dp = data.ctypes.data_as(C.POINTER(C.c_double))
    for rows, cols, nnz in ((1, 1, INT32_MAX), (INT32_MAX, 1, 1), (1, INT32_MAX, 1)):
        st = lib().gs_csr_from_host(ctx.handle, rows, cols, nnz, ip, jp, dp, C.byref(out))
        assert st == _lib.GS_ERR_TOO_LARGE and not out.value, (rows, cols, nnz)
    st = lib().gs_csr_from_host(ctx.handle, 1, 1, 1, ip, jp, dp, C.byref(out))
    assert st == _lib.GS_OK and out.value
    lib().gs_csr_destroy(out)
