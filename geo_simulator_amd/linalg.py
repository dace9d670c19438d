"""Device-resident CSR matrix and Jacobi-preconditioned CG.

Mirrors `CsMat<f64>` + `impl LinearOperator for CsMat<f64>`
(src/linalg/solver.rs:61-92), `SolverStats` (src/linalg/solver.rs:5-20),
`JacobiPreconditioner` (src/linalg/preconditioner.rs:29-62),
`ILUPreconditioner` (src/linalg/preconditioner.rs:78-204),
`AMGPreconditioner` (src/linalg/amg.rs) and
`ConjugateGradient` (src/linalg/iterative.rs:7-147).  Vectors may be numpy
arrays (host, like Vec<f64>: results come back as numpy) or torch tensors on
the device (results stay on the device).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import lib
from .runtime import default_context, dptr


def _is_torch(x):
    try:
        import torch
        return isinstance(x, torch.Tensor)
    except ImportError:  # pragma: no cover
        return False


class CsrMatrix:
    """A CSR matrix owned by libgeofem on one device (gs_csr)."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.handle = handle
        r, c, z = C.c_int64(), C.c_int64(), C.c_int64()
        ctx.check(lib().gs_csr_info(handle, C.byref(r), C.byref(c), C.byref(z)),
                  "gs_csr_info")
        self.shape = (r.value, c.value)
        self._nnz = z.value

    @classmethod
    def from_scipy(cls, A, ctx=None):
        ctx = ctx or default_context()
        A = sp.csr_matrix(A)
        A.sum_duplicates()
        A.sort_indices()
        indptr = np.ascontiguousarray(A.indptr, dtype=np.int64)
        indices = np.ascontiguousarray(A.indices, dtype=np.int64)
        data = np.ascontiguousarray(A.data, dtype=np.float64)
        h = C.c_void_p()
        ctx.check(lib().gs_csr_from_host(
            ctx.handle, A.shape[0], A.shape[1], A.nnz,
            indptr.ctypes.data_as(C.POINTER(C.c_int64)),
            indices.ctypes.data_as(C.POINTER(C.c_int64)),
            data.ctypes.data_as(C.POINTER(C.c_double)), C.byref(h)),
            "gs_csr_from_host")
        return cls(ctx, h)

    # sprs-like accessors
    def rows(self) -> int:
        return self.shape[0]

    def cols(self) -> int:
        return self.shape[1]

    def nnz(self) -> int:
        return self._nnz

    def to_scipy(self) -> sp.csr_matrix:
        """Export in the sprs shape: int64 indptr/indices, sorted columns."""
        indptr = np.empty(self.shape[0] + 1, dtype=np.int64)
        indices = np.empty(self._nnz, dtype=np.int64)
        data = np.empty(self._nnz, dtype=np.float64)
        self.ctx.check(lib().gs_csr_export(
            self.ctx.handle, self.handle,
            indptr.ctypes.data_as(C.POINTER(C.c_int64)),
            indices.ctypes.data_as(C.POINTER(C.c_int64)),
            data.ctypes.data_as(C.POINTER(C.c_double))), "gs_csr_export")
        return sp.csr_matrix((data, indices, indptr), shape=self.shape)

    def apply(self, v):
        """LinearOperator::apply: y = A v (src/linalg/solver.rs:62-73)."""
        ctx = self.ctx
        host = not _is_torch(v)
        x = ctx.to_device(v)
        if x.numel() != self.shape[1]:
            raise ValueError("vector length differs from the matrix columns")
        y = ctx.empty(self.shape[0])
        ctx.sync_torch()
        ctx.check(lib().gs_csr_spmv(ctx.handle, self.handle, dptr(x), dptr(y)),
                  "gs_csr_spmv")
        return y.cpu().numpy() if host else y

    def jacobi_diag_inv(self):
        d = self.ctx.empty(self.shape[0])
        self.ctx.sync_torch()
        self.ctx.check(lib().gs_jacobi_diag_inv(self.ctx.handle, self.handle,
                                                dptr(d)), "gs_jacobi_diag_inv")
        return d

    def set_spmv_shape(self, lanes: int = 0, workgroups: int = 0, flags: int = 0):
        """Launch-shape override for this matrix's SpMV (0 = automatic)."""
        st = lib().gs_csr_set_spmv_shape(self.handle, lanes, workgroups, flags)
        if st != 0:
            raise _lib.GeoFemError(st, "gs_csr_set_spmv_shape: invalid shape")

    def time_spmv(self, precision: str = "fp64", reps: int = 20) -> float:
        """Average duration (ms) of the CG SpMV kernel over `reps` launches."""
        out = C.c_double()
        prec = _lib.GS_PREC_FP64 if precision == "fp64" else _lib.GS_PREC_MIXED
        self.ctx.check(lib().gs_time_spmv(self.ctx.handle, self.handle, prec,
                                          reps, C.byref(out)), "gs_time_spmv")
        return out.value

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                lib().gs_csr_destroy(self.handle)
            self.handle = None
        except Exception:
            pass


@dataclass
class SolverStats:
    iterations: int
    residual_norm: float
    relative_residual: float
    converged: bool
    solve_time: float
    outer_iterations: int = 0
    spmv_count: int = 0


class JacobiPreconditioner:
    """M = diag(A): d_i = 1/A_ii if |A_ii| > 1e-14 else 1
    (src/linalg/preconditioner.rs:37-51)."""

    def __init__(self, A: CsrMatrix):
        self.diag_inv = A.jacobi_diag_inv()

    def apply(self, r):
        if _is_torch(r):
            return r * self.diag_inv
        return np.asarray(r) * self.diag_inv.cpu().numpy()


class ILUPreconditioner:
    """ILU(0) with zero fill (src/linalg/preconditioner.rs:78-204), factored
    and applied on the device, bit-identical to the reference's loops (rows
    in the caller's numbering).  `ILUPreconditioner(A)` is
    `ILUPreconditioner::new(&A)`; where the reference returns Err (a row
    without a diagonal, a pivot below 1e-15) this raises GeoFemError
    (GS_ERR_ZERO_PIVOT) with the reference's message."""

    def __init__(self, A: CsrMatrix):
        self.ctx = A.ctx
        self.handle = None
        h = C.c_void_p()
        self.ctx.sync_torch()
        self.ctx.check(lib().gs_ilu_create(self.ctx.handle, A.handle, C.byref(h)),
                       "ILUPreconditioner::new")
        self.handle = h
        n, z, lo, up = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        lib().gs_ilu_info(h, C.byref(n), C.byref(z), C.byref(lo), C.byref(up))
        self.shape = (n.value, n.value)
        self._nnz = z.value
        self.levels = (lo.value, up.value)

    @classmethod
    def new(cls, A: CsrMatrix) -> "ILUPreconditioner":
        return cls(A)

    def update(self, A: CsrMatrix) -> None:
        """ILUPreconditioner::update (:110-124): refactor with A's values."""
        self.ctx.sync_torch()
        self.ctx.check(lib().gs_ilu_update(self.ctx.handle, self.handle, A.handle),
                       "ILUPreconditioner::update")

    def apply(self, r):
        """z = U^-1 L^-1 r (:168-203); numpy in, numpy out; a device tensor
        in, a device tensor out."""
        ctx = self.ctx
        host = not _is_torch(r)
        x = ctx.to_device(r)
        if x.numel() != self.shape[0]:
            raise ValueError("vector length differs from the factored matrix")
        z = ctx.empty(self.shape[0])
        ctx.sync_torch()
        ctx.check(lib().gs_ilu_apply(ctx.handle, self.handle, dptr(x), dptr(z)),
                  "ILUPreconditioner::apply")
        return z.cpu().numpy() if host else z

    def factors(self) -> sp.csr_matrix:
        """The `factors` field: L (unit diagonal implied) below, U on and above
        the diagonal, on A's pattern."""
        n = self.shape[0]
        indptr = np.empty(n + 1, dtype=np.int64)
        indices = np.empty(self._nnz, dtype=np.int64)
        data = np.empty(self._nnz, dtype=np.float64)
        self.ctx.check(lib().gs_ilu_export(
            self.ctx.handle, self.handle, indptr.ctypes.data_as(C.POINTER(C.c_int64)),
            indices.ctypes.data_as(C.POINTER(C.c_int64)),
            data.ctypes.data_as(C.POINTER(C.c_double))), "gs_ilu_export")
        return sp.csr_matrix((data, indices, indptr), shape=self.shape)

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                lib().gs_ilu_destroy(self.handle)
            self.handle = None
        except Exception:
            pass


class AMGPreconditioner:
    """Algebraic multigrid (src/linalg/amg.rs): `AMGPreconditioner(A,
    max_levels, coarse_size, strength_threshold)` is
    `AMGPreconditioner::new(&A, ...)` (:44-129; defaults 10, 500, 0.25), the
    hierarchy built as the reference builds it (in A's caller numbering,
    bit-identical levels); `apply(r)` is one V-cycle (:132-192) on the
    device, bit-identical to the reference's."""

    def __init__(self, A: CsrMatrix, max_levels: int = 10, coarse_size: int = 500,
                 strength_threshold: float = 0.25):
        self.ctx = A.ctx
        self.handle = None
        h = C.c_void_p()
        self.ctx.sync_torch()
        self.ctx.check(lib().gs_amg_create(self.ctx.handle, A.handle, int(max_levels),
                                           int(coarse_size), float(strength_threshold),
                                           C.byref(h)), "AMGPreconditioner::new")
        self.handle = h
        nl = C.c_int64()
        lib().gs_amg_levels(h, C.byref(nl))
        self.levels = []
        for l in range(nl.value):
            v = [C.c_int64() for _ in range(5)]
            lib().gs_amg_level_info(h, l, *[C.byref(x) for x in v])
            self.levels.append(dict(rows=v[0].value, nnz=v[1].value, coarse=v[2].value,
                                    nnz_p=v[3].value, sweep_levels=v[4].value))
        self.shape = (A.shape[0], A.shape[0])

    @classmethod
    def new(cls, A: CsrMatrix, max_levels: int = 10, coarse_size: int = 500,
            strength_threshold: float = 0.25) -> "AMGPreconditioner":
        return cls(A, max_levels, coarse_size, strength_threshold)

    def apply(self, r):
        """z = v_cycle(r, 0); numpy in, numpy out; a device tensor in, a
        device tensor out."""
        ctx = self.ctx
        host = not _is_torch(r)
        x = ctx.to_device(r)
        if x.numel() != self.shape[0]:
            raise ValueError("vector length differs from the matrix")
        z = ctx.empty(self.shape[0])
        ctx.sync_torch()
        ctx.check(lib().gs_amg_apply(ctx.handle, self.handle, dptr(x), dptr(z)),
                  "AMGPreconditioner::apply")
        return z.cpu().numpy() if host else z

    def _export(self, level, which):
        info = self.levels[level]
        rows = info["rows"]
        if which == 2:
            d = np.empty(rows)
            self.ctx.check(lib().gs_amg_export(self.ctx.handle, self.handle, level, 2, None,
                                               None, d.ctypes.data_as(C.POINTER(C.c_double))),
                           "gs_amg_export")
            return d
        nnz = info["nnz"] if which == 0 else info["nnz_p"]
        cols = rows if which == 0 else info["coarse"]
        indptr = np.empty(rows + 1, dtype=np.int64)
        indices = np.empty(nnz, dtype=np.int64)
        data = np.empty(nnz, dtype=np.float64)
        self.ctx.check(lib().gs_amg_export(
            self.ctx.handle, self.handle, level, which,
            indptr.ctypes.data_as(C.POINTER(C.c_int64)),
            indices.ctypes.data_as(C.POINTER(C.c_int64)),
            data.ctypes.data_as(C.POINTER(C.c_double))), "gs_amg_export")
        return sp.csr_matrix((data, indices, indptr), shape=(rows, cols))

    def operator(self, level: int) -> sp.csr_matrix:
        """AMGLevel::a of `level` (caller numbering on level 0)."""
        return self._export(level, 0)

    def prolongation(self, level: int):
        """AMGLevel::p of `level` (None where the level has none)."""
        return self._export(level, 1) if self.levels[level]["coarse"] > 0 else None

    def diag_inv(self, level: int) -> np.ndarray:
        """AMGLevel::diag_inv of `level`."""
        return self._export(level, 2)

    def __del__(self):
        try:
            if self.handle and self.ctx.handle:
                lib().gs_amg_destroy(self.handle)
            self.handle = None
        except Exception:
            pass


class IdentityPreconditioner:
    """M = I (src/linalg/preconditioner.rs:65-71)."""

    def apply(self, r):
        return r.clone() if _is_torch(r) else np.array(r, dtype=np.float64)


class LinearOperator:
    """A caller's operator for `solve_with_operator` (the LinearOperator trait,
    src/linalg/solver.rs:29-59): `apply(v)` maps a device tensor of `cols`
    doubles to one of `rows`.  Any object with `apply` and `shape` works;
    this class wraps a function."""

    def __init__(self, fn, n):
        self.fn = fn
        self.shape = (int(n), int(n))

    def apply(self, v):
        return self.fn(v)

    def rows(self):
        return self.shape[0]

    def cols(self):
        return self.shape[1]


def _operator_callback(obj, tin, tout, errors):
    """An OPERATOR_FN that applies obj to the caller buffer tin into tout,
    synchronously (the library waits for nothing else)."""
    import torch

    def run(_user):
        try:
            y = obj.apply(tin)
            if not _is_torch(y):
                y = torch.as_tensor(np.asarray(y, dtype=np.float64), device=tout.device)
            if y.shape != tout.shape:
                raise ValueError(f"operator returned shape {tuple(y.shape)}, "
                                 f"expected {tuple(tout.shape)}")
            tout.copy_(y)
            torch.cuda.synchronize(tout.device)
            return 0
        except Exception as e:  # reported after the C call returns
            errors.append(e)
            return 1
    return _lib.OPERATOR_FN(run)


class ConjugateGradient:
    """Builder and solver with the reference's defaults
    (max_iterations 1000, tolerance 1e-8, abs_tolerance 1e-12, Jacobi on).

    Extensions: `with_precision("mixed")` runs FP64 flexible CG around an
    FP32 Jacobi-PCG recursion (residual, solution and search directions in
    FP64; the FP32 recursion reduces its residual by `inner_tolerance` between
    FP64 steps; 0, the default, picks 1e-3 for the Jacobi-scaled fixed-point
    operator of a scalar matrix and 1e-2 otherwise), stopping on the same FP64 criterion ||b - A x|| < tol ||b||
    (or < abs_tol)."""

    name = "ConjugateGradient"

    def __init__(self):
        self.max_iterations = 1000
        self.tolerance = 1e-8
        self.abs_tolerance = 1e-12
        self.use_preconditioner = True
        self.precision = "fp64"
        self.inner_tolerance = 0.0
        self.max_inner_iterations = 0
        self.max_outer_iterations = 50
        self.relative_breakdown = False

    @classmethod
    def new(cls):
        return cls()

    def with_max_iterations(self, n):
        self.max_iterations = int(n)
        return self

    def with_tolerance(self, t):
        self.tolerance = float(t)
        return self

    def with_abs_tolerance(self, t):
        self.abs_tolerance = float(t)
        return self

    def with_preconditioner(self, use):
        self.use_preconditioner = bool(use)
        return self

    def with_precision(self, precision: str):
        if precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.precision = precision
        return self

    def with_inner_tolerance(self, t):
        self.inner_tolerance = float(t)
        return self

    def with_max_outer_iterations(self, n):
        self.max_outer_iterations = int(n)
        return self

    def with_max_inner_iterations(self, n):
        self.max_inner_iterations = int(n)
        return self

    def with_relative_breakdown(self, on=True):
        """FP64: stop on p.Ap == 0 or a non-finite step instead of the
        reference's |p.Ap| < 1e-30, whose meaning depends on the units of b
        (the mixed path always uses the scale-free test)."""
        self.relative_breakdown = bool(on)
        return self

    def opts(self) -> _lib.CgOpts:
        o = _lib.CgOpts()
        lib().gs_cg_opts_default(C.byref(o))
        o.max_iterations = self.max_iterations
        o.tolerance = self.tolerance
        o.abs_tolerance = self.abs_tolerance
        o.use_preconditioner = 1 if self.use_preconditioner else 0
        o.precision = _lib.GS_PREC_FP64 if self.precision == "fp64" else _lib.GS_PREC_MIXED
        o.inner_tolerance = self.inner_tolerance
        o.max_inner_iterations = self.max_inner_iterations
        o.max_outer_iterations = self.max_outer_iterations
        o.relative_breakdown = 1 if self.relative_breakdown else 0
        return o

    def solve(self, A, b):
        """Solver::solve (src/linalg/iterative.rs:123-131) -> (x, SolverStats)."""
        return _krylov_solve("gs_cg_solve", self.opts(), A, b)

    def solve_with_operator(self, A, b, precond=None):
        """Solver::solve_with_operator (src/linalg/solver.rs:112-120,
        src/linalg/iterative.rs:133-140): the reference's CG loop on any
        operator `A` (a CsrMatrix, a LinearOperator, or anything with `apply`
        on device tensors and `shape`) with any preconditioner (`apply`; None
        or IdentityPreconditioner: M = I).  The vector work and the dot
        products run in the library on the device; A and precond are called
        back on device tensors.  FP64; `with_preconditioner` does not apply
        (precond decides).  -> (x, SolverStats); x is a numpy array for a host
        right-hand side, else a device tensor."""
        if self.precision != "fp64":
            raise ValueError("solve_with_operator runs in FP64")
        ctx = A.ctx if isinstance(A, CsrMatrix) else default_context()
        n = int(A.shape[0])
        if int(A.shape[1]) != n:
            raise ValueError("CG needs a square operator")
        host = not _is_torch(b)
        bd = ctx.to_device(b)
        if bd.numel() != n:
            raise ValueError("right-hand side length differs from the operator rows")
        x = ctx.empty(n)
        tin, tout = ctx.empty(n), ctx.empty(n)
        errors = []
        f_apply = _operator_callback(A, tin, tout, errors)
        f_prec = (_lib.OPERATOR_FN()  # NULL: the identity, applied by the library
                  if precond is None or isinstance(precond, IdentityPreconditioner)
                  else _operator_callback(precond, tin, tout, errors))
        st = _lib.SolverStatsC()
        o = self.opts()
        ctx.sync_torch()
        status = lib().gs_cg_solve_operator(ctx.handle, n, f_apply, None, f_prec, None,
                                            dptr(tin), dptr(tout), dptr(bd), C.byref(o),
                                            dptr(x), C.byref(st))
        if errors:
            raise errors[0]
        ctx.check(status, "gs_cg_solve_operator")
        stats = SolverStats(int(st.iterations), float(st.residual_norm),
                            float(st.relative_residual), bool(st.converged),
                            float(st.solve_time), 0, int(st.spmv_count))
        return (x.cpu().numpy() if host else x), stats

    # Solver trait accessors (src/linalg/solver.rs:122-135)
    def set_tolerance(self, t):
        self.tolerance = float(t)

    def set_abs_tolerance(self, t):
        self.abs_tolerance = float(t)


def time_stream(nbytes: int, kind: str = "read", reps: int = 10, ctx=None) -> float:
    """Average milliseconds of one streaming pass over nbytes of device memory:
    "read" sums the values (16-byte loads), "copy" copies them."""
    ctx = ctx or default_context()
    out = C.c_double()
    ctx.check(lib().gs_time_stream(ctx.handle, int(nbytes), 0 if kind == "read" else 1,
                                   int(reps), C.byref(out)), "gs_time_stream")
    return out.value


def norm(v) -> float:
    """SolverUtils::norm (src/linalg/solver.rs:153-155)."""
    return float(np.sqrt(np.sum(np.asarray(v, dtype=np.float64) ** 2)))


def _krylov_solve(entry, o, A, b, *extra):
    """Run the C ABI solver `entry` -> (x, SolverStats); x is a numpy array
    for a host right-hand side, else a device tensor."""
    if not isinstance(A, CsrMatrix):
        A = CsrMatrix.from_scipy(A)
    ctx = A.ctx
    host = not _is_torch(b)
    bd = ctx.to_device(b)
    if bd.numel() != A.shape[0]:
        raise ValueError("right-hand side length differs from the matrix rows")
    x = ctx.empty(A.shape[0])
    st = _lib.SolverStatsC()
    ctx.sync_torch()
    ctx.check(getattr(lib(), entry)(ctx.handle, A.handle, dptr(bd), C.byref(o),
                                    *extra, dptr(x), C.byref(st)), entry)
    stats = SolverStats(int(st.iterations), float(st.residual_norm),
                        float(st.relative_residual), bool(st.converged),
                        float(st.solve_time), int(st.outer_iterations),
                        int(st.spmv_count))
    return (x.cpu().numpy() if host else x), stats


class BiCGSTAB:
    """BiCGSTAB for general square matrices with the reference's builder and
    defaults (max_iterations 1000, tolerance 1e-8, abs_tolerance 1e-12,
    Jacobi on; src/linalg/iterative.rs:149-176), FP64 on the SELL operator."""

    name = "BiCGSTAB"

    def __init__(self):
        self.max_iterations = 1000
        self.tolerance = 1e-8
        self.abs_tolerance = 1e-12
        self.use_preconditioner = True
        self.verbose = False

    @classmethod
    def new(cls):
        return cls()

    def with_max_iterations(self, n):
        self.max_iterations = int(n)
        return self

    def with_tolerance(self, t):
        self.tolerance = float(t)
        return self

    def with_abs_tolerance(self, t):
        self.abs_tolerance = float(t)
        return self

    def with_preconditioner(self, use):
        self.use_preconditioner = bool(use)
        return self

    def with_verbose(self, verbose):
        """Accepted for the reference's builder; progress is not printed."""
        self.verbose = bool(verbose)
        return self

    def opts(self) -> _lib.CgOpts:
        o = _lib.CgOpts()
        lib().gs_cg_opts_default(C.byref(o))
        o.max_iterations = self.max_iterations
        o.tolerance = self.tolerance
        o.abs_tolerance = self.abs_tolerance
        o.use_preconditioner = 1 if self.use_preconditioner else 0
        return o

    def solve(self, A, b):
        """Solver::solve (src/linalg/iterative.rs:283-291) -> (x, SolverStats)."""
        return _krylov_solve("gs_bicgstab_solve", self.opts(), A, b)

    # Solver trait accessors (src/linalg/iterative.rs:293-299)
    def set_tolerance(self, t):
        self.tolerance = float(t)

    def set_abs_tolerance(self, t):
        self.abs_tolerance = float(t)


class GMRES(BiCGSTAB):
    """Restarted GMRES with the reference's builder and defaults (restart 10,
    max_iterations 1000, tolerance 1e-8, abs_tolerance 1e-12, Jacobi on;
    src/linalg/iterative.rs:302-331), FP64 on the SELL operator; iterations
    count inner steps over all cycles."""

    name = "GMRES"

    def __init__(self):
        super().__init__()
        self.restart = 10

    def with_restart(self, m):
        self.restart = int(m)
        return self

    def solve(self, A, b):
        """Solver::solve (src/linalg/iterative.rs:476-486) -> (x, SolverStats)."""
        return _krylov_solve("gs_gmres_solve", self.opts(), A, b, int(self.restart))
