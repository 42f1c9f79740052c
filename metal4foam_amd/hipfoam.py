"""Python host layer over the C-ABI (include/hipfoam.h).

``Context`` owns one library context (one per mesh and field, like the
reference's backend cache, src/metal4Foam/solvers/metalFoamSolver.C:57-74).
``hipPCG`` / ``hipPBiCGStab`` mirror the reference's run-time-selected solver
classes (metalPCG.H:47-92, metalPBiCGStab.H:36-81): constructed from a field
name, a matrix and the fvSolution controls, ``solve(psi, source)`` updates psi
in place and returns the solverPerformance.
"""

from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np

from . import _lib
from ._lib import Controls, Performance, Schedule, Stats

PCG, PBICGSTAB = _lib.PCG, _lib.PBICGSTAB
FP32, FP64 = _lib.FP32, _lib.FP64
TOPO_POINTERS, TOPO_CONTENT = _lib.TOPO_POINTERS, _lib.TOPO_CONTENT


class HipFoamError(RuntimeError):
    pass


def _ptr(a, t=None):
    if a is None:
        return None
    if t is None:
        return C.c_void_p(a.ctypes.data)
    return a.ctypes.data_as(t)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def default_controls(**kw) -> Controls:
    c = Controls()
    _lib.load().hipfoam_default_controls(C.byref(c))
    for k, v in kw.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


class DeviceArray:
    """A device buffer owned through the library (hipfoam_device_alloc)."""

    def __init__(self, ctx: "Context", shape, dtype):
        self.ctx = ctx
        self.shape = tuple(np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        ctx._check(ctx.lib.hipfoam_device_alloc(ctx.h, self.nbytes, C.byref(p)),
                   "device_alloc")
        self.ptr = p

    @classmethod
    def from_host(cls, ctx, a):
        a = np.ascontiguousarray(a)
        d = cls(ctx, a.shape, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        self.ctx._check(self.ctx.lib.hipfoam_copy_to_device(
            self.ctx.h, self.ptr, C.c_void_p(a.ctypes.data), self.nbytes), "copy_to_device")

    def download(self):
        out = np.empty(self.shape, self.dtype)
        self.ctx._check(self.ctx.lib.hipfoam_copy_to_host(
            self.ctx.h, C.c_void_p(out.ctypes.data), self.ptr, self.nbytes), "copy_to_host")
        return out

    def copy_from(self, other: "DeviceArray"):
        """Device-to-device copy on the context stream (asynchronous)."""
        assert other.nbytes == self.nbytes
        self.ctx._check(self.ctx.lib.hipfoam_copy_on_device(
            self.ctx.h, self.ptr, other.ptr, self.nbytes), "copy_on_device")

    def free(self):
        if self.ptr:
            self.ctx.lib.hipfoam_device_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One hipfoam context (metalFoamBackend equivalent)."""

    def __init__(self, device: int = 0, require: bool = True):
        self.lib = _lib.load()
        h = C.c_void_p()
        rc = self.lib.hipfoam_create(device, C.byref(h))
        self.h = h
        self.n = 0
        self._keep = []
        if require and not self.good:
            raise HipFoamError(f"hipfoam_create({device}) failed rc={rc}: "
                               f"{self.description}")

    # -- status -----------------------------------------------------------
    @property
    def good(self) -> bool:
        return bool(self.h) and bool(self.lib.hipfoam_good(self.h))

    @property
    def description(self) -> str:
        return self.lib.hipfoam_description(self.h).decode()

    @property
    def last_error(self) -> str:
        return self.lib.hipfoam_last_error(self.h).decode()

    def _check(self, rc, what):
        if rc != 0:
            raise HipFoamError(f"{what} failed rc={rc}: {self.last_error}")

    def close(self):
        if self.h:
            self.lib.hipfoam_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- backend interface ------------------------------------------------
    def set_topology(self, m):
        arrs = [_i32(m.lower_addr), _i32(m.upper_addr), _i32(m.owner_start),
                _i32(m.losort), _i32(m.losort_start)]
        self._keep = arrs  # pointer identity is the cache key
        lo, up, os_, ls, lss = arrs
        i32p = C.POINTER(C.c_int32)
        self._check(self.lib.hipfoam_set_topology(
            self.h, m.n, lo.shape[0], _ptr(lo, i32p), _ptr(up, i32p),
            _ptr(os_, i32p), _ptr(ls, i32p), _ptr(lss, i32p)), "set_topology")
        self.n = m.n

    def set_topology_check(self, mode):
        """TOPO_CONTENT (default): set_topology re-uploads whenever any
        addressing entry changed; TOPO_POINTERS: only when the sizes or the
        lowAddr/uppAddr pointers changed (the reference's key)."""
        self._check(self.lib.hipfoam_set_topology_check(self.h, int(mode)),
                    "set_topology_check")

    def set_coeffs(self, diag, upper, lower=None):
        f64p = C.POINTER(C.c_double)
        d, u, lo = _f64(diag), _f64(upper), _f64(lower)
        t = C.c_double()
        self._check(self.lib.hipfoam_set_coeffs(self.h, _ptr(d, f64p), _ptr(u, f64p),
                                                _ptr(lo, f64p), C.byref(t)), "set_coeffs")
        return t.value

    def set_matrix(self, diag, upper, lower=None):
        d, u, lo = _f64(diag), _f64(upper), _f64(lower)
        self._check(self.lib.hipfoam_set_matrix(self.h, _ptr(d), _ptr(u), _ptr(lo)),
                    "set_matrix")

    def pcg(self, r, target, max_iter, ips=1, order=8, beta=2.0, ratio=30.0) -> Stats:
        st = Stats()
        r = _f64(r)
        self._check(self.lib.hipfoam_pcg(self.h, _ptr(r, C.POINTER(C.c_double)), target,
                                         max_iter, ips, order, beta, ratio, C.byref(st)),
                    "pcg")
        return st

    def bicgstab(self, r, target, max_iter) -> Stats:
        st = Stats()
        r = _f64(r)
        self._check(self.lib.hipfoam_bicgstab(self.h, _ptr(r, C.POINTER(C.c_double)),
                                              target, max_iter, C.byref(st)), "bicgstab")
        return st

    def delta(self):
        out = np.empty(self.n, np.float32)
        self._check(self.lib.hipfoam_delta(self.h, _ptr(out, C.POINTER(C.c_float))),
                    "delta")
        return out

    # -- solver interface -------------------------------------------------
    def solve(self, solver, diag, upper, lower, psi, source, controls=None,
              residual=False, fp64=False):
        """Refined (or plain-FP64) solve on host arrays; psi updated in place."""
        controls = controls or default_controls()
        d, u, lo = _f64(diag), _f64(upper), _f64(lower)
        assert psi.dtype == np.float64 and psi.flags.c_contiguous
        b = _f64(source)
        res = np.empty(self.n) if residual else None
        perf = Performance()
        fn = self.lib.hipfoam_solve_fp64 if fp64 else self.lib.hipfoam_solve
        self._check(fn(self.h, solver, _ptr(d), _ptr(u), _ptr(lo), _ptr(psi), _ptr(b),
                       C.byref(controls), C.byref(perf), _ptr(res)),
                    "solve_fp64" if fp64 else "solve")
        return (perf, res) if residual else perf

    def solve_device(self, solver, diag, upper, lower, psi, source, controls=None,
                     fp64=False, residual=None):
        """Same with DeviceArray arguments (data stays in HBM)."""
        controls = controls or default_controls()
        perf = Performance()
        fn = self.lib.hipfoam_solve_fp64_device if fp64 else self.lib.hipfoam_solve_device
        self._check(fn(self.h, solver, diag.ptr, upper.ptr, None if lower is None else lower.ptr,
                       psi.ptr, source.ptr, C.byref(controls), C.byref(perf),
                       None if residual is None else residual.ptr), "solve_device")
        return perf

    def solve_device_noalloc(self, solver, diag, upper, lower, psi, source, controls=None,
                             residual=None):
        """hipfoam_solve_device_noalloc: the refined solve of solve_device with
        every per-solve buffer in the library's fixed scratch array (no device
        allocation); arguments are objects with a ``ptr`` (DeviceArray)."""
        controls = controls or default_controls()
        perf = Performance()
        self._check(self.lib.hipfoam_solve_device_noalloc(
            self.h, solver, diag.ptr, upper.ptr, None if lower is None else lower.ptr, psi.ptr,
            source.ptr, C.byref(controls), C.byref(perf),
            None if residual is None else residual.ptr), "solve_device_noalloc")
        return perf

    def adjoint(self, solver, diag, upper, lower, x, g, controls=None, precision=FP32):
        """Transpose solve A^T lambda = g and the gradients of L through A x = b.

        precision FP32: refined transpose solve; FP64: plain-FP64 transpose solve.
        Returns (perf, lambda, dUpper, dLower or None, dDiag, dSource)."""
        controls = controls or default_controls()
        d, u, lo = _f64(diag), _f64(upper), _f64(lower)
        x, g = _f64(x), _f64(g)
        nf = u.shape[0]
        lam = np.empty(self.n)
        du = np.empty(nf)
        dl = np.empty(nf) if lo is not None else None
        dd = np.empty(self.n)
        db = np.empty(self.n)
        perf = Performance()
        self._check(self.lib.hipfoam_adjoint(self.h, solver, precision, _ptr(d), _ptr(u),
                                             _ptr(lo),
                                             _ptr(x), _ptr(g), _ptr(lam), _ptr(du), _ptr(dl),
                                             _ptr(dd), _ptr(db), C.byref(controls),
                                             C.byref(perf)), "adjoint")
        return perf, lam, du, dl, dd, db

    def adjoint_device(self, solver, diag, upper, lower, x, g, lam, d_upper, d_lower, d_diag,
                       d_source, controls=None, precision=FP32):
        """hipfoam_adjoint_device: every array argument is an object with a
        ``ptr`` (a device address, or None to skip that gradient)."""
        controls = controls or default_controls()
        perf = Performance()

        def p(a):
            return None if a is None else a.ptr

        self._check(self.lib.hipfoam_adjoint_device(
            self.h, solver, precision, p(diag), p(upper), p(lower), p(x), p(g), p(lam),
            p(d_upper), p(d_lower), p(d_diag), p(d_source), C.byref(controls), C.byref(perf)),
            "adjoint_device")
        return perf

    # -- primitives -------------------------------------------------------
    def amul(self, x, precision=FP64):
        dt = np.float64 if precision == FP64 else np.float32
        x = np.ascontiguousarray(x, dtype=dt)
        y = np.empty(self.n, dt)
        self._check(self.lib.hipfoam_amul(self.h, precision, _ptr(x), _ptr(y)), "amul")
        return y

    def amul_device(self, x, y, precision=FP64):
        self._check(self.lib.hipfoam_amul_device(self.h, precision, x.ptr, y.ptr),
                    "amul_device")

    def synchronize(self):
        self._check(self.lib.hipfoam_synchronize(self.h), "synchronize")

    # -- schedule / timing ------------------------------------------------
    def get_schedule(self) -> Schedule:
        s = Schedule()
        self._check(self.lib.hipfoam_get_schedule(self.h, C.byref(s)), "get_schedule")
        return s

    def active_schedule(self) -> int:
        """SpMV schedule the current mesh's inner operator runs (0, 1 or 2)."""
        v = C.c_int32()
        self._check(self.lib.hipfoam_active_schedule(self.h, C.byref(v)), "active_schedule")
        return v.value

    def slot_layout(self) -> tuple:
        """(rows in patterned tiles, coefficient planes of a symmetric matrix)
        of the current mesh's row-slot layout; (0, 0) without row slots."""
        rows, planes = C.c_int64(), C.c_int32()
        self._check(self.lib.hipfoam_slot_layout(self.h, C.byref(rows), C.byref(planes)),
                    "slot_layout")
        return rows.value, planes.value

    def slot_tiles(self) -> tuple:
        """(patterned, explicit, face-loop) rows of the current mesh's row-slot
        layout (every row a face-loop row without one)."""
        p, e, f = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.lib.hipfoam_slot_tiles(self.h, C.byref(p), C.byref(e), C.byref(f)),
                    "slot_tiles")
        return p.value, e.value, f.value

    def set_schedule(self, **kw):
        s = self.get_schedule()
        for k, v in kw.items():
            setattr(s, k, v)
        self._check(self.lib.hipfoam_set_schedule(self.h, C.byref(s)), "set_schedule")
        return s

    def autotune(self) -> Schedule:
        s = Schedule()
        self._check(self.lib.hipfoam_autotune(self.h, C.byref(s)), "autotune")
        return s

    def persistent_stats(self) -> tuple:
        """(inner loops run as one persistent launch, of which abandoned and
        run again on the launched loop) since the context was created."""
        return self.persistent_abort()[:2]

    def persistent_abort(self) -> tuple:
        """persistent_stats() and the abort word of the last abandoned launch
        (0x10000 | workgroup << 4 | XCD id, 0x20000 | barrier index, 0)."""
        a, b, w = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.lib.hipfoam_persistent_stats(self.h, C.byref(a), C.byref(b),
                                                      C.byref(w)), "persistent_stats")
        return a.value, b.value, w.value

    def simulate_abandon(self, count: int = 1):
        """Report the next `count` persistent launches as abandoned after they
        ran (hipfoam_simulate_abandon): exercises the driver's recovery."""
        self._check(self.lib.hipfoam_simulate_abandon(self.h, int(count)), "simulate_abandon")

    def local_launches(self) -> int:
        """Inner loops run in one workgroup (schedule field localLoop) since
        the context was created."""
        a = C.c_int64()
        self._check(self.lib.hipfoam_local_stats(self.h, C.byref(a)), "local_stats")
        return a.value

    def set_kernel_timing(self, enable, kclass=0):
        self._check(self.lib.hipfoam_set_kernel_timing(self.h, int(enable), kclass),
                    "set_kernel_timing")

    def kernel_timing(self):
        n = C.c_int64()
        s = C.c_double()
        self._check(self.lib.hipfoam_kernel_timing(self.h, C.byref(n), C.byref(s)),
                    "kernel_timing")
        return n.value, s.value


class Batch:
    """Many LDU systems x nrhs right-hand sides solved together (hipfoam_batch_*)."""

    def __init__(self, ctx: Context, solver, systems, nrhs):
        self.ctx = ctx
        self.nrhs = nrhs
        self.systems = systems
        self._keep = []
        arr = (_lib.System * len(systems))()
        i32p, f64p = C.POINTER(C.c_int32), C.POINTER(C.c_double)
        for k, m in enumerate(systems):
            keep = [_i32(m.lower_addr), _i32(m.upper_addr), _i32(m.owner_start),
                    _i32(m.losort), _i32(m.losort_start), _f64(m.diag), _f64(m.upper),
                    _f64(m.lower)]
            self._keep.append(keep)
            lo, up, os_, ls, lss, d, u, lw = keep
            arr[k] = _lib.System(m.n, lo.shape[0], _ptr(lo, i32p), _ptr(up, i32p),
                                 _ptr(os_, i32p), _ptr(ls, i32p), _ptr(lss, i32p),
                                 _ptr(d, f64p), _ptr(u, f64p),
                                 None if lw is None else _ptr(lw, f64p))
        h = C.c_void_p()
        ctx._check(ctx.lib.hipfoam_batch_create(ctx.h, solver, len(systems), arr, nrhs,
                                                C.byref(h)), "batch_create")
        self.h = h
        self._keep = None  # the library copied everything

    def _rhs_block(self, a, n, layout, what):
        """[nrhs, n] contiguous FP64 copy of a right-hand-side block given in
        ``layout``: "rhs_major" = [nrhs, n] (the C-ABI's), "cell_major" =
        [n, nrhs] (the generators').  A 1-D array of n values is accepted when
        nrhs == 1, where both layouts coincide."""
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 1 and self.nrhs == 1 and a.shape[0] == n:
            return np.ascontiguousarray(a.reshape(1, n))
        if layout == "rhs_major":
            want = (self.nrhs, n)
        elif layout == "cell_major":
            want = (n, self.nrhs)
        else:
            raise ValueError(f"layout must be 'rhs_major' or 'cell_major', not {layout!r}")
        if a.shape != want:
            raise ValueError(f"{what} of system has shape {a.shape}, expected {want} "
                             f"({layout})")
        return np.ascontiguousarray(a.T if layout == "cell_major" else a)

    def set_rhs(self, k, source, psi=None, layout="rhs_major"):
        """source, psi (optional start): [nrhs, n] with layout "rhs_major"
        (the C-ABI's), or [n, nrhs] with layout "cell_major" (the generators'
        ``LduMatrix.source``); both arrays are read in the same layout."""
        n = self.systems[k].n
        src = self._rhs_block(source, n, layout, "source")
        ps = None if psi is None else self._rhs_block(psi, n, layout, "psi")
        assert src.size == self.nrhs * n and (ps is None or ps.size == self.nrhs * n)
        self.ctx._check(self.ctx.lib.hipfoam_batch_set_rhs(self.h, k, _ptr(src), _ptr(ps)),
                        "batch_set_rhs")

    def reset_psi(self):
        self.ctx._check(self.ctx.lib.hipfoam_batch_reset_psi(self.h), "batch_reset_psi")

    def solve(self, controls=None):
        controls = controls or default_controls()
        perf = (Performance * (len(self.systems) * self.nrhs))()
        self.ctx._check(self.ctx.lib.hipfoam_batch_solve(self.h, C.byref(controls), perf),
                        "batch_solve")
        return list(perf)

    def set_rhs_group(self, group):
        """Right-hand sides per workgroup of the iteration kernels (1..8)."""
        self.ctx._check(self.ctx.lib.hipfoam_batch_set_rhs_group(self.h, int(group)),
                        "batch_set_rhs_group")

    def autotune(self):
        """Times the batched SpMV over schedules and RHS groups; keeps the fastest
        (schedule in the context, group in the batch).  Returns (schedule, group)."""
        s = Schedule()
        g = C.c_int32()
        self.ctx._check(self.ctx.lib.hipfoam_batch_autotune(self.h, C.byref(s), C.byref(g)),
                        "batch_autotune")
        return s, g.value

    def psi(self, k):
        out = np.empty((self.nrhs, self.systems[k].n))
        self.ctx._check(self.ctx.lib.hipfoam_batch_get_psi(self.h, k, _ptr(out)),
                        "batch_get_psi")
        return out

    def close(self):
        if self.h:
            self.ctx.lib.hipfoam_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# Reference-interface mirror: run-time-selected solvers
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class SolverPerformance:
    """OpenFOAM solverPerformance subset returned by solve()."""

    solverName: str
    fieldName: str
    initialResidual: float = 0.0
    finalResidual: float = 0.0
    nIterations: int = 0
    converged: bool = False
    singular: bool = False
    detail: dict = dataclasses.field(default_factory=dict)


_BACKENDS: dict = {}


def _backend(field_name, mesh_key, device=0) -> Context:
    """One context per (field, mesh), as metalFoamSolver::backend()."""
    key = f"{field_name}@{mesh_key}"
    if key not in _BACKENDS:
        _BACKENDS[key] = Context(device)
    return _BACKENDS[key]


def _controls_from_dict(d: dict) -> Controls:
    """fvSolution solver dictionary + optional "hip" sub-dictionary."""
    c = default_controls()
    c.tolerance = float(d.get("tolerance", c.tolerance))
    c.relTol = float(d.get("relTol", c.relTol))
    c.maxIter = int(d.get("maxIter", c.maxIter))
    c.minIter = int(d.get("minIter", c.minIter))
    hip = d.get("hip", {}) or {}
    deflate = hip.get("deflate", "auto")
    c.deflate = {"off": 0, "on": 1, "auto": 2}[deflate]
    c.chebyshevOrder = int(hip.get("chebyshevOrder", c.chebyshevOrder))
    c.chebyshevRatio = float(hip.get("chebyshevRatio", c.chebyshevRatio))
    c.maxRestarts = int(hip.get("maxRestarts", c.maxRestarts))
    c.iterationsPerSync = int(hip.get("iterationsPerSync", c.iterationsPerSync))
    c.verbose = int(bool(hip.get("verbose", False)))
    return c


class _HipSolver:
    typeName = ""
    solver = PCG
    sym_only = True

    def __init__(self, fieldName, matrix, interfaceBouCoeffs=None, interfaceIntCoeffs=None,
                 interfaces=None, solverControls=None, device=0):
        self.fieldName = fieldName
        self.matrix = matrix
        self.interfaces = interfaces or []
        self.controls = solverControls or {}
        self.device = device

    def solve(self, psi: np.ndarray, source: np.ndarray, cmpt: int = 0) -> SolverPerformance:
        # coupled interfaces are rejected (metalFoamSolver.C:86-99)
        if any(i is not None for i in self.interfaces):
            raise HipFoamError(f"{self.typeName} does not support coupled boundaries "
                               "(processor/cyclic/AMI)")
        m = self.matrix
        if self.sym_only and m.lower is not None:
            raise HipFoamError(f"{self.typeName} is registered for symmetric matrices only")
        ctx = _backend(self.fieldName, id(m), self.device)
        ctx.set_topology(m)
        ctl = _controls_from_dict(self.controls)
        perf = ctx.solve(self.solver, m.diag, m.upper,
                         m.lower if self.solver == PBICGSTAB else None, psi, source, ctl)
        return SolverPerformance(self.typeName, self.fieldName, perf.initialResidual,
                                 perf.finalResidual, perf.nIterations, bool(perf.converged),
                                 bool(perf.singular), perf.as_dict())


class hipPCG(_HipSolver):
    """Mirror of metalPCG (symmetric table only, metalPCG.C:16-19)."""

    typeName = "hipPCG"
    solver = PCG
    sym_only = True


class hipPBiCGStab(_HipSolver):
    """Mirror of metalPBiCGStab (symmetric and asymmetric tables, metalPBiCGStab.C:16-22)."""

    typeName = "hipPBiCGStab"
    solver = PBICGSTAB
    sym_only = False


SOLVERS = {"hipPCG": hipPCG, "hipPBiCGStab": hipPBiCGStab}


def New(fieldName, matrix, solverControls, **kw):
    """lduMatrix::solver::New: select by the "solver" keyword."""
    name = solverControls["solver"]
    if name not in SOLVERS:
        raise HipFoamError(f"Unknown solver {name}; valid: {sorted(SOLVERS)}")
    cls = SOLVERS[name]
    if matrix.lower is not None and cls.sym_only:
        raise HipFoamError(f"{name} is not registered for asymmetric matrices")
    return cls(fieldName, matrix, solverControls=solverControls, **kw)
