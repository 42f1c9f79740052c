"""ctypes front-end of the CPU oracle (oracle/ldu_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  The product library never
imports, links or calls anything here.

Each function cites the reference code it restates in ldu_oracle.c.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libldu_oracle.so")

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)


class Addr(C.Structure):
    _fields_ = [("n", C.c_int32), ("nf", C.c_int32), ("low", _i32p),
                ("upp", _i32p), ("ownStart", _i32p), ("losort", _i32p),
                ("losortStart", _i32p)]


class InnerStats(C.Structure):
    _fields_ = [("nIterations", C.c_int), ("finalResSum", C.c_double),
                ("converged", C.c_int), ("singular", C.c_int),
                ("stagnated", C.c_int)]


class Controls(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("relTol", C.c_double),
                ("maxIter", C.c_int), ("minIter", C.c_int),
                ("deflate", C.c_int), ("chebyshevOrder", C.c_int),
                ("chebyshevRatio", C.c_double), ("maxRestarts", C.c_int),
                ("iterationsPerSync", C.c_int)]


class Perf(C.Structure):
    _fields_ = [("initialResidual", C.c_double), ("finalResidual", C.c_double),
                ("normFactor", C.c_double), ("nIterations", C.c_int),
                ("restarts", C.c_int), ("converged", C.c_int),
                ("singular", C.c_int), ("deflated", C.c_int),
                ("polished", C.c_int), ("chebBeta", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


PRE_DIAGONAL, PRE_DIC, PRE_CHEBYSHEV = 0, 1, 2

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def declare(L):
    """Declares the signatures of an oracle library build (a CDLL)."""
    A = C.POINTER(Addr)
    L.orc_amul.argtypes = [A, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_tmul.argtypes = [A, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_sumA.argtypes = [A, _f64p, _f64p, _f64p, _f64p]
    L.orc_norm_factor.argtypes = [A, _f64p, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_norm_factor.restype = C.c_double
    L.orc_spmv_f.argtypes = [A, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.orc_cheb_constants.argtypes = [C.c_int, C.c_double, C.c_double,
                                     _f32p, _f32p, _f32p]
    L.orc_cheb_apply_f.argtypes = [A, _f32p, _f32p, _f32p, _f32p, C.c_int,
                                   C.c_double, C.c_double, _f32p, _f32p,
                                   _f32p, _f32p]
    L.orc_cheb_apply_d.argtypes = [A, _f64p, _f64p, _f64p, _f64p, C.c_int,
                                   C.c_double, C.c_double, _f64p, _f64p,
                                   _f64p, _f64p]
    L.orc_inner_pcg_f.argtypes = [A, _f64p, _f64p, _f64p, C.c_double,
                                  C.c_int, C.c_int, C.c_int, C.c_double,
                                  C.c_double, _f32p]
    L.orc_inner_pcg_f.restype = InnerStats
    L.orc_inner_bicgstab_f.argtypes = [A, _f64p, _f64p, _f64p, _f64p,
                                       C.c_double, C.c_int, _f32p]
    L.orc_inner_bicgstab_f.restype = InnerStats
    L.orc_gershgorin.argtypes = [A, _f64p, _f64p]
    L.orc_gershgorin.restype = C.c_double
    L.orc_refined_solve.argtypes = [A, _f64p, _f64p, _f64p, C.c_int,
                                    C.POINTER(Controls), _f64p, _f64p,
                                    C.POINTER(Perf)]
    L.orc_pcg.argtypes = [A, _f64p, _f64p, C.c_int, C.c_int, C.c_double,
                          C.c_double, C.c_double, C.c_double, C.c_int,
                          C.c_int, _f64p, _f64p, C.POINTER(Perf)]
    L.orc_pbicgstab.argtypes = [A, _f64p, _f64p, _f64p, C.c_int,
                                C.c_double, C.c_double, C.c_int, C.c_int,
                                _f64p, _f64p, C.POINTER(Perf)]
    L.orc_adjoint_grads.argtypes = [A, _f64p, _f64p, _f64p, _f64p, _f64p]
    L.orc_pcg_diag_par.argtypes = [A, _f64p, _f64p, C.c_double, C.c_double, C.c_int,
                                   C.c_int, C.c_int, _f64p, _f64p, C.POINTER(Perf)]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = declare(C.CDLL(_LIB))
    return _lib


def build_native(outdir):
    """Builds the CPU-baseline variant of the oracle on this host
    (-O3 -march=native, OpenMP) into outdir and returns it as a declared
    CDLL, or None when the build fails (no compiler)."""
    os.makedirs(outdir, exist_ok=True)
    out = os.path.join(outdir, "libldu_oracle_native.so")
    try:
        subprocess.run(["make", "-s", "-C", _HERE, "native", f"NATIVE_OUT={out}"], check=True,
                       capture_output=True, timeout=300)
    except (OSError, subprocess.SubprocessError):
        return None
    return declare(C.CDLL(out))


def swap_library(L):
    """Makes every function of this module call the oracle build L (a CDLL
    passed through declare()) and returns the one it replaces.  Used by the
    mutation tests, which load deliberately broken builds of ldu_oracle.c to
    show that the pinning checks reject them."""
    global _lib
    prev = lib()
    _lib = L
    return prev


def _p(a, t):
    return a.ctypes.data_as(t)


def _c64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _c32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


class _AddrHolder:
    """Keeps the numpy arrays alive while the C struct points at them."""

    def __init__(self, m):
        self.arrs = [np.ascontiguousarray(x, dtype=np.int32) for x in
                     (m.lower_addr, m.upper_addr, m.owner_start, m.losort,
                      m.losort_start)]
        lo, up, os_, ls, lss = self.arrs
        self.s = Addr(m.n, lo.shape[0], _p(lo, _i32p), _p(up, _i32p),
                      _p(os_, _i32p), _p(ls, _i32p), _p(lss, _i32p))

    @property
    def ref(self):
        return C.byref(self.s)


def _ptr64(a):
    return None if a is None else _p(a, _f64p)


def _ptr32(a):
    return None if a is None else _p(a, _f32p)


def amul(m, x, upper=None, lower=None, diag=None, transpose=False):
    """y = A x (or A^T x) in FP64, lduMatrix face-loop order."""
    h = _AddrHolder(m)
    d = _c64(m.diag if diag is None else diag)
    u = _c64(m.upper if upper is None else upper)
    lo = _c64(m.lower if lower is None else lower)
    x = _c64(x)
    y = np.empty(m.n)
    fn = lib().orc_tmul if transpose else lib().orc_amul
    fn(h.ref, _ptr64(d), _ptr64(u), _ptr64(lo), _ptr64(x), _ptr64(y))
    return y


def sumA(m):
    h = _AddrHolder(m)
    u = np.empty(m.n)
    lib().orc_sumA(h.ref, _ptr64(_c64(m.diag)), _ptr64(_c64(m.upper)),
                   _ptr64(_c64(m.lower)), _ptr64(u))
    return u


def norm_factor(m, psi, b):
    h = _AddrHolder(m)
    psi, b = _c64(psi), _c64(b)
    Apsi = amul(m, psi)
    return lib().orc_norm_factor(h.ref, _ptr64(_c64(m.diag)),
                                 _ptr64(_c64(m.upper)), _ptr64(_c64(m.lower)),
                                 _ptr64(psi), _ptr64(b), _ptr64(Apsi))


def spmv_f(m, x, diag=None, upper=None, lower=None):
    """FP32 row-gather SpMV (reference kernel term order)."""
    h = _AddrHolder(m)
    d = _c32(m.diag if diag is None else diag)
    u = _c32(m.upper if upper is None else upper)
    lo = m.lower if lower is None else lower
    lo = _c32(lo)
    x = _c32(x)
    y = np.empty(m.n, dtype=np.float32)
    lib().orc_spmv_f(h.ref, _ptr32(d), _ptr32(u), _ptr32(lo), _ptr32(x),
                     _ptr32(y))
    return y


def cheb_constants(order, beta, ratio):
    c1 = C.c_float()
    c2 = np.zeros(max(order - 1, 1), dtype=np.float32)
    c3 = np.zeros(max(order - 1, 1), dtype=np.float32)
    lib().orc_cheb_constants(order, beta, ratio, C.byref(c1), _ptr32(c2),
                             _ptr32(c3))
    return c1.value, c2[: order - 1], c3[: order - 1]


def cheb_apply_f(m, r, order, beta, ratio, diag=None):
    h = _AddrHolder(m)
    d64 = _c64(m.diag if diag is None else diag)
    d = d64.astype(np.float32)
    rD = (1.0 / d64).astype(np.float32)
    u = _c32(m.upper)
    lo = _c32(m.lower)
    r = _c32(r)
    z = np.empty(m.n, np.float32)
    w = np.empty(m.n, np.float32)
    v = np.empty(m.n, np.float32)
    lib().orc_cheb_apply_f(h.ref, _ptr32(d), _ptr32(u), _ptr32(lo), _ptr32(rD),
                           order, beta, ratio, _ptr32(r), _ptr32(z), _ptr32(w),
                           _ptr32(v))
    return z


def cheb_apply_d(m, r, order, beta, ratio, diag=None):
    h = _AddrHolder(m)
    d = _c64(m.diag if diag is None else diag)
    rD = 1.0 / d
    r = _c64(r)
    z, w, v = np.empty(m.n), np.empty(m.n), np.empty(m.n)
    lib().orc_cheb_apply_d(h.ref, _ptr64(d), _ptr64(_c64(m.upper)),
                           _ptr64(_c64(m.lower)), _ptr64(rD), order, beta,
                           ratio, _ptr64(r), _ptr64(z), _ptr64(w), _ptr64(v))
    return z


def gershgorin(m, diag_star=None):
    h = _AddrHolder(m)
    d = _c64(m.diag if diag_star is None else diag_star)
    return lib().orc_gershgorin(h.ref, _ptr64(d), _ptr64(_c64(m.upper)))


def inner_pcg_f(m, r, target, max_iter, ips=1, cheb_order=8, cheb_beta=2.0,
                cheb_ratio=30.0, diag=None):
    h = _AddrHolder(m)
    d = _c64(m.diag if diag is None else diag)
    r = _c64(r)
    delta = np.empty(m.n, np.float32)
    st = lib().orc_inner_pcg_f(h.ref, _ptr64(d), _ptr64(_c64(m.upper)),
                               _ptr64(r), target, max_iter, ips, cheb_order,
                               cheb_beta, cheb_ratio, _ptr32(delta))
    return st, delta


def inner_bicgstab_f(m, r, target, max_iter):
    h = _AddrHolder(m)
    r = _c64(r)
    delta = np.empty(m.n, np.float32)
    st = lib().orc_inner_bicgstab_f(h.ref, _ptr64(_c64(m.diag)),
                                    _ptr64(_c64(m.upper)),
                                    _ptr64(_c64(m.lower)), _ptr64(r), target,
                                    max_iter, _ptr32(delta))
    return st, delta


def controls(tolerance=1e-6, relTol=0.0, maxIter=1000, minIter=0,
             deflate=2, chebyshevOrder=8, chebyshevRatio=30.0, maxRestarts=8,
             iterationsPerSync=1):
    return Controls(tolerance, relTol, maxIter, minIter, deflate,
                    chebyshevOrder, chebyshevRatio, maxRestarts,
                    iterationsPerSync)


def refined_solve(m, bicg=False, ctl=None, psi=None, source=None):
    """metalFoamSolver::solveImpl restated (FP64 outer, FP32 inner)."""
    h = _AddrHolder(m)
    ctl = ctl or controls()
    psi = np.zeros(m.n) if psi is None else _c64(psi).copy()
    b = _c64(m.source if source is None else source)
    perf = Perf()
    lower = _c64(m.lower) if bicg else None
    lib().orc_refined_solve(h.ref, _ptr64(_c64(m.diag)), _ptr64(_c64(m.upper)),
                            _ptr64(lower), int(bicg), C.byref(ctl),
                            _ptr64(psi), _ptr64(b), C.byref(perf))
    return psi, perf


def pcg(m, precon=PRE_DIC, tol=1e-6, relTol=0.0, maxIter=100000, minIter=0,
        order=8, beta=2.0, ratio=30.0, psi=None, source=None):
    """OpenFOAM PCG in FP64 (DIC / diagonal / Chebyshev-Jacobi)."""
    h = _AddrHolder(m)
    psi = np.zeros(m.n) if psi is None else _c64(psi).copy()
    b = _c64(m.source if source is None else source)
    perf = Perf()
    lib().orc_pcg(h.ref, _ptr64(_c64(m.diag)), _ptr64(_c64(m.upper)), precon,
                  order, beta, ratio, tol, relTol, maxIter, minIter,
                  _ptr64(psi), _ptr64(b), C.byref(perf))
    return psi, perf


def pbicgstab(m, precon=PRE_DIC, tol=1e-6, relTol=0.0, maxIter=100000,
              minIter=0, psi=None, source=None, lower=None, upper=None):
    """OpenFOAM PBiCGStab in FP64 (DILU / diagonal)."""
    h = _AddrHolder(m)
    psi = np.zeros(m.n) if psi is None else _c64(psi).copy()
    b = _c64(m.source if source is None else source)
    up = _c64(m.upper if upper is None else upper)
    lo = _c64(m.lower if lower is None else lower)
    perf = Perf()
    lib().orc_pbicgstab(h.ref, _ptr64(_c64(m.diag)), _ptr64(up), _ptr64(lo),
                        precon, tol, relTol, maxIter, minIter, _ptr64(psi),
                        _ptr64(b), C.byref(perf))
    return psi, perf


def pcg_diag_par(m, tol=1e-6, relTol=0.0, maxIter=100000, minIter=0, threads=0, psi=None,
                 source=None):
    """OpenFOAM PCG(diagonal) in FP64 over OpenMP threads (0: the runtime's
    default); one thread when the library was built without OpenMP."""
    h = _AddrHolder(m)
    psi = np.zeros(m.n) if psi is None else _c64(psi).copy()
    b = _c64(m.source if source is None else source)
    perf = Perf()
    lib().orc_pcg_diag_par(h.ref, _ptr64(_c64(m.diag)), _ptr64(_c64(m.upper)), tol, relTol,
                           maxIter, minIter, threads, _ptr64(psi), _ptr64(b), C.byref(perf))
    return psi, perf


def adjoint_grads(m, x, lam):
    h = _AddrHolder(m)
    x, lam = _c64(x), _c64(lam)
    du, dl, dd = np.empty(m.nfaces), np.empty(m.nfaces), np.empty(m.n)
    lib().orc_adjoint_grads(h.ref, _ptr64(x), _ptr64(lam), _ptr64(du),
                            _ptr64(dl), _ptr64(dd))
    return du, dl, dd
