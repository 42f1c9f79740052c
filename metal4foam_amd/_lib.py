"""ctypes binding of libhipfoam.so (the C-ABI declared in include/hipfoam.h).

Loading fails loudly when the HIP library has not been built: there is no CPU
fallback behind this binding.
"""

from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libhipfoam.so")
HEADER = os.path.join(ROOT, "include", "hipfoam.h")

OK, ERR_ARG, ERR_DEVICE, ERR_STATE, ERR_SIZE = 0, 1, 2, 3, 4
PCG, PBICGSTAB = 0, 1
FP32, FP64 = 0, 1
TOPO_POINTERS, TOPO_CONTENT = 0, 1

_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)
_vp = C.c_void_p


class Stats(C.Structure):
    _fields_ = [("nIterations", C.c_int32), ("finalResSum", C.c_double),
                ("converged", C.c_int32), ("singular", C.c_int32),
                ("stagnated", C.c_int32), ("tGpuSeconds", C.c_double),
                ("tWallSeconds", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Controls(C.Structure):
    _fields_ = [("tolerance", C.c_double), ("relTol", C.c_double),
                ("maxIter", C.c_int32), ("minIter", C.c_int32),
                ("deflate", C.c_int32), ("chebyshevOrder", C.c_int32),
                ("chebyshevRatio", C.c_double), ("maxRestarts", C.c_int32),
                ("iterationsPerSync", C.c_int32), ("verbose", C.c_int32)]


class Performance(C.Structure):
    _fields_ = [("initialResidual", C.c_double), ("finalResidual", C.c_double),
                ("normFactor", C.c_double), ("nIterations", C.c_int32),
                ("restarts", C.c_int32), ("converged", C.c_int32),
                ("singular", C.c_int32), ("deflated", C.c_int32),
                ("polished", C.c_int32), ("chebBeta", C.c_double),
                ("tConvert", C.c_double), ("tInner", C.c_double),
                ("tGpu", C.c_double), ("tTotal", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Schedule(C.Structure):
    _fields_ = [("spmvSchedule", C.c_int32), ("blockSize", C.c_int32),
                ("xcdRemap", C.c_int32), ("graphChunk", C.c_int32),
                ("gridLimit", C.c_int32), ("rowForm", C.c_int32),
                ("persistent", C.c_int32),
                ("localLoop", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class System(C.Structure):
    _fields_ = [("nCells", C.c_int32), ("nFaces", C.c_int32), ("lowAddr", _i32p),
                ("uppAddr", _i32p), ("ownStart", _i32p), ("losort", _i32p),
                ("losortStart", _i32p), ("diag", _f64p), ("upper", _f64p),
                ("lower", _f64p)]


_P = C.c_void_p  # hipfoam_context*

SIGNATURES = {
    "hipfoam_default_controls": (None, [C.POINTER(Controls)]),
    "hipfoam_create": (C.c_int32, [C.c_int32, C.POINTER(_P)]),
    "hipfoam_destroy": (None, [_P]),
    "hipfoam_good": (C.c_int32, [_P]),
    "hipfoam_description": (C.c_char_p, [_P]),
    "hipfoam_last_error": (C.c_char_p, [_P]),
    "hipfoam_set_topology": (C.c_int32, [_P, C.c_int32, C.c_int32, _i32p, _i32p,
                                         _i32p, _i32p, _i32p]),
    "hipfoam_set_topology_check": (C.c_int32, [_P, C.c_int32]),
    "hipfoam_set_coeffs": (C.c_int32, [_P, _f64p, _f64p, _f64p, _f64p]),
    "hipfoam_pcg": (C.c_int32, [_P, _f64p, C.c_double, C.c_int32, C.c_int32,
                                C.c_int32, C.c_double, C.c_double,
                                C.POINTER(Stats)]),
    "hipfoam_bicgstab": (C.c_int32, [_P, _f64p, C.c_double, C.c_int32,
                                     C.POINTER(Stats)]),
    "hipfoam_delta": (C.c_int32, [_P, _f32p]),
    "hipfoam_solve": (C.c_int32, [_P, C.c_int32, _vp, _vp, _vp, _vp, _vp,
                                  C.POINTER(Controls), C.POINTER(Performance), _vp]),
    "hipfoam_solve_device": (C.c_int32, [_P, C.c_int32, _vp, _vp, _vp, _vp, _vp,
                                         C.POINTER(Controls),
                                         C.POINTER(Performance), _vp]),
    "hipfoam_solve_device_noalloc": (C.c_int32, [_P, C.c_int32, _vp, _vp, _vp, _vp, _vp,
                                                 C.POINTER(Controls),
                                                 C.POINTER(Performance), _vp]),
    "hipfoam_solve_fp64": (C.c_int32, [_P, C.c_int32, _vp, _vp, _vp, _vp, _vp,
                                       C.POINTER(Controls),
                                       C.POINTER(Performance), _vp]),
    "hipfoam_solve_fp64_device": (C.c_int32, [_P, C.c_int32, _vp, _vp, _vp, _vp,
                                              _vp, C.POINTER(Controls),
                                              C.POINTER(Performance), _vp]),
    "hipfoam_adjoint": (C.c_int32, [_P, C.c_int32, C.c_int32, _vp, _vp, _vp, _vp, _vp,
                                    _vp, _vp, _vp, _vp, _vp, C.POINTER(Controls),
                                    C.POINTER(Performance)]),
    "hipfoam_adjoint_device": (C.c_int32, [_P, C.c_int32, C.c_int32, _vp, _vp, _vp, _vp,
                                           _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(Controls),
                                           C.POINTER(Performance)]),
    "hipfoam_batch_create": (C.c_int32, [_P, C.c_int32, C.c_int32, C.POINTER(System),
                                         C.c_int32, C.POINTER(_vp)]),
    "hipfoam_batch_set_rhs": (C.c_int32, [_vp, C.c_int32, _vp, _vp]),
    "hipfoam_batch_reset_psi": (C.c_int32, [_vp]),
    "hipfoam_batch_solve": (C.c_int32, [_vp, C.POINTER(Controls), C.POINTER(Performance)]),
    "hipfoam_batch_get_psi": (C.c_int32, [_vp, C.c_int32, _vp]),
    "hipfoam_batch_destroy": (None, [_vp]),
    "hipfoam_batch_set_rhs_group": (C.c_int32, [_vp, C.c_int32]),
    "hipfoam_batch_autotune": (C.c_int32, [_vp, C.POINTER(Schedule), C.POINTER(C.c_int32)]),
    "hipfoam_amul": (C.c_int32, [_P, C.c_int32, _vp, _vp]),
    "hipfoam_amul_device": (C.c_int32, [_P, C.c_int32, _vp, _vp]),
    "hipfoam_set_matrix": (C.c_int32, [_P, _vp, _vp, _vp]),
    "hipfoam_set_matrix_device": (C.c_int32, [_P, _vp, _vp, _vp]),
    "hipfoam_device_alloc": (C.c_int32, [_P, C.c_uint64, C.POINTER(_vp)]),
    "hipfoam_device_free": (C.c_int32, [_P, _vp]),
    "hipfoam_copy_to_device": (C.c_int32, [_P, _vp, _vp, C.c_uint64]),
    "hipfoam_copy_to_host": (C.c_int32, [_P, _vp, _vp, C.c_uint64]),
    "hipfoam_copy_on_device": (C.c_int32, [_P, _vp, _vp, C.c_uint64]),
    "hipfoam_synchronize": (C.c_int32, [_P]),
    "hipfoam_get_schedule": (C.c_int32, [_P, C.POINTER(Schedule)]),
    "hipfoam_active_schedule": (C.c_int32, [_P, C.POINTER(C.c_int32)]),
    "hipfoam_slot_layout": (C.c_int32, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int32)]),
    "hipfoam_slot_tiles": (C.c_int32, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int64)]),
    "hipfoam_set_schedule": (C.c_int32, [_P, C.POINTER(Schedule)]),
    "hipfoam_local_stats": (C.c_int32, [_P, C.POINTER(C.c_int64)]),
    "hipfoam_simulate_abandon": (C.c_int32, [_P, C.c_int32]),
    "hipfoam_persistent_stats": (C.c_int32, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                             C.POINTER(C.c_int64)]),
    "hipfoam_autotune": (C.c_int32, [_P, C.POINTER(Schedule)]),
    "hipfoam_set_kernel_timing": (C.c_int32, [_P, C.c_int32, C.c_int32]),
    "hipfoam_kernel_timing": (C.c_int32, [_P, C.POINTER(C.c_int64),
                                          C.POINTER(C.c_double)]),
}

_lib = None


def header_symbols(path: str = HEADER) -> list[str]:
    """Function names declared in include/hipfoam.h."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hipfoam_\w+)\s*\(", text)))


def load(path: str = LIB_PATH):
    """Load libhipfoam.so and attach the C signatures (raises if missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP library first "
            "(python -m metal4foam_amd.build); there is no CPU fallback")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib
