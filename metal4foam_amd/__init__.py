"""hipFoam: MI355X-native LDU Krylov solvers (hipPCG / hipPBiCGStab).

The product is the HIP library ``libhipfoam.so`` behind the C-ABI in
``include/hipfoam.h``; this package holds its sources (``csrc/``), the build
step, a ctypes binding and a thin Python host layer mirroring the reference
solver interface.  Synthetic LDU workloads live in ``meshgen``.
"""

__all__ = ["meshgen"]
