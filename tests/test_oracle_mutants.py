"""Mutation tests of the oracle's Krylov recurrences.

Each case compiles oracle/ldu_oracle.c with one plausible mistake planted in
its text (on the CPU, into a temporary directory), loads that build in place
of the oracle and runs the pinning checks of krylov_pins.py.  The mistake
must make the named check fail: a check that every mutant survives pins
nothing.  The first three are the beta recurrences that the convergence-only
checks let through (a degraded Krylov method still converges, given
iterations): the inner BiCGStab beta without alpha/omega
(metalFoamKernels.metal:348-354), OpenFOAM's PCG with beta = 0 (steepest
descent) and OpenFOAM's PBiCGStab beta without alpha/omega.
"""

import ctypes as C
import os
import shutil
import subprocess

import pytest

import krylov_pins as kp

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "oracle", "ldu_oracle.c")

# (id, text in ldu_oracle.c, replacement, pin that must fail)
MUTANTS = [
    ("inner_bicgstab_beta_without_alpha_over_omega",
     "beta = first ? 0.f : (t3 / rA0rAold) * (alpha / omega);",
     "beta = first ? 0.f : (t3 / rA0rAold);",
     kp.pin_inner_bicgstab_two_eigenvalues),
    ("openfoam_pcg_beta_zero",
     "const double b = wArA / wArAold;\n                for (int32_t i = 0; i < n; ++i) "
     "pA[i] = wA[i] + b * pA[i];",
     "const double b = 0.0 * (wArA / wArAold);\n                for (int32_t i = 0; i < n; "
     "++i) pA[i] = wA[i] + b * pA[i];",
     kp.pin_pcg_two_eigenvalues),
    ("openfoam_pbicgstab_beta_without_alpha_over_omega",
     "const double beta = (rA0rA / rA0rAold) * (alpha / omega);",
     "const double beta = (rA0rA / rA0rAold);",
     kp.pin_pbicgstab_two_eigenvalues),
    ("inner_pcg_beta_zero",
     "beta = first ? 0.f : t0 / wArAold;",
     "beta = 0.f;",
     kp.pin_inner_pcg_two_eigenvalues),
    ("inner_pcg_beta_inverted",
     "beta = first ? 0.f : t0 / wArAold;",
     "beta = first ? 0.f : wArAold / t0;",
     kp.pin_inner_pcg_two_eigenvalues),
    ("inner_bicgstab_p_update_sign",
     "p[i] = rA[i] + beta * (p[i] - omega * AyA[i]);",
     "p[i] = rA[i] + beta * (p[i] + omega * AyA[i]);",
     kp.pin_inner_bicgstab_two_eigenvalues),
    ("inner_bicgstab_omega_wrong_denominator",
     "omega = tAsA / tAtA;",
     "omega = tAsA / (tAtA + tAsA);",
     kp.pin_inner_bicgstab_iterates_match_scipy),
    ("openfoam_pcg_unpreconditioned_direction",
     "const double b = wArA / wArAold;\n                for (int32_t i = 0; i < n; ++i) "
     "pA[i] = wA[i] + b * pA[i];",
     "const double b = wArA / wArAold;\n                for (int32_t i = 0; i < n; ++i) "
     "pA[i] = rA[i] + b * pA[i];",
     kp.pin_pcg_iterates_match_scipy),
    ("openfoam_pbicgstab_omega_from_s",
     "omega = dotd(n, tA, sA) / tAtA;",
     "omega = dotd(n, tA, sA) / dotd(n, sA, sA);",
     kp.pin_pbicgstab_iterates_match_scipy),
    ("openfoam_pbicgstab_residual_with_alpha",
     "rA[i] = sA[i] - omega * tA[i];\n            }\n            perf->finalResidual",
     "rA[i] = sA[i] - alpha * tA[i];\n            }\n            perf->finalResidual",
     kp.pin_pbicgstab_two_eigenvalues),
]


@pytest.fixture(scope="module")
def gcc():
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not found")
    return cc


@pytest.mark.parametrize("mid,old,new,pin", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_mutant_is_rejected(orc, gcc, tmp_path, mid, old, new, pin):
    text = open(SRC).read()
    assert text.count(old) == 1, f"mutation site of {mid} not unique"
    src = tmp_path / "mutant.c"
    src.write_text(text.replace(old, new))
    so = tmp_path / "libmutant.so"
    subprocess.run([gcc, "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(so),
                    str(src), "-lm"], check=True)
    mutant = orc.declare(C.CDLL(str(so)))
    prev = orc.swap_library(mutant)
    try:
        with pytest.raises(AssertionError):
            pin(orc)
    finally:
        orc.swap_library(prev)
    # the unmutated oracle passes the same check
    pin(orc)
