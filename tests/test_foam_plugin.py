"""The OpenFOAM plugin (metal4foam_amd/openfoam: hipPCG, hipPBiCGStab) built
from its own sources against the API stand-in in tests/foam_shim and driven
like fvMatrix::solve: lduMatrix::solver::New by name, then solve(psi, source).

CPU tests cover the run-time selection tables and the FatalError paths
(metalPCG.C:16-19, metalPBiCGStab.C:16-22, metalFoamSolver.C:86-99,
139-144).  GPU tests check that the plugin hands the library exactly the
arrays and controls a direct C-ABI call gets (bitwise-identical psi), that
the result meets the oracle's bar, and the residual fields
(metalFoamSolver.C:112-117, 429-434).

OpenFOAM itself is not installed here: behaviour of the real lduMatrix
classes beyond what the stand-in restates is parity unpinned.
"""

import os
import struct
import subprocess

import numpy as np
import pytest

from metal4foam_amd import meshgen as mg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "tests", "foam_shim")
DRIVER = os.path.join(SHIM, "_build", "foam_driver")
LIB = os.path.join(ROOT, "metal4foam_amd", "libhipfoam.so")


def build_driver():
    """Compile the plugin + driver (host g++ only) when missing or stale."""
    if not os.path.exists(LIB):
        pytest.fail("metal4foam_amd/libhipfoam.so is not built (python metal4foam_amd/build.py)")
    r = subprocess.run(["make", "-s", "-C", SHIM], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return DRIVER


@pytest.fixture(scope="module")
def driver():
    return build_driver()


def write_case(path, m, psi, *, n_interfaces=0, coupled=-1, want_initial=False,
               want_final=False):
    hdr = struct.pack("<8i", m.n, m.nfaces, 0 if m.lower is None else 1,
                      n_interfaces, coupled, int(want_initial), int(want_final), 0)
    with open(path, "wb") as f:
        f.write(hdr)
        for a in (m.lower_addr, m.upper_addr, m.owner_start, m.losort, m.losort_start):
            f.write(np.ascontiguousarray(a, dtype="<i4").tobytes())
        arrays = [m.diag, m.upper] + ([] if m.lower is None else [m.lower])
        for a in arrays + [m.source, psi]:
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())


def write_controls(path, solver, tolerance=1e-6, rel_tol=0.0, hip=None, **extra):
    lines = [f"solver {solver};", f"tolerance {tolerance!r};", f"relTol {rel_tol!r};"]
    lines += [f"{k} {v};" for k, v in extra.items()]
    if hip:
        lines.append("hip {")
        lines += [f"    {k} {v};" for k, v in hip.items()]
        lines.append("}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def run_driver(driver, tmp_path, m, psi, solver, n_solves=1, case_kw=None, **ctl):
    case, dic, out = (str(tmp_path / x) for x in ("case.bin", "controls", "out.bin"))
    write_case(case, m, psi, **(case_kw or {}))
    write_controls(dic, solver, **ctl)
    print(f"[foam_driver] {solver} n={m.n} {m.name}", flush=True)
    r = subprocess.run([driver, case, dic, out, str(n_solves)], capture_output=True,
                       text=True, timeout=300)
    results = []
    for line in r.stdout.splitlines():
        if line.startswith("RESULT "):
            parts = line.split()
            kv = dict(p.split("=") for p in parts[2:])
            results.append({"solver": parts[1], "initial": float(kv["initial"]),
                            "final": float(kv["final"]), "iters": int(kv["iters"]),
                            "converged": bool(int(kv["converged"]))})
    fatal = [line[6:] for line in r.stdout.splitlines() if line.startswith("FATAL ")]
    arrays = None
    if r.returncode == 0:
        raw = np.fromfile(out, dtype="<f8")
        arrays = raw.reshape(-1, m.n)
    return r, results, fatal, arrays


def small_poisson():
    return mg.poisson(12, 10, 8, seed=31, lognormal=True)


def small_convdiff():
    return mg.convection_diffusion(10, 9, 8, seed=32)


# ---------------------------------------------------------------------------
# CPU: build, run-time selection, FatalError paths
# ---------------------------------------------------------------------------

def test_plugin_compiles_against_openfoam_api(driver):
    assert os.access(driver, os.X_OK)


def test_unknown_solver_lists_registered_symmetric(driver, tmp_path):
    r, _, fatal, _ = run_driver(driver, tmp_path, small_poisson(),
                                np.zeros(small_poisson().n), "hipFoo")
    assert r.returncode == 3 and len(fatal) == 1
    assert "Unknown symmetric matrix solver hipFoo" in fatal[0]
    assert "hipPBiCGStab hipPCG" in fatal[0]


def test_hipPCG_not_registered_for_asymmetric(driver, tmp_path):
    m = small_convdiff()
    r, _, fatal, _ = run_driver(driver, tmp_path, m, np.zeros(m.n), "hipPCG")
    assert r.returncode == 3
    assert "Unknown asymmetric matrix solver hipPCG" in fatal[0]
    valid = fatal[0].split("Valid solvers:")[1].split()
    assert valid == ["hipPBiCGStab"]


@pytest.mark.parametrize("solver", ["hipPCG", "hipPBiCGStab"])
def test_coupled_interface_is_fatal(driver, tmp_path, solver):
    m = small_poisson()
    r, res, fatal, _ = run_driver(driver, tmp_path, m, np.zeros(m.n), solver,
                                  case_kw=dict(n_interfaces=3, coupled=1))
    assert r.returncode == 3 and not res
    assert f"{solver} does not support coupled boundaries" in fatal[0]


def test_uncoupled_patches_are_accepted_before_the_device(driver, tmp_path):
    """Null interface entries (plain patches) pass the coupled check; without
    a device the next failure is the backend's."""
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present: the no-device path cannot be shown")
    m = small_poisson()
    r, _, fatal, _ = run_driver(driver, tmp_path, m, np.zeros(m.n), "hipPCG",
                                case_kw=dict(n_interfaces=4, coupled=-1))
    assert r.returncode == 3
    assert "HIP backend unavailable" in fatal[0] and "no HIP device" in fatal[0]


def test_unknown_topology_check_is_fatal(driver, tmp_path):
    """hip { topologyCheck ...; } accepts content (the default) and pointers
    (the reference's key) only."""
    m = small_poisson()
    r, res, fatal, _ = run_driver(driver, tmp_path, m, np.zeros(m.n), "hipPCG",
                                  hip={"topologyCheck": "sometimes"})
    assert r.returncode == 3 and not res
    assert "hip.topologyCheck must be content or pointers, not sometimes" in fatal[0]


@pytest.mark.parametrize("key,value,message", [
    ("persistent", "sometimes", "hip.persistent must be auto, on or off, not sometimes"),
    ("localLoop", "on", "hip.localLoop must be auto or off, not on"),
])
def test_unknown_loop_keys_are_fatal(driver, tmp_path, key, value, message):
    """hip { persistent auto|on|off; localLoop auto|off; } only."""
    m = small_poisson()
    r, res, fatal, _ = run_driver(driver, tmp_path, m, np.zeros(m.n), "hipPCG",
                                  hip={key: value})
    assert r.returncode == 3 and not res
    assert message in fatal[0]


# ---------------------------------------------------------------------------
# GPU: the plugin solve
# ---------------------------------------------------------------------------

def _direct_solve(ctx, m, solver, psi0, **ctl):
    from metal4foam_amd import hipfoam as H

    ctx.set_topology(m)
    psi = psi0.copy()
    lower = m.lower if solver == H.PBICGSTAB else None
    perf, res = ctx.solve(solver, m.diag, m.upper, lower, psi, m.source,
                          H.default_controls(**ctl), residual=True)
    return psi, perf, res


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["pcg_poisson", "bicg_poisson", "bicg_convdiff"])
def test_plugin_solve_matches_c_abi_bitwise(driver, tmp_path, ctx, orc, case):
    from metal4foam_amd import hipfoam as H

    solver = "hipPCG" if case == "pcg_poisson" else "hipPBiCGStab"
    m = mg.poisson(24, 20, 16, seed=41) if "poisson" in case else \
        mg.convection_diffusion(20, 18, 16, seed=42)
    psi0 = np.zeros(m.n)
    r, res, fatal, arrays = run_driver(
        driver, tmp_path, m, psi0, solver, tolerance=1e-7, rel_tol=0.0, maxIter=500,
        hip=dict(maxRestarts=6, chebyshevOrder=6, profile="true"))
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout)
    assert "profile: cells=" in r.stdout
    (out,) = res
    assert out["solver"] == solver and out["converged"]
    psi_d, perf, _ = _direct_solve(
        ctx, m, H.PBICGSTAB if solver == "hipPBiCGStab" else H.PCG, psi0,
        tolerance=1e-7, relTol=0.0, maxIter=500, maxRestarts=6, chebyshevOrder=6)
    np.testing.assert_array_equal(arrays[0], psi_d)
    assert out["iters"] == perf.nIterations
    assert out["initial"] == perf.initialResidual and out["final"] == perf.finalResidual
    # the oracle's bar: true FP64 residual of the returned psi
    tr = np.abs(m.source - orc.amul(m, arrays[0])).sum() / perf.normFactor
    assert tr == pytest.approx(out["final"], rel=1e-6) and tr < 1e-7


@pytest.mark.gpu
def test_plugin_residual_fields(driver, tmp_path, orc):
    """Initial and final FP64 residual fields when a solverInfo-style registry
    asks for them (metalFoamSolver.C:112-117, 429-434)."""
    m = mg.convection_diffusion(16, 16, 12, seed=43)
    psi0 = np.random.default_rng(5).standard_normal(m.n)
    r, res, _, arrays = run_driver(
        driver, tmp_path, m, psi0, "hipPBiCGStab", tolerance=1e-8,
        case_kw=dict(want_initial=True, want_final=True))
    assert r.returncode == 0, r.stdout + r.stderr
    assert res[0]["converged"] and res[0]["iters"] > 0
    psi, r_init, r_final = arrays
    scale = np.abs(m.source).max() + np.abs(m.diag).max() * np.abs(psi0).max()
    np.testing.assert_allclose(r_init, m.source - orc.amul(m, psi0), rtol=0,
                               atol=1e-13 * scale)
    np.testing.assert_allclose(r_final, m.source - orc.amul(m, psi), rtol=0,
                               atol=1e-12 * scale)
    # fields not requested: only psi is written
    r, _, _, arrays = run_driver(driver, tmp_path, m, psi0, "hipPBiCGStab",
                                 tolerance=1e-8)
    assert r.returncode == 0 and arrays.shape == (1, m.n)
    np.testing.assert_array_equal(arrays[0], psi)


@pytest.mark.gpu
def test_plugin_context_reused_across_solves(driver, tmp_path):
    """Three solves in one process, the solver object re-created for each (as
    fvMatrix::solve does), the library context created once; later solves
    start from the converged psi."""
    m = mg.convection_diffusion(16, 16, 12, seed=43)
    psi0 = np.random.default_rng(5).standard_normal(m.n)
    r, res, _, _ = run_driver(driver, tmp_path, m, psi0, "hipPBiCGStab", n_solves=3,
                              tolerance=1e-8)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("hipfoam") == 1  # one context description line
    assert len(res) == 3 and all(x["converged"] for x in res)
    assert res[0]["iters"] > max(res[1]["iters"], res[2]["iters"])
    assert max(res[1]["initial"], res[2]["initial"]) < 1e-3 * res[0]["initial"]


@pytest.mark.gpu
def test_plugin_min_and_max_iter_controls(driver, tmp_path):
    m = mg.poisson(20, 20, 20, seed=44)
    psi0 = np.zeros(m.n)
    r, res, _, _ = run_driver(driver, tmp_path, m, psi0, "hipPCG", tolerance=1e-12,
                              maxIter=7, hip=dict(maxRestarts=20))
    assert r.returncode == 0, r.stdout + r.stderr
    assert res[0]["iters"] <= 7 and not res[0]["converged"]
    # converged from the start, but minIter forces iterations
    r, res, _, _ = run_driver(driver, tmp_path, m, psi0, "hipPCG", tolerance=1e3,
                              minIter=3)
    assert r.returncode == 0, r.stdout + r.stderr
    assert res[0]["iters"] >= 3 and res[0]["converged"]


@pytest.mark.gpu
def test_plugin_deflate_pinned_keyword(driver, tmp_path):
    """hip { deflate pinned; } selects the projection along A 1: a pinned
    Poisson whose source has a non-zero sum converges in one restart."""
    m = mg.poisson(16, 14, 12, seed=45, lognormal=True, zero_mean_rhs=False, rhs="normal")
    r, res, _, arrays = run_driver(driver, tmp_path, m, np.zeros(m.n), "hipPCG",
                                   tolerance=1e-9, hip=dict(deflate="pinned", profile="true",
                                                            maxRestarts=8))
    assert r.returncode == 0, r.stdout + r.stderr
    assert res[0]["converged"]
    assert "restarts=1" in r.stdout or "restarts=0" in r.stdout
    x = np.linalg.solve(mg.dense(m), m.source)
    assert np.abs(arrays[0] - x).max() <= 1e-6 * np.abs(x).max()


@pytest.mark.gpu
@pytest.mark.parametrize("check", ["content", "pointers"])
def test_plugin_topology_check_keyword(driver, tmp_path, ctx, check):
    """hip { topologyCheck content|pointers; }: both keys give the direct
    C-ABI solves' psi bit for bit on an unchanged mesh, over two solves of
    one context (the second starting from the first's psi)."""
    from metal4foam_amd import hipfoam as H

    m = small_poisson()
    psi0 = np.random.default_rng(7).standard_normal(m.n)
    r, res, _, arrays = run_driver(driver, tmp_path, m, psi0, "hipPCG", n_solves=2,
                                   tolerance=1e-7, rel_tol=0.0, maxIter=500,
                                   hip=dict(topologyCheck=check, maxRestarts=6,
                                            chebyshevOrder=6))
    assert r.returncode == 0, r.stdout + r.stderr
    assert len(res) == 2 and all(x["converged"] for x in res)
    ctl = dict(tolerance=1e-7, relTol=0.0, maxIter=500, maxRestarts=6, chebyshevOrder=6)
    psi1, _, _ = _direct_solve(ctx, m, H.PCG, psi0, **ctl)
    psi2, _, _ = _direct_solve(ctx, m, H.PCG, psi1, **ctl)
    assert np.array_equal(arrays[0], psi2)


@pytest.mark.gpu
@pytest.mark.parametrize("persistent,local", [("auto", "auto"), ("off", "off"), ("on", "off")])
def test_plugin_loop_keywords(driver, tmp_path, ctx, persistent, local):
    """hip { persistent ...; localLoop ...; }: every choice of how the inner
    loop runs gives the direct C-ABI solve's psi bit for bit."""
    from metal4foam_amd import hipfoam as H

    m = small_poisson()
    psi0 = np.zeros(m.n)
    r, res, _, arrays = run_driver(driver, tmp_path, m, psi0, "hipPCG", tolerance=1e-7,
                                   rel_tol=0.0, maxIter=500,
                                   hip=dict(persistent=persistent, localLoop=local,
                                            maxRestarts=6))
    assert r.returncode == 0, r.stdout + r.stderr
    assert len(res) == 1 and res[0]["converged"]
    psi1, _, _ = _direct_solve(ctx, m, H.PCG, psi0, tolerance=1e-7, relTol=0.0, maxIter=500,
                               maxRestarts=6)
    assert np.array_equal(arrays[0], psi1)
