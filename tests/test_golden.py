"""The committed golden vectors of tests/golden/fixtures_tiny.npz (dense numpy
results on tiny LDU systems, written by tests/golden/make_fixtures.py) against
the oracle (CPU) and the HIP library through the C-ABI (GPU)."""

import os

import numpy as np
import pytest

from metal4foam_amd import meshgen as mg

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "fixtures_tiny.npz")
CASES = ["poisson_3x3x1", "poisson_4x4x4", "convdiff_4x3x2", "irregular_14"]


@pytest.fixture(scope="module")
def fx():
    return np.load(FIX, allow_pickle=False)


def case(fx, name):
    g = lambda k: fx[f"{name}.{k}"]  # noqa: E731
    lower = g("lower")
    m = mg.LduMatrix(int(g("n")), g("low"), g("upp"), g("own"), g("los"), g("loss"),
                     g("diag"), g("upper"), lower if lower.size else None, g("b"), (), name)
    return m, g


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


# ---------------------------------------------------------------------------
# CPU: the fixtures themselves and the oracle
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name", CASES)
def test_fixture_addressing_is_lduaddressing(fx, name):
    m, g = case(fx, name)
    own, los, loss = mg.derived_addressing(m.n, m.lower_addr, m.upper_addr)
    np.testing.assert_array_equal(g("own"), own)
    np.testing.assert_array_equal(g("los"), los)
    np.testing.assert_array_equal(g("loss"), loss)
    assert np.all(m.lower_addr < m.upper_addr)


@pytest.mark.parametrize("name", CASES)
def test_oracle_operator_against_fixture(fx, orc, name):
    m, g = case(fx, name)
    assert rel(orc.amul(m, g("x_probe")), g("Ax")) < 1e-14
    assert rel(orc.amul(m, g("x_probe"), transpose=True), g("ATx")) < 1e-14
    assert rel(orc.sumA(m), g("sumA")) < 1e-14
    assert orc.norm_factor(m, g("x_probe"), m.source) == pytest.approx(float(g("nf_probe")),
                                                                         rel=1e-13)


@pytest.mark.parametrize("name", CASES)
def test_oracle_solvers_against_fixture(fx, orc, name):
    m, g = case(fx, name)
    if m.lower is None:
        x, perf = orc.pcg(m, precon=orc.PRE_DIC, tol=1e-14)
    else:
        x, perf = orc.pbicgstab(m, precon=orc.PRE_DIC, tol=1e-14)
    assert rel(x, g("x")) < 1e-10


@pytest.mark.parametrize("name", CASES)
def test_oracle_adjoint_against_fixture(fx, orc, name):
    m, g = case(fx, name)
    du, dl, dd = orc.adjoint_grads(m, g("x"), g("lam"))
    if m.lower is None:
        assert rel(du + dl, g("dUpper")) < 1e-13
    else:
        assert rel(du, g("dUpper")) < 1e-13 and rel(dl, g("dLower")) < 1e-13
    assert rel(dd, g("dDiag")) < 1e-13


# ---------------------------------------------------------------------------
# GPU: the HIP library through the C-ABI
# ---------------------------------------------------------------------------

@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_hip_operator_against_fixture(fx, ctx, name):
    from metal4foam_amd import hipfoam as H

    m, g = case(fx, name)
    ctx.set_topology(m)
    ctx.set_matrix(m.diag, m.upper, m.lower)
    y = ctx.amul(g("x_probe"), H.FP64)
    assert rel(y, g("Ax")) < 1e-14
    # the FP32 inner operator: per-row rounding bound of an FP32 SpMV
    y32 = ctx.amul(g("x_probe").astype(np.float32), H.FP32).astype(np.float64)
    a = np.abs(mg.dense(m))
    bound = 8 * 2.0 ** -24 * (a @ np.abs(g("x_probe"))) + 1e-30
    assert np.all(np.abs(y32 - g("Ax")) <= bound * 2)


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_hip_solves_against_fixture(fx, ctx, name):
    from metal4foam_amd import hipfoam as H

    m, g = case(fx, name)
    ctx.set_topology(m)
    solver = H.PCG if m.lower is None else H.PBICGSTAB
    # plain FP64 GPU solve to 1e-14
    psi = np.zeros(m.n)
    perf = ctx.solve(solver, m.diag, m.upper, m.lower, psi, m.source,
                     H.default_controls(tolerance=1e-14, maxIter=5000), fp64=True)
    assert perf.converged and rel(psi, g("x")) < 1e-10
    # refined solve (FP32 inner, FP64 refinement) to 1e-11
    psi = np.zeros(m.n)
    perf = ctx.solve(solver, m.diag, m.upper, m.lower, psi, m.source,
                     H.default_controls(tolerance=1e-11, maxRestarts=30, maxIter=5000))
    print(f"{name}: refined it={perf.nIterations} restarts={perf.restarts} "
          f"res={perf.finalResidual:.2e} rel err {rel(psi, g('x')):.2e}")
    assert perf.converged and rel(psi, g("x")) < 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("name", CASES)
def test_hip_adjoint_against_fixture(fx, ctx, name):
    from metal4foam_amd import hipfoam as H

    m, g = case(fx, name)
    ctx.set_topology(m)
    solver = H.PCG if m.lower is None else H.PBICGSTAB
    x = g("x")
    perf, lam, du, dl, dd, db = ctx.adjoint(
        solver, m.diag, m.upper, m.lower, x, x - g("xt"),
        H.default_controls(tolerance=1e-14, maxIter=5000), precision=H.FP64)
    assert perf.converged
    assert rel(lam, g("lam")) < 1e-9 and rel(db, g("dSource")) < 1e-9
    assert rel(du, g("dUpper")) < 1e-9 and rel(dd, g("dDiag")) < 1e-9
    if m.lower is not None:
        assert rel(dl, g("dLower")) < 1e-9
