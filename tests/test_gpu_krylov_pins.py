"""The GPU Krylov loops on the matrices that pin the oracle's recurrences.

On a matrix whose Jacobi-scaled operator has two distinct eigenvalues, CG
and BiCGStab terminate at their second iteration (krylov_pins.py).  The
inner FP32 loops (hipfoam_pcg, hipfoam_bicgstab: metalFoamBackend.C:568-959)
must reach FP32 accuracy there, the plain-FP64 loops (hipfoam_solve_fp64)
FP64 accuracy, and one iteration must not be enough.  The row-slot layout
runs the one-face-per-row pair blocks; the complete graphs (rows of 11
faces) run the face-loop rows."""

import numpy as np
import pytest

import krylov_pins as kp
from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _default_schedule(ctx):
    ctx.set_schedule(spmvSchedule=2, blockSize=0, xcdRemap=2, graphChunk=-1, gridLimit=0,
                     rowForm=0)
    yield


def rel_l1(m, x):
    return np.abs(m.source - mg.dense(m) @ np.asarray(x, np.float64)).sum() / np.abs(
        m.source).sum()


@pytest.mark.parametrize("m", kp.two_eigenvalue_cases(symmetric_only=True),
                         ids=lambda m: m.name)
@pytest.mark.parametrize("order", [1, 8])
def test_inner_pcg_terminates_at_two(ctx, orc, m, order):
    print(f"[pins] inner pcg {m.name} order {order}", flush=True)
    ctx.set_topology(m)
    ctx.set_coeffs(m.diag, m.upper, None)
    beta = orc.gershgorin(m)
    st = ctx.pcg(m.source, 0.0, 1, order=order, beta=beta)
    assert st.nIterations == 1 and rel_l1(m, ctx.delta()) > 1e-4
    st = ctx.pcg(m.source, 0.0, 2, order=order, beta=beta)
    d = ctx.delta()
    assert st.nIterations == 2 and not st.singular
    assert rel_l1(m, d) < 2e-6, rel_l1(m, d)
    _, dref = orc.inner_pcg_f(m, m.source, 0.0, 2, cheb_order=order, cheb_beta=beta)
    assert np.abs(d - dref).max() <= 1e-5 * np.abs(dref).max()


@pytest.mark.parametrize("m", kp.two_eigenvalue_cases(), ids=lambda m: m.name)
def test_inner_bicgstab_terminates_at_two(ctx, orc, m):
    print(f"[pins] inner bicgstab {m.name}", flush=True)
    ctx.set_topology(m)
    ctx.set_coeffs(m.diag, m.upper, m.lower)
    st = ctx.bicgstab(m.source, 0.0, 1)
    assert st.nIterations == 1 and rel_l1(m, ctx.delta()) > 1e-4
    st = ctx.bicgstab(m.source, 0.0, 2)
    d = ctx.delta()
    assert st.nIterations == 2 and not st.singular
    assert rel_l1(m, d) < 2e-6, rel_l1(m, d)
    _, dref = orc.inner_bicgstab_f(m, m.source, 0.0, 2)
    assert np.abs(d - dref).max() <= 1e-5 * np.abs(dref).max()


@pytest.mark.parametrize("m", kp.two_eigenvalue_cases(), ids=lambda m: m.name)
@pytest.mark.parametrize("order", [1, 8])
def test_fp64_loops_terminate_at_two(ctx, m, order):
    solvers = [H.PBICGSTAB] if m.lower is not None else [H.PCG, H.PBICGSTAB]
    ctx.set_topology(m)
    for solver in solvers:
        if solver == H.PBICGSTAB and order != 1:
            continue
        print(f"[pins] fp64 {m.name} solver {solver} order {order}", flush=True)
        lower = m.lower if solver == H.PBICGSTAB else None
        for k, ok in ((1, False), (2, True)):
            psi = np.zeros(m.n)
            ctl = H.default_controls(tolerance=1e-13, maxIter=k, chebyshevOrder=order)
            perf = ctx.solve(solver, m.diag, m.upper, lower, psi, m.source, ctl, fp64=True)
            assert perf.nIterations == k
            if ok:
                assert perf.converged and rel_l1(m, psi) < 1e-12, rel_l1(m, psi)
            else:
                assert not perf.converged and rel_l1(m, psi) > 1e-4
