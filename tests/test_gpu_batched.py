"""GPU batched solves: every (system, rhs) segment must reproduce the
single-system refined solve (same controls) and reach the tolerance."""

import numpy as np
import pytest

from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _default_schedule(ctx):
    ctx.set_schedule(spmvSchedule=0, blockSize=256, xcdRemap=1, graphChunk=4, gridLimit=0,
                     rowForm=0)
    yield


def _systems(kind, sides, nrhs, seed):
    out = []
    for k, s in enumerate(sides):
        if kind == "sym":
            out.append(mg.poisson(s, s + 1, s - 1, seed=seed + k, lognormal=True,
                                  zero_mean_rhs=False, rhs="normal", nrhs=nrhs))
        else:
            out.append(mg.convection_diffusion(s, s + 1, s - 1, seed=seed + k, nrhs=nrhs))
    return out


@pytest.mark.parametrize("kind,solver", [("sym", H.PCG), ("asym", H.PBICGSTAB),
                                         ("sym", H.PBICGSTAB)])
def test_batched_matches_single_system_solves(ctx, orc, kind, solver):
    nrhs = 3
    systems = _systems(kind, [9, 14, 6, 11], nrhs, seed=300)
    batch = H.Batch(ctx, solver, systems, nrhs)
    for k, m in enumerate(systems):
        batch.set_rhs(k, m.source, layout="cell_major")
    ctl = H.default_controls(maxRestarts=6)
    perfs = batch.solve(ctl)
    for k, m in enumerate(systems):
        psi_b = batch.psi(k)
        ctx.set_topology(m)
        for r in range(nrhs):
            p = perfs[k * nrhs + r]
            b = m.source[:, r].copy()
            psi = np.zeros(m.n)
            ps = ctx.solve(solver, m.diag, m.upper, m.lower if solver == H.PBICGSTAB else None,
                           psi, b, ctl)
            tr = np.abs(b - orc.amul(m, psi_b[r])).sum() / p.normFactor
            print(f"sys{k} rhs{r}: batch it={p.nIterations} res={p.finalResidual:.3e} "
                  f"(true {tr:.3e}) single it={ps.nIterations} res={ps.finalResidual:.3e}")
            assert p.converged and tr < 1e-6
            assert tr == pytest.approx(p.finalResidual, rel=1e-6)
            assert p.normFactor == pytest.approx(ps.normFactor, rel=1e-12)
            assert p.deflated == ps.deflated and p.polished == ps.polished
            assert abs(p.nIterations - ps.nIterations) <= max(3, ps.nIterations // 10)
            np.testing.assert_allclose(psi_b[r], psi, rtol=0,
                                       atol=1e-4 * max(np.abs(psi).max(), 1.0))
    batch.close()


def test_batched_uneven_sizes_and_repeat(ctx, orc):
    nrhs = 2
    systems = _systems("sym", [5, 17, 8], nrhs, seed=400)
    batch = H.Batch(ctx, H.PCG, systems, nrhs)
    for k, m in enumerate(systems):
        batch.set_rhs(k, m.source, layout="cell_major")
    first = batch.solve()
    sol1 = [batch.psi(k) for k in range(len(systems))]
    # resolve from zero: same result bit for bit (deterministic reductions)
    batch.reset_psi()
    second = batch.solve()
    for k in range(len(systems)):
        np.testing.assert_array_equal(batch.psi(k), sol1[k])
    assert [p.nIterations for p in first] == [p.nIterations for p in second]
    # warm start from the solution: converged without iterations
    third = batch.solve()
    assert all(p.converged and p.nIterations == 0 for p in third)
    batch.close()


def test_batch_rejects_mixed_symmetry(ctx):
    a = mg.poisson(4, 4, 4, seed=1, nrhs=1)
    b = mg.convection_diffusion(4, 4, 4, seed=1, nrhs=1)
    with pytest.raises(H.HipFoamError):
        H.Batch(ctx, H.PBICGSTAB, [a, b], 1)
    with pytest.raises(H.HipFoamError):
        H.Batch(ctx, H.PCG, [b], 1)


@pytest.mark.parametrize("schedule", [0, 2])
def test_segments_of_one_system_on_different_diagonals(ctx, schedule):
    """The RHS of one system share its FP32 diagonal planes (exact and
    deflated) and pick one each.  On the pinned Poisson a zero-mean RHS stays
    deflated, while a RHS with a mean stalls under deflation and is polished
    on the exact diagonal: both planes are read within one system in the
    same launches, and every segment must still equal its own
    single-system solve."""
    ctx.set_schedule(spmvSchedule=schedule, blockSize=0, xcdRemap=2, graphChunk=4,
                     gridLimit=0, rowForm=0)
    nrhs = 3
    m0 = mg.poisson(12, 13, 11, seed=510, lognormal=True, zero_mean_rhs=True, rhs="normal",
                    nrhs=nrhs)
    m0.source[:, 1] += 1.0
    m1 = mg.poisson(9, 8, 10, seed=511, lognormal=True, zero_mean_rhs=False, rhs="normal",
                    nrhs=nrhs)
    systems = [m0, m1]
    batch = H.Batch(ctx, H.PCG, systems, nrhs)
    for k, m in enumerate(systems):
        batch.set_rhs(k, m.source, layout="cell_major")
    ctl = H.default_controls(maxRestarts=8)
    perfs = batch.solve(ctl)
    pol0 = [perfs[r].polished for r in range(nrhs)]
    print("system 0 deflated", [perfs[r].deflated for r in range(nrhs)], "polished", pol0)
    assert all(perfs[r].deflated for r in range(nrhs))
    assert not pol0[0] and pol0[1]
    for k, m in enumerate(systems):
        psi_b = batch.psi(k)
        ctx.set_topology(m)
        for r in range(nrhs):
            p = perfs[k * nrhs + r]
            psi = np.zeros(m.n)
            ps = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source[:, r].copy(), ctl)
            assert p.converged and p.finalResidual < 1e-6
            assert (p.deflated, p.polished) == (ps.deflated, ps.polished)
            assert abs(p.nIterations - ps.nIterations) <= max(3, ps.nIterations // 10)
            np.testing.assert_allclose(psi_b[r], psi, rtol=0,
                                       atol=1e-4 * max(np.abs(psi).max(), 1.0))
    batch.close()


def _solve_all(batch, systems, ctl):
    batch.reset_psi()
    perfs = batch.solve(ctl)
    return perfs, [batch.psi(k) for k in range(len(systems))]


@pytest.mark.parametrize("kind,solver", [("sym", H.PCG), ("asym", H.PBICGSTAB)])
def test_rhs_groups_give_identical_solves(ctx, kind, solver):
    """The RHS group only sets how many right-hand sides a workgroup of the
    iteration kernels serves (5 RHS: groups 3 and 2 leave a partial group);
    every segment's partials keep one slot per (tile, rhs), so the solves
    are bit-identical for every group size."""
    ctx.set_schedule(spmvSchedule=2, blockSize=0, xcdRemap=2, graphChunk=4, gridLimit=0,
                     rowForm=0)
    nrhs = 5
    systems = _systems(kind, [10, 7, 13], nrhs, seed=620)
    batch = H.Batch(ctx, solver, systems, nrhs)
    for k, m in enumerate(systems):
        batch.set_rhs(k, m.source, layout="cell_major")
    ctl = H.default_controls(maxRestarts=6)
    ref_p, ref_psi = _solve_all(batch, systems, ctl)
    assert all(p.converged for p in ref_p)
    for g in (1, 2, 3, 5, 8):
        batch.set_rhs_group(g)
        perfs, psi = _solve_all(batch, systems, ctl)
        print(f"group {g}: iterations {[p.nIterations for p in perfs]}")
        assert [p.nIterations for p in perfs] == [p.nIterations for p in ref_p]
        assert [p.finalResidual for p in perfs] == [p.finalResidual for p in ref_p]
        for a, b in zip(psi, ref_psi):
            np.testing.assert_array_equal(a, b)
    for bad in (0, 9):
        with pytest.raises(H.HipFoamError):
            batch.set_rhs_group(bad)
    batch.close()


def test_batch_autotune_keeps_results(ctx):
    """The batch autotuner picks a schedule (context) and an RHS group
    (batch) from timings alone: the solve after it is bit-identical."""
    keep = ctx.get_schedule()
    nrhs = 3
    systems = _systems("sym", [12, 9], nrhs, seed=640)
    batch = H.Batch(ctx, H.PCG, systems, nrhs)
    for k, m in enumerate(systems):
        batch.set_rhs(k, m.source, layout="cell_major")
    ctl = H.default_controls(maxRestarts=6)
    ref_p, ref_psi = _solve_all(batch, systems, ctl)
    sched, group = batch.autotune()
    print(f"batch autotune: schedule {sched.spmvSchedule} rowForm {sched.rowForm} "
          f"xcdRemap {sched.xcdRemap} group {group}")
    assert sched.spmvSchedule in (0, 1, 2) and sched.rowForm in (0, 1)
    assert sched.xcdRemap in (0, 1) and 1 <= group <= nrhs
    got = ctx.get_schedule()
    assert (got.spmvSchedule, got.rowForm, got.xcdRemap) == (
        sched.spmvSchedule, sched.rowForm, sched.xcdRemap)
    perfs, psi = _solve_all(batch, systems, ctl)
    assert [p.nIterations for p in perfs] == [p.nIterations for p in ref_p]
    for a, b in zip(psi, ref_psi):
        np.testing.assert_array_equal(a, b)
    batch.close()
    ctx.set_schedule(spmvSchedule=keep.spmvSchedule, blockSize=keep.blockSize,
                     xcdRemap=keep.xcdRemap, graphChunk=keep.graphChunk,
                     gridLimit=keep.gridLimit, rowForm=keep.rowForm)
