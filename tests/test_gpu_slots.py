"""The row-slot SpMV schedule (spmvSchedule 2): the same terms in the same
order as the face-loop schedules 0 and 1, so SpMV results, inner solves and
refined solves must agree bit for bit; meshes that do not allow the layout
(a row with more than three owner or neighbour faces, or a column offset
beyond int16) fall back to schedule 0."""

import numpy as np
import pytest

from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _restore_schedule(ctx):
    keep = ctx.get_schedule().as_dict()
    yield
    ctx.set_schedule(**keep)


def irregular(n, nextra, seed):
    """A chain plus random long-range faces: rows with 4+ faces."""
    rng = np.random.default_rng(seed)
    pairs = {(i, i + 1) for i in range(n - 1)}
    while len(pairs) < n - 1 + nextra:
        a, b = sorted(rng.choice(n, 2, replace=False))
        pairs.add((int(a), int(b)))
    pairs = sorted(pairs)
    low = np.array([p[0] for p in pairs], dtype=np.int32)
    upp = np.array([p[1] for p in pairs], dtype=np.int32)
    own, los, loss = mg.derived_addressing(n, low, upp)
    upper = -rng.uniform(0.5, 1.5, len(pairs))
    lower = -rng.uniform(0.5, 1.5, len(pairs))
    diag = -mg.row_offdiag_sums(n, low, upp, upper, lower) + 0.5
    m = mg.LduMatrix(n, low, upp, own, los, loss, diag, upper, lower,
                     rng.standard_normal(n), (n,), "irregular")
    return m


def spmv_all_schedules(ctx, m, x, lower):
    out = []
    for s in (0, 1, 2):
        ctx.set_schedule(spmvSchedule=s, blockSize=256, xcdRemap=1, graphChunk=8,
                         gridLimit=0, rowForm=0)
        ctx.set_matrix(m.diag, m.upper, lower)
        out.append(ctx.amul(x, H.FP32))
    return out


@pytest.mark.parametrize("case", ["poisson", "convdiff", "wide", "irregular"])
def test_spmv_schedules_bitwise_equal(ctx, case):
    if case == "poisson":
        m = mg.poisson(33, 21, 17, seed=5)
    elif case == "convdiff":
        m = mg.convection_diffusion(29, 18, 13, seed=6)
    elif case == "wide":   # z offset 200*170 = 34000 > int16: layout not possible
        m = mg.convection_diffusion(200, 170, 3, seed=7)
    else:
        m = irregular(3000, 2500, 8)
    print(f"{case}: n={m.n} faces={m.nfaces}", flush=True)
    ctx.set_topology(m)
    x = np.random.default_rng(9).standard_normal(m.n).astype(np.float32)
    for lower in (None, m.lower):
        ys = spmv_all_schedules(ctx, m, x, lower)
        assert ctx.active_schedule() == (2 if case in ("poisson", "convdiff") else 0)
        np.testing.assert_array_equal(ys[2], ys[0])
        np.testing.assert_array_equal(ys[1], ys[0])
        # and against an FP64 product, to the FP32 rounding bound
        mm = mg.LduMatrix(m.n, m.lower_addr, m.upper_addr, m.owner_start, m.losort,
                          m.losort_start, m.diag, m.upper, lower)
        ref = mg.dense(mm).astype(np.float32).astype(np.float64) @ x.astype(np.float64) \
            if m.n <= 4000 else None
        if ref is not None:
            absrow = np.abs(mg.dense(mm)) @ np.abs(x.astype(np.float64))
            assert np.all(np.abs(ys[2] - ref) <= 8 * 2.0 ** -24 * absrow + 1e-30)


@pytest.mark.parametrize("solver", [H.PCG, H.PBICGSTAB])
def test_refined_solve_identical_across_schedules(ctx, solver):
    m = mg.poisson(28, 24, 20, seed=11) if solver == H.PCG else \
        mg.convection_diffusion(26, 22, 18, seed=12)
    lower = None if solver == H.PCG else m.lower
    ctx.set_topology(m)
    results = []
    for s in (0, 2):
        ctx.set_schedule(spmvSchedule=s, blockSize=256, xcdRemap=1, graphChunk=8,
                         gridLimit=0, rowForm=0)
        psi = np.zeros(m.n)
        perf = ctx.solve(solver, m.diag, m.upper, lower, psi, m.source,
                         H.default_controls(maxRestarts=6))
        assert perf.converged
        results.append((psi, perf.nIterations, perf.finalResidual))
    np.testing.assert_array_equal(results[1][0], results[0][0])
    assert results[1][1:] == results[0][1:]


def test_batched_slots_identical_to_face_loops(ctx):
    systems = [mg.poisson(14, 12, 10, seed=21, nrhs=2, rhs="normal"),
               mg.poisson(9, 9, 9, seed=22, nrhs=2, rhs="normal")]
    outs = []
    for s in (0, 2):
        ctx.set_schedule(spmvSchedule=s, blockSize=256, xcdRemap=1, graphChunk=8,
                         gridLimit=0, rowForm=0)
        b = H.Batch(ctx, H.PCG, systems, 2)
        for k, m in enumerate(systems):
            b.set_rhs(k, m.source.T.copy())
        perf = b.solve(H.default_controls(maxRestarts=6))
        outs.append([b.psi(k) for k in range(len(systems))])
        b.close()
    for a, c in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(c, a)
