"""The refined solve's two rarely taken branches against the oracle.

* Hybrid polish (metalFoamSolver.C:367-391): on a pinned Neumann Poisson a
  right-hand side with a non-zero mean stalls under constant-mode deflation;
  once a restart reduces the residual by less than 1.5x the solve switches to
  the exact operator.
* Deflation bail (metalFoamSolver.C:192-209): a positive face coefficient
  makes diag - A 1 change sign against diag in its two rows, so deflation is
  refused even when requested.

The GPU solve must take the same branches as orc_refined_solve on the same
inputs, after the same number of restarts and inner iterations, and reach a
solution within the refined solve's tolerance of the oracle's."""

import numpy as np
import pytest

from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def polish_case():
    m = mg.poisson(12, 13, 11, seed=510, lognormal=True, zero_mean_rhs=True, rhs="normal",
                   nrhs=3)
    m.source[:, 1] += 1.0
    m.source = m.source[:, 1].copy()
    m.name = "polish_mean_rhs"
    return m


def polish_case_positive_faces():
    m = mg.poisson(10, 9, 8, seed=3, lognormal=True)
    m.upper = m.upper.copy()
    m.upper[[7, 100, 333]] = [0.8, 0.5, 1.2]
    m.name = "polish_positive_faces"
    return m


def bail_case():
    m = mg.poisson(6, 6, 1, seed=1, lognormal=False)
    m.upper = m.upper.copy()
    m.upper[7] = 5.0
    m.name = "bail_positive_face"
    return m


# (case, deflate control, (deflated, polished) the oracle takes); deflate 3
# (projection along A 1, no reference counterpart) solves the RHS with a mean
# without polishing
CASES = [(polish_case, 2, (1, 1)), (polish_case, 3, (1, 0)),
         (polish_case_positive_faces, 1, (1, 1)),
         (polish_case_positive_faces, 2, (1, 1)), (bail_case, 1, (0, 0)),
         (bail_case, 2, (0, 0))]


@pytest.mark.parametrize("schedule", [0, 2])
@pytest.mark.parametrize("make,deflate,flags", CASES,
                         ids=[f"{c[0].__name__}-deflate{c[1]}" for c in CASES])
def test_branch_matches_oracle(ctx, orc, make, deflate, flags, schedule):
    m = make()
    print(f"[branches] {m.name} deflate={deflate} schedule={schedule}", flush=True)
    ctx.set_schedule(spmvSchedule=schedule, blockSize=0, xcdRemap=2, graphChunk=-1,
                     gridLimit=0, rowForm=0)
    ctx.set_topology(m)
    ctl = H.default_controls(maxRestarts=8, deflate=deflate)
    psi = np.zeros(m.n)
    perf = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source, ctl)
    ref_psi, ref = orc.refined_solve(m, ctl=orc.controls(maxRestarts=8, deflate=deflate))
    print(f"  gpu: deflated={perf.deflated} polished={perf.polished} restarts={perf.restarts} "
          f"it={perf.nIterations} res={perf.finalResidual:.3e}; oracle: deflated={ref.deflated} "
          f"polished={ref.polished} restarts={ref.restarts} it={ref.nIterations} "
          f"res={ref.finalResidual:.3e}", flush=True)
    assert (ref.deflated, ref.polished) == flags
    assert (perf.deflated, perf.polished) == (ref.deflated, ref.polished)
    assert perf.restarts == ref.restarts
    assert perf.nIterations == ref.nIterations
    assert perf.chebBeta == pytest.approx(ref.chebBeta, rel=1e-12)
    assert perf.normFactor == pytest.approx(ref.normFactor, rel=1e-12)
    assert perf.converged and ref.converged
    true_res = np.abs(m.source - orc.amul(m, psi)).sum() / ref.normFactor
    assert true_res < 1e-6 and true_res == pytest.approx(perf.finalResidual, rel=1e-6)
    # both solutions meet the 1e-6 residual: their difference is bounded by
    # what that tolerance leaves, measured against a tight FP64 solve
    x = np.linalg.solve(mg.dense(m), m.source)
    scale = np.abs(x).max()
    err_gpu = np.abs(psi - x).max() / scale
    err_ref = np.abs(ref_psi - x).max() / scale
    print(f"  error vs dense: gpu {err_gpu:.2e} oracle {err_ref:.2e}", flush=True)
    assert err_gpu < max(5 * err_ref, 1e-6)
    assert np.abs(psi - ref_psi).max() / scale < max(5 * err_ref, 1e-6)


def test_batch_branches_match_oracle(ctx, orc):
    """The same branches per segment of a batch: one system whose three RHS
    stay deflated, polish and stay deflated, and the bail matrix, whose
    segments are never deflated."""
    ctx.set_schedule(spmvSchedule=2, blockSize=0, xcdRemap=2, graphChunk=-1, gridLimit=0,
                     rowForm=0)
    nrhs = 3
    m0 = mg.poisson(12, 13, 11, seed=510, lognormal=True, zero_mean_rhs=True, rhs="normal",
                    nrhs=nrhs)
    m0.source[:, 1] += 1.0
    m1 = bail_case()
    m1.source = np.random.default_rng(77).standard_normal((m1.n, nrhs))
    systems = [m0, m1]
    batch = H.Batch(ctx, H.PCG, systems, nrhs)
    for k, m in enumerate(systems):
        batch.set_rhs(k, m.source, layout="cell_major")
    perfs = batch.solve(H.default_controls(maxRestarts=8))
    want = {(0, 0): (1, 0), (0, 1): (1, 1), (0, 2): (1, 0)}
    for k, m in enumerate(systems):
        psi_b = batch.psi(k)
        for r in range(nrhs):
            p = perfs[k * nrhs + r]
            ref_psi, ref = orc.refined_solve(m, ctl=orc.controls(maxRestarts=8),
                                             source=m.source[:, r].copy())
            print(f"  sys{k} rhs{r}: gpu deflated={p.deflated} polished={p.polished} "
                  f"restarts={p.restarts} it={p.nIterations}; oracle {ref.deflated} "
                  f"{ref.polished} {ref.restarts} {ref.nIterations}", flush=True)
            assert (p.deflated, p.polished) == want.get((k, r), (0, 0))
            assert (p.deflated, p.polished, p.restarts, p.nIterations) == (
                ref.deflated, ref.polished, ref.restarts, ref.nIterations)
            tr = np.abs(m.source[:, r] - orc.amul(m, psi_b[r])).sum() / ref.normFactor
            assert p.converged and tr < 1e-6
    batch.close()
