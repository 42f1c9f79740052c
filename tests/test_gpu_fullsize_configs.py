"""The batched (64 uneven systems x 5 RHS, 8k-512k cells each) and adjoint
(128^3 Poisson) workloads at their full size on the GPU.

Checked through properties that do not depend on size: every segment's true
FP64 residual recomputed on the CPU (oracle Amul), the deflation decision per
system kind, agreement with single-system solves, the gradients against a
plain-FP64 transpose solve, the adjoint identity and central finite
differences."""

import numpy as np
import pytest

from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


@pytest.fixture(autouse=True)
def _default_schedule(ctx):
    ctx.set_schedule(spmvSchedule=2, blockSize=0, xcdRemap=2, graphChunk=-1, gridLimit=0,
                     rowForm=0)
    yield


def _segment_residual(orc, m, psi, b, nf):
    mm = mg.LduMatrix(m.n, m.lower_addr, m.upper_addr, m.owner_start, m.losort,
                      m.losort_start, m.diag, m.upper, m.lower, b)
    return np.abs(b - orc.amul(mm, psi)).sum() / nf


def _run_batches(ctx, systems, ctl):
    nrhs = systems[0].source.shape[1]
    sym = [m for m in systems if m.lower is None]
    asym = [m for m in systems if m.lower is not None]
    out = []
    for solver, group in ((H.PCG, sym), (H.PBICGSTAB, asym)):
        bt = H.Batch(ctx, solver, group, nrhs)
        for k, m in enumerate(group):
            bt.set_rhs(k, m.source, layout="cell_major")
        perfs = bt.solve(ctl)
        out.append((solver, group, perfs, [bt.psi(k) for k in range(len(group))]))
        bt.close()
    return nrhs, out


def test_batch64_full_size(ctx, orc):
    systems = mg.config3()
    cells = sum(m.n for m in systems)
    print(f"[c3] 64 systems, {cells:,} cells", flush=True)
    assert len(systems) == 64 and 12_000_000 < cells < 14_000_000
    ctl = H.default_controls(maxRestarts=8, maxIter=20000)
    nrhs, runs = _run_batches(ctx, systems, ctl)
    assert nrhs == 5
    nseg = 0
    worst = 0.0
    for solver, group, perfs, psis in runs:
        for k, m in enumerate(group):
            for r in range(nrhs):
                p = perfs[k * nrhs + r]
                tr = _segment_residual(orc, m, psis[k][r], m.source[:, r].copy(), p.normFactor)
                worst = max(worst, tr)
                nseg += 1
                assert p.converged and tr < 1e-6, (m.name, r, p.as_dict(), tr)
                assert tr == pytest.approx(p.finalResidual, rel=1e-5, abs=1e-12)
                # deflation exactly on the pinned Poisson half
                assert p.deflated == (1 if solver == H.PCG else 0), (m.name, r)
    print(f"[c3] {nseg} segments converged, worst true residual {worst:.2e}", flush=True)
    assert nseg == 320
    # a few segments against their single-system solves (smallest and largest
    # system of each kind, first and last right-hand side)
    for solver, group, perfs, psis in runs:
        sizes = [m.n for m in group]
        for k in (int(np.argmin(sizes)), int(np.argmax(sizes))):
            m = group[k]
            ctx.set_topology(m)
            for r in (0, nrhs - 1):
                psi = np.zeros(m.n)
                ps = ctx.solve(solver, m.diag, m.upper,
                               m.lower if solver == H.PBICGSTAB else None, psi,
                               m.source[:, r].copy(), ctl)
                p = perfs[k * nrhs + r]
                print(f"[c3] {m.name} rhs {r}: batch it={p.nIterations} "
                      f"single it={ps.nIterations}", flush=True)
                assert (p.deflated, p.polished) == (ps.deflated, ps.polished)
                assert p.normFactor == pytest.approx(ps.normFactor, rel=1e-12)
                assert abs(p.nIterations - ps.nIterations) <= max(3, ps.nIterations // 10)
                scale = max(np.abs(psi).max(), 1.0)
                assert np.abs(psis[k][r] - psi).max() <= 1e-4 * scale


def test_batch64_as_specified_rhs_with_pinned_projection(ctx, orc):
    """The batch with the Poisson right-hand sides as drawn (N(0,1), non-zero
    sum), solved with the pinned projection (deflate 3)."""
    systems = mg.config3(zero_mean_poisson=False)
    sym = [m for m in systems if m.lower is None]
    assert all(abs(m.source[:, r].sum()) > 1.0 for m in sym for r in range(5))
    ctl = H.default_controls(maxRestarts=8, maxIter=20000, deflate=3)
    nrhs, runs = _run_batches(ctx, systems, ctl)
    conv = 0
    iters = []
    for solver, group, perfs, psis in runs:
        for k, m in enumerate(group):
            for r in range(nrhs):
                p = perfs[k * nrhs + r]
                tr = _segment_residual(orc, m, psis[k][r], m.source[:, r].copy(), p.normFactor)
                if p.converged and tr < 1e-6:
                    conv += 1
                iters.append(p.nIterations)
    print(f"[c3 as specified, deflate 3] converged {conv}/320, inner iterations "
          f"min/median/max {min(iters)}/{int(np.median(iters))}/{max(iters)}", flush=True)
    assert conv == 320


def test_adjoint128_full_size(ctx, orc):
    m = mg.config1()
    xt = mg.config4_target(m.n)
    print(f"[c4] 128^3 adjoint, {m.n:,} cells", flush=True)
    ctx.set_topology(m)
    ctl = H.default_controls(maxRestarts=8, maxIter=20000)
    x = np.zeros(m.n)
    pf = ctx.solve(H.PCG, m.diag, m.upper, None, x, m.source, ctl)
    assert pf.converged
    g = x - xt
    pb, lam, du, dl, dd, db = ctx.adjoint(H.PCG, m.diag, m.upper, None, x, g,
                                          H.default_controls(maxRestarts=8, maxIter=20000,
                                                             deflate=3))
    print(f"[c4] forward it={pf.nIterations}, adjoint (fp32 refined, deflate 3) "
          f"it={pb.nIterations} res={pb.finalResidual:.2e}", flush=True)
    assert pb.converged and dl is None
    # reference gradients: plain-FP64 transpose solve, tight
    p64, lam64, du64, _, dd64, db64 = ctx.adjoint(
        H.PCG, m.diag, m.upper, None, x, g, H.default_controls(tolerance=1e-12, maxIter=20000),
        precision=H.FP64)
    assert p64.finalResidual < 1e-11
    diffs = {k: float(np.abs(a - b).max() / np.abs(b).max())
             for k, a, b in (("lambda", lam, lam64), ("dUpper", du, du64),
                             ("dDiag", dd, dd64), ("dSource", db, db64))}
    print(f"[c4] gradients vs FP64 transpose solve: {diffs}", flush=True)
    assert max(diffs.values()) <= 1e-6
    # dL/db = lambda; gradient formulas against the oracle's from the same lambda
    np.testing.assert_array_equal(db, lam)
    ou, ol, od = orc.adjoint_grads(m, x, lam)
    np.testing.assert_allclose(du, ou + ol, rtol=0, atol=1e-13 * np.abs(du).max())
    np.testing.assert_allclose(dd, od, rtol=0, atol=1e-13 * np.abs(dd).max())
    # adjoint identity <lambda, A dx> = <A^T lambda, dx> = <g, dx>, FP64.
    # lambda carries a large constant mode (g has a non-zero sum on the pinned
    # matrix), so these sums cancel heavily: each side is exact up to FP64
    # rounding, at most (log2 n + row length) eps sum_i |lambda_i| (|A| |dx|)_i
    # for numpy's pairwise dot over an SpMV (64 eps with margin), and
    # <A^T lambda - g, dx> is bounded by the transpose solve's own residual,
    # ||g - A^T lambda||_1 max|dx| = finalResidual * normFactor * max|dx|.
    rng = np.random.default_rng(1004)
    absA = mg.LduMatrix(m.n, m.lower_addr, m.upper_addr, m.owner_start, m.losort,
                        m.losort_start, np.abs(m.diag), np.abs(m.upper), None)
    for lam_k, perf_k in ((lam64, p64), (lam, pb)):
        for _ in range(2):
            dx = rng.standard_normal(m.n)
            lhs = np.dot(lam_k, orc.amul(m, dx))
            rhs = np.dot(orc.amul(m, lam_k, transpose=True), dx)
            gdx = np.dot(g, dx)
            rounding = 64 * np.finfo(np.float64).eps * np.dot(np.abs(lam_k),
                                                              orc.amul(absA, np.abs(dx)))
            solve_bound = perf_k.finalResidual * perf_k.normFactor * np.abs(dx).max()
            print(f"[c4] identity: <lam, A dx> = {lhs:.9e}, <A^T lam, dx> = {rhs:.9e}, "
                  f"<g, dx> = {gdx:.9e}; rounding bound {rounding:.2e}, solve bound "
                  f"{solve_bound:.2e}", flush=True)
            assert abs(lhs - rhs) <= rounding
            assert abs(lhs - gdx) <= solve_bound + rounding
    # central differences on a few faces and cells, tight plain-FP64 solves
    tight = H.default_controls(tolerance=1e-14, maxIter=20000)
    x64 = np.zeros(m.n)
    ctx.solve(H.PCG, m.diag, m.upper, None, x64, m.source, tight, fp64=True)
    _, _, duT, _, ddT, _ = ctx.adjoint(H.PCG, m.diag, m.upper, None, x64, x64 - xt, tight,
                                       precision=H.FP64)
    h = 1e-6

    def solve(d, u):
        y = np.zeros(m.n)
        ctx.solve(H.PCG, d, u, None, y, m.source, tight, fp64=True)
        return y

    def dloss(yp, ym):
        return 0.5 * np.dot(yp - ym, yp + ym - 2.0 * xt)

    worst = 0.0
    scale = max(np.abs(duT).max(), np.abs(ddT).max())
    for f in rng.choice(m.nfaces, 4, replace=False):
        u1, u2 = m.upper.copy(), m.upper.copy()
        u1[f] += h
        u2[f] -= h
        gfd = dloss(solve(m.diag, u1), solve(m.diag, u2)) / (2 * h)
        worst = max(worst, abs(gfd - duT[f]) / max(abs(duT[f]), 1e-3 * scale))
    for i in rng.choice(m.n, 4, replace=False):
        d1, d2 = m.diag.copy(), m.diag.copy()
        d1[i] += h
        d2[i] -= h
        gfd = dloss(solve(d1, m.upper), solve(d2, m.upper)) / (2 * h)
        worst = max(worst, abs(gfd - ddT[i]) / max(abs(ddT[i]), 1e-3 * scale))
    print(f"[c4] FD check on 4 faces + 4 cells: max relative error {worst:.2e}", flush=True)
    assert worst < 1e-4
