"""deflate = 3 (pinned): the deflated inner right-hand side projected along
u = A 1 instead of the constant.  No reference counterpart; on a Neumann
matrix with one pinned cell it makes right-hand sides with a non-zero sum
converge (the reference's zero-mean projection leaves a uniform residual
that the constant-mode shift cannot remove).  Solutions are checked against
dense FP64 solves and the plain-FP64 GPU solve."""

import numpy as np
import pytest

from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def true_res(m, psi, b, nf):
    lower = m.upper if m.lower is None else m.lower
    ax = m.diag * psi
    ax += np.bincount(m.lower_addr, weights=m.upper * psi[m.upper_addr], minlength=m.n)
    ax += np.bincount(m.upper_addr, weights=lower * psi[m.lower_addr], minlength=m.n)
    return np.abs(b - ax).sum() / nf


@pytest.mark.parametrize("shape", [(12, 10, 8), (24, 24, 24)])
def test_pinned_projection_converges_with_nonzero_sum(ctx, shape):
    m = mg.poisson(*shape, seed=51, lognormal=True, zero_mean_rhs=False, rhs="normal")
    assert abs(m.source.sum()) > 1.0
    ctx.set_topology(m)
    ctl = dict(tolerance=1e-9, maxRestarts=8, maxIter=20000)
    ref = np.zeros(m.n)
    perf64 = ctx.solve(H.PCG, m.diag, m.upper, None, ref, m.source,
                       H.default_controls(tolerance=1e-13, maxIter=50000), fp64=True)
    assert perf64.converged
    # reference projection: left as it ends (it stalls on this right-hand side)
    psi2 = np.zeros(m.n)
    p2 = ctx.solve(H.PCG, m.diag, m.upper, None, psi2, m.source, H.default_controls(**ctl))
    psi3 = np.zeros(m.n)
    p3 = ctx.solve(H.PCG, m.diag, m.upper, None, psi3, m.source,
                   H.default_controls(deflate=3, **ctl))
    print(f"{shape}: deflate=2 it={p2.nIterations} restarts={p2.restarts} "
          f"res={p2.finalResidual:.2e}; deflate=3 it={p3.nIterations} restarts={p3.restarts} "
          f"res={p3.finalResidual:.2e}; fp64 it={perf64.nIterations}")
    assert p3.converged and p3.deflated
    assert true_res(m, psi3, m.source, p3.normFactor) == pytest.approx(p3.finalResidual,
                                                                        rel=1e-6)
    assert np.abs(psi3 - ref).max() <= 1e-6 * np.abs(ref).max()
    if shape == (24, 24, 24):
        assert p3.nIterations < p2.nIterations


def test_pinned_projection_on_compatible_rhs_matches_auto(ctx):
    """With a zero-sum right-hand side both projections solve the same
    deflated systems; the iteration counts stay close."""
    m = mg.poisson(24, 20, 16, seed=52, lognormal=True)
    ctx.set_topology(m)
    out = []
    for d in (2, 3):
        psi = np.zeros(m.n)
        p = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source,
                      H.default_controls(deflate=d, tolerance=1e-8, maxRestarts=8))
        assert p.converged
        out.append((p.nIterations, psi))
    assert abs(out[0][0] - out[1][0]) <= max(3, out[0][0] // 10)
    assert np.abs(out[0][1] - out[1][1]).max() <= 1e-6 * np.abs(out[0][1]).max()


def test_pinned_projection_batched(ctx):
    nrhs = 3
    systems = [mg.poisson(s, s + 1, s - 1, seed=60 + s, lognormal=True, zero_mean_rhs=False,
                          rhs="normal", nrhs=nrhs) for s in (9, 13, 7)]
    b = H.Batch(ctx, H.PCG, systems, nrhs)
    for k, m in enumerate(systems):
        b.set_rhs(k, m.source, layout="cell_major")
    perfs = b.solve(H.default_controls(deflate=3, tolerance=1e-8, maxRestarts=8))
    for k, m in enumerate(systems):
        psi = b.psi(k)
        x = np.linalg.solve(mg.dense(m), m.source)
        for r in range(nrhs):
            p = perfs[k * nrhs + r]
            assert p.converged and p.deflated
            assert np.abs(psi[r] - x[:, r]).max() <= 1e-5 * np.abs(x[:, r]).max()
    b.close()


def test_refined_adjoint_with_pinned_projection(ctx):
    """g = x - x_t has a non-zero sum: the refined (FP32 inner) transpose
    solve converges with deflate = 3 and gives the FP64 adjoint's
    gradients."""
    m = mg.poisson(20, 18, 16, seed=53, lognormal=True)
    ctx.set_topology(m)
    x = np.zeros(m.n)
    ctx.solve(H.PCG, m.diag, m.upper, None, x, m.source,
              H.default_controls(tolerance=1e-12, maxIter=20000), fp64=True)
    xt = np.random.default_rng(54).standard_normal(m.n)
    g = x - xt
    ref = ctx.adjoint(H.PCG, m.diag, m.upper, None, x, g,
                      H.default_controls(tolerance=1e-12, maxIter=20000), precision=H.FP64)
    got = ctx.adjoint(H.PCG, m.diag, m.upper, None, x, g,
                      H.default_controls(tolerance=1e-9, deflate=3, maxRestarts=8,
                                         maxIter=20000), precision=H.FP32)
    print(f"adjoint: refined pinned it={got[0].nIterations} restarts={got[0].restarts} "
          f"res={got[0].finalResidual:.2e}; fp64 it={ref[0].nIterations}")
    assert got[0].converged
    for a, b in zip(got[1:], ref[1:]):
        if a is None:
            continue
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()
