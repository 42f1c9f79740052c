"""GPU adjoint: transpose solve A^T lambda = g on the same LDU addressing and
the face / diagonal / source gradients, against the oracle (dense transpose
solves, FP64 gradient formula) and central finite differences."""

import numpy as np
import pytest

from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = pytest.mark.gpu

# 1e-12: the reference method's absolute inner target keeps the FP32 inner
# solves from driving a non-zero-mean transpose RHS much below that (the
# oracle's restatement stalls at ~3e-13 on the same system)
TIGHT = dict(tolerance=1e-12, maxRestarts=12, maxIter=20000)


@pytest.fixture(autouse=True)
def _default_schedule(ctx):
    ctx.set_schedule(spmvSchedule=0, blockSize=256, xcdRemap=1, graphChunk=4, gridLimit=0,
                     rowForm=0)
    yield


def _loss(x, xt):
    return 0.5 * np.sum((x - xt) ** 2)


# Finite-difference losses use the plain-FP64 GPU solver at a fixed tight
# tolerance: its error varies smoothly with the perturbed coefficient, so the
# errors of the +h and -h solves cancel in the central difference (the refined
# solver's restart decisions can differ between the two).
FD = dict(tolerance=1e-14, maxIter=20000)


def _solve(ctx, m, bicg, diag=None, upper=None, lower=None, source=None, fp64=False):
    psi = np.zeros(m.n)
    perf = ctx.solve(H.PBICGSTAB if bicg else H.PCG, m.diag if diag is None else diag,
                     m.upper if upper is None else upper,
                     (m.lower if lower is None else lower) if bicg else None, psi,
                     m.source if source is None else source,
                     H.default_controls(**(FD if fp64 else TIGHT)), fp64=fp64)
    assert perf.converged or (fp64 and perf.finalResidual < 1e-12), perf.as_dict()
    return psi


@pytest.mark.parametrize("bicg", [False, True])
def test_adjoint_matches_dense_oracle(ctx, orc, bicg):
    if bicg:
        m = mg.convection_diffusion(10, 9, 8, seed=41)
    else:
        m = mg.poisson(10, 9, 8, seed=41, lognormal=True)
    ctx.set_topology(m)
    xt = np.random.default_rng(1004).standard_normal(m.n)
    x = _solve(ctx, m, bicg)
    g = x - xt
    perf, lam, du, dl, dd, db = ctx.adjoint(H.PBICGSTAB if bicg else H.PCG, m.diag, m.upper,
                                            m.lower if bicg else None, x, g,
                                            H.default_controls(**TIGHT))
    assert perf.converged
    a = mg.dense(m)
    lam_ref = np.linalg.solve(a.T, g)
    np.testing.assert_allclose(lam, lam_ref, rtol=0, atol=1e-10 * np.abs(lam_ref).max())
    ou, ol, od = orc.adjoint_grads(m, x, lam_ref)
    scale = np.abs(lam_ref).max() * np.abs(x).max()
    if bicg:
        np.testing.assert_allclose(du, ou, rtol=0, atol=1e-10 * scale)
        np.testing.assert_allclose(dl, ol, rtol=0, atol=1e-10 * scale)
    else:
        assert dl is None
        np.testing.assert_allclose(du, ou + ol, rtol=0, atol=1e-10 * scale)
    np.testing.assert_allclose(dd, od, rtol=0, atol=1e-10 * scale)
    np.testing.assert_array_equal(db, lam)
    # identity <lambda, A dx> = <A^T lambda, dx> = <g, dx>
    dx = np.random.default_rng(5).standard_normal(m.n)
    assert np.dot(lam, orc.amul(m, dx)) == pytest.approx(np.dot(g, dx), rel=1e-9)


@pytest.mark.parametrize("bicg", [False, True])
def test_adjoint_matches_finite_differences(ctx, orc, bicg):
    if bicg:
        m = mg.convection_diffusion(16, 14, 12, seed=42)
    else:
        m = mg.poisson(16, 14, 12, seed=42, lognormal=True)
    ctx.set_topology(m)
    rng = np.random.default_rng(1004)
    xt = rng.standard_normal(m.n)
    x = _solve(ctx, m, bicg, fp64=True)
    _, lam, du, dl, dd, db = ctx.adjoint(H.PBICGSTAB if bicg else H.PCG, m.diag, m.upper,
                                         m.lower if bicg else None, x, x - xt,
                                         H.default_controls(**TIGHT))
    # the plain-FP64 transpose solve gives the same gradients
    p64, lam64, du64, _, dd64, _ = ctx.adjoint(H.PBICGSTAB if bicg else H.PCG, m.diag, m.upper,
                                               m.lower if bicg else None, x, x - xt,
                                               H.default_controls(**FD), precision=H.FP64)
    assert p64.finalResidual < 1e-12
    np.testing.assert_allclose(lam64, lam, rtol=0, atol=1e-8 * np.abs(lam).max())
    np.testing.assert_allclose(du64, du, rtol=0, atol=1e-8 * np.abs(du).max())
    h = 1e-6
    faces = rng.choice(m.nfaces, 6, replace=False)
    cells = rng.choice(m.n, 6, replace=False)

    def fd(kind, idx):
        # L(+h) - L(-h) = 0.5 sum (y+ - y-)(y+ + y- - 2 x_t): no cancellation of L
        ys = []
        for sgn in (1, -1):
            d, u, lo, b = m.diag.copy(), m.upper.copy(), None if m.lower is None else \
                m.lower.copy(), m.source.copy()
            if kind == "upper":
                u[idx] += sgn * h
            elif kind == "lower":
                lo[idx] += sgn * h
            elif kind == "diag":
                d[idx] += sgn * h
            else:
                b[idx] += sgn * h
            ys.append(_solve(ctx, m, bicg, d, u, lo, b, fp64=True))
        return 0.5 * np.dot(ys[0] - ys[1], ys[0] + ys[1] - 2 * xt) / (2 * h)

    worst = 0.0
    for f in faces:
        g_fd = fd("upper", f)
        worst = max(worst, abs(g_fd - du[f]) / max(abs(du[f]), 1e-3))
        if bicg:
            g_fd = fd("lower", f)
            worst = max(worst, abs(g_fd - dl[f]) / max(abs(dl[f]), 1e-3))
    for i in cells:
        worst = max(worst, abs(fd("diag", i) - dd[i]) / max(abs(dd[i]), 1e-3))
        worst = max(worst, abs(fd("source", i) - db[i]) / max(abs(db[i]), 1e-3))
    print(f"bicg={bicg}: worst relative FD mismatch {worst:.2e}")
    assert worst < 1e-4
