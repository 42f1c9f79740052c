"""hipfoam_solve_device_noalloc against hipfoam_solve_device, through the
C-ABI, on device arrays: two calls in a row on one GPU, the second with a
smaller mesh (the shared scratch array still holds the first call's
contents), symmetric hipPCG and asymmetric hipPBiCGStab.  The two entry
points run the same kernels on the same inputs, so psi, the residual field
and every solverPerformance field must agree bit for bit.  Also: the
context's own staged operator survives the call, and a mesh beyond the
scratch array's capacity is refused before any launch."""

import numpy as np
import pytest

from metal4foam_amd import _lib
from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = pytest.mark.gpu

PERF_FIELDS = ("initialResidual", "finalResidual", "normFactor", "nIterations", "restarts",
               "converged", "singular", "deflated", "polished", "chebBeta")


def _dev(ctx, a):
    return H.DeviceArray.from_host(ctx, np.ascontiguousarray(a, dtype=np.float64))


def _solve_both(ctx, m, solver):
    lower = m.lower if solver == H.PBICGSTAB else None
    d, u = _dev(ctx, m.diag), _dev(ctx, m.upper)
    lo = None if lower is None else _dev(ctx, lower)
    b = _dev(ctx, m.source)
    ctl = H.default_controls(maxRestarts=8)
    out = []
    for noalloc in (False, True):
        psi = _dev(ctx, np.zeros(m.n))
        res = _dev(ctx, np.zeros(m.n))
        if noalloc:
            perf = ctx.solve_device_noalloc(solver, d, u, lo, psi, b, ctl, residual=res)
        else:
            perf = ctx.solve_device(solver, d, u, lo, psi, b, ctl, residual=res)
        out.append((perf, psi.download(), res.download()))
    return out


CASES = [
    ("poisson_44x40x36", lambda: mg.poisson(44, 40, 36, seed=701, lognormal=True), H.PCG),
    ("convdiff_18x17x16", lambda: mg.convection_diffusion(18, 17, 16, seed=702), H.PBICGSTAB),
    ("poisson_16x15x14", lambda: mg.poisson(16, 15, 14, seed=703, lognormal=True), H.PCG),
]


def test_noalloc_matches_solve_device_two_calls(ctx, orc):
    ctx.set_schedule(spmvSchedule=2, blockSize=0, xcdRemap=2, graphChunk=-1, gridLimit=0,
                     rowForm=0)
    sizes = []
    for name, make, solver in CASES:
        m = make()
        sizes.append(m.n)
        print(f"[noalloc] {name}: {m.n} cells, solver {solver}", flush=True)
        ctx.set_topology(m)
        (p_ref, psi_ref, r_ref), (p_na, psi_na, r_na) = _solve_both(ctx, m, solver)
        assert p_ref.converged and p_na.converged
        for f in PERF_FIELDS:
            assert getattr(p_na, f) == getattr(p_ref, f), (name, f)
        np.testing.assert_array_equal(psi_na, psi_ref)
        np.testing.assert_array_equal(r_na, r_ref)
        tr = np.abs(m.source - orc.amul(m, psi_na)).sum() / p_na.normFactor
        assert tr < 1e-6
    # the calls ran on meshes of decreasing size after the largest
    assert sizes[0] > sizes[1] and sizes[0] > sizes[2]


def test_noalloc_keeps_context_operator_and_rejects_large_mesh(ctx):
    ctx.set_schedule(spmvSchedule=2, blockSize=0, xcdRemap=2, graphChunk=-1, gridLimit=0,
                     rowForm=0)
    m = mg.poisson(20, 19, 18, seed=704, lognormal=True)
    ctx.set_topology(m)
    ctx.set_matrix(m.diag, m.upper, None)
    x = np.random.default_rng(705).standard_normal(m.n)
    y0 = ctx.amul(x, H.FP64)
    # an allocation-free solve of another matrix on the same mesh
    other = m.upper * 1.5
    d, u = _dev(ctx, m.diag - 0.5 * (m.diag - m.diag.min())), _dev(ctx, other)
    b, psi = _dev(ctx, m.source), _dev(ctx, np.zeros(m.n))
    ctx.solve_device_noalloc(H.PCG, d, u, None, psi, b, H.default_controls(deflate=0))
    np.testing.assert_array_equal(ctx.amul(x, H.FP64), y0)
    # 64^3 cells need more than the 64 MiB array: refused with HIPFOAM_ERR_SIZE
    big = mg.poisson(64, 64, 64, seed=706, lognormal=False)
    ctx.set_topology(big)
    d2, u2 = _dev(ctx, big.diag), _dev(ctx, big.upper)
    b2, psi2 = _dev(ctx, big.source), _dev(ctx, np.zeros(big.n))
    rc = ctx.lib.hipfoam_solve_device_noalloc(ctx.h, H.PCG, d2.ptr, u2.ptr, None, psi2.ptr,
                                             b2.ptr, H.default_controls(), None, None)
    assert rc == _lib.ERR_SIZE, rc
    assert "scratch" in ctx.last_error
    assert np.all(psi2.download() == 0.0)
