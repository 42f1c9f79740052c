"""The persistent inner loop (k_pcg_persistent: the whole Chebyshev PCG loop
of a small mesh as one launch on one XCD) against the launched loop, through
the C-ABI.

It runs the same rows, tiles, partials and fold order as the launched
kernels on the per-tile launch shape, so every solve must agree bit for bit:
psi, residual, iteration counts, restarts and the refinement flags, for the
refined solve (single and several inner solves, polish), the plain-FP64
solve, and the reference-backend inner entry (Chebyshev orders 2 and 8,
iterationsPerSync > 1, maxIter exits).  The persistent counters prove the
persistent kernel ran and was never abandoned."""

import numpy as np
import pytest

from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _auto_schedule(ctx):
    # the one-workgroup loop (localLoop) would take the meshes of at most
    # 3,072 rows: off here, so every case runs the one-XCD loop
    ctx.set_schedule(spmvSchedule=2, blockSize=0, xcdRemap=2, graphChunk=-1, gridLimit=0,
                     rowForm=0, persistent=-1, localLoop=0)
    yield
    ctx.set_schedule(persistent=-1, localLoop=-1)


def polish_case():
    m = mg.poisson(12, 13, 11, seed=510, lognormal=True, zero_mean_rhs=True, rhs="normal",
                   nrhs=3)
    m.source = m.source[:, 1] + 1.0
    m.name = "polish_mean_rhs"
    return m


MESHES = {
    "c0_100x100": mg.config0,
    "poisson24_lognormal": lambda: mg.poisson(24, 24, 24, seed=11, lognormal=True),
    "poisson_7x9x5": lambda: mg.poisson(7, 9, 5, seed=12, lognormal=True),
    "poisson_17x13x11_unpinned_rhs": lambda: mg.poisson(17, 13, 11, seed=13, lognormal=True,
                                                        zero_mean_rhs=False),
    "poisson32_128_tiles": lambda: mg.poisson(32, 32, 32, seed=14, lognormal=True),
    "poisson20_random_numbering": lambda: mg.random_renumbering(
        mg.poisson(20, 20, 20, seed=15, lognormal=True), seed=16),
    "polish": polish_case,
}


def _solve(ctx, m, persistent, fp64=False, **ctl):
    ctx.set_schedule(persistent=persistent)
    ctx.set_topology(m)
    psi = np.zeros(m.n)
    before = ctx.persistent_stats()
    perf, res = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source,
                          H.default_controls(**ctl), residual=True, fp64=fp64)
    after = ctx.persistent_stats()
    return psi, res, perf, (after[0] - before[0], after[1] - before[1])


def _same_perf(a, b):
    for k in ("nIterations", "restarts", "converged", "singular", "deflated", "polished"):
        assert getattr(a, k) == getattr(b, k), k
    for k in ("initialResidual", "finalResidual", "normFactor", "chebBeta"):
        assert getattr(a, k) == getattr(b, k), k


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("fp64", [False, True], ids=["refined", "fp64"])
def test_persistent_equals_launched(ctx, name, fp64):
    m = MESHES[name]()
    ctl = dict(maxRestarts=6) if not fp64 else dict(maxIter=2000)
    print(f"[persistent] {name}: {m.n} cells, {'fp64' if fp64 else 'refined'}", flush=True)
    psi0, res0, perf0, used0 = _solve(ctx, m, 0, fp64, **ctl)
    psi1, res1, perf1, used1 = _solve(ctx, m, 1, fp64, **ctl)
    assert used0 == (0, 0)
    assert used1[0] >= 1 and used1[1] == 0, (used1, hex(ctx.persistent_abort()[2]))
    _same_perf(perf0, perf1)
    np.testing.assert_array_equal(psi1, psi0)
    np.testing.assert_array_equal(res1, res0)
    if name != "poisson_17x13x11_unpinned_rhs":
        assert perf1.converged, (perf1.finalResidual, perf1.nIterations)
    if name == "polish" and not fp64:
        assert perf1.polished == 1


ASYM = {
    "convdiff_20": lambda: mg.convection_diffusion(20, 20, 20, seed=41),
    "convdiff_13x11x9_peclet8": lambda: mg.convection_diffusion(13, 11, 9, seed=42, umax=8.0),
    "convdiff_100x100x1": lambda: mg.convection_diffusion(100, 100, 1, seed=43),
    "convdiff_24_random_numbering": lambda: mg.random_renumbering(
        mg.convection_diffusion(24, 24, 24, seed=44), seed=45),
}


@pytest.mark.parametrize("name", list(ASYM))
@pytest.mark.parametrize("fp64", [False, True], ids=["refined", "fp64"])
def test_persistent_bicgstab_equals_launched(ctx, name, fp64):
    """hipPBiCGStab on asymmetric convection-diffusion: persistent and
    launched loops bit for bit."""
    m = ASYM[name]()
    print(f"[persistent] {name}: {m.n} cells, bicgstab {'fp64' if fp64 else 'refined'}",
          flush=True)
    out = []
    for p in (0, 1):
        ctx.set_schedule(persistent=p)
        ctx.set_topology(m)
        psi = np.zeros(m.n)
        before = ctx.persistent_stats()
        perf, res = ctx.solve(H.PBICGSTAB, m.diag, m.upper, m.lower, psi, m.source,
                              H.default_controls(maxRestarts=6), residual=True, fp64=fp64)
        after = ctx.persistent_stats()
        out.append((psi, res, perf, (after[0] - before[0], after[1] - before[1])))
    (psi0, res0, perf0, used0), (psi1, res1, perf1, used1) = out
    assert used0 == (0, 0)
    assert used1[0] >= 1 and used1[1] == 0, (used1, hex(ctx.persistent_abort()[2]))
    _same_perf(perf0, perf1)
    np.testing.assert_array_equal(psi1, psi0)
    np.testing.assert_array_equal(res1, res0)
    assert perf1.converged, (perf1.finalResidual, perf1.nIterations)


@pytest.mark.parametrize("max_iter", [300, 4])
def test_persistent_bicgstab_inner_entry(ctx, max_iter):
    """hipfoam_bicgstab (metalFoamBackend::bicgstab): stats and delta bit for
    bit, through the |sA| exit (a loose target) and a maxIter exit."""
    m = mg.convection_diffusion(18, 17, 16, seed=46)
    ctx.set_topology(m)
    ctx.set_coeffs(m.diag, m.upper, m.lower)
    r = m.source.copy()
    for target in (1e-6 * np.abs(r).sum(), 0.5 * np.abs(r).sum()):
        out = []
        for p in (0, 1):
            ctx.set_schedule(persistent=p)
            before = ctx.persistent_stats()[0]
            st = ctx.bicgstab(r, target, max_iter)
            out.append((st, ctx.delta(), ctx.persistent_stats()[0] - before))
        (s0, d0, u0), (s1, d1, u1) = out
        assert (u0, u1) == (0, 1)
        for k in ("nIterations", "finalResSum", "converged", "singular", "stagnated"):
            assert getattr(s0, k) == getattr(s1, k), k
        np.testing.assert_array_equal(d1, d0)


REFINE_BRANCHES = [
    ("deflate_off", dict(deflate=0)),
    ("deflate_forced", dict(deflate=1)),
    ("pinned_projection", dict(deflate=3)),
    ("no_restarts", dict(maxRestarts=0)),
    ("min_iter", dict(minIter=40)),
    ("max_iter_cap", dict(maxIter=7)),
    ("ips3_order5", dict(iterationsPerSync=3, chebyshevOrder=5)),
    ("rel_tol", dict(relTol=1e-3)),
]


@pytest.mark.parametrize("name,ctl", REFINE_BRANCHES, ids=[b[0] for b in REFINE_BRANCHES])
def test_persistent_refined_branches(ctx, name, ctl):
    """Every branch of the refined solve's driver (deflation on/off/forced,
    pinned projection, restarts, minIter, the maxIter cap, sync batches,
    relTol) gives the same solve with persistent inner loops."""
    m = mg.poisson(19, 17, 13, seed=61, lognormal=True, zero_mean_rhs=False)
    out = [_solve(ctx, m, p, **ctl) for p in (0, 1)]
    (psi0, res0, perf0, used0), (psi1, res1, perf1, used1) = out
    assert used0 == (0, 0) and used1[0] >= 1 and used1[1] == 0, used1
    _same_perf(perf0, perf1)
    np.testing.assert_array_equal(psi1, psi0)
    np.testing.assert_array_equal(res1, res0)


def test_persistent_refined_converged_at_start(ctx):
    """psi already solves the system: no inner solve, the same zero restarts."""
    m = mg.poisson(16, 15, 14, seed=62, lognormal=True)
    ctx.set_topology(m)
    psi = np.zeros(m.n)
    ctx.set_schedule(persistent=0)
    ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source,
              H.default_controls(tolerance=1e-10, maxRestarts=12))
    outs = []
    for p in (0, 1):
        ctx.set_schedule(persistent=p)
        x = psi.copy()
        perf = ctx.solve(H.PCG, m.diag, m.upper, None, x, m.source, H.default_controls())
        outs.append((x, perf))
    (x0, p0), (x1, p1) = outs
    assert p0.nIterations == 0 and p0.converged
    _same_perf(p0, p1)
    np.testing.assert_array_equal(x1, x0)


def test_persistent_auto_limit(ctx):
    """Auto (-1) runs the persistent loop up to 160 tiles (FP32 inner loop) or
    96 (plain FP64) and the launched loop beyond; the explicit block shape of
    512 threads is not the per-tile shape."""
    small = mg.poisson(16, 16, 160, seed=21, lognormal=True)   # 160 tiles
    large = mg.poisson(16, 16, 161, seed=21, lognormal=True)   # 161 tiles
    small64 = mg.poisson(16, 16, 96, seed=21, lognormal=True)  # 96 tiles
    large64 = mg.poisson(16, 16, 97, seed=21, lognormal=True)  # 97 tiles
    used = _solve(ctx, small, -1)[3]
    assert used[0] >= 1 and used[1] == 0
    assert _solve(ctx, large, -1)[3] == (0, 0)
    used = _solve(ctx, large, 1)[3]
    assert used[0] >= 1 and used[1] == 0
    used = _solve(ctx, small64, -1, fp64=True)[3]
    assert used[0] == 1 and used[1] == 0
    assert _solve(ctx, large64, -1, fp64=True)[3] == (0, 0)
    assert _solve(ctx, large64, 1, fp64=True)[3] == (1, 0)
    ctx.set_schedule(blockSize=512)
    assert _solve(ctx, small, 1)[3] == (0, 0)
    ctx.set_schedule(blockSize=0)
    # Jacobi PCG (order 1) keeps the launched loop
    assert _solve(ctx, small, 1, chebyshevOrder=1)[3] == (0, 0)


@pytest.mark.parametrize("order,ips,max_iter", [(8, 1, 200), (2, 1, 300), (8, 3, 200),
                                                (8, 1, 5), (3, 4, 7)])
def test_persistent_inner_entry(ctx, orc, order, ips, max_iter):
    """hipfoam_pcg (metalFoamBackend::pcg): stats and delta bit for bit, and
    the oracle's inner FP32 loop reaches the same iteration count."""
    m = mg.poisson(18, 18, 18, seed=31, lognormal=True)
    ctx.set_topology(m)
    ctx.set_coeffs(m.diag, m.upper, None)
    r = m.source - m.source.mean()
    target = 1e-5 * np.abs(r).sum()
    out = []
    for p in (0, 1):
        ctx.set_schedule(persistent=p)
        before = ctx.persistent_stats()[0]
        st = ctx.pcg(r, target, max_iter, ips=ips, order=order, beta=2.0, ratio=30.0)
        out.append((st, ctx.delta(), ctx.persistent_stats()[0] - before))
    (s0, d0, u0), (s1, d1, u1) = out
    assert (u0, u1) == (0, 1)
    for k in ("nIterations", "finalResSum", "converged", "singular", "stagnated"):
        assert getattr(s0, k) == getattr(s1, k), k
    np.testing.assert_array_equal(d1, d0)
    if max_iter < 10:
        assert s1.nIterations == max_iter and not s1.converged


def test_persistent_against_oracle_c0(ctx, orc):
    m = mg.config0()
    ctx.set_schedule(persistent=1)
    ctx.set_topology(m)
    psi = np.zeros(m.n)
    perf = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source,
                     H.default_controls(maxRestarts=4))
    ref_psi, ref = orc.refined_solve(m, ctl=orc.controls(maxRestarts=4))
    true_res = np.abs(m.source - orc.amul(m, psi)).sum() / ref.normFactor
    print(f"[persistent] c0: gpu iters={perf.nIterations} restarts={perf.restarts} "
          f"res={perf.finalResidual:.3e}; oracle iters={ref.nIterations} "
          f"res={ref.finalResidual:.3e}", flush=True)
    assert perf.converged and true_res < 1e-6
    assert abs(perf.nIterations - ref.nIterations) <= max(4, ref.nIterations // 10)
    assert np.abs(psi - ref_psi).max() <= 1e-5 * np.abs(ref_psi).max()


RECOVERY = [("pcg", 1, False), ("pcg", 2, False), ("pcg", 3, False), ("bicgstab", 1, False),
            ("bicgstab", 2, False), ("pcg", 1, True), ("bicgstab", 1, True)]


@pytest.mark.parametrize("kind,count,fp64", RECOVERY,
                         ids=[f"{k}-{c}{'-fp64' if f else ''}" for k, c, f in RECOVERY])
def test_persistent_abandon_recovery(ctx, kind, count, fp64):
    """Persistent launches reported abandoned (hipfoam_simulate_abandon: each
    runs to its end, then its state is set back to running and its abort word
    set, as a launch that gave up leaves them).  The refinement queued behind
    the launch writes nothing, the driver solves again -- persistent, then on
    the launched loop when that launch is abandoned too (count 2), and again
    at the next restart (count 3) -- and every solve is bit-identical to the
    undisturbed one."""
    m = mg.config0() if kind == "pcg" else mg.convection_diffusion(100, 100, 1, seed=43)
    solver = H.PCG if kind == "pcg" else H.PBICGSTAB
    ctl = dict(maxRestarts=6) if not fp64 else dict(maxIter=2000)
    print(f"[persistent] recovery {kind} x{count} {'fp64' if fp64 else 'refined'}", flush=True)
    ctx.set_schedule(persistent=-1)
    ctx.set_topology(m)
    out = []
    try:
        for sim in (0, count):
            ctx.simulate_abandon(sim)
            before = ctx.persistent_abort()
            psi = np.zeros(m.n)
            perf, res = ctx.solve(solver, m.diag, m.upper, m.lower, psi, m.source,
                                  H.default_controls(**ctl), residual=True, fp64=fp64)
            after = ctx.persistent_abort()
            out.append((psi, res, perf, after[1] - before[1], after[2]))
    finally:
        ctx.simulate_abandon(0)
    (psi0, res0, perf0, ab0, _), (psi1, res1, perf1, ab1, code) = out
    assert ab0 == 0
    assert ab1 == count and code == 0x40000, (ab1, hex(code))
    _same_perf(perf0, perf1)
    np.testing.assert_array_equal(psi1, psi0)
    np.testing.assert_array_equal(res1, res0)
    assert perf1.converged
