"""The one-workgroup inner loops (k_pcg_local, k_bicg_local: the whole
Chebyshev PCG or BiCGStab loop of a mesh of at most 2,816 rows (BiCGStab: 1,536) in
one workgroup, its vectors in LDS; PCG meshes of more than 2,048 rows run a
six-pass build) against the launched loop, through the
C-ABI.

It runs the same rows, tiles, partials and fold order as the launched
kernels on the per-tile launch shape, so every solve must agree bit for bit:
psi, residual, iteration counts, restarts and the refinement flags, for the
refined solve (every branch of its driver), the plain-FP64 solve (at most
1,536 rows for PCG, 1,280 for BiCGStab) and the reference-backend inner
entries (Chebyshev orders 2, 3 and 8, iterationsPerSync > 1, maxIter exits,
BiCGStab's |sA| exit).  The launch counter proves the
one-workgroup kernel ran, and the limits keep it off where it does not
apply."""

import numpy as np
import pytest

from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = pytest.mark.gpu

MAX_ROWS = {False: 2816, True: 1536}  # PCG: FP32 inner loop, plain FP64
MAX_ROWS_BICG = {False: 1536, True: 1280}


@pytest.fixture(autouse=True)
def _auto_schedule(ctx):
    ctx.set_schedule(spmvSchedule=2, blockSize=0, xcdRemap=2, graphChunk=-1, gridLimit=0,
                     rowForm=0, persistent=-1, localLoop=-1)
    yield
    ctx.set_schedule(persistent=-1, localLoop=-1)


def polish_case():
    m = mg.poisson(12, 13, 11, seed=510, lognormal=True, zero_mean_rhs=True, rhs="normal",
                   nrhs=3)
    m.source = m.source[:, 1] + 1.0
    m.name = "polish_mean_rhs"
    return m


MESHES = {
    "cavity_20x20": lambda: mg.poisson(20, 20, 1, seed=1000, lognormal=False),
    "cavity_45x45": lambda: mg.poisson(45, 45, 1, seed=1000, lognormal=False),
    "cavity_50x50_six_passes": lambda: mg.poisson(50, 50, 1, seed=1000, lognormal=False),
    "cavity_54x54_over_the_limit": lambda: mg.poisson(54, 54, 1, seed=1000, lognormal=False),
    "poisson_7x9x5": lambda: mg.poisson(7, 9, 5, seed=12, lognormal=True),
    "poisson12_lognormal": lambda: mg.poisson(12, 12, 12, seed=11, lognormal=True),
    "poisson_32x64_four_pass_max": lambda: mg.poisson(32, 64, 1, seed=17, lognormal=True),
    "poisson_41x50_six_pass_min": lambda: mg.poisson(41, 50, 1, seed=22, lognormal=True),
    "poisson_44x64_max_rows": lambda: mg.poisson(44, 64, 1, seed=17, lognormal=True),
    "poisson_11x10x9_unpinned_rhs": lambda: mg.poisson(11, 10, 9, seed=13, lognormal=True,
                                                       zero_mean_rhs=False),
    "poisson12_rcm_explicit_tiles": lambda: mg.rcm_renumbering(mg.random_renumbering(
        mg.poisson(12, 12, 12, seed=15, lognormal=True), seed=16)),
    "polish": polish_case,
    # pass and tile boundaries of the workgroup: 6 rows, one wave, a one-row
    # last tile, a second pass of one row
    "poisson_3x2x1": lambda: mg.poisson(3, 2, 1, seed=18, lognormal=True),
    "chain_64": lambda: mg.poisson(64, 1, 1, seed=19, lognormal=False),
    "chain_257": lambda: mg.poisson(257, 1, 1, seed=20, lognormal=False),
    "chain_513": lambda: mg.poisson(513, 1, 1, seed=21, lognormal=False),
}


def _solve(ctx, m, local, fp64=False, kind=H.PCG, **ctl):
    """One solve on the launched loop (local False) or with the one-workgroup
    loop allowed; the one-XCD persistent loop is off in both."""
    ctx.set_schedule(persistent=0, localLoop=-1 if local else 0)
    ctx.set_topology(m)
    psi = np.zeros(m.n)
    before = ctx.local_launches()
    perf, res = ctx.solve(kind, m.diag, m.upper, m.lower, psi, m.source,
                          H.default_controls(**ctl), residual=True, fp64=fp64)
    return psi, res, perf, ctx.local_launches() - before


def _same_perf(a, b):
    for k in ("nIterations", "restarts", "converged", "singular", "deflated", "polished"):
        assert getattr(a, k) == getattr(b, k), k
    for k in ("initialResidual", "finalResidual", "normFactor", "chebBeta"):
        assert getattr(a, k) == getattr(b, k), k


@pytest.mark.parametrize("name", list(MESHES))
@pytest.mark.parametrize("fp64", [False, True], ids=["refined", "fp64"])
def test_local_equals_launched(ctx, name, fp64):
    m = MESHES[name]()
    ctl = dict(maxRestarts=6) if not fp64 else dict(maxIter=2000)
    print(f"[local] {name}: {m.n} cells, {'fp64' if fp64 else 'refined'}", flush=True)
    ctx.set_topology(m)
    pat, expl, faces = ctx.slot_tiles()
    assert faces == 0
    if name == "poisson12_rcm_explicit_tiles":
        assert expl > 0
    psi0, res0, perf0, used0 = _solve(ctx, m, False, fp64, **ctl)
    psi1, res1, perf1, used1 = _solve(ctx, m, True, fp64, **ctl)
    assert used0 == 0
    if m.n <= MAX_ROWS[fp64]:
        assert used1 >= 1
    else:
        assert used1 == 0
    _same_perf(perf0, perf1)
    np.testing.assert_array_equal(psi1, psi0)
    np.testing.assert_array_equal(res1, res0)
    if name != "poisson_11x10x9_unpinned_rhs":
        assert perf1.converged, (perf1.finalResidual, perf1.nIterations)
    if name == "polish" and not fp64:
        assert perf1.polished == 1


REFINE_BRANCHES = [
    ("deflate_off", dict(deflate=0)),
    ("deflate_forced", dict(deflate=1)),
    ("pinned_projection", dict(deflate=3)),
    ("no_restarts", dict(maxRestarts=0)),
    ("min_iter", dict(minIter=40)),
    ("max_iter_cap", dict(maxIter=7)),
    ("ips3_order5", dict(iterationsPerSync=3, chebyshevOrder=5)),
    ("order2", dict(chebyshevOrder=2)),
    ("rel_tol", dict(relTol=1e-3)),
]


@pytest.mark.parametrize("name,ctl", REFINE_BRANCHES, ids=[b[0] for b in REFINE_BRANCHES])
def test_local_refined_branches(ctx, name, ctl):
    """Every branch of the refined solve's driver gives the same solve with
    one-workgroup inner loops."""
    m = mg.poisson(13, 12, 11, seed=61, lognormal=True, zero_mean_rhs=False)
    (psi0, res0, perf0, used0), (psi1, res1, perf1, used1) = [
        _solve(ctx, m, local, **ctl) for local in (False, True)]
    assert used0 == 0 and used1 >= 1
    _same_perf(perf0, perf1)
    np.testing.assert_array_equal(psi1, psi0)
    np.testing.assert_array_equal(res1, res0)


@pytest.mark.parametrize("order,ips,max_iter", [(8, 1, 200), (2, 1, 300), (8, 3, 200),
                                                (8, 1, 5), (3, 4, 7)])
def test_local_inner_entry(ctx, order, ips, max_iter):
    """hipfoam_pcg (metalFoamBackend::pcg): stats and delta bit for bit."""
    m = mg.poisson(12, 12, 12, seed=31, lognormal=True)
    ctx.set_topology(m)
    ctx.set_coeffs(m.diag, m.upper, None)
    r = m.source - m.source.mean()
    target = 1e-5 * np.abs(r).sum()
    out = []
    for local in (0, -1):
        ctx.set_schedule(persistent=0, localLoop=local)
        before = ctx.local_launches()
        st = ctx.pcg(r, target, max_iter, ips=ips, order=order, beta=2.0, ratio=30.0)
        out.append((st, ctx.delta(), ctx.local_launches() - before))
    (s0, d0, u0), (s1, d1, u1) = out
    assert (u0, u1) == (0, 1)
    for k in ("nIterations", "finalResSum", "converged", "singular", "stagnated"):
        assert getattr(s0, k) == getattr(s1, k), k
    np.testing.assert_array_equal(d1, d0)
    if max_iter < 10:
        assert s1.nIterations == max_iter and not s1.converged


def test_local_limits(ctx):
    """The one-workgroup loop takes at most 2,816 rows (FP32 inner loop) or
    1,536 (plain FP64), only row-slot tiles, only the per-tile launch shape,
    only Chebyshev PCG, and never under kernel timing; elsewhere the one-XCD
    loop or the launched loop runs, with the same result."""
    at = mg.poisson(44, 64, 1, seed=21, lognormal=True)     # 2,816 rows
    over = mg.poisson(313, 9, 1, seed=21, lognormal=True)   # 2,817 rows
    at64 = mg.poisson(48, 32, 1, seed=21, lognormal=True)   # 1,536 rows
    over64 = mg.poisson(53, 29, 1, seed=21, lognormal=True)  # 1,537 rows
    assert _solve(ctx, at, True)[3] >= 1
    assert _solve(ctx, over, True)[3] == 0
    assert _solve(ctx, at64, True, fp64=True)[3] == 1
    assert _solve(ctx, over64, True, fp64=True)[3] == 0
    assert _solve(ctx, over64, True)[3] >= 1
    # face tiles: a random numbering puts rows of more than 3 owner faces in
    # every tile
    rnd = mg.random_renumbering(mg.poisson(12, 12, 12, seed=15, lognormal=True), seed=16)
    ctx.set_topology(rnd)
    assert ctx.slot_tiles()[2] > 0
    a = _solve(ctx, rnd, False)
    b = _solve(ctx, rnd, True)
    assert b[3] == 0
    np.testing.assert_array_equal(b[0], a[0])
    ctx.set_schedule(blockSize=512)
    assert _solve(ctx, at, True)[3] == 0
    ctx.set_schedule(blockSize=0)
    assert _solve(ctx, at, True, chebyshevOrder=1)[3] == 0
    # BiCGStab: 1,536 rows (FP32 inner loop) or 1,280 (plain FP64)
    cd_at = mg.convection_diffusion(32, 48, 1, seed=41)     # 1,536 rows
    cd_over = mg.convection_diffusion(53, 29, 1, seed=41)   # 1,537 rows
    cd_at64 = mg.convection_diffusion(32, 40, 1, seed=41)   # 1,280 rows
    cd_over64 = mg.convection_diffusion(61, 21, 1, seed=41)  # 1,281 rows
    assert _solve(ctx, cd_at, True, kind=H.PBICGSTAB)[3] >= 1
    assert _solve(ctx, cd_over, True, kind=H.PBICGSTAB)[3] == 0
    assert _solve(ctx, cd_at64, True, fp64=True, kind=H.PBICGSTAB)[3] == 1
    assert _solve(ctx, cd_over64, True, fp64=True, kind=H.PBICGSTAB)[3] == 0
    assert _solve(ctx, cd_over64, True, kind=H.PBICGSTAB)[3] >= 1
    ctx.set_kernel_timing(True)
    try:
        assert _solve(ctx, at, True)[3] == 0
    finally:
        ctx.set_kernel_timing(False)


def test_local_against_oracle_cavity(ctx, orc):
    """The 20x20 2D pressure Poisson on the one-workgroup loop against the
    oracle's refined solve."""
    m = mg.poisson(20, 20, 1, seed=1000, lognormal=False)
    ctx.set_schedule(persistent=-1, localLoop=-1)
    ctx.set_topology(m)
    psi = np.zeros(m.n)
    before = ctx.local_launches()
    perf = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source,
                     H.default_controls(maxRestarts=4))
    assert ctx.local_launches() - before >= 1
    ref_psi, ref = orc.refined_solve(m, ctl=orc.controls(maxRestarts=4))
    true_res = np.abs(m.source - orc.amul(m, psi)).sum() / ref.normFactor
    print(f"[local] cavity 20x20: gpu iters={perf.nIterations} restarts={perf.restarts} "
          f"res={perf.finalResidual:.3e}; oracle iters={ref.nIterations} "
          f"res={ref.finalResidual:.3e}", flush=True)
    assert perf.converged and true_res < 1e-6
    assert abs(perf.nIterations - ref.nIterations) <= max(4, ref.nIterations // 10)
    assert np.abs(psi - ref_psi).max() <= 1e-5 * np.abs(ref_psi).max()


ASYM = {
    "convdiff_20x20": lambda: mg.convection_diffusion(20, 20, 1, seed=41),
    "convdiff_11x11x10": lambda: mg.convection_diffusion(11, 11, 10, seed=42),
    "convdiff_30x30_peclet8": lambda: mg.convection_diffusion(30, 30, 1, seed=43, umax=8.0),
    "convdiff_45x45_over_the_limit": lambda: mg.convection_diffusion(45, 45, 1, seed=44),
    "momentum_20x20": lambda: mg.convection_diffusion(20, 20, 1, seed=1005, shift=1.0, umax=2.0),
    "convdiff_32x40_fp64_max_rows": lambda: mg.convection_diffusion(32, 40, 1, seed=45),
    "convdiff_32x48_max_rows": lambda: mg.convection_diffusion(32, 48, 1, seed=48),
    "convdiff10_rcm_explicit_tiles": lambda: mg.rcm_renumbering(mg.random_renumbering(
        mg.convection_diffusion(10, 10, 10, seed=46), seed=47)),
    "convdiff_3x2x1": lambda: mg.convection_diffusion(3, 2, 1, seed=49),
    "convdiff_257x2x1": lambda: mg.convection_diffusion(257, 2, 1, seed=50),
    "convdiff_513x2x1": lambda: mg.convection_diffusion(513, 2, 1, seed=51),
}


@pytest.mark.parametrize("name", list(ASYM))
@pytest.mark.parametrize("fp64", [False, True], ids=["refined", "fp64"])
def test_local_bicgstab_equals_launched(ctx, name, fp64):
    """hipPBiCGStab on asymmetric convection-diffusion: the one-workgroup and
    the launched loops bit for bit."""
    m = ASYM[name]()
    print(f"[local] {name}: {m.n} cells, bicgstab {'fp64' if fp64 else 'refined'}", flush=True)
    ctx.set_topology(m)
    assert ctx.slot_tiles()[2] == 0
    ctl = dict(maxRestarts=6)
    psi0, res0, perf0, used0 = _solve(ctx, m, False, fp64, kind=H.PBICGSTAB, **ctl)
    psi1, res1, perf1, used1 = _solve(ctx, m, True, fp64, kind=H.PBICGSTAB, **ctl)
    assert used0 == 0
    if m.n <= MAX_ROWS_BICG[fp64]:
        assert used1 >= 1
    else:
        assert used1 == 0
    _same_perf(perf0, perf1)
    np.testing.assert_array_equal(psi1, psi0)
    np.testing.assert_array_equal(res1, res0)
    assert perf1.converged, (perf1.finalResidual, perf1.nIterations)


@pytest.mark.parametrize("max_iter", [300, 4])
def test_local_bicgstab_inner_entry(ctx, max_iter):
    """hipfoam_bicgstab (metalFoamBackend::bicgstab): stats and delta bit for
    bit, through the |sA| exit (a loose target) and a maxIter exit."""
    m = mg.convection_diffusion(11, 11, 10, seed=48)
    ctx.set_topology(m)
    ctx.set_coeffs(m.diag, m.upper, m.lower)
    r = m.source.copy()
    for target in (1e-6 * np.abs(r).sum(), 0.5 * np.abs(r).sum()):
        out = []
        for local in (0, -1):
            ctx.set_schedule(persistent=0, localLoop=local)
            before = ctx.local_launches()
            st = ctx.bicgstab(r, target, max_iter)
            out.append((st, ctx.delta(), ctx.local_launches() - before))
        (s0, d0, u0), (s1, d1, u1) = out
        assert (u0, u1) == (0, 1)
        for k in ("nIterations", "finalResSum", "converged", "singular", "stagnated"):
            assert getattr(s0, k) == getattr(s1, k), k
        np.testing.assert_array_equal(d1, d0)
