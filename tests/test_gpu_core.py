"""GPU parity of the LDU SpMV, the inner FP32 solvers, the refined solve and
the plain-FP64 comparator, through the C-ABI, against the CPU oracle."""

import numpy as np
import pytest

from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = pytest.mark.gpu

SCHEDULES = [dict(spmvSchedule=s, blockSize=b, xcdRemap=r, gridLimit=g, rowForm=f)
             for s in (0, 1, 2) for b in (64, 256, 1024) for r in (0, 1) for g in (0, 7)
             for f in (0, 1)]


def rand_asym(shape, seed):
    rng = np.random.default_rng(seed)
    n, low, upp, own, los, loss = mg.structured_addressing(*shape)
    nf = low.shape[0]
    upper = -rng.uniform(0.1, 2.0, nf)
    lower = -rng.uniform(0.1, 2.0, nf)
    diag = -mg.row_offdiag_sums(n, low, upp, upper, lower) + rng.uniform(0.5, 1.5, n)
    b = rng.standard_normal(n)
    return mg.LduMatrix(n, low, upp, own, los, loss, diag, upper, lower, b, shape, "rand")


@pytest.fixture(autouse=True)
def _default_schedule(ctx):
    ctx.set_schedule(spmvSchedule=2, blockSize=256, xcdRemap=1, graphChunk=4, gridLimit=0,
                     rowForm=0)
    yield


def gamma(k, u):
    return k * u / (1 - k * u)


# ---------------------------------------------------------------------------
# SpMV
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("sched", SCHEDULES, ids=lambda s: "s{spmvSchedule}b{blockSize}r{xcdRemap}g{gridLimit}f{rowForm}".format(**s))
def test_spmv_fp32_integer_inputs_bit_exact(ctx, sched):
    m = rand_asym((37, 23, 11), 1)
    rng = np.random.default_rng(2)
    m.diag = rng.integers(-8, 9, m.n).astype(float)
    m.upper = rng.integers(-8, 9, m.nfaces).astype(float)
    m.lower = rng.integers(-8, 9, m.nfaces).astype(float)
    x = rng.integers(-16, 17, m.n).astype(np.float32)
    ctx.set_topology(m)
    ctx.set_schedule(**sched)
    ctx.set_matrix(m.diag, m.upper, m.lower)
    y = ctx.amul(x, H.FP32)
    import oracle
    np.testing.assert_array_equal(y, oracle.spmv_f(m, x))
    # symmetric path (lower aliased to upper)
    ctx.set_matrix(m.diag, m.upper, None)
    m.lower = None
    np.testing.assert_array_equal(ctx.amul(x, H.FP32), oracle.spmv_f(m, x))


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_spmv_full_size_rounding_bounds(ctx, orc, cfg):
    m = mg.config1() if cfg == "c1" else mg.config2()
    x = np.random.default_rng(3).standard_normal(m.n)
    ctx.set_topology(m)
    ctx.set_matrix(m.diag, m.upper, m.lower)
    # FP64: row-gather order vs the oracle's face loop
    y64 = ctx.amul(x, H.FP64)
    ref = orc.amul(m, x)
    absrow = orc.amul(mg.LduMatrix(m.n, m.lower_addr, m.upper_addr, m.owner_start, m.losort,
                                   m.losort_start, np.abs(m.diag), np.abs(m.upper),
                                   None if m.lower is None else np.abs(m.lower)), np.abs(x))
    assert np.all(np.abs(y64 - ref) <= 2 * gamma(8, 2.0 ** -53) * absrow)
    # FP32: against exact arithmetic on the FP32-rounded operands
    x32 = x.astype(np.float32)
    m32 = mg.LduMatrix(m.n, m.lower_addr, m.upper_addr, m.owner_start, m.losort,
                       m.losort_start, m.diag.astype(np.float32).astype(float),
                       m.upper.astype(np.float32).astype(float),
                       None if m.lower is None else m.lower.astype(np.float32).astype(float))
    y32 = ctx.amul(x32, H.FP32).astype(np.float64)
    exact = orc.amul(m32, x32.astype(np.float64))
    absrow32 = orc.amul(mg.LduMatrix(m.n, m.lower_addr, m.upper_addr, m.owner_start, m.losort,
                                     m.losort_start, np.abs(m32.diag), np.abs(m32.upper),
                                     None if m32.lower is None else np.abs(m32.lower)),
                        np.abs(x32.astype(np.float64)))
    assert np.all(np.abs(y32 - exact) <= gamma(8, 2.0 ** -24) * absrow32)


# ---------------------------------------------------------------------------
# inner FP32 solvers (metalFoamBackend::pcg / ::bicgstab)
# ---------------------------------------------------------------------------

def test_inner_pcg_matches_oracle(ctx, orc):
    m = mg.config0()
    u = orc.sumA(m)
    d0 = m.diag - u  # deflated operator, as the refined solve builds it
    r = m.source - m.source.mean()
    beta = orc.gershgorin(m, d0)
    nf = orc.norm_factor(m, np.zeros(m.n), m.source)
    target = 1e-6 * nf
    ctx.set_topology(m)
    ctx.set_coeffs(d0, m.upper, None)
    for ips in (1, 4):
        st = ctx.pcg(r, target, 1000, ips=ips, order=8, beta=beta, ratio=30.0)
        delta = ctx.delta()
        so, do = orc.inner_pcg_f(m, r, target, 1000, ips=ips, cheb_order=8, cheb_beta=beta,
                                 diag=d0)
        print(f"ips={ips} gpu it={st.nIterations} res={st.finalResSum:.4g} "
              f"oracle it={so.nIterations} res={so.finalResSum:.4g}")
        assert st.converged == so.converged == 1
        assert abs(st.nIterations - so.nIterations) <= max(3, 3 * ips)
        if ips > 1:
            assert st.nIterations % ips == 0
        dd = delta - delta.mean()
        ddo = do - do.mean()
        assert np.abs(dd - ddo).max() <= 2e-3 * np.abs(ddo).max()


def test_inner_pcg_flags(ctx, orc):
    m = mg.poisson(16, 16, 1, seed=3, lognormal=False)
    ctx.set_topology(m)
    # deflated (exactly singular) operator with a non-zero-mean RHS: A p has zero
    # mean, so the mean of r never decays and the L1 residual must stagnate
    d0 = m.diag - orc.sumA(m)
    ctx.set_coeffs(d0, m.upper, None)
    r1 = m.source + 1.0
    st = ctx.pcg(r1, 0.0, 100000, order=8, beta=2.0)
    so, _ = orc.inner_pcg_f(m, r1, 0.0, 100000, cheb_order=8, cheb_beta=2.0, diag=d0)
    print("gpu", st.as_dict(), "oracle", so.nIterations, so.stagnated, so.singular)
    assert st.stagnated and so.stagnated and not st.converged
    ctx.set_coeffs(m.diag, m.upper, None)
    st = ctx.pcg(m.source, 0.0, 3, order=8, beta=2.0)
    assert st.nIterations == 3 and not (st.converged or st.stagnated or st.singular)
    r = m.source.copy()
    r[5] = np.nan
    st = ctx.pcg(r, 1.0, 50, order=8, beta=2.0)
    assert st.singular and st.nIterations == 1
    st = ctx.pcg(m.source, 1.0, 0, order=8, beta=2.0)
    assert st.nIterations == 0 and st.finalResSum == 0.0
    # Jacobi (order 1)
    st = ctx.pcg(m.source, 1e-4 * np.abs(m.source).sum(), 1000, order=1)
    so, _ = orc.inner_pcg_f(m, m.source, 1e-4 * np.abs(m.source).sum(), 1000, cheb_order=1)
    assert st.converged and abs(st.nIterations - so.nIterations) <= 3


def test_inner_bicgstab_matches_oracle(ctx, orc):
    m = mg.convection_diffusion(24, 24, 24, seed=5)
    nf = orc.norm_factor(m, np.zeros(m.n), m.source)
    ctx.set_topology(m)
    ctx.set_coeffs(m.diag, m.upper, m.lower)
    st = ctx.bicgstab(m.source, 1e-6 * nf, 1000)
    delta = ctx.delta()
    so, do = orc.inner_bicgstab_f(m, m.source, 1e-6 * nf, 1000)
    print(f"gpu it={st.nIterations} res={st.finalResSum:.4g} "
          f"oracle it={so.nIterations} res={so.finalResSum:.4g}")
    assert st.converged == so.converged == 1
    assert abs(st.nIterations - so.nIterations) <= 2
    assert np.abs(delta - do).max() <= 1e-3 * np.abs(do).max()


# ---------------------------------------------------------------------------
# refined solve (metalFoamSolver::solveImpl with FP64 on the GPU)
# ---------------------------------------------------------------------------

def _true_residual(orc, m, psi, nf):
    return np.abs(m.source - orc.amul(m, psi)).sum() / nf


@pytest.mark.parametrize("case", ["c0", "poisson3d_32", "convdiff_24"])
def test_refined_solve_matches_oracle(ctx, orc, case):
    if case == "c0":
        m, bicg = mg.config0(), False
    elif case == "poisson3d_32":
        m, bicg = mg.poisson(32, 32, 32, seed=21, lognormal=True), False
    else:
        m, bicg = mg.convection_diffusion(24, 24, 24, seed=22), True
    ctl = H.default_controls(maxRestarts=4)
    psi = np.zeros(m.n)
    ctx.set_topology(m)
    perf, res = ctx.solve(H.PBICGSTAB if bicg else H.PCG, m.diag, m.upper,
                          m.lower if bicg else None, psi, m.source, ctl, residual=True)
    po_psi, po = orc.refined_solve(m, bicg=bicg, ctl=orc.controls(maxRestarts=4))
    print(f"{case}: gpu it={perf.nIterations} restarts={perf.restarts} "
          f"res={perf.finalResidual:.3e} defl={perf.deflated} polish={perf.polished} | "
          f"oracle it={po.nIterations} restarts={po.restarts} res={po.finalResidual:.3e}")
    assert perf.converged and po.converged
    assert perf.initialResidual == pytest.approx(po.initialResidual, rel=1e-12)
    assert perf.normFactor == pytest.approx(po.normFactor, rel=1e-12)
    assert perf.deflated == po.deflated and perf.polished == po.polished
    assert perf.chebBeta == pytest.approx(po.chebBeta, rel=1e-12)
    assert abs(perf.nIterations - po.nIterations) <= max(4, 0.1 * po.nIterations)
    tr = _true_residual(orc, m, psi, perf.normFactor)
    assert tr == pytest.approx(perf.finalResidual, rel=1e-6) and tr < 1e-6
    np.testing.assert_allclose(res, m.source - orc.amul(m, psi), rtol=0,
                               atol=1e-9 * np.abs(m.source).max())
    # error against a tight FP64 solve
    if bicg:
        x, _ = orc.pbicgstab(m, precon=orc.PRE_DIC, tol=1e-13)
    else:
        x, _ = orc.pcg(m, precon=orc.PRE_DIC, tol=1e-13)
    err = np.abs(psi - x).max() / np.abs(x).max()
    err_o = np.abs(po_psi - x).max() / np.abs(x).max()
    print(f"{case}: max rel err vs FP64 oracle: gpu {err:.3e} oracle-refined {err_o:.3e}")
    assert err < max(10 * err_o, 1e-5)


@pytest.mark.parametrize("kind", ["pcg", "bicg", "pcg_fp64", "bicg_fp64"])
def test_launch_shapes(ctx, kind):
    """768 256-row tiles: the automatic launch shape runs 384 workgroups of 512
    threads with every second stage merged into a row kernel, and an explicit
    512-thread schedule with the same grid gives the same bits.  Other grids
    group the reductions differently: 256-thread workgroups without a limit
    (768, separate finalise kernels), 1024-thread ones (192, merged), a
    512-workgroup limit (merged).  Their FP64 sums may round differently in
    the last bits, so they agree to 1e-9 with the same iteration count."""
    bicg = kind.startswith("bicg")
    fp64 = kind.endswith("fp64")
    m = rand_asym((64, 64, 48), 11) if bicg else mg.poisson(64, 64, 48, seed=11, lognormal=True)
    ctx.set_topology(m)
    lower = m.lower if bicg else None
    solver = H.PBICGSTAB if bicg else H.PCG
    outs = {}
    for bs, gl in ((0, 0), (512, 0), (256, 0), (1024, 0), (256, 512)):
        ctx.set_schedule(blockSize=bs, gridLimit=gl)
        psi = np.zeros(m.n)
        perf = ctx.solve(solver, m.diag, m.upper, lower, psi, m.source, fp64=fp64)
        assert perf.converged
        outs[(bs, gl)] = (psi, perf.nIterations, perf.finalResidual)
    ref, it0, res0 = outs[(0, 0)]
    psi, it, res = outs[(512, 0)]
    np.testing.assert_array_equal(psi, ref)
    assert it == it0 and res == res0
    for key in ((256, 0), (1024, 0), (256, 512)):
        psi, it, res = outs[key]
        np.testing.assert_allclose(psi, ref, rtol=0, atol=1e-9 * np.abs(ref).max())
        assert it == it0
        assert res == pytest.approx(res0, rel=1e-6)
    ctx.set_schedule(blockSize=256, gridLimit=0)


def test_solve_device_entry_and_graph_modes_agree_bitwise(ctx):
    m = mg.poisson(24, 20, 16, seed=7, lognormal=True)
    ctx.set_topology(m)
    outs = []
    # graph chunks, eager launches, the looped row form and every block-to-XCD
    # mapping all give the same bits
    for chunk, form, remap in ((4, 0, 1), (0, 0, 1), (2, 0, 1), (8, 0, 1), (-1, 0, 1),
                               (4, 1, 1), (4, 0, 0), (4, 0, 2)):
        ctx.set_schedule(graphChunk=chunk, rowForm=form, xcdRemap=remap)
        psi = np.zeros(m.n)
        perf = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source)
        outs.append((psi, perf.nIterations))
    for psi, it in outs[1:]:
        np.testing.assert_array_equal(psi, outs[0][0])
        assert it == outs[0][1]
    ctx.set_schedule(graphChunk=4, xcdRemap=1)
    D = H.DeviceArray
    d, u, p, b = (D.from_host(ctx, m.diag), D.from_host(ctx, m.upper),
                  D.from_host(ctx, np.zeros(m.n)), D.from_host(ctx, m.source))
    perf = ctx.solve_device(H.PCG, d, u, None, p, b)
    np.testing.assert_array_equal(p.download(), outs[0][0])


def test_refined_solve_failure_modes(ctx, orc):
    m = mg.poisson(12, 12, 4, seed=9, lognormal=False)
    ctx.set_topology(m)
    # NaN in the source: breakdown reported, no exception
    b = m.source.copy()
    b[3] = np.nan
    perf = ctx.solve(H.PCG, m.diag, m.upper, None, np.zeros(m.n), b)
    assert not perf.converged
    # maxIter = 1 caps the total inner iterations
    perf = ctx.solve(H.PCG, m.diag, m.upper, None, np.zeros(m.n), m.source,
                     H.default_controls(maxIter=1))
    assert perf.nIterations == 1 and not perf.converged
    # tolerance 0: never converges, stops at maxRestarts
    perf = ctx.solve(H.PCG, m.diag, m.upper, None, np.zeros(m.n), m.source,
                     H.default_controls(tolerance=0.0, maxRestarts=2, maxIter=100000))
    assert not perf.converged and perf.restarts == 3
    # already converged: no iterations
    x, _ = orc.pcg(m, precon=orc.PRE_DIC, tol=1e-14)
    psi = x.copy()
    perf = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source)
    assert perf.nIterations == 0 and perf.converged
    np.testing.assert_array_equal(psi, x)


# ---------------------------------------------------------------------------
# plain FP64 comparator (OpenFOAM PCG / PBiCGStab loops on the GPU)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("bicg", [False, True])
def test_fp64_solve_matches_openfoam_loop(ctx, orc, bicg):
    if bicg:
        m = mg.convection_diffusion(20, 18, 16, seed=31)
    else:
        m = mg.poisson(20, 18, 16, seed=31, lognormal=True)
    ctx.set_topology(m)
    psi = np.zeros(m.n)
    perf, res = ctx.solve(H.PBICGSTAB if bicg else H.PCG, m.diag, m.upper,
                          m.lower if bicg else None, psi, m.source, H.default_controls(),
                          residual=True, fp64=True)
    if bicg:
        xo, po = orc.pbicgstab(m, precon=orc.PRE_DIAGONAL, tol=1e-6)
    else:
        beta = orc.gershgorin(m)
        xo, po = orc.pcg(m, precon=orc.PRE_CHEBYSHEV, order=8, beta=beta, ratio=30.0, tol=1e-6)
    print(f"fp64 bicg={bicg}: gpu it={perf.nIterations} res={perf.finalResidual:.4e} "
          f"oracle it={po.nIterations} res={po.finalResidual:.4e}")
    assert perf.converged and po.converged
    tr = np.abs(m.source - orc.amul(m, psi)).sum() / perf.normFactor
    assert tr == pytest.approx(perf.finalResidual, rel=1e-6)
    if not bicg:
        # FP64 CG: same trajectory up to summation order
        assert abs(perf.nIterations - po.nIterations) <= 1
        np.testing.assert_allclose(psi, xo, rtol=0, atol=1e-9 * np.abs(xo).max())
    else:
        # FP64 BiCGStab on this matrix is order-sensitive near the threshold: a
        # permuted summation alone moves the oracle between 44 and 45 iterations
        assert abs(perf.nIterations - po.nIterations) <= max(2, po.nIterations // 20)
        xt, _ = orc.pbicgstab(m, precon=orc.PRE_DIC, tol=1e-13)
        e_gpu = np.abs(psi - xt).max() / np.abs(xt).max()
        e_orc = np.abs(xo - xt).max() / np.abs(xt).max()
        print(f"max rel err vs tight FP64: gpu {e_gpu:.3e} oracle {e_orc:.3e}")
        assert e_gpu < 10 * e_orc + 1e-9


def test_kernel_timing_counts_launches(ctx):
    m = mg.poisson(32, 32, 32, seed=4, lognormal=True)
    ctx.set_topology(m)
    ctx.set_kernel_timing(True, 0)
    psi = np.zeros(m.n)
    perf = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source)
    n, s = ctx.kernel_timing()
    ctx.set_kernel_timing(False, 0)
    assert perf.converged
    # order 8: seven fused Chebyshev SpMV steps per iteration, of which five
    # are interior.  An iteration's update is applied by the next iteration's
    # first step and checked after its last, so every inner solve (restarts
    # is the index of the last one) also launches the steps of the iteration
    # that finds it done; on grids small enough for merged stages the state
    # reaches the steps one iteration later still (they return at entry).
    inner = perf.restarts + 1
    assert 5 * (perf.nIterations + inner) <= n <= 5 * (perf.nIterations + 2 * inner)
    assert 0 < s < perf.tTotal


def test_autotune_picks_a_valid_schedule(ctx):
    m = mg.poisson(48, 48, 48, seed=4, lognormal=True)
    ctx.set_topology(m)
    ctx.set_matrix(m.diag, m.upper, None)
    s = ctx.autotune()
    assert s.spmvSchedule in (0, 1, 2) and s.blockSize in (0, 128, 256, 512, 1024)
    psi = np.zeros(m.n)
    perf = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source)
    assert perf.converged
