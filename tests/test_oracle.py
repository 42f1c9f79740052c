"""Pins the CPU oracle (oracle/ldu_oracle.c) against independent facts:
dense numpy algebra on tiny meshes, closed forms, textbook polynomials and
exact-termination properties.  Each check is chosen so that a plausible
mistake (a swapped operand, a dropped term, a wrong sign) fails it.
"""

import numpy as np
import pytest

from metal4foam_amd import meshgen as mg


def rand_asym(shape, seed):
    """Random asymmetric LDU matrix with strict row dominance."""
    rng = np.random.default_rng(seed)
    n, low, upp, own, los, loss = mg.structured_addressing(*shape)
    nf = low.shape[0]
    upper = -rng.uniform(0.1, 2.0, nf)
    lower = -rng.uniform(0.1, 2.0, nf)
    diag = -mg.row_offdiag_sums(n, low, upp, upper, lower) + rng.uniform(0.5, 1.5, n)
    b = rng.standard_normal(n)
    return mg.LduMatrix(n, low, upp, own, los, loss, diag, upper, lower, b, shape, "rand")


# ---------------------------------------------------------------------------
# operator primitives
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("shape", [(3, 3, 1), (4, 3, 2)])
def test_amul_tmul_suma_match_dense(orc, shape):
    m = rand_asym(shape, 1)
    a = mg.dense(m)
    x = np.random.default_rng(2).standard_normal(m.n)
    np.testing.assert_allclose(orc.amul(m, x), a @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(orc.amul(m, x, transpose=True), a.T @ x, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(orc.sumA(m), a.sum(axis=1), rtol=1e-13, atol=1e-13)
    # the dense check is sensitive to a transposed operand
    assert not np.allclose(orc.amul(m, x), a.T @ x)


def test_spmv_f_integer_inputs_are_exact(orc):
    # small integers: every product and partial sum is exact in FP32
    m = rand_asym((5, 4, 3), 3)
    rng = np.random.default_rng(4)
    m.diag = rng.integers(-8, 9, m.n).astype(float)
    m.upper = rng.integers(-8, 9, m.nfaces).astype(float)
    m.lower = rng.integers(-8, 9, m.nfaces).astype(float)
    x = rng.integers(-16, 17, m.n).astype(float)
    y = orc.spmv_f(m, x)
    np.testing.assert_array_equal(y.astype(np.float64), mg.dense(m) @ x)


def test_spmv_f_rounding_bound(orc):
    m = rand_asym((6, 5, 4), 5)
    x = np.random.default_rng(6).standard_normal(m.n)
    x32 = x.astype(np.float32).astype(np.float64)
    m32 = mg.LduMatrix(m.n, m.lower_addr, m.upper_addr, m.owner_start, m.losort,
                       m.losort_start, m.diag.astype(np.float32).astype(float),
                       m.upper.astype(np.float32).astype(float),
                       m.lower.astype(np.float32).astype(float))
    a = mg.dense(m32)
    y = orc.spmv_f(m, x).astype(np.float64)
    exact = a @ x32
    k = 8  # row length + 1 (7-point stencil)
    u = 2.0 ** -24
    gamma = k * u / (1 - k * u)
    bound = gamma * (np.abs(a) @ np.abs(x32))
    assert np.all(np.abs(y - exact) <= bound)


def test_norm_factor_closed_forms(orc):
    m = rand_asym((4, 3, 2), 7)
    b = m.source
    # psi = 0: sum|b| + small
    assert orc.norm_factor(m, np.zeros(m.n), b) == pytest.approx(np.abs(b).sum() + 1e-20,
                                                                  rel=1e-14)
    # psi = c 1: A psi = c u, so the first term vanishes and the second is |b - c u|
    c = 0.37
    u = mg.dense(m).sum(axis=1)
    assert orc.norm_factor(m, np.full(m.n, c), b) == pytest.approx(
        np.abs(b - c * u).sum() + 1e-20, rel=1e-13)


# ---------------------------------------------------------------------------
# Chebyshev-Jacobi preconditioner
# ---------------------------------------------------------------------------

def test_cheb_constants_first_step(orc):
    beta, ratio = 2.0, 30.0
    c1, c2, c3 = orc.cheb_constants(8, beta, ratio)
    alpha = beta / ratio
    theta = 0.5 * (beta + alpha)
    assert c1 == pytest.approx(1.0 / theta, rel=1e-7)
    assert len(c2) == 7 and len(c3) == 7


@pytest.mark.parametrize("order", [1, 2, 5, 8])
def test_cheb_apply_is_textbook_chebyshev_polynomial(orc, order):
    """z = p(M) D^-1 r with 1 - lambda p(lambda) = T_k((theta-lambda)/delta)/T_k(theta/delta)."""
    m = mg.poisson(4, 3, 3, seed=11, lognormal=True)
    beta = orc.gershgorin(m)
    ratio = 30.0
    a = mg.dense(m)
    d = np.diag(a)
    dh = 1.0 / np.sqrt(d)
    s = dh[:, None] * a * dh[None, :]
    lam, v = np.linalg.eigh(s)
    alpha = beta / ratio
    theta, delta = 0.5 * (beta + alpha), 0.5 * (beta - alpha)

    def T(k, x):
        return np.polynomial.chebyshev.chebval(x, [0] * k + [1])

    R = T(order, (theta - lam) / delta) / T(order, theta / delta)
    p = (1.0 - R) / lam
    r = np.random.default_rng(3).standard_normal(m.n)
    z_expect = dh * (v @ (p * (v.T @ (dh * r))))
    z = orc.cheb_apply_d(m, r, order, beta, ratio)
    np.testing.assert_allclose(z, z_expect, rtol=1e-10, atol=1e-12)
    zf = orc.cheb_apply_f(m, r, order, beta, ratio).astype(np.float64)
    np.testing.assert_allclose(zf, z_expect, rtol=2e-5, atol=2e-5 * np.abs(z_expect).max())


def test_gershgorin_closed_form(orc):
    # unit Poisson, deflated diagonal: every row has sum|offdiag| = |diag0| -> beta = 2
    m = mg.poisson(5, 4, 1, seed=0, lognormal=False)
    u = orc.sumA(m)
    assert orc.gershgorin(m, m.diag - u) == 2.0
    # asymmetric: uses |upper| only (metalFoamSolver.C:231-240)
    a = rand_asym((3, 3, 2), 9)
    so = np.bincount(a.lower_addr, np.abs(a.upper), a.n) + np.bincount(
        a.upper_addr, np.abs(a.upper), a.n)
    assert orc.gershgorin(a) == pytest.approx((1 + so / np.abs(a.diag)).max(), rel=1e-14)


# ---------------------------------------------------------------------------
# OpenFOAM built-in solvers (CPU baseline)
# ---------------------------------------------------------------------------

def chain(n, seed, asym=False):
    """1D chain: DIC / DILU are exact factorisations (no fill-in)."""
    rng = np.random.default_rng(seed)
    nn, low, upp, own, los, loss = mg.structured_addressing(n, 1, 1)
    upper = -rng.uniform(0.5, 1.5, nn - 1)
    lower = -rng.uniform(0.5, 1.5, nn - 1) if asym else None
    diag = -mg.row_offdiag_sums(nn, low, upp, upper, upper if lower is None else lower) + 0.3
    b = rng.standard_normal(nn)
    return mg.LduMatrix(nn, low, upp, own, los, loss, diag, upper, lower, b, (n, 1, 1))


def test_pcg_dic_exact_on_tree(orc):
    m = chain(40, 1)
    psi, perf = orc.pcg(m, precon=orc.PRE_DIC, tol=1e-12)
    assert perf.nIterations == 1 and perf.converged
    np.testing.assert_allclose(psi, np.linalg.solve(mg.dense(m), m.source), rtol=1e-10)


def test_pbicgstab_dilu_exact_on_tree(orc):
    m = chain(40, 2, asym=True)
    psi, perf = orc.pbicgstab(m, precon=orc.PRE_DIC, tol=1e-12)
    assert perf.nIterations == 1 and perf.converged
    np.testing.assert_allclose(psi, np.linalg.solve(mg.dense(m), m.source), rtol=1e-10)


@pytest.mark.parametrize("precon", [0, 1, 2])
def test_pcg_solution_matches_dense(orc, precon):
    m = mg.poisson(5, 4, 3, seed=4, lognormal=True)
    beta = orc.gershgorin(m)
    psi, perf = orc.pcg(m, precon=precon, tol=1e-12, beta=beta)
    assert perf.converged and perf.initialResidual == pytest.approx(1.0)
    x = np.linalg.solve(mg.dense(m), m.source)
    np.testing.assert_allclose(psi, x, rtol=1e-8, atol=1e-8 * np.abs(x).max())
    # final residual is the true normalised L1 residual
    r = m.source - mg.dense(m) @ psi
    assert np.abs(r).sum() / perf.normFactor < 1e-12


@pytest.mark.parametrize("precon", [0, 1])
def test_pbicgstab_solution_matches_dense(orc, precon):
    m = mg.convection_diffusion(5, 4, 3, seed=5)
    psi, perf = orc.pbicgstab(m, precon=precon, tol=1e-12)
    assert perf.converged
    x = np.linalg.solve(mg.dense(m), m.source)
    np.testing.assert_allclose(psi, x, rtol=1e-8, atol=1e-8 * np.abs(x).max())


def test_pcg_max_iter_and_min_iter(orc):
    m = mg.poisson(6, 6, 1, seed=2, lognormal=False)
    _, perf = orc.pcg(m, precon=0, tol=1e-12, maxIter=3)
    assert perf.nIterations == 3 and not perf.converged
    _, perf = orc.pcg(m, precon=0, tol=1e30, minIter=4)
    assert perf.nIterations == 4


# ---------------------------------------------------------------------------
# the reference's refined solver
# ---------------------------------------------------------------------------

def test_inner_pcg_f_on_jacobi_exact_system(orc):
    # diagonal-only matrix: D^-1 A = I, Jacobi CG converges in one iteration
    n = 17
    m = mg.LduMatrix(n, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(n + 1, np.int32),
                     np.zeros(0, np.int32), np.zeros(n + 1, np.int32),
                     np.linspace(1, 3, n), np.zeros(0), None)
    r = np.random.default_rng(0).standard_normal(n)
    st, delta = orc.inner_pcg_f(m, r, target=1e-3, max_iter=10, cheb_order=1)
    assert st.nIterations == 1 and st.converged
    np.testing.assert_allclose(delta, r / m.diag, rtol=1e-6)


def test_inner_pcg_f_returns_best_iterate_and_flags(orc):
    m = mg.poisson(8, 8, 1, seed=3, lognormal=False)
    r = m.source
    # unreachable target: stagnation after 50 non-improving iterations
    st, _ = orc.inner_pcg_f(m, r, target=0.0, max_iter=100000, cheb_order=8, cheb_beta=2.0)
    assert st.stagnated and not st.converged
    st, _ = orc.inner_pcg_f(m, r, target=0.0, max_iter=3, cheb_order=8, cheb_beta=2.0)
    assert st.nIterations == 3 and not (st.converged or st.stagnated or st.singular)
    r_nan = r.copy()
    r_nan[3] = np.nan
    st, _ = orc.inner_pcg_f(m, r_nan, target=1e-6, max_iter=10, cheb_order=8, cheb_beta=2.0)
    assert st.singular and st.nIterations == 1


@pytest.mark.parametrize("bicg", [False, True])
def test_refined_solve_reaches_tolerance(orc, bicg):
    if bicg:
        m = mg.convection_diffusion(8, 7, 6, seed=8)
    else:
        m = mg.poisson(9, 8, 7, seed=8, lognormal=True)
    psi, perf = orc.refined_solve(m, bicg=bicg, ctl=orc.controls(maxRestarts=4))
    a = mg.dense(m)
    r = m.source - a @ psi
    assert perf.converged and perf.finalResidual < 1e-6
    assert np.abs(r).sum() / perf.normFactor == pytest.approx(perf.finalResidual, rel=1e-6)
    assert perf.deflated == (0 if bicg else 1)
    x = np.linalg.solve(a, m.source)
    assert np.abs(psi - x).max() / np.abs(x).max() < 1e-4


def test_refined_solve_deflation_rules(orc):
    m = mg.poisson(6, 6, 1, seed=1, lognormal=False)
    _, p_auto = orc.refined_solve(m, ctl=orc.controls(deflate=2))
    _, p_off = orc.refined_solve(m, ctl=orc.controls(deflate=0))
    assert p_auto.deflated == 1 and p_off.deflated == 0
    assert p_auto.chebBeta == 2.0  # closed form for the deflated unit Poisson
    # a pinned-everywhere matrix (u ~ diag) is not deflated in auto mode
    m2 = mg.poisson(6, 6, 1, seed=1, lognormal=False)
    m2.diag = m2.diag + 1.0
    _, p2 = orc.refined_solve(m2, ctl=orc.controls(deflate=2))
    assert p2.deflated == 0
    # a positive face coefficient makes diag - u flip sign in its two rows:
    # deflation bails out (metalFoamSolver.C:199-207) and the solve still converges
    m3 = mg.poisson(6, 6, 1, seed=1, lognormal=False)
    m3.upper = m3.upper.copy()
    m3.upper[7] = 5.0
    u3 = orc.sumA(m3)
    assert np.any((m3.diag - u3) * m3.diag <= 0)
    _, p3 = orc.refined_solve(m3, ctl=orc.controls(deflate=1))
    assert p3.deflated == 0
    _, p3a = orc.refined_solve(m3, ctl=orc.controls(deflate=2))
    assert p3a.deflated == 0


def test_refined_solve_already_converged_and_min_iter(orc):
    m = mg.poisson(5, 5, 1, seed=1, lognormal=False)
    x = np.linalg.solve(mg.dense(m), m.source)
    psi, perf = orc.refined_solve(m, psi=x)
    assert perf.nIterations == 0 and perf.converged and perf.restarts == 0
    psi, perf = orc.refined_solve(m, psi=x, ctl=orc.controls(minIter=2))
    assert perf.nIterations >= 1


# ---------------------------------------------------------------------------
# adjoint
# ---------------------------------------------------------------------------

def test_adjoint_grads_match_finite_differences(orc):
    m = rand_asym((4, 3, 2), 12)
    rng = np.random.default_rng(13)
    xt = rng.standard_normal(m.n)

    def loss(mm):
        x = np.linalg.solve(mg.dense(mm), mm.source)
        return 0.5 * np.sum((x - xt) ** 2)

    a = mg.dense(m)
    x = np.linalg.solve(a, m.source)
    lam = np.linalg.solve(a.T, x - xt)
    du, dl, dd = orc.adjoint_grads(m, x, lam)
    h = 1e-6
    for f in [0, 3, 7, m.nfaces - 1]:
        for arr, g in ((m.upper, du), (m.lower, dl)):
            old = arr[f]
            arr[f] = old + h
            lp = loss(m)
            arr[f] = old - h
            lm = loss(m)
            arr[f] = old
            assert (lp - lm) / (2 * h) == pytest.approx(g[f], rel=1e-5, abs=1e-9)
    for i in [0, 5, m.n - 1]:
        old = m.diag[i]
        m.diag[i] = old + h
        lp = loss(m)
        m.diag[i] = old - h
        lm = loss(m)
        m.diag[i] = old
        assert (lp - lm) / (2 * h) == pytest.approx(dd[i], rel=1e-5, abs=1e-9)


def test_refined_solve_pinned_projection(orc):
    """deflate 3 (no reference counterpart): the inner right-hand side is
    projected along u = A 1, r - (sum r / sum u) u, which sums to zero and so
    lies in the range of the deflated (singular, symmetric) operator whatever
    sum(b) is.  On a pinned Neumann Poisson with a RHS of non-zero sum it
    converges without polishing, where the constant projection (deflate 2)
    stalls until the polish; both reach the dense solution."""
    m = mg.poisson(12, 10, 8, seed=51, lognormal=True, zero_mean_rhs=False, rhs="normal")
    assert abs(m.source.sum()) > 1.0
    x = np.linalg.solve(mg.dense(m), m.source)
    ctl = dict(tolerance=1e-9, maxRestarts=8, maxIter=20000)
    psi3, p3 = orc.refined_solve(m, ctl=orc.controls(deflate=3, **ctl))
    psi2, p2 = orc.refined_solve(m, ctl=orc.controls(deflate=2, **ctl))
    assert p3.converged and p3.deflated and not p3.polished
    assert p2.converged and p2.deflated and p2.polished
    assert p3.restarts < p2.restarts
    for psi in (psi3, psi2):
        assert np.abs(psi - x).max() <= 1e-8 * np.abs(x).max()
    # on a zero-sum RHS the two projections coincide for the first inner solve
    mz = mg.poisson(12, 10, 8, seed=51, lognormal=True, zero_mean_rhs=True, rhs="normal")
    _, pz3 = orc.refined_solve(mz, ctl=orc.controls(deflate=3, **ctl))
    _, pz2 = orc.refined_solve(mz, ctl=orc.controls(deflate=2, **ctl))
    assert pz3.converged and pz2.converged and not pz3.polished
    assert abs(pz3.nIterations - pz2.nIterations) <= max(3, pz2.nIterations // 10)


def test_pcg_diag_par_matches_serial_loop(orc):
    """The all-cores CPU baseline (OpenFOAM PCG with the diagonal
    preconditioner, row-gather SpMV over threads) follows the serial loop:
    same normFactor, iteration count within one, same solution."""
    m = mg.poisson(11, 9, 8, seed=31, lognormal=True)
    x1, p1 = orc.pcg(m, precon=orc.PRE_DIAGONAL, tol=1e-10)
    x2, p2 = orc.pcg_diag_par(m, tol=1e-10)
    assert p2.converged and abs(p1.nIterations - p2.nIterations) <= 1
    assert p2.normFactor == pytest.approx(p1.normFactor, rel=1e-13)
    assert np.abs(x1 - x2).max() <= 1e-8 * np.abs(x1).max()
    r = m.source - mg.dense(m) @ x2
    assert np.abs(r).sum() / p2.normFactor == pytest.approx(p2.finalResidual, rel=1e-6)
