"""The row-slot SpMV schedule (spmvSchedule 2): the same terms in the same
order as the face-loop schedules 0 and 1, so SpMV results, inner solves and
refined solves must agree bit for bit; meshes that do not allow the layout
(a row with more than three owner or neighbour faces, or one slot's column
offsets spanning more than an int16 delta within a 256-row tile) run those
tiles on the face loops, and fall back to schedule 0 when no tile fits."""

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


def scattered(n, seed):
    """A chain plus one long-range face per cell pair of a random matching:
    at most three faces per row, but offsets scattered over the whole mesh."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(n)
    pairs = {(i, i + 1) for i in range(n - 1)}
    for k in range(0, n - 1, 2):
        a, b = sorted((int(perm[k]), int(perm[k + 1])))
        if b - a > 1:
            pairs.add((a, b))
    pairs = sorted(pairs)
    low = np.array([p[0] for p in pairs], dtype=np.int32)
    upp = np.array([p[1] for p in pairs], dtype=np.int32)
    own, los, loss = mg.derived_addressing(n, low, upp)
    upper = -rng.uniform(0.5, 1.5, len(pairs))
    lower = -rng.uniform(0.5, 1.5, len(pairs))
    diag = -mg.row_offdiag_sums(n, low, upp, upper, lower) + 0.5
    return mg.LduMatrix(n, low, upp, own, los, loss, diag, upper, lower,
                        rng.standard_normal(n), (n,), "scattered")


def blocks(seed, gen=mg.convection_diffusion):
    """Three structured blocks on one block-diagonal mesh: tiles that straddle
    two blocks see two offset patterns and keep per-row offsets (explicit
    tiles); the others are patterned."""
    parts = [gen(20, 20, 20, seed=seed), gen(23, 23, 23, seed=seed + 1),
             gen(17, 19, 21, seed=seed + 2)]
    n0, low, upp, d, up, lo, src = 0, [], [], [], [], [], []
    for p in parts:
        low.append(p.lower_addr + n0)
        upp.append(p.upper_addr + n0)
        d.append(p.diag)
        up.append(p.upper)
        lo.append(p.lower_or_upper())
        src.append(p.source)
        n0 += p.n
    low = np.concatenate(low).astype(np.int32)
    upp = np.concatenate(upp).astype(np.int32)
    own, los, loss = mg.derived_addressing(n0, low, upp)
    return mg.LduMatrix(n0, low, upp, own, los, loss, np.concatenate(d), np.concatenate(up),
                        np.concatenate(lo), np.concatenate(src), (n0,), "blocks")


def face_order_shuffled(seed):
    """A structured block whose faces of one owner are not in neighbour
    order: every tile keeps per-row offsets."""
    m = mg.convection_diffusion(15, 13, 11, seed=seed)
    rng = np.random.default_rng(seed)
    order = np.lexsort((rng.random(m.nfaces), m.lower_addr))
    low, upp = m.lower_addr[order].astype(np.int32), m.upper_addr[order].astype(np.int32)
    own, los, loss = mg.derived_addressing(m.n, low, upp)
    return mg.LduMatrix(m.n, low, upp, own, los, loss, m.diag, m.upper[order],
                        m.lower[order], m.source, m.shape, "face_order")


def spmv_all_schedules(ctx, m, x, lower):
    out = []
    for s in (0, 1, 2):
        ctx.set_schedule(spmvSchedule=s, blockSize=256, xcdRemap=1, graphChunk=4,
                         gridLimit=0, rowForm=0)
        ctx.set_matrix(m.diag, m.upper, lower)
        out.append(ctx.amul(x, H.FP32))
    return out


@pytest.mark.parametrize("case", ["poisson", "convdiff", "wide", "irregular", "scattered",
                                  "blocks", "face_order"])
def test_spmv_schedules_bitwise_equal(ctx, case):
    if case == "poisson":
        m = mg.poisson(33, 21, 17, seed=5)
    elif case == "convdiff":
        m = mg.convection_diffusion(29, 18, 13, seed=6)
    elif case == "wide":   # z offset 200*170 = 34000: an int16 delta from the tile base
        m = mg.convection_diffusion(200, 170, 3, seed=7)
    elif case == "irregular":
        m = irregular(3000, 2500, 8)
    elif case == "blocks":
        m = blocks(30)
    elif case == "face_order":
        m = face_order_shuffled(31)
    else:
        m = scattered(200000, 9)
    print(f"{case}: n={m.n} faces={m.nfaces}", flush=True)
    ctx.set_topology(m)
    patterned, planes = ctx.slot_layout()
    if case in ("poisson", "convdiff", "wide"):
        assert (patterned, planes) == (m.n, 3)
    elif case == "blocks":
        assert 0 < m.n - patterned < m.n // 10
    elif case == "face_order":
        assert patterned == 0 and planes == 3
    elif case == "scattered":
        # offsets beyond int16 in most tiles: those run the face loops
        assert patterned == 0 and planes == 3
        _, explicit, faces = ctx.slot_tiles()
        assert explicit > 0 and faces > m.n // 2
    else:
        assert (patterned, planes) == (0, 0)
    x = np.random.default_rng(9).standard_normal(m.n).astype(np.float32)
    for lower in (None, m.lower):
        ys = spmv_all_schedules(ctx, m, x, lower)
        assert ctx.active_schedule() == (0 if case == "irregular" else 2)
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


@pytest.mark.parametrize("solver,mesh", [(H.PCG, "box"), (H.PBICGSTAB, "box"),
                                         (H.PCG, "blocks"), (H.PBICGSTAB, "blocks")])
def test_refined_solve_identical_across_schedules(ctx, solver, mesh):
    if mesh == "blocks":
        m = blocks(13, mg.poisson) if solver == H.PCG else blocks(14)
    else:
        m = mg.poisson(28, 24, 20, seed=11) if solver == H.PCG else \
            mg.convection_diffusion(26, 22, 18, seed=12)
    lower = None if solver == H.PCG else m.lower
    ctx.set_topology(m)
    results = []
    for s in (0, 2):
        ctx.set_schedule(spmvSchedule=s, blockSize=256, xcdRemap=1, graphChunk=4,
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
        ctx.set_schedule(spmvSchedule=s, blockSize=256, xcdRemap=1, graphChunk=4,
                         gridLimit=0, rowForm=0)
        b = H.Batch(ctx, H.PCG, systems, 2)
        for k, m in enumerate(systems):
            b.set_rhs(k, m.source.T.copy())
        perf = b.solve(H.default_controls(maxRestarts=6))
        outs.append([b.psi(k) for k in range(len(systems))])
        b.close()
    for a, c in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(c, a)


@pytest.mark.parametrize("sched", [0, 1, 2])
@pytest.mark.parametrize("sym", [True, False])
def test_spmv_fp32_bit_exact_against_oracle(ctx, orc, sched, sym):
    """Without FMA contraction every schedule rounds each row term exactly like
    the oracle's FP32 row gather, so arbitrary float inputs match bit for bit."""
    m = mg.convection_diffusion(23, 19, 11, seed=40)
    if sym:
        m.lower = None
    rng = np.random.default_rng(41)
    x = rng.standard_normal(m.n).astype(np.float32)
    ctx.set_topology(m)
    ctx.set_schedule(spmvSchedule=sched)
    ctx.set_matrix(m.diag, m.upper, m.lower)
    np.testing.assert_array_equal(ctx.amul(x, H.FP32), orc.spmv_f(m, x))


@pytest.mark.parametrize("bicg", [False, True])
def test_inner_solve_against_oracle_bits(ctx, orc, bicg):
    """The inner FP32 iterations against the oracle's.  The SpMV and the
    Chebyshev and BiCGStab vector updates round exactly as the oracle's do;
    the scalars are double sums rounded once to float, whose summation order
    differs from the oracle's, which could move a scalar by one float ulp.  On
    these inputs no scalar moves, and the corrections agree bit for bit."""
    m = mg.convection_diffusion(20, 16, 12, seed=42) if bicg else mg.poisson(20, 16, 12, seed=43)
    ctx.set_topology(m)
    ctx.set_schedule(spmvSchedule=2)
    lower = m.lower if bicg else None
    ctx.set_coeffs(m.diag, m.upper, lower)
    r = m.source
    target = 1e-4 * np.abs(r).sum()
    if bicg:
        st = ctx.bicgstab(r, target, 200)
        so, do = orc.inner_bicgstab_f(m, r, target, 200)
    else:
        st = ctx.pcg(r, target, 200, order=8, beta=2.5)
        so, do = orc.inner_pcg_f(m, r, target, 200, cheb_order=8, cheb_beta=2.5)
    d = ctx.delta()
    same = np.mean(d == do)
    ulps = np.abs(d.view(np.int32).astype(np.int64) - do.view(np.int32).astype(np.int64)).max()
    print(f"{'bicg' if bicg else 'pcg'}: it gpu {st.nIterations} oracle {so.nIterations}, "
          f"identical {100 * same:.1f} %, max ulp distance {ulps}")
    assert st.nIterations == so.nIterations
    np.testing.assert_array_equal(d, do)
