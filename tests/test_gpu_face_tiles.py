"""Face tiles of the row-slot layout: a 256-row tile holding a row of more
than three owner or neighbour faces (or column offsets no base reaches with
int16 deltas) runs its rows on the LDU face loops, while the rest of the mesh
stays on slots; a slot tile whose neighbour faces are owned by face-tile rows
stores those coefficients in its own planes (symmetric operators included).
The face loops and the slots add the same terms in the same order, so every
result must equal the schedule-0 (face loops everywhere) result bit for bit:
SpMV in FP32 and FP64 (and the oracle's FP32 row sums), refined and plain-FP64
solves, and a batch.  Renumbered meshes: a random cell order leaves every tile
a face tile (schedule 0 runs), reverse Cuthill-McKee brings the slots back."""

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


def long_rows(nx, ny, nz, extra, seed, sym):
    """A structured block plus, for `extra` random cells, three faces to cells
    a few thousand rows away: those rows have more than three faces a side."""
    n, low, upp, *_ = mg.structured_addressing(nx, ny, nz)
    rng = np.random.default_rng(seed)
    pairs = set(zip(low.tolist(), upp.tolist()))
    for _ in range(extra):
        a = int(rng.integers(0, n))
        for d in (1, 2, 3):
            b = (a + d * 4099) % n
            if a != b:
                pairs.add((min(a, b), max(a, b)))
    pairs = sorted(pairs)
    low = np.array([p[0] for p in pairs], dtype=np.int32)
    upp = np.array([p[1] for p in pairs], dtype=np.int32)
    own, los, loss = mg.derived_addressing(n, low, upp)
    nf = low.shape[0]
    if sym:
        upper = -np.exp(rng.standard_normal(nf))
        lower = None
        diag = -mg.row_offdiag_sums(n, low, upp, upper, upper)
        diag[0] += 1.0
    else:
        upper = -rng.uniform(0.5, 1.5, nf)
        lower = -rng.uniform(0.5, 1.5, nf)
        diag = -mg.row_offdiag_sums(n, low, upp, upper, lower) + 0.3
    b = rng.standard_normal(n)
    if sym:
        b -= b.mean()
    return mg.LduMatrix(n, low, upp, own, los, loss, diag, upper, lower, b, (nx, ny, nz),
                        f"long_rows_{'sym' if sym else 'asym'}")


def _set(ctx, schedule):
    ctx.set_schedule(spmvSchedule=schedule, blockSize=0, xcdRemap=2, graphChunk=-1,
                     gridLimit=0, rowForm=0)


@pytest.mark.parametrize("sym", [True, False], ids=["sym", "asym"])
def test_face_tiles_spmv_bitwise(ctx, orc, sym):
    m = long_rows(44, 42, 40, 4, 21 if sym else 22, sym)
    ctx.set_topology(m)
    pat, expl, faces = ctx.slot_tiles()
    print(f"[face tiles] {m.name}: patterned {pat}, explicit {expl}, face-loop rows {faces}",
          flush=True)
    assert 0 < faces < m.n // 10 and pat > m.n // 2
    x32 = np.random.default_rng(3).standard_normal(m.n).astype(np.float32)
    x64 = np.random.default_rng(4).standard_normal(m.n)
    out = {}
    for s in (0, 2):
        _set(ctx, s)
        ctx.set_matrix(m.diag, m.upper, m.lower)
        assert ctx.active_schedule() == s
        out[s] = (ctx.amul(x32, H.FP32), ctx.amul(x64, H.FP64))
    np.testing.assert_array_equal(out[2][0], out[0][0])
    np.testing.assert_array_equal(out[2][1], out[0][1])
    np.testing.assert_array_equal(out[2][0], orc.spmv_f(m, x32))
    np.testing.assert_allclose(out[2][1], orc.amul(m, x64), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("sym", [True, False], ids=["sym", "asym"])
def test_face_tiles_solves_bitwise(ctx, orc, sym):
    m = long_rows(40, 38, 36, 3, 23 if sym else 24, sym)
    solver = H.PCG if sym else H.PBICGSTAB
    ctx.set_topology(m)
    assert ctx.slot_tiles()[2] > 0
    res = {}
    for s in (0, 2):
        _set(ctx, s)
        psi = np.zeros(m.n)
        p = ctx.solve(solver, m.diag, m.upper, m.lower, psi, m.source,
                      H.default_controls(maxRestarts=8))
        psi64 = np.zeros(m.n)
        p64 = ctx.solve(solver, m.diag, m.upper, m.lower, psi64, m.source,
                        H.default_controls(tolerance=1e-9, maxIter=5000), fp64=True)
        res[s] = (psi, p.nIterations, p.finalResidual, psi64, p64.nIterations)
        assert p.converged and p64.converged
    for a, b in zip(res[2], res[0]):
        np.testing.assert_array_equal(a, b)
    tr = np.abs(m.source - orc.amul(m, res[2][0])).sum() / np.abs(m.source).sum()
    assert tr < 2e-6


def test_face_tiles_batch_bitwise(ctx):
    nrhs = 2
    systems = [long_rows(24, 22, 20, 2, 30 + k, True) for k in range(3)]
    for m in systems:
        m.source = np.random.default_rng(m.n).standard_normal((m.n, nrhs))
        m.source -= m.source.mean(axis=0)
    out = []
    for s in (0, 2):
        _set(ctx, s)
        bt = H.Batch(ctx, H.PCG, systems, nrhs)
        for k, m in enumerate(systems):
            bt.set_rhs(k, m.source, layout="cell_major")
        perfs = bt.solve(H.default_controls(maxRestarts=8))
        assert all(p.converged for p in perfs)
        out.append(([bt.psi(k) for k in range(len(systems))], [p.nIterations for p in perfs]))
        bt.close()
    assert out[0][1] == out[1][1]
    for a, b in zip(out[0][0], out[1][0]):
        np.testing.assert_array_equal(a, b)


def test_renumbered_meshes(ctx, orc):
    base = mg.poisson(36, 34, 32, seed=41, lognormal=True)
    rnd = mg.random_renumbering(base, 42)
    rcm = mg.rcm_renumbering(rnd)
    sols = {}
    for m, want in ((rnd, 0), (rcm, 2)):
        ctx.set_topology(m)
        _set(ctx, 2)
        assert ctx.active_schedule() == want, m.name
        psi = np.zeros(m.n)
        p = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source,
                      H.default_controls(maxRestarts=8))
        assert p.converged
        tr = np.abs(m.source - orc.amul(m, psi)).sum() / p.normFactor
        assert tr < 1e-6
        sols[m.name] = (psi, p)
        print(f"[renumbered] {m.name}: schedule {ctx.active_schedule()}, tiles "
              f"{ctx.slot_tiles()}, it={p.nIterations}", flush=True)
