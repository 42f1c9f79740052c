"""Edge cases of the solve path on the GPU: empty meshes, meshes without
faces (diagonal matrices), one cell, disconnected blocks, and a 256^3 mesh
(16.8 M cells, column offsets up to 65536 held as per-tile bases plus int16
deltas) checked through the true FP64 residual."""

import numpy as np
import pytest

from metal4foam_amd import hipfoam as H
from metal4foam_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def diagonal_system(n, seed):
    rng = np.random.default_rng(seed)
    e = np.zeros(0, dtype=np.int32)
    own = np.zeros(n + 1, dtype=np.int32)
    return mg.LduMatrix(n, e, e, own, e, own.copy(), rng.uniform(1.0, 3.0, n), np.zeros(0),
                        None, rng.standard_normal(n), (n,), "diagonal")


def block_diagonal(a, b):
    """Two independent systems as one LDU matrix (cells of b after a)."""
    low = np.concatenate([a.lower_addr, b.lower_addr + a.n]).astype(np.int32)
    upp = np.concatenate([a.upper_addr, b.upper_addr + a.n]).astype(np.int32)
    n = a.n + b.n
    own, los, loss = mg.derived_addressing(n, low, upp)
    lower = None if a.lower is None else np.concatenate([a.lower, b.lower])
    return mg.LduMatrix(n, low, upp, own, los, loss, np.concatenate([a.diag, b.diag]),
                        np.concatenate([a.upper, b.upper]), lower,
                        np.concatenate([a.source, b.source]), (n,), "blocks")


def true_res(m, psi, nf):
    lower = m.upper if m.lower is None else m.lower
    ax = m.diag * psi
    ax += np.bincount(m.lower_addr, weights=m.upper * psi[m.upper_addr], minlength=m.n)
    ax += np.bincount(m.upper_addr, weights=lower * psi[m.lower_addr], minlength=m.n)
    return np.abs(m.source - ax).sum() / nf


@pytest.mark.parametrize("solver", [H.PCG, H.PBICGSTAB])
def test_empty_mesh(ctx, solver):
    m = diagonal_system(0, 1)
    ctx.set_topology(m)
    psi = np.zeros(0)
    perf = ctx.solve(solver, m.diag, m.upper, None, psi, m.source)
    assert perf.converged and perf.nIterations == 0
    perf = ctx.solve(solver, m.diag, m.upper, None, psi, m.source, fp64=True)
    assert perf.converged and perf.nIterations == 0


@pytest.mark.parametrize("solver", [H.PCG, H.PBICGSTAB])
@pytest.mark.parametrize("n", [1, 300])
def test_diagonal_matrix(ctx, solver, n):
    m = diagonal_system(n, 2)
    ctx.set_topology(m)
    psi = np.zeros(n)
    perf = ctx.solve(solver, m.diag, m.upper, None, psi, m.source,
                     H.default_controls(tolerance=1e-12))
    assert perf.converged
    np.testing.assert_allclose(psi, m.source / m.diag, rtol=1e-12, atol=0)
    psi = np.zeros(n)
    perf = ctx.solve(solver, m.diag, m.upper, None, psi, m.source,
                     H.default_controls(tolerance=1e-12), fp64=True)
    assert perf.converged and perf.nIterations <= 1
    np.testing.assert_allclose(psi, m.source / m.diag, rtol=1e-14, atol=0)


@pytest.mark.parametrize("kind", ["sym", "asym"])
def test_disconnected_blocks_match_separate_solves(ctx, orc, kind):
    if kind == "sym":
        a = mg.poisson(9, 8, 7, seed=3, pin=1.0)
        b = mg.poisson(6, 5, 4, seed=4, pin=1.0)
        solver = H.PCG
    else:
        a = mg.convection_diffusion(9, 8, 7, seed=5)
        b = mg.convection_diffusion(6, 5, 4, seed=6)
        solver = H.PBICGSTAB
    m = block_diagonal(a, b)
    ctx.set_topology(m)
    psi = np.zeros(m.n)
    perf = ctx.solve(solver, m.diag, m.upper, m.lower, psi, m.source,
                     H.default_controls(tolerance=1e-10, maxRestarts=20, maxIter=5000))
    assert perf.converged
    x = np.concatenate([np.linalg.solve(mg.dense(a), a.source),
                        np.linalg.solve(mg.dense(b), b.source)])
    assert np.abs(psi - x).max() <= 1e-6 * np.abs(x).max()
    assert true_res(m, psi, perf.normFactor) < 1e-10


@pytest.mark.slow
def test_256_cubed_poisson(ctx):
    m = mg.poisson(256, 256, 256, seed=1001, lognormal=True)
    print(f"[edge] 256^3: n={m.n} faces={m.nfaces}", flush=True)
    ctx.set_topology(m)
    keep = ctx.get_schedule().as_dict()
    x = np.random.default_rng(3).standard_normal(m.n).astype(np.float32)
    ys = []
    for sched in (0, 2):
        ctx.set_schedule(spmvSchedule=sched)
        ctx.set_matrix(m.diag, m.upper, None)
        ys.append(ctx.amul(x, H.FP32))
    ctx.set_schedule(**keep)
    assert ctx.active_schedule() == 2
    # z offset 65,536: every tile still patterned, on bases beyond int16
    assert ctx.slot_layout() == (m.n, 3)
    np.testing.assert_array_equal(ys[1], ys[0])
    psi = np.zeros(m.n)
    perf = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source,
                     H.default_controls(maxRestarts=8, maxIter=20000))
    tr = true_res(m, psi, perf.normFactor)
    print(f"[edge] 256^3 refined: it={perf.nIterations} restarts={perf.restarts} "
          f"res={perf.finalResidual:.3e} true={tr:.3e} t={perf.tTotal:.3f}s", flush=True)
    assert perf.converged and tr < 1e-6
    assert tr == pytest.approx(perf.finalResidual, rel=1e-6)


def test_staged_host_copies_round_trip(ctx):
    """Host <-> device copies of 4 MiB and more go through pinned 8 MiB
    chunks (three in a ring) with the host side split over threads: sizes
    below, at and beyond the threshold, chunk-size multiples and ragged
    tails all arrive intact, and a copy issued after a device-side copy sees
    its result."""
    rng = np.random.default_rng(5)
    chunk = 8 << 20
    for nbytes in (1000, 4 << 20, (4 << 20) + 3, chunk, chunk + 5, 3 * chunk + 7,
                   5 * chunk + 11, (50 << 20) + 13):
        a = rng.integers(0, 256, nbytes, dtype=np.uint8)
        d = H.DeviceArray.from_host(ctx, a)
        e = H.DeviceArray(ctx, (nbytes,), np.uint8)
        e.copy_from(d)
        back = e.download()
        assert back.tobytes() == a.tobytes(), nbytes
        d.free()
        e.free()


def test_host_array_solve_matches_device_solve(ctx):
    """A solve through host arrays (staged copies in and out) gives the same
    bits as the same solve on device arrays."""
    m = mg.poisson(96, 96, 64, seed=3, lognormal=True)  # psi, source 4.5 MiB; upper 13.4 MiB
    ctx.set_topology(m)
    psi = np.zeros(m.n)
    perf_h, res_h = ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source, residual=True)
    D = H.DeviceArray
    dd, du, ds = D.from_host(ctx, m.diag), D.from_host(ctx, m.upper), D.from_host(ctx, m.source)
    dp, dr = D.from_host(ctx, np.zeros(m.n)), D(ctx, (m.n,), np.float64)
    perf_d = ctx.solve_device(H.PCG, dd, du, None, dp, ds, residual=dr)
    np.testing.assert_array_equal(dp.download(), psi)
    np.testing.assert_array_equal(dr.download(), res_h)
    assert perf_h.nIterations == perf_d.nIterations
    for a in (dd, du, ds, dp, dr):
        a.free()


def _sampled(length, k=1024):
    """Indices a fingerprint of k evenly spaced entries (first and last
    included) reads: the in-place rewrite below avoids all of them."""
    if length <= k:
        return set(range(length))
    return {j * (length - 1) // (k - 1) for j in range(k)}


def _rewrite_unsampled(m):
    """Swap the neighbours of two owner faces of one row, in place, choosing
    a row whose changed upperAddr and losort entries are all unsampled.
    Returns the positions changed in upperAddr and losort."""
    su, sl = _sampled(m.nfaces), _sampled(m.nfaces)
    for r in range(m.n // 2, m.n):
        f0, f1 = m.owner_start[r], m.owner_start[r + 1]
        if f1 - f0 < 3:
            continue
        p1, p2 = f0 + 1, f0 + 2
        upp = m.upper_addr.copy()
        upp[p1], upp[p2] = upp[p2], upp[p1]
        own, los, loss = mg.derived_addressing(m.n, m.lower_addr, upp)
        ql = np.nonzero(los != m.losort)[0]
        if {p1, p2} & su or set(ql.tolist()) & sl:
            continue
        assert np.array_equal(own, m.owner_start) and np.array_equal(loss, m.losort_start)
        m.upper_addr[:] = upp  # same buffers
        m.losort[:] = los
        return [p1, p2], ql
    raise AssertionError("no row with unsampled faces")


def test_topology_rewritten_in_place_is_detected(ctx, orc):
    """The default topology check hashes every addressing entry: a mesh
    rewritten in place (same buffers, same sizes) at entries that a sampled
    fingerprint never reads is uploaded again, and the solve matches the
    oracle on the rewritten mesh.  The reference's key (sizes and pointers,
    HIPFOAM_TOPO_POINTERS) keeps the stale addressing, which this shows."""
    m = mg.poisson(20, 18, 16, seed=11, lognormal=True)
    ctx.set_topology_check(H.TOPO_CONTENT)
    ctx.set_topology(m)
    ctl = H.default_controls(maxRestarts=8)
    psi_a = np.zeros(m.n)
    assert ctx.solve(H.PCG, m.diag, m.upper, None, psi_a, m.source, ctl).converged
    bufs = [m.lower_addr, m.upper_addr, m.owner_start, m.losort, m.losort_start]
    old = [b.copy() for b in bufs]
    pu, pl = _rewrite_unsampled(m)
    assert all(b is H._i32(b) for b in bufs)  # the arrays passed are the same buffers
    print(f"[topology] rewrote upperAddr at {pu} and losort at {pl.tolist()}", flush=True)
    ctx.set_topology(m)
    psi_b = np.zeros(m.n)
    perf = ctx.solve(H.PCG, m.diag, m.upper, None, psi_b, m.source, ctl)
    ref_psi, ref = orc.refined_solve(m, ctl=orc.controls(maxRestarts=8))
    x = np.linalg.solve(mg.dense(m), m.source)
    scale = np.abs(x).max()
    print(f"[topology] gpu it={perf.nIterations} oracle it={ref.nIterations}; "
          f"|psi_b - x|/|x| = {np.abs(psi_b - x).max() / scale:.2e}, "
          f"|psi_a - x|/|x| = {np.abs(psi_a - x).max() / scale:.2e}", flush=True)
    assert perf.converged and perf.nIterations == ref.nIterations
    assert perf.normFactor == pytest.approx(ref.normFactor, rel=1e-12)
    assert np.abs(psi_b - ref_psi).max() <= 1e-5 * scale
    # a 1e-6 residual leaves the pinned matrix's constant mode loose: the GPU's
    # error against the dense solution is bounded by the oracle's own
    assert np.abs(psi_b - x).max() <= max(5 * np.abs(ref_psi - x).max(), 1e-6 * scale)
    assert np.abs(psi_a - x).max() > 1e-3 * scale  # the rewrite changed the solution
    # the reference's key: the rewrite back to the first mesh goes unseen
    ctx.set_topology_check(H.TOPO_POINTERS)
    for b, o in zip(bufs, old):
        b[:] = o
    ctx.set_topology(m)
    psi_stale = np.zeros(m.n)
    ctx.solve(H.PCG, m.diag, m.upper, None, psi_stale, m.source, ctl)
    assert np.abs(psi_stale - psi_a).max() > 1e-3 * scale
    # back to the content check: the same call now uploads the first mesh
    ctx.set_topology_check(H.TOPO_CONTENT)
    ctx.set_topology(m)
    psi_again = np.zeros(m.n)
    ctx.solve(H.PCG, m.diag, m.upper, None, psi_again, m.source, ctl)
    np.testing.assert_array_equal(psi_again, psi_a)


def test_same_mesh_at_new_addresses_is_not_uploaded_again(ctx):
    """The content key ignores where the arrays live: the same mesh passed in
    fresh copies gives the same solve (and the check itself must not fail)."""
    m = mg.poisson(9, 8, 7, seed=12, lognormal=True)
    ctx.set_topology(m)
    psi = np.zeros(m.n)
    ctx.solve(H.PCG, m.diag, m.upper, None, psi, m.source)
    m2 = mg.LduMatrix(m.n, m.lower_addr.copy(), m.upper_addr.copy(), m.owner_start.copy(),
                      m.losort.copy(), m.losort_start.copy(), m.diag, m.upper, None, m.source)
    ctx.set_topology(m2)
    psi2 = np.zeros(m.n)
    ctx.solve(H.PCG, m.diag, m.upper, None, psi2, m.source)
    np.testing.assert_array_equal(psi2, psi)
    with pytest.raises(H.HipFoamError):
        ctx.set_topology_check(7)
