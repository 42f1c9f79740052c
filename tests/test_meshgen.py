"""Synthetic LDU generators: addressing invariants and workload conventions."""

import numpy as np
import pytest

from metal4foam_amd import meshgen as mg


def _loops_owner_start(n, own):
    # lduAddressing::calcOwnerStart as a plain loop
    out = [len(own)] * (n + 1)
    out[0] = 0
    nOwnStart, i = 0, 1
    for f, o in enumerate(own):
        if o > nOwnStart:
            while i <= o:
                out[i] = f
                i += 1
            nOwnStart = o
    return np.array(out)


def _loops_losort(n, nbr):
    # lduAddressing::calcLosort as a plain loop (scatter by neighbour, gather)
    cell_faces = [[] for _ in range(n)]
    for f, c in enumerate(nbr):
        cell_faces[c].append(f)
    losort = [f for c in range(n) for f in cell_faces[c]]
    start = [0] * (n + 1)
    k = 0
    for c in range(n):
        start[c] = k
        k += len(cell_faces[c])
    start[n] = len(nbr)
    return np.array(losort), np.array(start)


@pytest.mark.parametrize("shape", [(1, 1, 1), (3, 1, 1), (3, 3, 1), (4, 3, 2), (5, 4, 3)])
def test_structured_addressing_matches_lduaddressing_loops(shape):
    n, low, upp, own, los, loss = mg.structured_addressing(*shape)
    nx, ny, nz = shape
    assert n == nx * ny * nz
    assert low.shape[0] == (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    assert low.dtype == np.int32 and upp.dtype == np.int32
    assert np.all(low < upp)
    # upper-triangular order: sorted by owner, then neighbour
    key = low.astype(np.int64) * n + upp
    assert np.all(np.diff(key) > 0)
    np.testing.assert_array_equal(own, _loops_owner_start(n, low))
    ls, lss = _loops_losort(n, upp)
    np.testing.assert_array_equal(los, ls)
    np.testing.assert_array_equal(loss, lss)


def test_neighbours_are_lexicographic_stencil():
    n, low, upp, *_ = mg.structured_addressing(4, 3, 2)
    d = upp - low
    assert set(np.unique(d)) == {1, 4, 12}


def test_config0_conventions():
    m = mg.config0()
    assert m.n == 10000 and m.nfaces == 19800 and m.symmetric
    assert np.all(m.upper == -1.0)
    nbrs = np.bincount(m.lower_addr, minlength=m.n) + np.bincount(m.upper_addr, minlength=m.n)
    expect = nbrs.astype(float)
    expect[0] += 1.0
    np.testing.assert_array_equal(m.diag, expect)
    assert abs(m.source.mean()) < 1e-15 and np.abs(m.source).max() <= 2.0


def test_poisson_rowsum_is_pin():
    m = mg.poisson(6, 5, 4, seed=1, lognormal=True)
    a = mg.dense(m)
    u = a.sum(axis=1)
    e0 = np.zeros(m.n)
    e0[0] = 1.0
    np.testing.assert_allclose(u, e0, atol=1e-12)
    np.testing.assert_array_equal(a, a.T)


def test_convection_diffusion_flux_is_divergence_free_and_matrix_dominant():
    m = mg.convection_diffusion(7, 6, 5, seed=3)
    phi = m.phi
    div = np.bincount(m.lower_addr, weights=phi, minlength=m.n) - np.bincount(
        m.upper_addr, weights=phi, minlength=m.n)
    assert np.abs(div).max() < 1e-12
    assert np.isclose(np.abs(phi).max(), 2.0)
    a = mg.dense(m)
    off = np.abs(a).sum(axis=1) - np.abs(np.diag(a))
    np.testing.assert_allclose(np.diag(a) - off, 0.1, atol=1e-12)
    assert not np.allclose(a, a.T)
    # upwind: upper picks the inflow part, lower the outflow part
    np.testing.assert_array_equal(m.upper, -1.0 - np.maximum(-phi, 0.0))
    np.testing.assert_array_equal(m.lower, -1.0 - np.maximum(phi, 0.0))


def test_config3_batch_shapes():
    ms = mg.config3(nsys=6, nrhs=5)
    assert len(ms) == 6
    for s, m in enumerate(ms):
        side = m.shape[0]
        assert 20 <= side <= 80 and m.shape == (side, side, side)
        assert m.source.shape == (m.n, 5)
        assert m.symmetric == (s % 2 == 0)


def test_generators_are_deterministic():
    a, b = mg.config0(), mg.config0()
    np.testing.assert_array_equal(a.source, b.source)
    c, d = mg.poisson(5, 5, 5, 7), mg.poisson(5, 5, 5, 7)
    np.testing.assert_array_equal(c.upper, d.upper)


def test_renumbering_is_a_symmetric_permutation():
    """renumber() gives P A P^T with the right-hand side permuted alike, in
    OpenFOAM's face order; RCM restores short rows after a random order."""
    m = mg.convection_diffusion(7, 6, 5, seed=3)
    perm = np.random.default_rng(5).permutation(m.n)
    r = mg.renumber(m, perm)
    P = np.zeros((m.n, m.n))
    P[perm, np.arange(m.n)] = 1.0
    np.testing.assert_allclose(mg.dense(r), P @ mg.dense(m) @ P.T, rtol=0, atol=0)
    np.testing.assert_array_equal(r.source[perm], m.source)
    assert np.all(r.lower_addr < r.upper_addr)
    assert np.all(np.diff(r.lower_addr) >= 0)
    x = np.linalg.solve(mg.dense(m), m.source)
    np.testing.assert_allclose(np.linalg.solve(mg.dense(r), r.source)[perm], x, rtol=1e-12)
    rnd = mg.random_renumbering(mg.poisson(12, 11, 10, seed=1), 2)
    rows = lambda q: np.diff(q.owner_start)  # noqa: E731
    assert (rows(rnd) > 3).any()
    rcm = mg.rcm_renumbering(rnd)
    assert (rows(rcm) <= 3).all() and (np.diff(rcm.losort_start) <= 3).all()
