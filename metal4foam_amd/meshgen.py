"""Synthetic LDU matrices in OpenFOAM's lduAddressing layout.

The mesh is a structured nx x ny x nz hex block with lexicographic cell order
``c = i + nx*(j + ny*k)``.  Internal faces are listed in OpenFOAM's
upper-triangular order (sorted by owner, then by neighbour: +x, +y, +z), so
``lowerAddr[f] < upperAddr[f]`` and ``lowerAddr`` is non-decreasing.  The
derived addressing follows ``lduAddressing``:

* ``ownerStart[i]``  first face owned by cell i (length n+1, last = nFaces)
* ``losort``         faces stably sorted by neighbour (``calcLosort``)
* ``losortStart[i]`` first losort slot whose neighbour is cell i (length n+1)

Coefficient conventions follow ``lduMatrix``: ``upper[f] = A[l, u]`` and
``lower[f] = A[u, l]`` with ``l = lowerAddr[f]``, ``u = upperAddr[f]``.

Every generator is deterministic: numpy ``Generator(MT19937(seed))``, with
the seeds 1000..1004 used by the benchmark workloads:

=====  ====================================================================
seed   workload
=====  ====================================================================
1000   2D Neumann pressure Poisson, 100x100x1, upper=lower=-1, cell 0 pinned
1001   3D pressure Poisson 128^3, face coefficients -exp(N(0,1)), cell 0 pinned
1002   3D convection-diffusion 96^3, first-order upwind, divergence-free flux
1003   batch of 64 uneven systems (32 Poisson + 32 convection-diffusion), 5 RHS
1004   adjoint target x_t ~ N(0,1) on the 128^3 Poisson matrix
=====  ====================================================================
"""

from __future__ import annotations

import dataclasses

import numpy as np

I32 = np.int32


@dataclasses.dataclass
class LduMatrix:
    """An LDU system: addressing (int32) + FP64 coefficients (+ RHS)."""

    n: int
    lower_addr: np.ndarray  # int32[F]  owner
    upper_addr: np.ndarray  # int32[F]  neighbour
    owner_start: np.ndarray  # int32[n+1]
    losort: np.ndarray  # int32[F]
    losort_start: np.ndarray  # int32[n+1]
    diag: np.ndarray  # f64[n]
    upper: np.ndarray  # f64[F]
    lower: np.ndarray | None  # f64[F] or None (symmetric: lower == upper)
    source: np.ndarray | None = None  # f64[n] (or [n, nrhs])
    shape: tuple = ()
    name: str = ""

    @property
    def nfaces(self) -> int:
        return int(self.lower_addr.shape[0])

    @property
    def symmetric(self) -> bool:
        return self.lower is None

    def lower_or_upper(self) -> np.ndarray:
        return self.upper if self.lower is None else self.lower


def structured_addressing(nx: int, ny: int, nz: int):
    """Internal-face addressing of a structured block, OpenFOAM face order.

    Returns (n, lowerAddr, upperAddr, ownerStart, losort, losortStart).
    """
    n = nx * ny * nz
    c = np.arange(n, dtype=np.int64)
    i = c % nx
    j = (c // nx) % ny
    k = c // (nx * ny)
    owners = []
    keys = []
    nbrs = []
    for d, (mask, off) in enumerate(
        ((i < nx - 1, 1), (j < ny - 1, nx), (k < nz - 1, nx * ny))
    ):
        o = c[mask]
        owners.append(o)
        nbrs.append(o + off)
        keys.append(o * 3 + d)  # owner-major, then +x,+y,+z == neighbour order
    key = np.concatenate(keys)
    order = np.argsort(key, kind="stable")
    low = np.concatenate(owners)[order].astype(I32)
    upp = np.concatenate(nbrs)[order].astype(I32)
    own_start, losort, losort_start = derived_addressing(n, low, upp)
    return n, low, upp, own_start, losort, losort_start


def derived_addressing(n: int, low: np.ndarray, upp: np.ndarray):
    """ownerStart, losort and losortStart exactly as lduAddressing builds them."""
    nf = low.shape[0]
    own_start = np.searchsorted(low, np.arange(n + 1), side="left").astype(I32)
    own_start[n] = nf
    losort = np.argsort(upp, kind="stable").astype(I32)
    counts = np.bincount(upp, minlength=n)
    losort_start = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=losort_start[1:])
    return own_start, losort, losort_start.astype(I32)


def _rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.MT19937(seed))


def row_offdiag_sums(n, low, upp, upper, lower):
    """Sum of the off-diagonal coefficients of every row (FP64)."""
    s = np.bincount(low, weights=upper, minlength=n)
    s += np.bincount(upp, weights=lower, minlength=n)
    return s


def poisson(nx, ny, nz, seed, lognormal=True, pin=1.0, zero_mean_rhs=True,
            rhs="uniform", nrhs=1, name="poisson"):
    """Neumann pressure Poisson with one pinned reference cell.

    Face coefficients ``upper = lower = -a_f`` with ``a_f = exp(N(0,1))``
    (lognormal conductivity) or ``a_f = 1``; ``diag_i = sum_f a_f`` over the
    faces of cell i, plus ``pin`` at cell 0 (the reference cell).
    """
    rng = _rng(seed)
    n, low, upp, own, los, loss = structured_addressing(nx, ny, nz)
    nf = low.shape[0]
    a = np.exp(rng.standard_normal(nf)) if lognormal else np.ones(nf)
    upper = -a
    diag = -row_offdiag_sums(n, low, upp, upper, upper)
    diag[0] += pin
    if rhs == "uniform":
        b = rng.uniform(-1.0, 1.0, size=(n, nrhs))
    else:
        b = rng.standard_normal((n, nrhs))
    if zero_mean_rhs:
        b -= b.mean(axis=0, keepdims=True)
    if nrhs == 1:
        b = b[:, 0].copy()
    return LduMatrix(n, low, upp, own, los, loss, diag, upper, None, b,
                     (nx, ny, nz), name)


def divergence_free_flux(nx, ny, nz, rng, umax=2.0):
    """Face fluxes (OpenFOAM face order) from the discrete curl of a random
    edge potential that vanishes on the boundary, scaled to max|phi| = umax.

    The discrete divergence of a discrete curl telescopes to zero, and the
    boundary-face fluxes vanish with the potential, so the internal fluxes
    alone satisfy sum_out(phi) - sum_in(phi) = 0 in every cell.
    """
    az = rng.standard_normal((nz, ny + 1, nx + 1))  # z-edges at (x node, y node), z cell
    ay = rng.standard_normal((nz + 1, ny, nx + 1))  # y-edges at (x node, z node), y cell
    ax = rng.standard_normal((nz + 1, ny + 1, nx))  # x-edges at (y node, z node), x cell
    az[:, :, 0] = az[:, :, nx] = 0.0
    az[:, 0, :] = az[:, ny, :] = 0.0
    ay[:, :, 0] = ay[:, :, nx] = 0.0
    ay[0, :, :] = ay[nz, :, :] = 0.0
    ax[:, 0, :] = ax[:, ny, :] = 0.0
    ax[0, :, :] = ax[nz, :, :] = 0.0
    # flux through the x-face at x node xi (between cells xi-1 and xi)
    phx = (az[:, 1:, :] - az[:, :-1, :]) - (ay[1:, :, :] - ay[:-1, :, :])  # (nz, ny, nx+1)
    phy = (ax[1:, :, :] - ax[:-1, :, :]) - (az[:, :, 1:] - az[:, :, :-1])  # (nz, ny+1, nx)
    phz = (ay[:, :, 1:] - ay[:, :, :-1]) - (ax[:, 1:, :] - ax[:, :-1, :])  # (nz+1, ny, nx)
    # internal faces only, keyed by owner cell and direction
    fx = phx[:, :, 1:nx]  # owner (k, j, i) for i < nx-1
    fy = phy[:, 1:ny, :]
    fz = phz[1:nz, :, :]
    n = nx * ny * nz
    c = np.arange(n, dtype=np.int64)
    i = c % nx
    j = (c // nx) % ny
    k = c // (nx * ny)
    vals, keys = [], []
    for d, (mask, arr) in enumerate(
        ((i < nx - 1, fx), (j < ny - 1, fy), (k < nz - 1, fz))
    ):
        vals.append(arr.reshape(-1))
        keys.append(c[mask] * 3 + d)
    key = np.concatenate(keys)
    order = np.argsort(key, kind="stable")
    phi = np.concatenate(vals)[order]
    m = np.abs(phi).max() if phi.size else 1.0
    return phi * (umax / m if m > 0 else 1.0)


def convection_diffusion(nx, ny, nz, seed, nrhs=1, name="convdiff", umax=2.0,
                         shift=0.1):
    """Asymmetric upwind convection-diffusion.

    ``upper_f = -1 - max(-phi_f, 0)``, ``lower_f = -1 - max(phi_f, 0)`` with a
    divergence-free face flux phi (owner -> neighbour positive), and
    ``diag_i = -sum(row off-diagonals) + shift``; RHS ~ N(0, 1).
    """
    rng = _rng(seed)
    n, low, upp, own, los, loss = structured_addressing(nx, ny, nz)
    phi = divergence_free_flux(nx, ny, nz, rng, umax)
    upper = -1.0 - np.maximum(-phi, 0.0)
    lower = -1.0 - np.maximum(phi, 0.0)
    diag = -row_offdiag_sums(n, low, upp, upper, lower) + shift
    b = rng.standard_normal((n, nrhs))
    if nrhs == 1:
        b = b[:, 0].copy()
    m = LduMatrix(n, low, upp, own, los, loss, diag, upper, lower, b,
                  (nx, ny, nz), name)
    m.phi = phi
    return m


# --------------------------------------------------------------------------
# benchmark workloads
# --------------------------------------------------------------------------

# --------------------------------------------------------------------------
# matrices whose Jacobi-scaled operator has exactly two eigenvalues
# --------------------------------------------------------------------------
#
# With a constant diagonal d, D^-1 A = A / d.  When A is diagonalisable with
# two distinct eigenvalues, every Krylov method whose iterates are the exact
# polynomial recurrences of CG / BiCG terminates in two iterations (its
# residual polynomial of degree 2 vanishes on both eigenvalues), whatever
# right-hand side it is given.  A method with a wrong beta (steepest descent,
# or BiCGStab without its alpha/omega factor) does not.  The same holds for
# any preconditioner that is a polynomial in D^-1 A (Chebyshev-Jacobi).

def pair_blocks(npairs, seed, diag=3.0, upper=-1.0, lower=None, name="pairs"):
    """``npairs`` disjoint, identical 2x2 blocks [[d, u], [l, d]] on cells
    (2k, 2k+1), one face each.  D^-1 A has the eigenvalues 1 +- sqrt(u l)/d
    (u l > 0); ``lower=None`` is the symmetric block l = u.  Every row has one
    face, so the mesh fits any row layout.  RHS ~ N(0, 1)."""
    rng = _rng(seed)
    n = 2 * npairs
    low = np.arange(0, n, 2, dtype=I32)
    upp = low + 1
    own, los, loss = derived_addressing(n, low, upp)
    up = np.full(npairs, float(upper))
    lo = None if lower is None else np.full(npairs, float(lower))
    d = np.full(n, float(diag))
    b = rng.standard_normal(n)
    return LduMatrix(n, low, upp, own, los, loss, d, up, lo, b, (n, 1, 1), name)


def complete_graph(n, seed, shift=1.0, weights=None, name="complete"):
    """Every pair of the n cells shares a face (n(n-1)/2 faces, owner-major
    order), so every row has n-1 faces.  A = (n + shift) I - a b^T with
    b_i = 1/a_i: the off-diagonals are upper_f = -a_l/a_u and
    lower_f = -a_u/a_l (the complete-graph Laplacian plus shift when a = 1,
    symmetric), the diagonal is n - 1 + shift.  A rank-one perturbation of a
    multiple of I with b^T a = n, so A has the eigenvalues n + shift (n-1
    times) and shift, and is diagonalisable.  ``weights`` = a (positive);
    None gives the symmetric matrix.  RHS ~ N(0, 1)."""
    rng = _rng(seed)
    ii, jj = np.triu_indices(n, 1)
    low = ii.astype(I32)
    upp = jj.astype(I32)
    own, los, loss = derived_addressing(n, low, upp)
    d = np.full(n, n - 1.0 + shift)
    if weights is None:
        up = -np.ones(low.shape[0])
        lo = None
    else:
        a = np.asarray(weights, dtype=np.float64)
        up = -a[low] / a[upp]
        lo = -a[upp] / a[low]
    b = rng.standard_normal(n)
    return LduMatrix(n, low, upp, own, los, loss, d, up, lo, b, (n, 1, 1), name)


def renumber(m: LduMatrix, perm, name=None) -> LduMatrix:
    """The same operator with cell c renamed perm[c]: the faces are listed
    again in OpenFOAM's upper-triangular order (owner = the lower new number,
    sorted by owner then neighbour), upper/lower swap where a face's
    orientation flips, and the diagonal and right-hand side move with their
    cells."""
    perm = np.asarray(perm, dtype=np.int64)
    a, b = perm[m.lower_addr], perm[m.upper_addr]
    flip = a > b
    low = np.where(flip, b, a)
    upp = np.where(flip, a, b)
    order = np.lexsort((upp, low))
    low = low[order].astype(I32)
    upp = upp[order].astype(I32)
    lo_src = m.lower_or_upper()
    upper = np.where(flip, lo_src, m.upper)[order]
    lower = None if m.lower is None else np.where(flip, m.upper, m.lower)[order]
    own, los, loss = derived_addressing(m.n, low, upp)
    diag = np.empty_like(m.diag)
    diag[perm] = m.diag
    src = None
    if m.source is not None:
        src = np.empty_like(m.source)
        src[perm] = m.source
    out = LduMatrix(m.n, low, upp, own, los, loss, diag, upper, lower, src, m.shape,
                    name or f"{m.name}_renumbered")
    return out


def random_renumbering(m: LduMatrix, seed) -> LduMatrix:
    """m with its cells in a random order (no locality left)."""
    return renumber(m, _rng(seed).permutation(m.n), f"{m.name}_random_order")


def rcm_renumbering(m: LduMatrix, start=None) -> LduMatrix:
    """m renumbered by reverse Cuthill-McKee (scipy.sparse.csgraph), the
    bandwidth-reducing order OpenFOAM's renumberMesh offers; applied to
    another renumbering it recovers a banded order."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    g = sp.coo_matrix((np.ones(m.nfaces), (m.lower_addr, m.upper_addr)), shape=(m.n, m.n))
    g = (g + g.T).tocsr()
    order = reverse_cuthill_mckee(g, symmetric_mode=True)
    perm = np.empty(m.n, dtype=np.int64)
    perm[order] = np.arange(m.n)
    return renumber(m, perm, f"{m.name}_rcm")


def config0():
    """2D Neumann Poisson 100x100x1: unit faces, cell 0 pinned, U[-1,1]."""
    return poisson(100, 100, 1, seed=1000, lognormal=False, name="poisson2d_100x100")


def config1(side=128):
    """3D lognormal Poisson side^3 (128^3 = 2,097,152 cells)."""
    return poisson(side, side, side, seed=1001, lognormal=True,
                   name=f"poisson3d_lognormal_{side}^3")


def config2(side=96):
    """3D upwind convection-diffusion side^3 (96^3 = 884,736 cells)."""
    return convection_diffusion(side, side, side, seed=1002,
                                name=f"convdiff3d_{side}^3")


def config3(nsys=64, nrhs=5, side_lo=20, side_hi=80, seed=1003, zero_mean_poisson=True):
    """Batch of uneven systems: half lognormal Poisson, half convection-
    diffusion, cube side ~ U{side_lo..side_hi}, ``nrhs`` N(0,1) RHS each.

    ``zero_mean_poisson=True`` projects the Poisson systems' N(0,1) right-hand
    sides to zero mean, as pressure equations like the 128^3 workload have (a
    Neumann system with one pinned cell and a RHS of non-zero sum has a
    solution dominated by the pinned mode, sum(b) at the reference cell).
    ``False`` keeps the plain N(0,1) draws of the specification; those need
    the pinned projection (deflate 3) to converge reliably.  The draws are the
    same either way: only the projection differs."""
    rng = _rng(seed)
    sides = rng.integers(side_lo, side_hi + 1, size=nsys)
    seeds = rng.integers(0, 2**31 - 1, size=nsys)
    out = []
    for s, (side, sd) in enumerate(zip(sides, seeds)):
        side = int(side)
        if s % 2 == 0:
            m = poisson(side, side, side, seed=int(sd), lognormal=True,
                        zero_mean_rhs=zero_mean_poisson, rhs="normal", nrhs=nrhs,
                        name=f"batch{s}_poisson_{side}^3")
        else:
            m = convection_diffusion(side, side, side, seed=int(sd), nrhs=nrhs,
                                     name=f"batch{s}_convdiff_{side}^3")
        out.append(m)
    return out


def config4_target(n, seed=1004):
    """Adjoint target x_t ~ N(0,1)."""
    return _rng(seed).standard_normal(n)


def dense(m: LduMatrix) -> np.ndarray:
    """Dense FP64 matrix of an LDU system (tests on tiny meshes only)."""
    a = np.zeros((m.n, m.n))
    a[np.arange(m.n), np.arange(m.n)] = m.diag
    a[m.lower_addr, m.upper_addr] = m.upper
    a[m.upper_addr, m.lower_addr] = m.lower_or_upper()
    return a
