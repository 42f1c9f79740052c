"""Writes tests/golden/fixtures_tiny.npz: golden vectors for tiny LDU systems.

Every expected value comes from dense numpy linear algebra on the matrix the
LDU arrays describe (numpy.linalg.solve, dense products), independently of
the oracle and of the GPU library, so the fixtures pin both.  Inputs are the
seeded generators of metal4foam_amd.meshgen plus one irregular graph whose
rows have more than three faces.

    python tests/golden/make_fixtures.py      # rewrites the .npz

Per case <c> the archive holds (np.load(..., allow_pickle=False)):
  <c>.n, .low, .upp, .own, .los, .loss (int32 addressing), .diag, .upper,
  .lower (empty when symmetric), .b, .x_probe, .Ax (A x_probe), .ATx
  (A^T x_probe), .sumA (A 1), .nf_probe (OpenFOAM normFactor of x_probe),
  .x (A^-1 b), .xt (adjoint target), .lam (A^-T (x - xt)), .dUpper,
  .dLower (symmetric: the combined gradient in dUpper, dLower empty),
  .dDiag, .dSource (gradients of 0.5|x - xt|^2).
"""

import os
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from metal4foam_amd import meshgen as mg  # noqa: E402


def irregular(n, nextra, seed):
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
    return mg.LduMatrix(n, low, upp, own, los, loss, diag, upper, lower,
                        rng.standard_normal(n), (n,), "irregular")


def cases():
    return {
        "poisson_3x3x1": mg.poisson(3, 3, 1, seed=11, lognormal=False),
        "poisson_4x4x4": mg.poisson(4, 4, 4, seed=12, lognormal=True),
        "convdiff_4x3x2": mg.convection_diffusion(4, 3, 2, seed=13),
        "irregular_14": irregular(14, 12, 14),
    }


def norm_factor(a, psi, b):
    """OpenFOAM's default normFactor: sum(|A psi - xbar A1| + |b - xbar A1|) + 1e-20."""
    xbar = psi.mean()
    t = (a @ np.ones(len(psi))) * xbar
    return np.abs(a @ psi - t).sum() + np.abs(b - t).sum() + 1e-20


def main():
    out = {}
    for name, m in cases().items():
        a = mg.dense(m)
        rng = np.random.default_rng(zlib.crc32(name.encode()))
        x_probe = rng.standard_normal(m.n)
        x = np.linalg.solve(a, m.source)
        xt = rng.standard_normal(m.n)
        lam = np.linalg.solve(a.T, x - xt)
        l, u = m.lower_addr, m.upper_addr
        du = -lam[l] * x[u]
        dl = -lam[u] * x[l]
        rec = {
            "n": np.array(m.n, dtype=np.int32), "low": m.lower_addr, "upp": m.upper_addr,
            "own": m.owner_start, "los": m.losort, "loss": m.losort_start,
            "diag": m.diag, "upper": m.upper,
            "lower": np.zeros(0) if m.lower is None else m.lower,
            "b": m.source, "x_probe": x_probe, "Ax": a @ x_probe, "ATx": a.T @ x_probe,
            "sumA": a @ np.ones(m.n), "nf_probe": np.array(norm_factor(a, x_probe, m.source)),
            "x": x, "xt": xt, "lam": lam,
            "dUpper": du + dl if m.lower is None else du,
            "dLower": np.zeros(0) if m.lower is None else dl,
            "dDiag": -lam * x, "dSource": lam,
        }
        for k, v in rec.items():
            out[f"{name}.{k}"] = np.asarray(v)
    path = os.path.join(HERE, "fixtures_tiny.npz")
    np.savez(path, **out)
    print(f"wrote {path}: {len(out)} arrays")


if __name__ == "__main__":
    main()
