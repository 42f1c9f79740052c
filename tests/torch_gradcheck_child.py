"""Child process of tests/test_gpu_torch_gradcheck.py: torch.autograd.gradcheck
of the adjoint through torch_ldu.LduSolve.

It runs in a process of its own because torch ships its own HIP runtime
(torch/lib/libamdhip64.so, the same soname as /opt/rocm's): importing torch
after the library has loaded /opt/rocm's runtime would map a second runtime
into the process, and torch's autograd initialises it (hipGetDeviceCount).
Here torch is imported first, so the library binds to torch's runtime and the
process holds one; this is checked in /proc/self/maps before any HIP call.

Prints one line per case and exits non-zero on the first failure."""

import os
import sys

import torch  # noqa: E402  (first: the library must bind to torch's HIP runtime)

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

from metal4foam_amd import _lib  # noqa: E402
from metal4foam_amd import hipfoam as H  # noqa: E402
from metal4foam_amd import meshgen as mg  # noqa: E402
from torch_ldu import LduSolve, SolveSpec  # noqa: E402

TIGHT64 = dict(tolerance=1e-14, maxIter=20000)
TIGHT32 = dict(tolerance=1e-12, maxRestarts=12, maxIter=20000)


def hip_runtimes():
    with open("/proc/self/maps") as f:
        return sorted({ln.split()[-1] for ln in f if "libamdhip64" in ln})


def case(ctx, kind, adjoint_precision, device):
    m = (mg.poisson(4, 4, 3, seed=91, lognormal=True) if kind == "sym"
         else mg.convection_diffusion(4, 3, 3, seed=92))
    solver = H.PCG if kind == "sym" else H.PBICGSTAB
    adj = (TIGHT64 if adjoint_precision == H.FP64
           else dict(TIGHT32, deflate=3 if kind == "sym" else 2))
    spec = SolveSpec(ctx, m, solver, forward_fp64=True,
                     forward_controls=H.default_controls(**TIGHT64),
                     adjoint_precision=adjoint_precision,
                     adjoint_controls=H.default_controls(**adj))

    def t(a):
        return torch.tensor(np.asarray(a, dtype=np.float64), device=device, requires_grad=True)

    diag, upper, src = t(m.diag), t(m.upper), t(m.source)
    lower = None if m.lower is None else t(m.lower)
    if lower is not None:
        inputs = (diag, upper, lower, src)

        def fn(d, u, lo, b):
            return LduSolve.apply(d, u, lo, b, spec)
    else:
        inputs = (diag, upper, src)

        def fn(d, u, b):
            return LduSolve.apply(d, u, None, b, spec)

    ok = torch.autograd.gradcheck(fn, inputs, eps=1e-6, atol=1e-6, rtol=1e-4)
    assert ok
    # scalar loss L = 0.5 |x - x_t|^2 through backward(), against dense FP64
    xt = np.random.default_rng(93).standard_normal(m.n)
    x = LduSolve.apply(diag, upper, lower, src, spec)
    L = 0.5 * ((x - torch.tensor(xt, device=device)) ** 2).sum()
    L.backward()
    a = mg.dense(m)
    xn = np.linalg.solve(a, m.source)
    lam = np.linalg.solve(a.T, xn - xt)
    np.testing.assert_allclose(src.grad.cpu().numpy(), lam, rtol=0,
                               atol=1e-9 * np.abs(lam).max())
    np.testing.assert_allclose(diag.grad.cpu().numpy(), -lam * xn, rtol=0,
                               atol=1e-9 * np.abs(lam * xn).max())
    ii, jj = m.lower_addr, m.upper_addr
    if lower is None:
        gu = -(lam[ii] * xn[jj] + lam[jj] * xn[ii])
    else:
        gu = -lam[ii] * xn[jj]
        np.testing.assert_allclose(lower.grad.cpu().numpy(), -lam[jj] * xn[ii], rtol=0,
                                   atol=1e-9 * np.abs(lam).max() * np.abs(xn).max())
    np.testing.assert_allclose(upper.grad.cpu().numpy(), gu, rtol=0,
                               atol=1e-9 * np.abs(lam).max() * np.abs(xn).max())


def main():
    _lib.load()
    rts = hip_runtimes()
    print(f"[gradcheck] torch {torch.__version__} (HIP {torch.version.hip}); "
          f"HIP runtimes mapped: {rts}", flush=True)
    if len(rts) != 1:
        print("[gradcheck] more than one HIP runtime in the process: not running", flush=True)
        return 3
    ctx = H.Context(0)
    print(f"[gradcheck] {ctx.description}", flush=True)
    devices = ["cpu"] + (["cuda"] if torch.cuda.is_available() else [])
    for device in devices:
        for kind in ("sym", "asym"):
            for prec in (H.FP64, H.FP32):
                print(f"[gradcheck] {kind} adjoint={'fp64' if prec == H.FP64 else 'fp32-refined'}"
                      f" tensors on {device} ...", flush=True)
                case(ctx, kind, prec, device)
                print("[gradcheck]   ok", flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
