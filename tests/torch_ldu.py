"""A thin torch.autograd.Function over the library's solve and adjoint, for
the gradient test harness only (the product path is the C-ABI; nothing in
metal4foam_amd/ imports torch).

``LduSolve.apply(diag, upper, lower, source, spec)`` returns x = A^-1 b for the
LDU matrix (diag, upper, lower) on the addressing ``spec.mesh`` (lower None:
symmetric, upper serves both triangles).  backward() solves the transpose
system A^T lambda = dL/dx with hipfoam_adjoint and returns
dL/ddiag = -lambda o x, dL/dupper, dL/dlower and dL/db = lambda.  CUDA float64
tensors go through the _device entry points (their memory stays in HBM:
torch's stream is synchronised first, the library synchronises its own
before returning); CPU float64 tensors go through the host entry points.
"""

from __future__ import annotations

import ctypes as C
import dataclasses

import torch

from metal4foam_amd import hipfoam as H


@dataclasses.dataclass
class SolveSpec:
    ctx: "H.Context"
    mesh: object  # an LduMatrix whose addressing the solve uses
    solver: int = H.PCG
    forward_fp64: bool = True  # plain-FP64 forward solve (else refined)
    forward_controls: object = None
    adjoint_precision: int = H.FP64
    adjoint_controls: object = None


class _Dev:
    """A device address as the library's _device wrappers take it."""

    def __init__(self, t):
        self.ptr = None if t is None else C.c_void_p(t.data_ptr())


def _contig(t):
    return None if t is None else t.detach().contiguous()


class LduSolve(torch.autograd.Function):
    @staticmethod
    def forward(fctx, diag, upper, lower, source, spec: SolveSpec):
        for t in (diag, upper, lower, source):
            if t is not None and t.dtype != torch.float64:
                raise TypeError("LduSolve takes float64 tensors")
        diag, upper, lower, source = map(_contig, (diag, upper, lower, source))
        hc = spec.ctx
        hc.set_topology(spec.mesh)
        ctl = spec.forward_controls or H.default_controls()
        x = torch.zeros_like(source)
        if source.is_cuda:
            torch.cuda.synchronize(source.device)
            perf = hc.solve_device(spec.solver, _Dev(diag), _Dev(upper),
                                   None if lower is None else _Dev(lower), _Dev(x),
                                   _Dev(source), ctl, fp64=spec.forward_fp64)
        else:
            xn = x.numpy()
            perf = hc.solve(spec.solver, diag.numpy(), upper.numpy(),
                            None if lower is None else lower.numpy(), xn, source.numpy(),
                            ctl, fp64=spec.forward_fp64)
        # a 1e-14 target may end a few ulps above it in plain FP64
        if not (perf.converged or perf.finalResidual < 1e-12):
            raise RuntimeError(f"forward solve did not converge: {perf.as_dict()}")
        fctx.spec = spec
        fctx.has_lower = lower is not None
        fctx.save_for_backward(diag, upper, lower if lower is not None else diag, x)
        return x

    @staticmethod
    def backward(fctx, gx):
        spec = fctx.spec
        diag, upper, lower, x = fctx.saved_tensors
        lower = lower if fctx.has_lower else None
        hc = spec.ctx
        hc.set_topology(spec.mesh)
        ctl = spec.adjoint_controls or H.default_controls()
        g = gx.detach().contiguous()
        lam = torch.empty_like(x)
        dd = torch.empty_like(diag)
        du = torch.empty_like(upper)
        dl = torch.empty_like(upper) if lower is not None else None
        if x.is_cuda:
            torch.cuda.synchronize(x.device)
            perf = hc.adjoint_device(spec.solver, _Dev(diag), _Dev(upper),
                                     None if lower is None else _Dev(lower), _Dev(x), _Dev(g),
                                     _Dev(lam), _Dev(du), None if dl is None else _Dev(dl),
                                     _Dev(dd), None, ctl, precision=spec.adjoint_precision)
        else:
            perf, lam_n, du_n, dl_n, dd_n, _ = hc.adjoint(
                spec.solver, diag.numpy(), upper.numpy(),
                None if lower is None else lower.numpy(), x.numpy(), g.numpy(), ctl,
                precision=spec.adjoint_precision)
            lam, du, dd = (torch.from_numpy(a) for a in (lam_n, du_n, dd_n))
            dl = None if dl_n is None else torch.from_numpy(dl_n)
        if not (perf.converged or perf.finalResidual < 1e-10):
            raise RuntimeError(f"adjoint solve did not converge: {perf.as_dict()}")
        return dd, du, dl, lam, None
