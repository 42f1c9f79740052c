"""torch.autograd.gradcheck of the adjoint, through the thin
torch.autograd.Function of the gradient harness (torch_ldu.LduSolve): the
analytical gradients of hipfoam_adjoint (transpose solve + face, diagonal and
source gradient kernels) against torch's central differences of the forward
solve, in float64, for every coefficient and source entry of a small
symmetric (hipPCG) and asymmetric (hipPBiCGStab) system, with the plain-FP64
and the refined transpose solve, on host and on device tensors.

The cases run in a child process (torch_gradcheck_child.py), where torch is
imported before the library so that one HIP runtime serves both."""

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

CHILD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "torch_gradcheck_child.py")


def test_gradcheck_in_child_process():
    print("[gradcheck] starting the child process (first torch import may take a minute)",
          flush=True)
    r = subprocess.run([sys.executable, "-u", CHILD], capture_output=True, text=True,
                       timeout=600)
    print(r.stdout[-6000:], flush=True)
    print(r.stderr[-3000:], flush=True)
    assert r.returncode == 0, r.returncode
    assert r.stdout.count("[gradcheck]   ok") == 8 or (
        r.stdout.count("[gradcheck]   ok") == 4 and "tensors on cuda" not in r.stdout)
