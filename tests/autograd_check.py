"""Test-only torch.autograd.Function over the library's SGNS loss kernels.

Run as a separate process by tests/test_gpu_zz_autograd.py: torch brings its
own HIP runtime, so this check keeps it out of the main test process.  Torch
only holds CPU tensors here; the kernels run through the library's C ABI on
device copies.  Exits 0 and prints OK when the gradients match torch's own
float64 autograd of the same loss."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "gdistger-re_amd"))

import numpy as np  # noqa: E402

import dger_sgns as G  # noqa: E402
from oracle import loss_oracle as L  # noqa: E402

G.lib()
import torch  # noqa: E402


class SGNSLoss(torch.autograd.Function):
    """Forward = dger_loss_fwd_f64, backward = dger_loss_bwd_f64."""

    @staticmethod
    def forward(ctx_, syn0, syn1, cidx, oidx):
        d0 = G.DeviceArray(syn0.detach().numpy())
        d1 = G.DeviceArray(syn1.detach().numpy())
        dc = G.DeviceArray(cidx.numpy())
        do = G.DeviceArray(oidx.numpy())
        ctx_.save_for_backward(syn0, syn1, cidx, oidx)
        v = G.loss_fwd(d0, d1, dc, do, cidx.shape[0], cidx.shape[1], oidx.shape[1])
        return torch.tensor(v, dtype=syn0.dtype)

    @staticmethod
    def backward(ctx_, gout):
        syn0, syn1, cidx, oidx = ctx_.saved_tensors
        d0 = G.DeviceArray(syn0.detach().numpy())
        d1 = G.DeviceArray(syn1.detach().numpy())
        dc = G.DeviceArray(cidx.numpy())
        do = G.DeviceArray(oidx.numpy())
        g0, g1 = G.loss_bwd(d0, d1, dc, do, cidx.shape[0], cidx.shape[1], oidx.shape[1], float(gout))
        return torch.from_numpy(g0), torch.from_numpy(g1), None, None


def main():
    s0, s1, ctx, out = L.gradcheck_workload(rows=60, dim=16, n_mb=8, seed=5)
    t0 = torch.tensor(s0, requires_grad=True)
    t1 = torch.tensor(s1, requires_grad=True)
    tc = torch.from_numpy(ctx)
    to = torch.from_numpy(out)
    y = SGNSLoss.apply(t0, t1, tc, to)
    (2.0 * y).backward()
    r0 = torch.tensor(s0, requires_grad=True)
    r1 = torch.tensor(s1, requires_grad=True)
    f = torch.einsum("mcd,mod->mco", r0[tc.long()], r1[to.long()])
    sign = torch.ones(out.shape[1], dtype=torch.float64)
    sign[0] = -1
    ref = torch.nn.functional.softplus(f * sign).sum()
    (2.0 * ref).backward()
    assert abs(float(y) - float(ref)) <= 1e-12 * abs(float(ref)), (float(y), float(ref))
    assert torch.allclose(t0.grad, r0.grad, rtol=1e-10, atol=1e-14)
    assert torch.allclose(t1.grad, r1.grad, rtol=1e-10, atol=1e-14)
    print("OK")


if __name__ == "__main__":
    main()
