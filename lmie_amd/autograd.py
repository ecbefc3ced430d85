"""torch.autograd.Function over the C library's reverse-mode pass.

``expansion(params, ...)`` returns the six expansion-coefficient sets of a
size-integrated solution as a differentiable function of the scalar inputs
(m_r, m_i and each mode's a1, a2): the forward calls lmie_adjoint_forward,
the backward lmie_adjoint_reverse with the incoming gradient as the
cotangent of the coefficients.  torch only carries the tensors; the
solution and its gradient are computed by the library.
"""
import numpy as np
import torch

from . import lmie


class _Expansion(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, lib, base_modes, weights, precision):
        # theta = [mr, mi, a1_0, a2_0, a1_1, a2_1, ...] (float64, CPU)
        t = theta.detach().cpu().numpy().astype(np.float64)
        modes = []
        for k, p in enumerate(base_modes):
            q = dict(p, mr=float(t[0]), mi=float(t[1]), a1=float(t[2 + 2 * k]), a2=float(t[3 + 2 * k]))
            modes.append(q)
        adj = lmie.Adjoint(lib, modes, weights, precision)
        out = adj.forward()
        ctx.adj = adj
        ctx.n_modes = len(modes)
        return torch.from_numpy(np.ascontiguousarray(out["gc"]))

    @staticmethod
    def backward(ctx, grad_gc):
        g = ctx.adj.reverse({"gc": grad_gc.detach().cpu().numpy().astype(np.float64)})
        ctx.adj.free()
        out = [g["mr"], g["mi"]]
        for k in range(ctx.n_modes):
            out += [g["modes"][k][0], g["modes"][k][1]]
        return torch.tensor(out, dtype=torch.float64), None, None, None, None


def expansion(theta, lib, base_modes, weights=None, precision="fp64"):
    """Differentiable gc [6][max_coef] as a function of theta =
    [mr, mi, a1_0, a2_0, a1_1, a2_1, ...] for the given base parameter dicts."""
    return _Expansion.apply(theta, lib, base_modes, weights, precision)
