"""torch.autograd.Function over the C library's reverse-mode pass.

``expansion(params, ...)`` returns the six expansion-coefficient sets of a
size-integrated solution as a differentiable function of the scalar inputs
(m_r, m_i, each mode's a1, a2 and, with grid=True, the wavelength and the
radius range r1, r2 shared by the modes): the forward calls lmie_adjoint_forward,
the backward lmie_adjoint_reverse with the incoming gradient as the
cotangent of the coefficients.  torch only carries the tensors; the
solution and its gradient are computed by the library.  The forward's device
state belongs to the graph node: backward may run more than once
(retain_graph=True), and the state is released when the node is freed.
"""
import numpy as np
import torch

from . import lmie


class _Expansion(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, lib, base_modes, weights, precision, grid):
        # theta = [mr, mi, a1_0, a2_0, a1_1, a2_1, ...] (+ [lambda, r1, r2] with grid), float64
        t = theta.detach().cpu().numpy().astype(np.float64)
        modes = []
        for k, p in enumerate(base_modes):
            q = dict(p, mr=float(t[0]), mi=float(t[1]), a1=float(t[2 + 2 * k]), a2=float(t[3 + 2 * k]))
            if grid:
                q.update({"lambda": float(t[-3]), "r1": float(t[-2]), "r2": float(t[-1])})
            modes.append(q)
        adj = lmie.Adjoint(lib, modes, weights, precision)
        out = adj.forward()
        ctx.adj = adj
        ctx.n_modes = len(modes)
        ctx.grid = grid
        return torch.from_numpy(np.ascontiguousarray(out["gc"]))

    @staticmethod
    def backward(ctx, grad_gc):
        # the device state stays with the graph node: a second backward
        # (retain_graph=True) runs another reverse pass on it, and it is
        # released with the node (Adjoint.__del__) when the graph is dropped
        g = ctx.adj.reverse({"gc": grad_gc.detach().cpu().numpy().astype(np.float64)})
        out = [g["mr"], g["mi"]]
        for k in range(ctx.n_modes):
            out += [g["modes"][k][0], g["modes"][k][1]]
        if ctx.grid:
            out += [g["lambda"], g["r1"], g["r2"]]
        return torch.tensor(out, dtype=torch.float64), None, None, None, None, None


def expansion(theta, lib, base_modes, weights=None, precision="fp64", grid=False):
    """Differentiable gc [6][max_coef] as a function of theta =
    [mr, mi, a1_0, a2_0, a1_1, a2_1, ...] for the given base parameter dicts;
    grid=True appends [lambda, r1, r2].  The coefficient count follows lambda
    and r2 (it is fixed near the point of evaluation)."""
    return _Expansion.apply(theta, lib, base_modes, weights, precision, grid)
