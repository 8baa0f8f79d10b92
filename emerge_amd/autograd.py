"""Differentiable assembly: A = K(mu_r^-1) - k0^2 M(eps_r) as a ``torch.autograd.Function``.

Forward runs the fused assembly kernel; backward runs the hand-written HIP backward
kernel (``em_ned2_backward``) which returns dL/d eps_r and dL/d mu_r^-1 per tet from the
upstream gradient on the COO values. Complex leaves follow PyTorch's convention
(grad = dL/dRe + i dL/dIm); real leaves receive the real part.

Leaves are tet-major: ``(T,)`` scalar per tet, ``(T, 3)`` diagonal or ``(T, 3, 3)`` full.
The gradient is computed in FP64 for either forward precision.
"""
from __future__ import annotations

import torch


class AssembleA(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps_r, mu_r_inv, plan, k0, precision="fp64"):
        er = plan.material(eps_r.detach(), layout="tet")
        ur = plan.material(mu_r_inv.detach(), inverse=True, layout="tet")
        A = plan.assemble_a(er, ur, float(k0), precision)
        ctx.plan = plan
        ctx.k0 = float(k0)
        ctx.kinds = (er.kind, ur.kind)
        ctx.leaf = ((eps_r.shape, eps_r.dtype), (mu_r_inv.shape, mu_r_inv.dtype))
        return A

    @staticmethod
    def backward(ctx, gA):
        plan = ctx.plan
        g = gA.to(torch.complex128).contiguous()
        outs = []
        for need, (shape, dtype) in zip(ctx.needs_input_grad[:2], ctx.leaf):
            outs.append(torch.zeros(shape, dtype=torch.complex128, device=plan.device)
                        if need else None)
        plan.backward(g, ctx.k0, ctx.kinds[0], outs[0], ctx.kinds[1], outs[1])
        res = []
        for o, (shape, dtype) in zip(outs, ctx.leaf):
            if o is not None and not dtype.is_complex:
                o = o.real.to(dtype)
            elif o is not None:
                o = o.to(dtype)
            res.append(o)
        return res[0], res[1], None, None, None


def assemble_a(plan, eps_r: torch.Tensor, mu_r_inv: torch.Tensor, k0: float,
               precision: str = "fp64") -> torch.Tensor:
    """COO values of A (differentiable in ``eps_r`` and ``mu_r_inv``)."""
    if eps_r.dim() == 0 or mu_r_inv.dim() == 0:
        raise ValueError("leaves must be per tet: (T,), (T,3) or (T,3,3)")
    return AssembleA.apply(eps_r, mu_r_inv, plan, k0, precision)


def bilinear_form(A: torch.Tensor, rows: torch.Tensor, cols: torch.Tensor, lam: torch.Tensor,
                  x: torch.Tensor) -> torch.Tensor:
    """L = Re(lam^H A x) for A given as COO values with (rows, cols)."""
    return torch.real(torch.sum(torch.conj(lam[rows]) * A * x[cols]))

