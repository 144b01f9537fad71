"""Differentiable angular-spectrum propagation over the C ABI.

The ASM operator A = crop . IFFT . diag(H) . FFT . pad is C-linear; its adjoint
A^H = crop . IFFT . diag(conj H) . FFT . pad is the same fused three-pass kernel
with the conjugate transfer function (olx_asm_c64 with OLX_ADJOINT).  With
PyTorch's complex-gradient convention (grad = dL/dRe + i dL/dIm) the
vector-Jacobian product of a C-linear map is A^H applied to the output gradient.

For a phase-only hologram u = exp(i phi):  dL/dphi = Im(conj(u) * grad_u).
"""
import torch

from . import ops


def _asm_copy(u, dy, dx, lam, z, bandlimit, adjoint):
    """AsmProp of a copy of u; leading dimensions are independent fields."""
    out = u.contiguous().clone()
    ny, nx = out.shape[-2], out.shape[-1]
    for plane in out.reshape(-1, ny, nx):
        ops.AsmProp(plane, ny, nx, dy, dx, lam, z, bandlimit=bandlimit, adjoint=adjoint)
    return out


class AsmPropFn(torch.autograd.Function):
    """out = AsmProp(u); backward applies the adjoint kernel."""

    @staticmethod
    def forward(ctx, u, dy, dx, lam, z, bandlimit):
        ctx.args = (dy, dx, lam, z, bandlimit)
        return _asm_copy(u, dy, dx, lam, z, bandlimit, False)

    @staticmethod
    def backward(ctx, g):
        dy, dx, lam, z, bandlimit = ctx.args
        return _asm_copy(g, dy, dx, lam, z, bandlimit, True), None, None, None, None, None


def asm_prop(u, dy, dx, lam, z, bandlimit=True):
    """Differentiable ASM propagation of a complex64 CUDA field (not in place)."""
    return AsmPropFn.apply(u, dy, dx, lam, z, bandlimit)


def normalized_intensity_mse(field, target, norm="max"):
    """MSE between the normalised intensity |field|^2 and the target.

    norm="max":  mean((I / max I - T)^2), T in [0, 1];
    norm="mean": mean((I / mean I - T / mean T)^2), invariant to the scale of
    either side.  The max form has a degenerate optimum at large fields (one
    bright pixel makes I / max I ~ 0 everywhere else, loss ~ mean T^2); the mean
    form does not, and is the default of the optimiser."""
    inten = field.real ** 2 + field.imag ** 2
    if norm == "max":
        return torch.mean((inten / inten.max() - target) ** 2)
    if norm == "mean":
        return torch.mean((inten / inten.mean() - target / target.mean()) ** 2)
    raise ValueError(f"unknown normalisation {norm!r}")


class MultiPlaneLossFn(torch.autograd.Function):
    """L(phi) = sum_p MSE(I_p / max I_p, T_p) (or the mean-normalised form),
    I_p = |A_p exp(i phi)|^2.

    Memory-lean: no plane field is kept.  The mean-normalised form runs natively
    (ops.multiplane_mse): when phi needs a gradient, forward computes it along with
    the loss, plane by plane (propagate, reduce, form dL/df, adjoint-propagate), and
    backward scales it.  The max form keeps phi; backward recomputes each plane's
    field, takes the intensity-to-loss gradient with torch (max subgradient
    included), applies the adjoint propagation and accumulates Im(conj(u) A_p^H G_p)."""

    @staticmethod
    def forward(ctx, phi, targets, zs, dy, dx, lam, bandlimit, norm, need_grad):
        ctx.cfg = (targets, zs, dy, dx, lam, bandlimit, norm)
        if norm == "mean":
            loss, g = ops.multiplane_mse(phi.detach().contiguous(), list(targets), list(zs), dy,
                                         dx, lam, bandlimit, grad=need_grad)
            ctx.save_for_backward(g)
            return loss.to(phi.dtype)
        ctx.save_for_backward(phi)
        u = torch.polar(torch.ones_like(phi), phi)
        loss = phi.new_zeros(())
        for t, z in zip(targets, zs):
            f = _asm_copy(u, dy, dx, lam, z, bandlimit, False)
            loss = loss + normalized_intensity_mse(f, t, norm)
        return loss

    @staticmethod
    def backward(ctx, gout):
        targets, zs, dy, dx, lam, bandlimit, norm = ctx.cfg
        if norm == "mean":
            (g,) = ctx.saved_tensors
            return g * gout, None, None, None, None, None, None, None, None
        (phi,) = ctx.saved_tensors
        u = torch.polar(torch.ones_like(phi), phi)
        gu = torch.zeros_like(u)
        for t, z in zip(targets, zs):
            f = _asm_copy(u, dy, dx, lam, z, bandlimit, False).requires_grad_(True)
            with torch.enable_grad():
                lp = normalized_intensity_mse(f, t, norm)
                (gf,) = torch.autograd.grad(lp, f)
            gu += _asm_copy(gf, dy, dx, lam, z, bandlimit, True)
        gphi = torch.imag(torch.conj(u) * gu) * gout
        return gphi, None, None, None, None, None, None, None, None


def multiplane_loss(phi, targets, zs, dy, dx, lam, bandlimit=True, norm="max"):
    need_grad = torch.is_grad_enabled() and phi.requires_grad
    return MultiPlaneLossFn.apply(phi, targets, zs, dy, dx, lam, bandlimit, norm, need_grad)


def multiplane_loss_composed(phi, targets, zs, dy, dx, lam, bandlimit=True, norm="max"):
    """The same loss composed from differentiable ASM propagations and torch
    operations (autograd through AsmPropFn): a cross-check of the native path."""
    u = torch.polar(torch.ones_like(phi), phi)
    loss = phi.new_zeros(())
    for t, z in zip(targets, zs):
        loss = loss + normalized_intensity_mse(asm_prop(u, dy, dx, lam, z, bandlimit), t, norm)
    return loss
