"""Multi-plane phase-only hologram optimisation with the adjoint ASM kernel.

A phase map phi (float32, ny x nx) is propagated with the angular-spectrum method
to several planes; the loss is the sum over planes of the MSE between the
normalised intensity and the normalised target (mean normalisation by default,
max normalisation on request).  Adam updates phi; gradients come from
MultiPlaneLossFn: for the mean normalisation one native pass per plane
(propagate, reduce, dL/df, hand-written adjoint propagation), for the max
normalisation a backward that recomputes each plane's field.
"""
import time

import numpy as np
import torch

from . import synth
from .autograd import multiplane_loss


def gaussian_targets(ny, nx, zs, seed=5, device="cuda"):
    """One seeded target per plane: a normalised sum of 32 random Gaussians."""
    return [torch.from_numpy(synth.gaussian_targets(ny, nx, seed + 101 * i)).to(device)
            for i in range(len(zs))]


def optimize_phase(targets, zs, dy=synth.PITCH, dx=synth.PITCH, lam=synth.LAMBDA, steps=200,
                   lr=0.05, seed=0, phi0=None, bandlimit=True, norm="mean", callback=None):
    """Run `steps` Adam iterations; returns (phi, losses, seconds)."""
    ny, nx = targets[0].shape
    dev = targets[0].device
    if phi0 is None:
        phi0 = torch.from_numpy(
            synth.uniform(seed, synth.S_PHASE, ny * nx, 0.0, 2 * np.pi).astype(np.float32)
            .reshape(ny, nx))
    phi = phi0.to(dev).clone().requires_grad_(True)
    # fused: the whole Adam update in one kernel on the device
    opt = torch.optim.Adam([phi], lr=lr, fused=phi.is_cuda)
    losses = []
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for it in range(steps):
        opt.zero_grad(set_to_none=True)
        loss = multiplane_loss(phi, targets, zs, dy, dx, lam, bandlimit, norm)
        loss.backward()
        opt.step()
        losses.append(loss.detach())
        if callback is not None:
            callback(it, loss)
    torch.cuda.synchronize(dev)
    dt = time.perf_counter() - t0
    return phi.detach(), [float(l) for l in losses], dt


def _mean_norm_loss_delta(F, dF, T):
    """L(F + dF) - L(F) for L = sum_p mean((I_p / mean I_p - T_p / mean T_p)^2), in FP64
    and without cancellation: I' - I = 2 Re(conj F dF) + |dF|^2 enters directly."""
    total = torch.zeros((), dtype=torch.float64, device=F[0].device)
    for Fp, dFp, Tp in zip(F, dF, T):
        I = Fp.real ** 2 + Fp.imag ** 2
        dI = 2 * (Fp.real * dFp.real + Fp.imag * dFp.imag) + dFp.real ** 2 + dFp.imag ** 2
        m, dm = I.mean(), dI.mean()
        t = Tp / Tp.mean()
        a0 = I / m - t
        da = dI / (m + dm) - I * dm / (m * (m + dm))  # (I + dI) / (m + dm) - I / m
        total = total + torch.mean(da * (2 * a0 + da))
    return total


def central_fd_gradient(phi, targets, zs, pixels, dy=synth.PITCH, dx=synth.PITCH,
                        lam=synth.LAMBDA, h=1e-3, bandlimit=True):
    """Central finite differences of the mean-normalised multi-plane loss at the flat
    pixel indices `pixels`.  A phase step at pixel k changes the field by
    u_k (e^{+-ih} - 1) times column k of each plane's propagation operator (the
    propagation of a unit impulse at k), so each difference is evaluated exactly, in
    FP64, from the propagated fields without recomputing them."""
    from . import ops
    ny, nx = phi.shape
    u = torch.exp(1j * phi.double()).to(torch.complex64)
    F = []
    for z in zs:
        f = u.clone()
        ops.AsmProp(f, ny, nx, dy, dx, lam, z, bandlimit=bandlimit)
        F.append(f.to(torch.complex128))
    Td = [t.double() for t in targets]
    phv = phi.reshape(-1).double().cpu().numpy()
    fd = []
    for k in pixels:
        d = torch.zeros((ny, nx), dtype=torch.complex64, device=phi.device)
        d.view(-1)[int(k)] = 1.0
        resp = []
        for z in zs:
            r = d.clone()
            ops.AsmProp(r, ny, nx, dy, dx, lam, z, bandlimit=bandlimit)
            resp.append(r.to(torch.complex128))
        uk = np.exp(1j * phv[int(k)])
        dl = []
        for s in (1.0, -1.0):
            c = complex(uk * (np.exp(1j * s * h) - 1.0))
            dl.append(_mean_norm_loss_delta(F, [c * r for r in resp], Td).item())
        fd.append((dl[0] - dl[1]) / (2 * h))
    return np.array(fd)

