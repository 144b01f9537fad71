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
