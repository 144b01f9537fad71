"""GPU checks of the adjoint-based autograd and the multi-plane optimisation.

* the VJP of AsmPropFn equals the oracle's FP64 adjoint;
* the multi-plane loss gradient equals an FP64 reference gradient built from the
  oracle's complex128 ASM and its adjoint (max-normalisation subgradient included),
  and central finite differences of the FP64 loss on 64 seeded pixels;
* the same for the mean-normalised loss;
* 200 Adam steps reduce the loss (both normalisations).
"""
import numpy as np
import pytest
import torch

import olref
from openol_amd import synth

pytestmark = pytest.mark.gpu
LAM, P = 532e-9, 8e-6


def ag():
    from openol_amd import autograd
    return autograd


def ref_loss_grad(phi, targets, zs, dy=P, dx=P, norm="max"):
    """FP64 loss and dL/dphi (oracle complex128 ASM + adjoint), with the
    normalisation's derivative written out by hand."""
    u = np.exp(1j * phi.astype(np.float64))
    L, gu = 0.0, np.zeros_like(u)
    for t, z in zip(targets, zs):
        f = olref.asm_prop(u, dy, dx, LAM, z, olref.BANDLIMIT)
        inten = np.abs(f) ** 2
        n = inten.size
        if norm == "max":
            k = np.unravel_index(np.argmax(inten), inten.shape)
            m = inten[k]
            r = inten / m - t
            dI = 2 * r / (m * n)
            dI[k] += -np.sum(2 * r * inten / (m * m * n))
        else:
            # r = I/mu - T/tau, mu = mean I: dr_j/dI_k = delta_jk/mu - I_j/(mu^2 n)
            mu = inten.mean()
            r = inten / mu - t / t.mean()
            dI = 2 * r / (mu * n) - 2 * np.sum(r * inten) / (mu * mu * n * n)
        L += np.mean(r ** 2)
        gu += olref.asm_prop(2 * dI * f, dy, dx, LAM, z, olref.BANDLIMIT | olref.ADJOINT)
    return L, np.imag(np.conj(u) * gu)


def test_asm_vjp_is_the_adjoint():
    ny, nx = 64, 128
    u0 = synth.random_field(ny, nx, seed=70)
    g0 = synth.random_field(ny, nx, seed=71)
    u = torch.from_numpy(u0).cuda().requires_grad_(True)
    y = ag().asm_prop(u, 8e-6, 6e-6, LAM, 0.02)
    g = torch.from_numpy(g0).cuda()
    (torch.conj(g) * y).real.sum().backward()
    ref = olref.asm_prop(g0.astype(np.complex128), 8e-6, 6e-6, LAM, 0.02,
                         olref.BANDLIMIT | olref.ADJOINT)
    assert olref.rel_l2(u.grad.cpu().numpy(), ref) < 1e-5
    # forward is the plain propagation
    fwd = olref.asm_prop(u0.astype(np.complex128), 8e-6, 6e-6, LAM, 0.02, olref.BANDLIMIT)
    assert olref.rel_l2(y.detach().cpu().numpy(), fwd) < 1e-5


@pytest.mark.parametrize("n", [64, 2048])
def test_multiplane_gradient_vs_fp64(n):
    zs = [0.10, 0.12, 0.15] if n == 2048 else [0.010, 0.012, 0.015]
    phi0 = synth.uniform(7, synth.S_PHASE, n * n, 0, 2 * np.pi).reshape(n, n).astype(np.float32)
    tg = [synth.gaussian_targets(n, n, 11 + i, smin=max(2.0, n / 100), smax=max(4.0, n / 10))
          for i in range(3)]
    phi = torch.from_numpy(phi0).cuda().requires_grad_(True)
    targets = [torch.from_numpy(t).cuda() for t in tg]
    loss = ag().multiplane_loss(phi, targets, zs, P, P, LAM)
    loss.backward()
    Lr, gr = ref_loss_grad(phi0, [t.astype(np.float64) for t in tg], zs)
    assert abs(loss.item() - Lr) < 1e-5 * Lr
    err = olref.rel_l2(phi.grad.cpu().numpy(), gr)
    print(f"multi-plane gradient {n}^2: rel L2 vs FP64 = {err:.3e}")
    assert err < 1e-4
    if n == 64:
        # central finite differences of the FP64 loss on 64 seeded pixels
        rng = np.random.default_rng(9)
        idx = rng.choice(n * n, 64, replace=False)
        eps = 1e-3
        fd = []
        for k in idx:
            p1 = phi0.astype(np.float64).copy().ravel()
            p2 = p1.copy()
            p1[k] += eps
            p2[k] -= eps
            l1, _ = ref_loss_grad(p1.reshape(n, n), [t.astype(np.float64) for t in tg], zs)
            l2, _ = ref_loss_grad(p2.reshape(n, n), [t.astype(np.float64) for t in tg], zs)
            fd.append((l1 - l2) / (2 * eps))
        fd = np.array(fd)
        got = phi.grad.cpu().numpy().ravel()[idx]
        assert olref.rel_l2(got, fd) < 1e-3, (got[:4], fd[:4])


@pytest.mark.parametrize("n", [64, 512])
def test_multiplane_gradient_mean_norm_vs_fp64(n):
    zs = [0.010, 0.012, 0.015] if n == 64 else [0.05, 0.06, 0.075]
    phi0 = synth.uniform(8, synth.S_PHASE, n * n, 0, 2 * np.pi).reshape(n, n).astype(np.float32)
    tg = [synth.gaussian_targets(n, n, 21 + i, smin=max(2.0, n / 100), smax=max(4.0, n / 10))
          for i in range(3)]
    phi = torch.from_numpy(phi0).cuda().requires_grad_(True)
    targets = [torch.from_numpy(t).cuda() for t in tg]
    loss = ag().multiplane_loss(phi, targets, zs, P, P, LAM, norm="mean")
    loss.backward()
    Lr, gr = ref_loss_grad(phi0, [t.astype(np.float64) for t in tg], zs, norm="mean")
    assert abs(loss.item() - Lr) < 1e-5 * Lr
    assert olref.rel_l2(phi.grad.cpu().numpy(), gr) < 1e-4
    if n == 64:
        rng = np.random.default_rng(10)
        idx = rng.choice(n * n, 32, replace=False)
        eps = 1e-3
        fd = []
        for k in idx:
            p1 = phi0.astype(np.float64).copy().ravel()
            p2 = p1.copy()
            p1[k] += eps
            p2[k] -= eps
            l1, _ = ref_loss_grad(p1.reshape(n, n), [t.astype(np.float64) for t in tg], zs, norm="mean")
            l2, _ = ref_loss_grad(p2.reshape(n, n), [t.astype(np.float64) for t in tg], zs, norm="mean")
            fd.append((l1 - l2) / (2 * eps))
        assert olref.rel_l2(phi.grad.cpu().numpy().ravel()[idx], np.array(fd)) < 1e-3


@pytest.mark.parametrize("shape", [(128, 128), (64, 256), (1024, 1024), (1024, 2048)])
def test_native_multiplane_matches_composed(shape):
    """The native mean-normalised loss and gradient (one pass per plane) against the
    same loss composed from differentiable ASM propagations and torch operations;
    gout scales the gradient; without grad mode the loss alone is computed.  Padded
    sides of 2048 and more take the tiled passes, whose last adjoint pass writes the
    phase gradient itself; smaller fields add it in a separate kernel."""
    ny, nx = shape
    zs = [0.010, 0.013]
    phi0 = synth.uniform(9, synth.S_PHASE, ny * nx, 0, 2 * np.pi).reshape(ny, nx).astype(np.float32)
    targets = [torch.from_numpy(synth.gaussian_targets(ny, nx, 31 + i, smin=3.0, smax=15.0)).cuda()
               for i in range(len(zs))]
    phi = torch.from_numpy(phi0).cuda().requires_grad_(True)
    loss = ag().multiplane_loss(phi, targets, zs, P, P, LAM, norm="mean")
    (2.5 * loss).backward()
    phi2 = torch.from_numpy(phi0).cuda().requires_grad_(True)
    loss2 = ag().multiplane_loss_composed(phi2, targets, zs, P, P, LAM, norm="mean")
    (2.5 * loss2).backward()
    assert abs(loss.item() - loss2.item()) < 1e-5 * loss2.item()
    assert olref.rel_l2(phi.grad.cpu().numpy(), phi2.grad.cpu().numpy()) < 1e-5
    with torch.no_grad():
        l3 = ag().multiplane_loss(phi, targets, zs, P, P, LAM, norm="mean")
    assert l3.item() == loss.item()
    # the composed form with the max normalisation still runs through the same ops
    l4 = ag().multiplane_loss(phi.detach(), targets, zs, P, P, LAM, norm="max")
    l5 = ag().multiplane_loss_composed(phi.detach(), targets, zs, P, P, LAM, norm="max")
    assert abs(l4.item() - l5.item()) < 1e-5 * l5.item()


def test_native_multiplane_rejects_bad_inputs():
    from openol_amd import ops
    from openol_amd._lib import OlxInvalid
    phi = torch.zeros((32, 32), device="cuda")
    t = torch.ones((32, 32), device="cuda")
    with pytest.raises(OlxInvalid):
        ops.multiplane_mse(phi.double(), [t], [0.01], P, P, LAM)
    with pytest.raises(OlxInvalid):
        ops.multiplane_mse(phi, [t, t], [0.01], P, P, LAM)
    with pytest.raises(OlxInvalid):
        ops.multiplane_mse(phi, [t[:16]], [0.01], P, P, LAM)


def test_adam_mean_norm_2048():
    """The optimisation workload's shape: the mean-normalised loss falls by half."""
    from openol_amd import optim
    n, zs = 2048, [0.10, 0.12, 0.15]
    targets = optim.gaussian_targets(n, n, zs, seed=5)
    phi, losses, dt = optim.optimize_phase(targets, zs, steps=60, lr=0.05, norm="mean")
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.6 * losses[0]


def test_adam_reduces_loss():
    from openol_amd import optim
    n = 256
    zs = [0.05, 0.06, 0.075]
    targets = [torch.from_numpy(synth.gaussian_targets(n, n, 3 + i, smin=4, smax=30)).cuda()
               for i in range(3)]
    phi, losses, dt = optim.optimize_phase(targets, zs, steps=200, lr=0.05, norm="max")
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f} in {dt:.2f} s")
    # three competing planes from one phase-only hologram: the loss must fall
    # clearly and keep falling (late average below early average)
    assert np.all(np.isfinite(losses))
    assert losses[-1] < 0.85 * losses[0]
    assert np.mean(losses[-20:]) < np.mean(losses[20:40])
