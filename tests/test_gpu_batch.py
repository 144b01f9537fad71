"""GPU checks of olx_prop_batch_c64 (include/olx.h): nb stacked fields propagated in
one call, each with a shared factor field multiplied in before (pre) and after
(post) the propagation.

* Without factors, field k of a batch is bit-identical to a one-field call on it
  (the same kernels on the same data: the batch only changes which rows a work-group
  reads).
* With factors, every field matches post * oracle(pre * u_k) within the one-field
  bar (relative L2 1e-5 against the complex64 oracle), on the untiled power-of-two
  path, the tiled path (padded sides >= 2048) and the mixed-radix path.
* A lens array fused in as pre agrees with the separate multiply then propagation
  that the reference runs (Lensarry::mult_complex then AsmProp) within 1e-6.
"""
import numpy as np
import pytest
import torch

import olref
import openol_amd as ol
from openol_amd import _lib, synth

pytestmark = pytest.mark.gpu
LAM, P = 532e-9, 8e-6


def stack(nb, ny, nx, seed):
    return np.stack([synth.random_field(ny, nx, seed=seed + k) for k in range(nb)])


def factor(ny, nx, seed):
    # unit-modulus phase mask with an amplitude taper: exercises both parts of the product
    rng = np.random.default_rng(seed)
    amp = 0.5 + 0.5 * rng.random((ny, nx))
    return (amp * np.exp(1j * 2 * np.pi * rng.random((ny, nx)))).astype(np.complex64)


@pytest.mark.parametrize("ny,nx,nb", [(64, 64, 3), (128, 256, 2), (1024, 1024, 3), (90, 150, 3),
                                      (1080, 64, 2)])
@pytest.mark.parametrize("mode", ["asm", "fresnel"])
def test_batch_matches_single_calls(ny, nx, nb, mode):
    print(f"batch {mode} {nb} x {ny} x {nx}", flush=True)
    u0 = stack(nb, ny, nx, 200)
    z = 0.02
    if mode == "asm":
        flags, pm = _lib.BANDLIMIT, ol.ASM
    else:
        flags, pm = _lib.FRESNEL_BL_CPU, ol.FRESNEL
    u = torch.from_numpy(u0).cuda()
    ol.prop_batch(u, P, P, LAM, z, pm, flags)
    got = u.cpu().numpy()
    for k in range(nb):
        v = torch.from_numpy(u0[k].copy()).cuda()
        if mode == "asm":
            ol.AsmProp(v, ny, nx, P, P, LAM, z, bandlimit=True)
        else:
            ol.FresnelPropBandLimit(v, ny, nx, P, P, LAM, z)
        assert np.array_equal(got[k], v.cpu().numpy()), k


@pytest.mark.parametrize("ny,nx,nb", [(64, 64, 3), (1024, 1024, 2), (90, 150, 3), (256, 128, 1)])
@pytest.mark.parametrize("mode", ["asm", "fresnel"])
def test_batch_factors_vs_oracle(ny, nx, nb, mode):
    print(f"batch factors {mode} {nb} x {ny} x {nx}", flush=True)
    u0 = stack(nb, ny, nx, 300)
    pre, post = factor(ny, nx, 7), factor(ny, nx, 8)
    z = 0.03
    u = torch.from_numpy(u0).cuda()
    if mode == "asm":
        ol.prop_batch(u, P, P, LAM, z, ol.ASM, _lib.BANDLIMIT, torch.from_numpy(pre).cuda(),
                      torch.from_numpy(post).cuda())
    else:
        ol.prop_batch(u, P, P, LAM, z, ol.FRESNEL, 0, torch.from_numpy(pre).cuda(),
                      torch.from_numpy(post).cuda())
    got = u.cpu().numpy()
    for k in range(nb):
        src = (u0[k].astype(np.complex128) * pre).astype(np.complex64)
        if mode == "asm":
            ref = olref.asm_prop(src, P, P, LAM, z, olref.BANDLIMIT)
        else:
            ref = olref.fresnel_prop(src, P, P, LAM, z, 0)
        ref = ref.astype(np.complex128) * post
        err = olref.rel_l2(got[k], ref)
        assert err < 1e-5, (k, err)


def test_batch_pre_only_and_post_only():
    ny, nx, nb = 128, 128, 2
    u0 = stack(nb, ny, nx, 400)
    f = factor(ny, nx, 9)
    for which in ("pre", "post"):
        u = torch.from_numpy(u0).cuda()
        kw = {which: torch.from_numpy(f).cuda()}
        ol.prop_batch(u, P, P, LAM, 0.01, ol.ASM, 0, **kw)
        got = u.cpu().numpy()
        for k in range(nb):
            src = (u0[k] * f).astype(np.complex64) if which == "pre" else u0[k]
            ref = olref.asm_prop(src, P, P, LAM, 0.01, 0).astype(np.complex128)
            if which == "post":
                ref = ref * f
            assert olref.rel_l2(got[k], ref) < 1e-5, (which, k)


def test_fused_lens_equals_separate_multiply():
    """Lensarry::mult_complex then AsmProp (include/Lensarry.h:262-270,
    include/olprop.h:227-248) as two passes against the lens fused into pass A."""
    ny = nx = 2048
    u0 = synth.random_field(ny, nx, seed=500)
    L = ol.Lensarry(ny, nx, P, 16 * P, 0.01, LAM)
    lens = L.gen_complex(ol.PLAY)
    sep = torch.from_numpy(u0).cuda()
    ol.mult(sep, lens, sep)
    ol.AsmProp(sep, ny, nx, P, P, LAM, 0.05, bandlimit=True)
    fused = torch.from_numpy(u0).cuda().view(1, ny, nx)
    ol.prop_batch(fused, P, P, LAM, 0.05, ol.ASM, _lib.BANDLIMIT, pre=lens)
    err = olref.rel_l2(fused[0].cpu().numpy(), sep.cpu().numpy().astype(np.complex128))
    assert err < 1e-6, err


def test_batch_errors():
    u = torch.zeros((2, 64, 64), dtype=torch.complex64, device="cuda")
    with pytest.raises(ValueError):
        ol.prop_batch(u, P, P, LAM, 0.01, ol.AUTO)
    with pytest.raises(ValueError):
        ol.prop_batch(u, P, P, LAM, 0.01, ol.ASM, pre=torch.zeros((64, 32), dtype=torch.complex64,
                                                                       device="cuda"))
    with pytest.raises(_lib.OlxError):  # 3 x 8192^2: a 3 GiB intermediate
        big = torch.empty((3, 8192, 8192), dtype=torch.complex64, device="cuda")
        ol.prop_batch(big, P, P, LAM, 0.01, ol.ASM)
    with pytest.raises(_lib.OlxError):  # 17: a padded side 34 = 2 * 17
        ol.prop_batch(torch.zeros((2, 17, 16), dtype=torch.complex64, device="cuda"), P, P, LAM,
                      0.01, ol.ASM)
