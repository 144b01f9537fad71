"""GPU checks of the fused 8-bit phase epilogues: olx_cgh_img8_c64 (the CGH kernel
writes phase8bit of its phase, include/olimg.h:205-213) and the d_img output of
olx_prop_batch_c64 (the propagation's last pass does).

The byte of a fused epilogue is the byte of the two-pass chain (phase written as
float, then olx_phase8bit), bit for bit: both take atan2 of the same float pair and
apply the same conversion.  Against the oracle (FP64 sums, FP64 phase, the
reference's phase8bit) the bytes agree except where the phase sits within rounding
of a byte boundary: those differ by one step, modulo 256.
"""
import numpy as np
import pytest
import torch

import olref
import openol_amd as ol
from openol_amd import _lib, synth

pytestmark = pytest.mark.gpu
LAM, P = 532e-9, 8e-6


def cloud(n, ny, nx, seed):
    pts = synth.random_points(n, seed=seed, xmax=nx * P, ymax=ny * P, zmin=0.1, zmax=0.2)
    return pts, ol.Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"], device=0)


def byte_steps(a, b):
    d = (a.astype(np.int16) - b.astype(np.int16)) % 256
    return np.minimum(d, 256 - d)


@pytest.mark.parametrize("kind", ["fresnel", "exact", "sliced", "prec_f32"])
@pytest.mark.parametrize("offset", [False, True])
def test_cgh_img8_equals_two_pass_chain(kind, offset):
    ny, nx = (96, 160) if kind != "sliced" else (64, 64)
    n = 300 if kind != "sliced" else 8192  # few tiles, many points: sliced sums
    print(f"cgh img8 {kind} offset={offset}", flush=True)
    pts, obj = cloud(n, ny, nx, 60)
    exact, f32 = kind == "exact", kind == "prec_f32"
    phase = torch.empty((ny, nx), dtype=torch.float32, device="cuda")
    ol.Cgh(obj, phase, ny, nx, P, P, LAM, exact_r=exact, prec_f32=f32)
    two = ol.phase8bit(phase, offset=offset).cpu().numpy()
    img = torch.empty((ny, nx), dtype=torch.uint8, device="cuda")
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ol.CghPhase8(obj, img, ny, nx, P, P, LAM, out=U, exact_r=exact, offset=offset, prec_f32=f32)
    fused = img.cpu().numpy()
    assert np.array_equal(fused, two)
    # the field stored beside the bytes is the plain CGH's
    U2 = torch.empty_like(U)
    ol.Cgh(obj, U2, ny, nx, P, P, LAM, exact_r=exact, prec_f32=f32)
    assert torch.equal(U, U2)
    if kind in ("fresnel", "exact"):
        ref = olref.cgh(pts["x"], pts["y"], pts["z"], ny, nx, P, P, LAM, pts["amp"], pts["phase"],
                        flags=olref.CGH_EXACT_R if exact else 0, dtype=np.complex128)
        rb = olref.phase8bit(np.angle(ref), offset=offset)
        steps = byte_steps(fused, rb)
        assert steps.max() <= 1
        assert np.mean(steps > 0) < 0.01


def test_cgh_img8_accumulate_and_errors():
    ny, nx = 64, 96
    _, obj = cloud(200, ny, nx, 61)
    base = torch.zeros((ny, nx), dtype=torch.complex64, device="cuda")
    ol.Cgh(obj, base, ny, nx, P, P, LAM)
    U = base.clone()
    img = torch.empty((ny, nx), dtype=torch.uint8, device="cuda")
    ol.CghPhase8(obj, img, ny, nx, P, P, LAM, out=U, accumulate=True)
    ph, _ = ol.encode(U)
    assert np.array_equal(img.cpu().numpy(), ol.phase8bit(ph).cpu().numpy())
    assert torch.allclose(U, 2 * base, rtol=1e-5, atol=1e-4)
    with pytest.raises(ValueError):
        ol.CghPhase8(obj, None, ny, nx, P, P, LAM)
    with pytest.raises(_lib.OlxError):  # accumulate needs the field
        ol.CghPhase8(obj, img, ny, nx, P, P, LAM, accumulate=True)


@pytest.mark.parametrize("ny,nx,nb", [(64, 64, 2), (1024, 1024, 1), (90, 150, 2)])
@pytest.mark.parametrize("offset", [False, True])
def test_prop_img8_equals_two_pass_chain(ny, nx, nb, offset):
    print(f"prop img8 {nb} x {ny} x {nx} offset={offset}", flush=True)
    u0 = np.stack([synth.random_field(ny, nx, seed=70 + k) for k in range(nb)])
    u = torch.from_numpy(u0).cuda()
    img = torch.empty((nb, ny, nx), dtype=torch.uint8, device="cuda")
    ol.prop_batch(u, P, P, LAM, 0.02, ol.ASM, _lib.BANDLIMIT, img=img, img_offset=offset)
    v = torch.from_numpy(u0).cuda()
    ol.prop_batch(v, P, P, LAM, 0.02, ol.ASM, _lib.BANDLIMIT)
    assert torch.equal(u, v)  # the field is unchanged by the extra output
    ph, _ = ol.encode(u)
    assert np.array_equal(img.cpu().numpy(), ol.phase8bit(ph, offset=offset).cpu().numpy())
    for k in range(nb):
        ref = olref.asm_prop(u0[k].astype(np.complex128), P, P, LAM, 0.02, olref.BANDLIMIT)
        steps = byte_steps(img[k].cpu().numpy(), olref.phase8bit(np.angle(ref), offset=offset))
        assert steps.max() <= 1 and np.mean(steps > 0) < 0.01
