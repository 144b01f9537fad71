"""GPU checks of the complex128 (FP64) propagation and FFT passes (olx_asm_c128,
olx_fresnel_c128, olx_prop_c128, olx_fft2d_c128): the reference's fftw3 path for
std::complex<double> fields.

Against the FP64 oracle (its complex128 restatement) and numpy's FP64 FFT the
results agree to rounding: relative L2 below 1e-12 for the FFT and Fresnel, 1e-9 for
ASM, whose phase of up to 1e5 cycles carries about 1e-11 cycle of FP64 rounding in
each implementation (the complex64 path's floor is about 2e-7).  Power-of-two and mixed-radix sides, the band-limit windows, the
adjoint, and the configuration-1 field (1024 x 1024) in complex128.
"""
import numpy as np
import pytest
import torch

import olref
import openol_amd as ol
from openol_amd import _lib, synth

pytestmark = pytest.mark.gpu
LAM, P = 532e-9, 8e-6
TOL = 1e-12
TOL_ASM = 1e-9


def field(ny, nx, seed):
    return synth.random_field(ny, nx, seed=seed, dtype=np.complex128)


@pytest.mark.parametrize("shape", [(8, 8), (64, 96), (90, 150), (256, 128), (15, 26), (1024, 1024)])
@pytest.mark.parametrize("bl,adj", [(False, False), (True, False), (True, True)])
def test_asm_c128_matches_fp64_oracle(shape, bl, adj):
    ny, nx = shape
    print(f"asm c128 {shape} bl={bl} adj={adj}", flush=True)
    u0 = field(ny, nx, 500 + ny)
    u = torch.from_numpy(u0).cuda()
    z = 0.05 if ny > 64 else 0.004
    ol.AsmProp(u, ny, nx, P, P, LAM, z, bandlimit=bl, adjoint=adj)
    flags = (olref.BANDLIMIT if bl else 0) | (olref.ADJOINT if adj else 0)
    ref = olref.asm_prop(u0, P, P, LAM, z, flags)
    assert olref.rel_l2(u.cpu().numpy(), ref) < TOL_ASM


@pytest.mark.parametrize("shape", [(16, 16), (90, 150), (256, 256), (2048, 64)])
@pytest.mark.parametrize("variant", [None, "cpu", "gpu"])
def test_fresnel_c128_matches_fp64_oracle(shape, variant):
    ny, nx = shape
    print(f"fresnel c128 {shape} {variant}", flush=True)
    u0 = field(ny, nx, 600 + nx)
    z = 0.3 if variant is None else 0.002
    u = torch.from_numpy(u0).cuda()
    if variant is None:
        ol.FresnelProp(u, ny, nx, P, 6e-6, LAM, z)
        flags = 0
    else:
        ol.FresnelPropBandLimit(u, ny, nx, P, 6e-6, LAM, z, variant=variant)
        flags = olref.FRESNEL_BL_CPU if variant == "cpu" else olref.FRESNEL_BL_GPU
    ref = olref.fresnel_prop(u0, P, 6e-6, LAM, z, flags)
    assert olref.rel_l2(u.cpu().numpy(), ref) < TOL


def test_fresnel_c128_adjoint_dot_product():
    ny, nx = 64, 80
    a, b = field(ny, nx, 70), field(ny, nx, 71)
    ua = torch.from_numpy(a).cuda()
    ol.FresnelProp(ua, ny, nx, P, P, LAM, 0.1)
    vb = torch.from_numpy(b).cuda()
    ol.FresnelProp(vb, ny, nx, P, P, LAM, 0.1, adjoint=True)
    lhs = np.vdot(b, ua.cpu().numpy())
    rhs = np.vdot(vb.cpu().numpy(), a)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12


def test_prop_and_masm_c128():
    ny, nx = 128, 128
    u0 = field(ny, nx, 72)
    for d in (0.05, 2.0):
        u = torch.from_numpy(u0).cuda()
        chosen = ol.Prop(u, ny, nx, P, P, LAM, d)
        ref, ch = olref.prop(u0.astype(np.complex64), P, P, LAM, d)
        assert chosen == {0: ol.ASM, 1: ol.FRESNEL, 2: ol.FRESNEL_BL}[ch]
        r64 = olref.asm_prop(u0, P, P, LAM, d) if ch == 0 else olref.fresnel_prop(
            u0, P, P, LAM, d, 0 if ch == 1 else olref.FRESNEL_BL_CPU)
        assert olref.rel_l2(u.cpu().numpy(), r64) < (TOL_ASM if ch == 0 else TOL)
    u = torch.from_numpy(u0).cuda()
    ol.mAsmProp(u, ny, nx, P, P, LAM, 0.05)
    ref = olref.asm_prop(u0, P, P, LAM, 0.05) / (4 * ny * nx)
    assert olref.rel_l2(u.cpu().numpy(), ref) < TOL_ASM


# 154 = 2 * 7 * 11, 1001 = 7 * 11 * 13 and 4004 = 4 * 7 * 11 * 13 put radix 7, 11 and 13 in
# stages after the first, where their twiddles come from the per-stage tables
@pytest.mark.parametrize("shape", [(8, 8), (90, 150), (1024, 512), (7, 13), (8192, 16),
                                   (154, 1001), (4004, 8)])
def test_fft_c128_matches_numpy(shape):
    u0 = field(*shape, 73)
    u = torch.from_numpy(u0).cuda()
    ol.fft(u)
    assert olref.rel_l2(u.cpu().numpy(), np.fft.fft2(u0)) < TOL
    ol.ifft(u)
    assert olref.rel_l2(u.cpu().numpy(), u0 * u0.size) < TOL


def test_c128_size_limits():
    u = torch.zeros((8192, 16), dtype=torch.complex128, device="cuda")
    with pytest.raises(_lib.OlxError):  # padded 16384 rows: beyond the FP64 passes
        ol.AsmProp(u, 8192, 16, P, P, LAM, 0.01)
    with pytest.raises(_lib.OlxError):  # 578 = 2 x 17^2: a prime above 13
        ol.fft(torch.zeros((578, 16), dtype=torch.complex128, device="cuda"))
    with pytest.raises(ValueError):
        ol.AsmProp(torch.zeros((16, 16), dtype=torch.complex128, device="cuda"), 16, 16, P, P,
                   LAM, 0.01, prec_f32=True)
