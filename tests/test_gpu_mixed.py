"""GPU parity of the mixed-radix path: field sides that are not powers of two.

The reference pads any ny x nx field to 2ny x 2nx and transforms it with FFTW /
cuFFT (include/olutils.h:19-66, include/golutils.h:14-53), so its own demo sizes
(1920 x 1080 and 3840 x 2160, tests/para.h:12-13, 41-42) run through mixed-radix
transforms of 3840 x 2160 and 7680 x 4320 (2^8 3 5 and 2^4 3^3 5; 2^9 3 5 and
2^5 3^3 5).  Tolerance: relative L2 <= 1e-5 against the oracle's complex64
restatement (PREC_T = double), as for the power-of-two path; the 2-D FFT is
compared with numpy's FP64 pocketfft.
"""
import numpy as np
import pytest
import torch

import olref
from openol_amd import synth

pytestmark = pytest.mark.gpu
LAM = 532e-9
TOL = 1e-5


def ol():
    import openol_amd
    return openol_amd


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.complex64)).cuda()


def rel(a, b):
    return olref.rel_l2(a, b)


@pytest.mark.parametrize("shape", [(12, 18), (30, 45), (7, 13), (2, 3), (1080, 1920),
                                   (4320, 30), (6, 16384), (15360, 4), (121, 169),
                                   (154, 1001), (81, 675)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft2d_mixed_matches_numpy(shape, sign):
    u0 = synth.random_field(*shape, seed=70)
    u = dev(u0)
    (ol().fft if sign < 0 else ol().ifft)(u)
    ref = np.fft.fft2(u0.astype(np.complex128)) if sign < 0 else \
        np.fft.ifft2(u0.astype(np.complex128)) * (shape[0] * shape[1])
    assert rel(u.cpu().numpy(), ref) < 3e-6, shape


@pytest.mark.parametrize("shape,d,z", [((1080, 1920), 8e-6, 0.1), ((135, 250), 8e-6, 0.02),
                                       ((9, 14), 6e-6, 0.001), ((5, 7), 8e-6, 0.0005),
                                       ((2160, 3840), 3.74e-6, 0.1), ((96, 1000), 1e-5, 0.05)])
@pytest.mark.parametrize("bl", [False, True])
def test_asm_mixed_matches_oracle(shape, d, z, bl):
    u0 = synth.random_field(*shape, seed=71)
    u = dev(u0)
    ol().AsmProp(u, *shape, d, d, LAM, z, bandlimit=bl)
    ref = olref.asm_prop(u0, d, d, LAM, z, olref.BANDLIMIT if bl else 0)
    assert rel(u.cpu().numpy(), ref) < TOL


def test_asm_mixed_adjoint_and_tiny_band():
    """Adjoint at a mixed size, and a distance whose band limit cuts the spectrum."""
    shape = (270, 480)
    u0 = synth.random_field(*shape, seed=72)
    u = dev(u0)
    ol().AsmProp(u, *shape, 8e-6, 8e-6, LAM, 0.3, adjoint=True)
    assert rel(u.cpu().numpy(), olref.asm_prop(u0, 8e-6, 8e-6, LAM, 0.3, olref.ADJOINT)) < TOL
    u = dev(u0)
    ol().AsmProp(u, *shape, 2e-6, 2e-6, LAM, 1.5, bandlimit=True)
    assert rel(u.cpu().numpy(), olref.asm_prop(u0, 2e-6, 2e-6, LAM, 1.5, olref.BANDLIMIT)) < TOL


@pytest.mark.parametrize("variant,flags", [(None, 0), ("cpu", olref.FRESNEL_BL_CPU),
                                           ("gpu", olref.FRESNEL_BL_GPU)])
@pytest.mark.parametrize("shape,d,z", [((1080, 1920), 8e-6, 0.5), ((150, 90), 6e-6, 0.02)])
def test_fresnel_mixed_matches_oracle(variant, flags, shape, d, z):
    u0 = synth.random_field(*shape, seed=73)
    u = dev(u0)
    if variant is None:
        ol().FresnelProp(u, *shape, d, d, LAM, z)
    else:
        ol().FresnelPropBandLimit(u, *shape, d, d, LAM, z, variant=variant)
    assert rel(u.cpu().numpy(), olref.fresnel_prop(u0, d, d, LAM, z, flags)) < TOL


def test_fresnel_4k_reference_call():
    """tests/Prop_test.cpp:46's call: FresnelProp(u, 2160, 3840, 3.74e-6, 3.74e-6,
    532e-9, 0.1) on a 3840 x 2160 field (padded 7680 x 4320)."""
    shape = (2160, 3840)
    u0 = synth.random_field(*shape, seed=74)
    u = dev(u0)
    ol().FresnelProp(u, *shape, 3.74e-6, 3.74e-6, LAM, 0.1)
    assert rel(u.cpu().numpy(), olref.fresnel_prop(u0, 3.74e-6, 3.74e-6, LAM, 0.1)) < TOL


@pytest.mark.parametrize("shape,z,s,oy,ox", [((1080, 1920), 0.15, 0.5, 0.0, 0.0),
                                             ((90, 150), -0.02, 0.25, 1e-4, -2e-4)])
def test_arss_mixed_matches_oracle(shape, z, s, oy, ox):
    u0 = synth.random_field(*shape, seed=75)
    src = dev(u0)
    dst = torch.empty_like(src)
    ol().ARSSFresnelProp(src, dst, *shape, 8e-6, 8e-6, LAM, z, s, oy, ox)
    assert rel(dst.cpu().numpy(), olref.arss_prop(u0, 8e-6, 8e-6, LAM, z, s, oy, ox)) < TOL


def test_shifted_asm_mixed_matches_oracle():
    shape = (180, 300)
    u0 = synth.random_field(*shape, seed=76)
    src = dev(u0)
    dst = torch.empty_like(src)
    ol().shiftedAsmProp(src, dst, *shape, 8e-6, 8e-6, LAM, 0.05, 2e-4, -1e-4)
    ref = olref.shifted_asm(u0, 8e-6, 8e-6, LAM, 0.05, 2e-4, -1e-4,
                            olref.SHIFT_WINDOW | olref.SHIFT_CPU_SCALE)
    assert rel(dst.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("shape,d,z", [((1080, 1920), 8e-6, 0.1), ((1080, 1920), 8e-6, 2.0),
                                       ((1080, 1920), 8e-6, 0.004)])
def test_prop_dispatcher_mixed(shape, d, z):
    u0 = synth.random_field(*shape, seed=77)
    u = dev(u0)
    chosen = ol().Prop(u, *shape, d, d, LAM, z)
    ref, ch = olref.prop(u0, d, d, LAM, z)
    # the library reports the band-limited Fresnel as FRESNEL_BL, the oracle as 2
    assert {ol().ASM: 0, ol().FRESNEL: 1, ol().FRESNEL_BL: 2}[chosen] == ch
    assert rel(u.cpu().numpy(), ref) < TOL


def test_unsupported_prime_is_rejected():
    """A side whose padded length has a prime factor above 13 (2 x 17) is refused."""
    u = torch.zeros((17, 64), dtype=torch.complex64, device="cuda")
    from openol_amd import _lib
    with pytest.raises(_lib.OlxError) as e:
        ol().AsmProp(u, 17, 64, 8e-6, 8e-6, LAM, 0.01)
    assert e.value.code == _lib.EUNSUPPORTED
