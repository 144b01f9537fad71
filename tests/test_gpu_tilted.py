"""Tilted-plane propagation on the GPU against the FP64 oracle (both forms, both
samplers, ASM and Fresnel, rectangular fields, angles of both signs), and at
4096^2 the size-independent property that angle 0 with bilinear sampling is the
plain propagation times 1/M^2."""
import numpy as np
import pytest
import torch

import olref
import openol_amd as ol
from openol_amd import synth

LAM, P = 532e-9, 8e-6
DEV = torch.device("cuda", 0)


def gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(64, 64), (128, 256), (1024, 1024)])
@pytest.mark.parametrize("kind", ["source", "reference"])
@pytest.mark.parametrize("interp", ["bilinear", "bicubic"])
@pytest.mark.parametrize("mode", ["asm", "fresnel"])
def test_tilted_matches_oracle(shape, kind, interp, mode):
    ny, nx = shape
    u0 = synth.random_field(ny, nx, seed=ny + 3 * nx)
    z = 0.004 if mode == "asm" else 0.12
    fn = ol.TiltedSourceProp if kind == "source" else ol.TiltedReferenceProp
    for deg in (3.0, -1.5):
        u = gpu(u0)
        fn(u, ny, nx, P, P, LAM, z, np.deg2rad(deg), ol.BILINEAR if interp == "bilinear" else ol.BICUBIC,
           ol.ASM if mode == "asm" else ol.FRESNEL)
        ref = olref.tilted(u0, P, P, LAM, z, np.deg2rad(deg), interp, mode, kind)
        assert olref.rel_l2(u.cpu().numpy(), ref) < 1e-5, deg


@pytest.mark.gpu
def test_tilted_4096_zero_angle_is_plain_asm():
    n = 4096
    u0 = synth.random_field(n, n, seed=9)
    yy, xx = np.mgrid[0:n, 0:n]
    u0 = (u0 * np.exp(-((xx - n / 2) ** 2 + (yy - n / 2) ** 2) / (2 * (n / 6) ** 2))).astype(np.complex64)
    a = gpu(u0)
    ol.TiltedReferenceProp(a, n, n, P, P, LAM, 0.05, 0.0)
    b = gpu(u0)
    ol.AsmProp(b, n, n, P, P, LAM, 0.05)
    M2 = (2 * n) ** 2
    assert olref.rel_l2(a.cpu().numpy() * M2, b.cpu().numpy()) < 1e-5


@pytest.mark.gpu
def test_tilted_errors():
    u = gpu(synth.random_field(64, 64, seed=1))
    with pytest.raises(Exception):
        ol.TiltedSourceProp(u, 64, 64, P, P, LAM, 0.01, 0.1, ol.BILINEAR, ol.AUTO)
    with pytest.raises(Exception):
        ol.TiltedSourceProp(u, 64, 64, P, P, LAM, 0.01, float("nan"))
