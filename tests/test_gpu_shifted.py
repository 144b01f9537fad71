"""Off-axis (shifted) ASM and split propagation on the GPU against the oracle:
every window branch, both scales, the adjoint, rectangular and tiled sizes,
aliasing source and destination, strided blocks accumulated by split
propagation, and at 4096^2 the size-independent property that the split form
sums to the whole-field propagation."""
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


FLAG_CASES = {
    "gpu": (dict(), 0),
    "window": (dict(bandlimit=True), olref.SHIFT_WINDOW),
    "adjoint": (dict(adjoint=True), olref.ADJOINT),
    "window_adjoint": (dict(bandlimit=True, adjoint=True), olref.SHIFT_WINDOW | olref.ADJOINT),
}


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(64, 64), (128, 256), (512, 512), (1024, 1024), (1024, 2048)])
@pytest.mark.parametrize("case", list(FLAG_CASES))
def test_shifted_asm_matches_oracle(shape, case):
    ny, nx = shape
    kw, flags = FLAG_CASES[case]
    u0 = synth.random_field(ny, nx, seed=ny + nx)
    # offsets inside and beyond half the padded width, on both axes and signs
    for oy, ox in ((0.0, 0.0), (0.3 * ny * P, -0.5 * nx * P), (-1.5 * ny * P, 1.25 * nx * P)):
        src = gpu(u0)
        dst = torch.empty_like(src)
        ol.gshiftedAsmProp(src, dst, ny, nx, P, P, LAM, 0.004, oy, ox, **kw)
        ref = olref.shifted_asm(u0, P, P, LAM, 0.004, oy, ox, flags)
        assert olref.rel_l2(dst.cpu().numpy(), ref) < 1e-5, (oy, ox)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [128, 2048])
def test_shifted_asm_cpu_form_in_place(n):
    """ol::shiftedAsmProp: window and net scale 1/M^2; source and destination alias."""
    u0 = synth.random_field(n, n, seed=n)
    u = gpu(u0)
    ol.shiftedAsmProp(u, u, n, n, P, P, LAM, 0.01, -0.2 * n * P, 0.7 * n * P)
    ref = olref.shifted_asm(u0, P, P, LAM, 0.01, -0.2 * n * P, 0.7 * n * P,
                            olref.SHIFT_WINDOW | olref.SHIFT_CPU_SCALE)
    assert olref.rel_l2(u.cpu().numpy(), ref) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("shape,blocks", [((256, 256), (2, 2)), ((512, 256), (4, 2)),
                                          ((256, 512), (1, 4)), ((2048, 2048), (2, 2))])
def test_split_matches_oracle(shape, blocks):
    ny, nx = shape
    by, bx = blocks
    u0 = synth.random_field(ny, nx, seed=7)
    for kind in ("gpu", "cpu", "fresnel"):
        u = gpu(u0)
        if kind == "gpu":
            ol.gsplitProp(u, ny, nx, P, P, LAM, 0.003, ol.ASM, by, bx)
            ref = olref.split_prop(u0, P, P, LAM, 0.003, by, bx, 0, 0)
        elif kind == "cpu":
            ol.splitAsmProp(u, ny, nx, P, P, LAM, 0.003, by, bx)
            ref = olref.split_prop(u0, P, P, LAM, 0.003, by, bx, 0,
                                   olref.SHIFT_WINDOW | olref.SHIFT_CPU_SCALE)
        else:
            ol.gsplitProp(u, ny, nx, P, P, LAM, 0.08, ol.FRESNEL, by, bx)
            ref = olref.split_prop(u0, P, P, LAM, 0.08, by, bx, 1, 0)
        assert olref.rel_l2(u.cpu().numpy(), ref) < 1e-5, kind
        if shape[0] >= 2048:
            break  # the oracle's windowed and Fresnel forms are covered at the smaller sizes


@pytest.mark.gpu
def test_split_4096_sums_to_whole_field():
    """4096^2 as 2 x 2 blocks of 2048^2 (the tiled kernels, strided blocks summed in
    pass C): equals one whole-field AsmProp for sources inside the field."""
    n = 4096
    yy, xx = np.mgrid[0:n, 0:n]
    u0 = np.zeros((n, n), np.complex64)
    rng = np.random.default_rng(3)
    for cy, cx in rng.integers(600, n - 600, size=(12, 2)):
        u0 += (np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * 6.0 ** 2))
               * np.exp(1j * rng.uniform(0, 6))).astype(np.complex64)
    a = gpu(u0)
    ol.gsplitProp(a, n, n, P, P, LAM, 0.01, ol.ASM, 2, 2)
    b = gpu(u0)
    ol.AsmProp(b, n, n, P, P, LAM, 0.01)
    assert olref.rel_l2(a.cpu().numpy(), b.cpu().numpy()) < 1e-5


@pytest.mark.gpu
def test_split_errors():
    u = gpu(synth.random_field(64, 64, seed=1))
    with pytest.raises(Exception):
        ol.gsplitProp(u, 64, 64, P, P, LAM, 0.01, ol.ASM, 3, 2)     # 3 does not divide 64
    with pytest.raises(Exception):
        ol.gsplitProp(u, 64, 64, P, P, LAM, 0.01, ol.AUTO, 2, 2)
    with pytest.raises(Exception):
        ol.gsplitProp(u, 64, 64, P, P, LAM, 0.01, ol.ASM, 64, 64)   # 1 x 1 blocks: too small
