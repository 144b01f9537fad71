"""Optical elements on the GPU against the oracle: Diffuser::random_phase (the
std::mt19937 stream applied on the device), the GPU diffuser kernels given the
same uniforms, the Philox uniform stream, and the square / band-limited /
circular lens arrays (gen and mult), at 64^2 to 4096^2.  Float tolerances: the
reference takes cos / sin of float phases with the float library functions, the
device rounds FP64 values to float (within an ulp of each other)."""
import numpy as np
import pytest
import torch

import olref
import openol_amd as ol

LAM = 532e-9
DEV = torch.device("cuda", 0)


def gpu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def field(ny, nx, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex64)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,step,oy,ox", [((64, 64), ol.PLAY, 0, 0), ((96, 200), ol.RECORD, 5, -7),
                                              ((1024, 1024), ol.PLAY, -300, 17)])
def test_random_phase_matches_oracle(shape, step, oy, ox):
    ny, nx = shape
    u0 = field(ny, nx, 1)
    d = ol.Diffuser(ny, nx)
    u = gpu(u0)
    d.random_phase(u, step, seed=42, range=float(np.float32(2 * np.pi)), oy=oy, ox=ox)
    ref = olref.diffuser_random_phase(u0, step, 42, float(np.float32(2 * np.pi)), oy, ox)
    assert np.abs(u.cpu().numpy() - ref).max() <= 4e-7 * np.abs(u0).max()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["phase", "xphase", "yphase", "bitrandom", "block", "blockstep", "bitphase"])
def test_diffuser_kernels_match_oracle(kind):
    ny, nx, bs = 256, 384, 8
    u0 = field(ny, nx, 2)
    b = ol.uniform(ny * nx, seed=5, device=0)
    u = gpu(u0)
    d = ol.Diffuser(ny, nx)
    d._apply(u, kind, b, 1.7, bs)
    ref = olref.diffuser_apply(u0, b.cpu().numpy(), 1.7, kind, bs)
    assert np.abs(u.cpu().numpy() - ref).max() <= 1e-6 * np.abs(u0).max()


@pytest.mark.gpu
def test_uniform_stream():
    a = ol.uniform(1_000_003, seed=9, device=0).cpu().numpy()
    b = ol.uniform(1_000_003, seed=9, device=0).cpu().numpy()
    c = ol.uniform(1_000_003, seed=10, device=0).cpu().numpy()
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() > 0 and a.max() <= 1
    assert abs(a.mean() - 0.5) < 2e-3 and abs(a.var() - 1 / 12) < 2e-3
    h, _ = np.histogram(a, bins=64, range=(0, 1))
    assert h.min() > 0.9 * len(a) / 64


@pytest.mark.gpu
@pytest.mark.parametrize("cls", ["square", "circle"])
@pytest.mark.parametrize("shape", [(64, 64), (200, 136), (4096, 4096)])
def test_lens_arrays_match_oracle(cls, shape):
    ny, nx = shape
    pitch, bw, f = 8e-6, 8e-6 * 20, 0.03
    L = (ol.LensarryCircle if cls == "circle" else ol.Lensarry)(ny, nx, pitch, bw, f, LAM)
    for step, oy, ox, pl in ((ol.PLAY, 0, 0, 0.0), (ol.RECORD, 3, -4, 0.0), (ol.PLAY, 1, 2, 50e-6)):
        got = L.gen_complex_limit(pl, step, oy, ox, device=0) if pl else L.gen_complex(step, oy, ox, device=0)
        ref = olref.lens_array(ny, nx, pitch, bw, f, LAM, step, oy, ox, pl, cls == "circle")
        g = got.cpu().numpy()
        assert np.array_equal(g == 0, ref == 0) and np.array_equal(g == 1, ref == 1)
        assert np.abs(g - ref).max() <= 4e-7
        if shape[0] <= 256:
            u0 = field(ny, nx, 3)
            u = gpu(u0)
            if pl:
                L.mult_complex_limit(u, pl, step, oy, ox)
            else:
                L.mult_complex(u, step, oy, ox)
            assert np.abs(u.cpu().numpy() - u0 * ref).max() <= 1e-6 * np.abs(u0).max()


@pytest.mark.gpu
def test_optics_errors():
    u = gpu(field(64, 64, 4))
    with pytest.raises(Exception):
        ol.Diffuser(64, 64).random_blockphase(u, ol.PLAY, blocksize=5)
    with pytest.raises(Exception):
        ol.Lensarry(64, 64, 8e-6, 1e-7, 0.01, LAM).gen_complex(device=0)
