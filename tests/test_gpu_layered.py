"""GPU parity of the layered Fresnel CGH on MFMA (through the C ABI).

Compared on the SAME layer-snapped cloud, so depth-binning error is separated
from arithmetic error: relative L2 <= 1e-3 against the FP64 oracle for FP16 and
split-BF16 factors; plain BF16 is reported (its 8-bit mantissa gives ~2e-3).
"""
import numpy as np
import pytest
import torch

import olref
from openol_amd import synth

pytestmark = pytest.mark.gpu
LAM = 532e-9
P = 8e-6


def ol():
    import openol_amd
    return openol_amd


def cloud(n, seed, ny, nx, zmin, zmax):
    pts = synth.random_points(n, seed=seed, xmax=nx * P, ymax=ny * P, zmin=zmin, zmax=zmax)
    for k in ("amp", "phase"):
        pts[k] = pts[k].astype(np.float32).astype(np.float64)
    return pts


def snapped(pts, zmin, zmax, nl):
    return np.array([ol().layer_depth(z, zmin, zmax, nl) for z in pts["z"]])


@pytest.mark.parametrize("prec,tol", [("f16", 1e-3), ("bf16x3", 1e-3), ("bf16", 6e-3)])
def test_layered_matches_fp64_on_snapped_cloud(prec, tol):
    ny, nx, n, nl = 128, 256, 700, 8
    zmin, zmax = 0.10, 0.20
    pts = cloud(n, 60, ny, nx, zmin, zmax)
    obj = ol().Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"])
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ol().CghLayered(obj, U, ny, nx, P, P, LAM, nl, zmin, zmax, precision=prec)
    zl = snapped(pts, zmin, zmax, nl)
    ref = olref.cgh(pts["x"], pts["y"], zl, ny, nx, P, P, LAM, pts["amp"], pts["phase"],
                    dtype=np.complex128)
    err = olref.rel_l2(U.cpu().numpy(), ref)
    print(f"layered {prec}: rel L2 vs FP64 (snapped) = {err:.3e}")
    assert err < tol
    # the direct FP32 kernel on the snapped cloud agrees with the same oracle
    objs = ol().Object(pts["x"], pts["y"], zl, pts["amp"], pts["phase"])
    D = torch.empty_like(U)
    ol().Cgh(objs, D, ny, nx, P, P, LAM)
    assert olref.rel_l2(D.cpu().numpy(), ref) < 1e-4


def test_layered_chunking_accumulate_and_empty():
    # more points than one factor chunk (32768), and a point count that is not a
    # multiple of the 32-point K step
    ny, nx, n, nl = 128, 128, 40013, 16
    zmin, zmax = 0.10, 0.20
    pts = cloud(n, 61, ny, nx, zmin, zmax)
    obj = ol().Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"])
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ol().CghLayered(obj, U, ny, nx, P, P, LAM, nl, zmin, zmax, precision="f16")
    zl = snapped(pts, zmin, zmax, nl)
    rng = np.random.default_rng(3)
    ph, pw = rng.integers(0, ny, 256), rng.integers(0, nx, 256)
    ref = olref.cgh_pixels(pts["x"], pts["y"], zl, ph, pw, P, P, LAM, pts["amp"], pts["phase"])
    got = U.cpu().numpy()[ph, pw]
    assert olref.rel_l2(got, ref) < 1e-3
    # accumulate: two halves add up to the whole
    a = ol().Object(pts["x"][:20000], pts["y"][:20000], pts["z"][:20000], pts["amp"][:20000], pts["phase"][:20000])
    b = ol().Object(pts["x"][20000:], pts["y"][20000:], pts["z"][20000:], pts["amp"][20000:], pts["phase"][20000:])
    V = torch.empty_like(U)
    ol().CghLayered(a, V, ny, nx, P, P, LAM, nl, zmin, zmax, precision="f16")
    ol().CghLayered(b, V, ny, nx, P, P, LAM, nl, zmin, zmax, precision="f16", accumulate=True)
    assert olref.rel_l2(V.cpu().numpy(), U.cpu().numpy()) < 1e-5
    e = ol().Object(np.zeros(0), np.zeros(0), np.zeros(0))
    W = torch.ones_like(U)
    ol().CghLayered(e, W, ny, nx, P, P, LAM, nl, zmin, zmax)
    assert torch.count_nonzero(W).item() == 0


@pytest.mark.parametrize("prec", ["f16", "bf16x3"])
@pytest.mark.parametrize("shape", [(256, 128), (512, 384)])
def test_layered_wide_tile(prec, shape):
    """Sides that are multiples of 256 x 128 run the 256 x 128 work-group tile.  Its
    sums run in the 128 x 128 kernel's K order, so the two agree bit for bit (more
    points than one factor chunk, accumulate on); the oracle bounds the error."""
    ny, nx = shape
    n, nl, zmin, zmax = 33000, 8, 0.10, 0.20
    pts = cloud(n, 62, ny, nx, zmin, zmax)
    obj = ol().Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"])
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ol().CghLayered(obj, U, ny, nx, P, P, LAM, nl, zmin, zmax, precision=prec)
    V = torch.empty_like(U)
    ol().CghLayered(obj, V, ny, nx, P, P, LAM, nl, zmin, zmax, precision=prec, tile128=True)
    assert torch.equal(U, V)
    zl = snapped(pts, zmin, zmax, nl)
    rng = np.random.default_rng(4)
    ph, pw = rng.integers(0, ny, 256), rng.integers(0, nx, 256)
    ref = olref.cgh_pixels(pts["x"], pts["y"], zl, ph, pw, P, P, LAM, pts["amp"], pts["phase"])
    assert olref.rel_l2(U.cpu().numpy()[ph, pw], ref) < 1e-3


@pytest.mark.parametrize("prec", ["f16", "bf16x3"])
@pytest.mark.parametrize("shape", [(100, 128), (1080, 200), (90, 333), (2160, 136)])
def test_layered_edge_tiles(prec, shape):
    """Sides that are not multiples of the work-group tile (the reference's 1080 and
    2160 rows among them): the GEMM covers whole tiles and stores the pixels inside
    the hologram only.  Sampled pixels, the last row and column included, against
    the FP64 oracle on the snapped cloud; the 128 x 128 tiles agree bit for bit."""
    ny, nx = shape
    n, nl, zmin, zmax = 3000, 8, 0.10, 0.20
    pts = cloud(n, 63, ny, nx, zmin, zmax)
    obj = ol().Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"])
    U = torch.empty((ny, nx), dtype=torch.complex64, device="cuda")
    ol().CghLayered(obj, U, ny, nx, P, P, LAM, nl, zmin, zmax, precision=prec)
    V = torch.empty_like(U)
    ol().CghLayered(obj, V, ny, nx, P, P, LAM, nl, zmin, zmax, precision=prec, tile128=True)
    assert torch.equal(U, V)
    zl = snapped(pts, zmin, zmax, nl)
    rng = np.random.default_rng(6)
    ph = np.concatenate([rng.integers(0, ny, 300), np.full(20, ny - 1), rng.integers(0, ny, 20)])
    pw = np.concatenate([rng.integers(0, nx, 300), rng.integers(0, nx, 20), np.full(20, nx - 1)])
    ref = olref.cgh_pixels(pts["x"], pts["y"], zl, ph, pw, P, P, LAM, pts["amp"], pts["phase"])
    assert olref.rel_l2(U.cpu().numpy()[ph, pw], ref) < 1e-3


def test_layered_edge_tiles_stay_inside():
    """Every store stays inside the hologram: a 100 x 130 output in the middle of a
    larger buffer leaves the buffer around it untouched."""
    ny, nx = 100, 130
    pts = cloud(500, 64, ny, nx, 0.10, 0.20)
    obj = ol().Object(pts["x"], pts["y"], pts["z"], pts["amp"], pts["phase"])
    buf = torch.full((ny * nx + 4096,), 3 - 5j, dtype=torch.complex64, device="cuda")
    U = buf[2048:2048 + ny * nx].view(ny, nx)
    ol().CghLayered(obj, U, ny, nx, P, P, LAM, 8, 0.10, 0.20)
    assert torch.all(buf[:2048] == 3 - 5j) and torch.all(buf[2048 + ny * nx:] == 3 - 5j)


def test_layered_rejects_unsupported_shapes():
    from openol_amd._lib import OlxError
    obj = ol().Object([1e-4], [1e-4], [0.15])
    U = torch.empty((100, 8200), dtype=torch.complex64, device="cuda")
    with pytest.raises(OlxError):
        ol().CghLayered(obj, U, 100, 8200, P, P, LAM, 8, 0.1, 0.2)
