"""GPU parity: fused ASM / Fresnel / ARSS propagation vs the oracle (through the C ABI).

Tolerances: relative L2 <= 1e-5 against the oracle's fftw3f-path restatement
(complex64 fields, PREC_T = double), the bar stated for the propagation path.
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


@pytest.mark.parametrize("shape", [(8, 8), (16, 64), (256, 128), (1024, 1024), (8192, 16)])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft2d_matches_numpy(shape, sign):
    u0 = synth.random_field(*shape, seed=21)
    u = dev(u0)
    (ol().fft if sign < 0 else ol().ifft)(u)
    ref = np.fft.fft2(u0.astype(np.complex128)) if sign < 0 else \
        np.fft.ifft2(u0.astype(np.complex128)) * (shape[0] * shape[1])
    assert rel(u.cpu().numpy(), ref) < 2e-6


@pytest.mark.parametrize("shape,dy,dx,z", [
    ((8, 8), 6e-6, 8e-6, 0.01), ((8, 4), 6e-6, 8e-6, 0.01), ((4, 16), 8e-6, 8e-6, 0.003),
    ((64, 64), 8e-6, 8e-6, 0.02), ((128, 256), 4e-6, 8e-6, 0.05), ((512, 512), 8e-6, 8e-6, 0.1)])
@pytest.mark.parametrize("flags", [0, olref.BANDLIMIT])
def test_asm_matches_oracle(shape, dy, dx, z, flags):
    u0 = synth.random_field(*shape, seed=22)
    u = dev(u0)
    ol().AsmProp(u, shape[0], shape[1], dy, dx, LAM, z, bandlimit=bool(flags))
    ref = olref.asm_prop(u0, dy, dx, LAM, z, flags)
    assert rel(u.cpu().numpy(), ref) < TOL


def test_asm_bandlimit_active_matches_golden():
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))
    for tag, shp in (("sq", (8, 8)), ("rect", (8, 4))):
        u = dev(g[f"field_{tag}"])
        ol().AsmProp(u, *shp, 1e-6, 1e-6, LAM, 1e-4, bandlimit=True)
        assert rel(u.cpu().numpy(), g[f"asm_bl_{tag}"]) < TOL
        u = dev(g[f"field_{tag}"])
        ol().AsmProp(u, *shp, 6e-6, 8e-6, LAM, 0.01)
        assert rel(u.cpu().numpy(), g[f"asm_{tag}"]) < TOL


def test_asm_1024_full_size():
    """1024^2 field, 8 um pitch, z = 0.10 m, padded to 2048^2, band-limited TF."""
    n = 1024
    u0 = synth.random_field(n, n, seed=1)
    u = dev(u0)
    ol().AsmProp(u, n, n, 8e-6, 8e-6, LAM, 0.10, bandlimit=True)
    got = u.cpu().numpy()
    assert rel(got, olref.asm_prop(u0, 8e-6, 8e-6, LAM, 0.10, olref.BANDLIMIT)) < TOL
    # fftw3 (complex128) path and FP64: the FP32 floor is ~2e-7
    ref64 = olref.asm_prop(u0.astype(np.complex128), 8e-6, 8e-6, LAM, 0.10, olref.BANDLIMIT)
    assert rel(got, ref64) < TOL
    # intensity output
    assert rel(np.abs(got) ** 2, np.abs(ref64) ** 2) < 2 * TOL


def test_asm_adjoint_matches_oracle_and_dot_product():
    ny, nx = 128, 64
    u0 = synth.random_field(ny, nx, seed=23)
    v0 = synth.random_field(ny, nx, seed=24)
    a = dev(u0)
    ol().AsmProp(a, ny, nx, 8e-6, 6e-6, LAM, 0.03, adjoint=True)
    assert rel(a.cpu().numpy(), olref.asm_prop(u0, 8e-6, 6e-6, LAM, 0.03, olref.ADJOINT)) < TOL
    Au = dev(u0)
    ol().AsmProp(Au, ny, nx, 8e-6, 6e-6, LAM, 0.03)
    AHv = dev(v0)
    ol().AsmProp(AHv, ny, nx, 8e-6, 6e-6, LAM, 0.03, adjoint=True)
    lhs = np.vdot(v0.astype(np.complex128), Au.cpu().numpy().astype(np.complex128))
    rhs = np.vdot(AHv.cpu().numpy().astype(np.complex128), u0.astype(np.complex128))
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)


def test_asm_full_size_properties():
    """4096^2 (the metric workload): z = 0 identity, linearity, adjoint dot product."""
    n = 4096
    u0 = synth.random_field(n, n, seed=25)
    u = dev(u0)
    ol().AsmProp(u, n, n, 8e-6, 8e-6, LAM, 0.0, bandlimit=True)
    assert rel(u.cpu().numpy(), u0) < 1e-6
    v0 = synth.random_field(n, n, seed=26)
    a, b = dev(u0), dev(v0)
    c = dev(u0 + 0.5j * v0)
    for t in (a, b, c):
        ol().AsmProp(t, n, n, 8e-6, 8e-6, LAM, 0.10, bandlimit=True)
    assert torch.linalg.norm(a + 0.5j * b - c) / torch.linalg.norm(c) < 1e-6
    vh = dev(v0)
    ol().AsmProp(vh, n, n, 8e-6, 8e-6, LAM, 0.10, bandlimit=True, adjoint=True)
    lhs = torch.vdot(torch.from_numpy(v0).cuda().to(torch.complex128).flatten(),
                     a.to(torch.complex128).flatten())
    rhs = torch.vdot(vh.to(torch.complex128).flatten(),
                     torch.from_numpy(u0).cuda().to(torch.complex128).flatten())
    assert abs((lhs - rhs) / lhs).item() < 1e-5


@pytest.mark.parametrize("n", [2048])
def test_asm_matches_oracle_large(n):
    u0 = synth.random_field(n, n, seed=27)
    u = dev(u0)
    ol().AsmProp(u, n, n, 8e-6, 8e-6, LAM, 0.12, bandlimit=True)
    assert rel(u.cpu().numpy(), olref.asm_prop(u0, 8e-6, 8e-6, LAM, 0.12, olref.BANDLIMIT)) < TOL


@pytest.mark.parametrize("shape,dy,dx,z", [
    ((1024, 2048), 8e-6, 6e-6, 0.08), ((2048, 1024), 6e-6, 8e-6, 0.11), ((8192, 1024), 8e-6, 8e-6, 0.1)])
def test_asm_tiled_rectangular(shape, dy, dx, z):
    """Padded sides >= 2048 run the tiled passes, whose work-group shapes follow each
    axis separately (2 to 16 rows / columns per group)."""
    u0 = synth.random_field(*shape, seed=35)
    for flags, adj in ((olref.BANDLIMIT, False), (olref.BANDLIMIT | olref.ADJOINT, True)):
        u = dev(u0)
        ol().AsmProp(u, *shape, dy, dx, LAM, z, bandlimit=True, adjoint=adj)
        assert rel(u.cpu().numpy(), olref.asm_prop(u0, dy, dx, LAM, z, flags)) < TOL


@pytest.mark.parametrize("shape", [(256, 256), (1024, 1024), (1024, 2048), (2048, 1024),
                                   (2048, 4096), (4096, 4096), (8192, 8192)])
def test_asm_bitwise_equals_shifted_at_zero_offset(shape):
    """On the tiled path ASM evaluates the transfer function once per mirror pair of
    spectral rows (v and -v, shared through LDS; row L/2 on its own); the shifted
    form evaluates every row.  At zero offset its phase, window and scale are the
    same FP64 arithmetic, so the fields agree bit for bit, window on and off, forward
    and adjoint."""
    ny, nx = shape
    g = torch.Generator(device="cuda").manual_seed(ny + 3 * nx)
    u0 = torch.randn(ny, nx, dtype=torch.complex64, device="cuda", generator=g)
    for bandlimit in (False, True):
        for adjoint in (False, True):
            u = u0.clone()
            ol().AsmProp(u, ny, nx, 8e-6, 6e-6, LAM, 0.09, bandlimit=bandlimit, adjoint=adjoint)
            v = torch.empty_like(u0)
            ol().gshiftedAsmProp(u0, v, ny, nx, 8e-6, 6e-6, LAM, 0.09, 0.0, 0.0,
                                 bandlimit=bandlimit, adjoint=adjoint)
            assert torch.equal(u, v), (bandlimit, adjoint)
            assert torch.isfinite(torch.view_as_real(u)).all()


@pytest.mark.parametrize("variant,flags", [(None, 0), ("cpu", olref.FRESNEL_BL_CPU)])
def test_fresnel_tiled(variant, flags):
    shape, d, z = (1024, 2048), 8e-6, 0.3
    u0 = synth.random_field(*shape, seed=36)
    u = dev(u0)
    if variant is None:
        ol().FresnelProp(u, *shape, d, d, LAM, z)
    else:
        ol().FresnelPropBandLimit(u, *shape, d, d, LAM, z, variant=variant)
    assert rel(u.cpu().numpy(), olref.fresnel_prop(u0, d, d, LAM, z, flags)) < TOL


def test_arss_tiled_offsets():
    shape = (2048, 1024)
    u0 = synth.random_field(*shape, seed=37)
    src = dev(u0)
    dst = torch.empty_like(src)
    ol().ARSSFresnelProp(src, dst, *shape, 8e-6, 8e-6, LAM, -0.12, 0.25, 1e-4, -2e-4)
    ref = olref.arss_prop(u0, 8e-6, 8e-6, LAM, -0.12, 0.25, 1e-4, -2e-4)
    assert rel(dst.cpu().numpy(), ref) < TOL


def test_asm_8192_properties():
    """Largest supported field (8192^2, padded 16384^2): z = 0 identity and linearity."""
    n = 8192
    u0 = synth.random_field(n, n, seed=38)
    u = dev(u0)
    ol().AsmProp(u, n, n, 8e-6, 8e-6, LAM, 0.0, bandlimit=True)
    assert rel(u.cpu().numpy(), u0) < 1e-6
    a = dev(u0)
    c = dev(u0 * (0.5 - 1j))
    for t in (a, c):
        ol().AsmProp(t, n, n, 8e-6, 8e-6, LAM, 0.2, bandlimit=True)
    assert (torch.linalg.norm(a * (0.5 - 1j) - c) / torch.linalg.norm(c)).item() < 1e-6


@pytest.mark.parametrize("variant,flags", [(None, 0), ("cpu", olref.FRESNEL_BL_CPU),
                                           ("gpu", olref.FRESNEL_BL_GPU)])
@pytest.mark.parametrize("shape,dy,dx,z", [((8, 8), 6e-6, 8e-6, 0.001), ((8, 4), 6e-6, 8e-6, 0.002),
                                           ((256, 128), 8e-6, 6e-6, 0.05), ((512, 512), 8e-6, 8e-6, 0.2)])
def test_fresnel_matches_oracle(variant, flags, shape, dy, dx, z):
    u0 = synth.random_field(*shape, seed=28)
    u = dev(u0)
    if variant is None:
        ol().FresnelProp(u, *shape, dy, dx, LAM, z)
    else:
        ol().FresnelPropBandLimit(u, *shape, dy, dx, LAM, z, variant=variant)
    ref = olref.fresnel_prop(u0, dy, dx, LAM, z, flags)
    assert rel(u.cpu().numpy(), ref) < TOL


def test_fresnel_adjoint():
    u0 = synth.random_field(64, 128, seed=29)
    u = dev(u0)
    ol().FresnelProp(u, 64, 128, 8e-6, 8e-6, LAM, 0.04, adjoint=True)
    ref = olref.fresnel_prop(u0, 8e-6, 8e-6, LAM, 0.04, olref.ADJOINT)
    assert rel(u.cpu().numpy(), ref) < TOL


@pytest.mark.parametrize("shape,dy,dx,z,s,oy,ox", [
    ((8, 8), 6e-6, 8e-6, 0.002, 0.5, 0.0, 0.0),
    ((8, 4), 12e-6, 16e-6, -0.0005, 0.5, 3e-5, -2e-5),
    ((256, 256), 8e-6, 8e-6, 0.15, 0.5, 0.0, 0.0),
    ((512, 256), 16e-6, 16e-6, -0.1, 0.25, 1e-4, -2e-4),
    ((128, 128), 8e-6, 8e-6, 0.05, 2.0, 0.0, 0.0),
    ((64, 64), 8e-6, 8e-6, 0.01, 1.0, 0.0, 0.0)])
def test_arss_matches_oracle(shape, dy, dx, z, s, oy, ox):
    u0 = synth.random_field(*shape, seed=30)
    src = dev(u0)
    dst = torch.empty_like(src)
    ol().ARSSFresnelProp(src, dst, *shape, dy, dx, LAM, z, s, oy, ox)
    ref = olref.arss_prop(u0, dy, dx, LAM, z, s, oy, ox)
    assert rel(dst.cpu().numpy(), ref) < TOL
    assert torch.equal(src.cpu(), torch.from_numpy(u0))  # source untouched
    # aliasing src == dst (tests/gProp_test.cu:67 calls it in place)
    ol().ARSSFresnelProp(src, src, *shape, dy, dx, LAM, z, s, oy, ox)
    assert rel(src.cpu().numpy(), ref) < TOL


def test_arss_wrp_4096_properties():
    """4096^2 WRP -> ARSS s = 0.5, z = 0.15 m (the WRP workload's shape): linearity and a
    2048^2 oracle comparison at the same optical parameters."""
    n = 2048
    u0 = synth.random_field(n, n, seed=31)
    src = dev(u0)
    ol().ARSSFresnelProp(src, src, n, n, 8e-6, 8e-6, LAM, 0.15, 0.5)
    assert rel(src.cpu().numpy(), olref.arss_prop(u0, 8e-6, 8e-6, LAM, 0.15, 0.5)) < TOL
    n = 4096
    a0 = synth.random_field(n, n, seed=32)
    b0 = synth.random_field(n, n, seed=33)
    a, b, c = dev(a0), dev(b0), dev(a0 - 2j * b0)
    for t in (a, b, c):
        ol().ARSSFresnelProp(t, t, n, n, 8e-6, 8e-6, LAM, 0.15, 0.5)
    assert torch.linalg.norm(a - 2j * b - c) / torch.linalg.norm(c) < 1e-6


def test_prop_dispatcher():
    n, dx = 64, 8e-6
    du = 1 / (dx * 2 * n)
    umax = du * n
    zmax = np.sqrt(1 / LAM ** 2 - 2 * umax ** 2) / (2 * umax * du)
    zmin = 2 * (dx * n) * dx / LAM
    u0 = synth.random_field(n, n, seed=34)
    for z, want in ((0.5 * zmax, 0), (2 * zmin, 1), (0.5 * (zmax + zmin), 3)):
        assert ol().AsmPropCheck(n, n, dx, dx, LAM, z) == olref.asm_prop_check(n, n, dx, dx, LAM, z)
        assert ol().FresnelPropCheck(n, n, dx, dx, LAM, z) == olref.fresnel_prop_check(n, n, dx, dx, LAM, z)
        u = dev(u0)
        ch = ol().Prop(u, n, n, dx, dx, LAM, z)
        assert ch == want
        ref, rch = olref.prop(u0, dx, dx, LAM, z, mode=2)
        assert rel(u.cpu().numpy(), ref) < TOL


def test_invalid_arguments_raise():
    from openol_amd._lib import OlxError, OlxInvalid
    u = dev(synth.random_field(17, 16, seed=1))
    with pytest.raises(OlxError):
        ol().AsmProp(u, 17, 16, 8e-6, 8e-6, LAM, 0.1)  # padded 34 = 2 x 17: prime above 13
    u = dev(synth.random_field(16, 16, seed=1))
    with pytest.raises(OlxInvalid):
        ol().AsmProp(u, 16, 16, 8e-6, 8e-6, -LAM, 0.1)
    with pytest.raises(OlxInvalid):
        ol().FresnelProp(u, 16, 16, 8e-6, 8e-6, LAM, 0.0)
    with pytest.raises(OlxInvalid):
        ol().AsmProp(u, 16, 8, 8e-6, 8e-6, LAM, 0.1)  # shape mismatch
    with pytest.raises(OlxInvalid):  # a real tensor (complex128 runs the FP64 passes)
        ol().AsmProp(u.real.contiguous(), 16, 16, 8e-6, 8e-6, LAM, 0.1)
    with pytest.raises(OlxInvalid):
        ol().Prop(u, 16, 16, 8e-6, 8e-6, LAM, 0.1, propmode=7)


def test_launch_timing_sampling():
    """olx_ctx_enable_timing(ctx, k): every k-th launch is bracketed by HIP events
    (the three passes of a tiled propagation in turn); 0 stops, a negative period is
    rejected."""
    from openol_amd import _lib
    L, c = _lib.load(), _lib.ctx()
    n = 1024
    u = dev(synth.random_field(n, n, seed=40))
    names = ("prop_rows_pad_fwd", "prop_cols_conv", "prop_rows_crop_inv")
    for period, props, per_pass in ((1, 4, 4), (4, 8, 2)):
        L.olx_ctx_reset_timing(c)
        assert L.olx_ctx_enable_timing(c, period) == 0
        for _ in range(props):
            ol().AsmProp(u, n, n, 8e-6, 8e-6, LAM, 0.1, bandlimit=True)
        L.olx_ctx_enable_timing(c, 0)
        rep = _lib.timing_report()
        assert all(rep[k][0] == per_pass and rep[k][1] > 0 for k in names), rep
    assert L.olx_ctx_enable_timing(c, -2) != 0


@pytest.mark.parametrize("mode", ["asm_bl", "asm_bl_adjoint", "asm", "fresnel", "fresnel_bl",
                                  "shifted", "shifted_window", "asm_prec_f32"])
def test_column_pass_8192_all_modes(mode):
    """Fields of 4096 rows (padded columns of 8192: one transform per LDS round, the
    pipelined exchanges, the forward transform with its remainder radix first): every
    transfer-function mode against the oracle, on 4096 x 1024 (fast to check on the
    CPU)."""
    ny, nx, d = 4096, 1024, 8e-6
    print(f"column pass 8192: {mode}", flush=True)
    u0 = synth.random_field(ny, nx, seed=41)
    u = dev(u0)
    if mode.startswith("asm"):
        bl, adj, f32 = "bl" in mode, "adjoint" in mode, "f32" in mode
        ol().AsmProp(u, ny, nx, d, d, LAM, 0.1, bandlimit=bl, adjoint=adj, prec_f32=f32)
        flags = (olref.BANDLIMIT if bl else 0) | (olref.ADJOINT if adj else 0) | \
            (olref.PREC_F32 if f32 else 0)
        ref = olref.asm_prop(u0, d, d, LAM, 0.1, flags)
        got = u
    elif mode.startswith("fresnel"):
        if mode == "fresnel":
            ol().FresnelProp(u, ny, nx, d, d, LAM, 0.4)
            ref = olref.fresnel_prop(u0, d, d, LAM, 0.4, 0)
        else:
            ol().FresnelPropBandLimit(u, ny, nx, d, d, LAM, 0.4)
            ref = olref.fresnel_prop(u0, d, d, LAM, 0.4, olref.FRESNEL_BL_CPU)
        got = u
    else:
        got = torch.empty_like(u)
        win = mode == "shifted_window"
        ol().gshiftedAsmProp(u, got, ny, nx, d, d, LAM, 0.1, 2e-4, -3e-4, bandlimit=win)
        ref = olref.shifted_asm(u0, d, d, LAM, 0.1, 2e-4, -3e-4, olref.SHIFT_WINDOW if win else 0)
    tol = 2e-5 if mode == "asm_prec_f32" else TOL
    assert rel(got.cpu().numpy(), ref) < tol


@pytest.mark.parametrize("variant,flags", [(None, 0), ("gpu", olref.FRESNEL_BL_GPU)])
def test_fresnel_16384_spectrum(variant, flags):
    """Padded columns of 16384: the row spectrum table is the FP64 FFT of the chirp
    split once into two 8192-point halves (8192 above the single-group limit).  8192 x 64
    against the oracle's direct sums."""
    ny, nx, d, z = 8192, 64, 8e-6, 0.05
    u0 = synth.random_field(ny, nx, seed=43)
    u = torch.from_numpy(u0).cuda()
    if variant is None:
        ol().FresnelProp(u, ny, nx, d, d, LAM, z)
    else:
        ol().FresnelPropBandLimit(u, ny, nx, d, d, LAM, z, variant=variant)
    assert rel(u.cpu().numpy(), olref.fresnel_prop(u0, d, d, LAM, z, flags)) < TOL


def test_column_pass_8192_batch():
    """Two stacked 4096-row fields in one call: the column pass takes field blockIdx.y."""
    ny, nx, d = 4096, 1024, 8e-6
    u0 = np.stack([synth.random_field(ny, nx, seed=42 + k) for k in range(2)])
    u = torch.from_numpy(u0).cuda()
    ol().prop_batch(u, d, d, LAM, 0.1, ol().ASM, olref.BANDLIMIT)
    for k in range(2):
        ref = olref.asm_prop(u0[k], d, d, LAM, 0.1, olref.BANDLIMIT)
        assert rel(u[k].cpu().numpy(), ref) < TOL
