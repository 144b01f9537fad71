"""Pins for the oracle (CPU restatement).  CPU-only, runs in the build container.

Every oracle function is checked against something it does not share code with:
the brute-force FP64 fixtures in tests/golden/golden.npz (explicit DFT matrices,
direct circular convolutions, direct sums), closed forms from wave optics
(paraxial Gaussian beam, scaled Fresnel direct sum, Gauss-sum and Pythagorean
CGH values), and identities (Parseval, inverse, adjoint dot products, cross-
function relations).
"""
import os

import numpy as np
import pytest

import olref
from openol_amd import synth

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden.npz"))
LAM = 532e-9


def rel(a, b):
    return olref.rel_l2(a, b)


# --------------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(8, 16), (16, 16), (6, 10), (32, 4), (12, 18), (30, 15),
                                   (27, 20), (7, 11), (26, 13), (49, 24)])
def test_fft_matches_dft_matrices(shape):
    ny, nx = shape
    u = synth.random_field(ny, nx, seed=3, dtype=np.complex128)
    Fy = np.exp(-2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    Fx = np.exp(-2j * np.pi * np.outer(np.arange(nx), np.arange(nx)) / nx)
    ref = Fy @ u @ Fx.T
    assert rel(olref.fft2(u, -1), ref) < 1e-13
    assert rel(olref.fft2(u.astype(np.complex64), -1), ref) < 2e-6
    refb = np.conj(Fy) @ u @ np.conj(Fx).T
    assert rel(olref.fft2(u, +1), refb) < 1e-13


@pytest.mark.parametrize("shape", [(270, 3840), (2160, 96), (4320, 30), (64, 7680), (1250, 486),
                                   (1001, 17)])
def test_fft_mixed_radix_matches_pocketfft(shape):
    """Non-power-of-two lengths (the reference's 1920 x 1080 and 3840 x 2160 demo
    fields pad to 3840 x 2160 and 7680 x 4320: 2^8 3 5, 2^4 3^3 5, 2^9 3 5, 2^5 3^3 5)
    and prime factors up to 17, against numpy's pocketfft, an independent mixed-radix
    implementation."""
    u = synth.random_field(*shape, seed=5, dtype=np.complex128)
    assert rel(olref.fft2(u, -1), np.fft.fft2(u)) < 1e-13
    assert rel(olref.fft2(u, +1), np.fft.ifft2(u) * u.size) < 1e-13
    assert rel(olref.fft2(u.astype(np.complex64), -1), np.fft.fft2(u)) < 3e-6


def test_fft_parseval_and_inverse():
    u = synth.random_field(64, 128, seed=4, dtype=np.complex128)
    U = olref.fft2(u, -1)
    assert abs(np.vdot(U, U).real / (64 * 128) - np.vdot(u, u).real) < 1e-10 * np.vdot(u, u).real
    back = olref.fft2(U, +1) / (64 * 128)
    assert rel(back, u) < 1e-14
    # single modes: delta -> constant, plane wave -> delta
    d = np.zeros((16, 16), np.complex128)
    d[0, 0] = 1
    assert np.allclose(olref.fft2(d), 1)
    k = (3, 5)
    pw = np.exp(2j * np.pi * (k[0] * np.arange(16)[:, None] + k[1] * np.arange(16)[None, :]) / 16)
    P = olref.fft2(pw)
    assert abs(P[k] - 256) < 1e-10 and np.sum(np.abs(P) > 1e-9) == 1


# ------------------------------------------------------------------- goldens
@pytest.mark.parametrize("tag", ["sq", "rect"])
def test_asm_golden(tag):
    u = GOLD[f"field_{tag}"]
    out = olref.asm_prop(u, 6e-6, 8e-6, LAM, 0.01)
    assert rel(out, GOLD[f"asm_{tag}"]) < 1e-12
    out32 = olref.asm_prop(u.astype(np.complex64), 6e-6, 8e-6, LAM, 0.01)
    assert rel(out32, GOLD[f"asm_{tag}"]) < 2e-6
    bl = olref.asm_prop(u, 1e-6, 1e-6, LAM, 1e-4, olref.BANDLIMIT)
    assert rel(bl, GOLD[f"asm_bl_{tag}"]) < 1e-12
    # the window is active there: without it the result differs
    assert rel(olref.asm_prop(u, 1e-6, 1e-6, LAM, 1e-4), GOLD[f"asm_bl_{tag}"]) > 1e-2


@pytest.mark.parametrize("tag", ["sq", "rect"])
def test_fresnel_golden(tag):
    u = GOLD[f"field_{tag}"]
    assert rel(olref.fresnel_prop(u, 6e-6, 8e-6, LAM, 0.002), GOLD[f"fresnel_{tag}"]) < 1e-12
    assert rel(olref.fresnel_prop(u, 6e-6, 8e-6, LAM, 0.001, olref.FRESNEL_BL_CPU),
               GOLD[f"fresnel_blcpu_{tag}"]) < 1e-12
    assert rel(olref.fresnel_prop(u, 6e-6, 8e-6, LAM, 0.001, olref.FRESNEL_BL_GPU),
               GOLD[f"fresnel_blgpu_{tag}"]) < 1e-12
    # the two windows differ when dy != dx
    assert rel(GOLD[f"fresnel_blcpu_{tag}"], GOLD[f"fresnel_blgpu_{tag}"]) > 1e-3
    out32 = olref.fresnel_prop(u.astype(np.complex64), 6e-6, 8e-6, LAM, 0.002)
    assert rel(out32, GOLD[f"fresnel_{tag}"]) < 2e-6


@pytest.mark.parametrize("tag", ["sq", "rect"])
def test_arss_golden(tag):
    u = GOLD[f"field_{tag}"]
    assert rel(olref.arss_prop(u, 6e-6, 8e-6, LAM, 0.002, 0.5), GOLD[f"arss_open_{tag}"]) < 1e-12
    assert rel(olref.arss_prop(u, 12e-6, 16e-6, LAM, -0.0005, 0.5, 3e-5, -2e-5),
               GOLD[f"arss_win_{tag}"]) < 1e-12
    out32 = olref.arss_prop(u.astype(np.complex64), 6e-6, 8e-6, LAM, 0.002, 0.5)
    assert rel(out32, GOLD[f"arss_open_{tag}"]) < 2e-6


def test_cgh_golden():
    g = {k: GOLD[f"cgh_{k}"] for k in ("x", "y", "z", "amp", "phase")}
    plain = olref.cgh(g["x"], g["y"], g["z"], 12, 16, 6e-6, 8e-6, LAM, dtype=np.complex128)
    assert rel(plain, GOLD["cgh_plain"]) < 1e-12
    plain32 = olref.cgh(g["x"], g["y"], g["z"], 12, 16, 6e-6, 8e-6, LAM)
    assert rel(plain32, GOLD["cgh_plain"]) < 1e-6
    ext = olref.cgh(g["x"], g["y"], g["z"], 12, 16, 6e-6, 8e-6, LAM, g["amp"], g["phase"],
                    dtype=np.complex128)
    assert rel(ext, GOLD["cgh_ext"]) < 1e-12
    ex = olref.cgh(g["x"], g["y"], g["z"], 12, 16, 6e-6, 8e-6, LAM, g["amp"], g["phase"],
                   flags=olref.CGH_EXACT_R, dtype=np.complex128)
    assert rel(ex, GOLD["cgh_exact"]) < 1e-9
    ph, pw = np.array([0, 5, 11]), np.array([3, 15, 0])
    pix = olref.cgh_pixels(g["x"], g["y"], g["z"], ph, pw, 6e-6, 8e-6, LAM, g["amp"], g["phase"])
    assert rel(pix, GOLD["cgh_ext"][ph, pw]) < 1e-12


def test_wrp_golden():
    g = {k: GOLD[f"wrp_{k}"] for k in ("x", "y", "z")}
    for sq, key in ((False, "wrp_circle"), (True, "wrp_square")):
        out = olref.wrp(g["x"], g["y"], g["z"], 24, 20, 8e-6, LAM, 0.15, square=sq,
                        dtype=np.complex128)
        assert rel(out, GOLD[key]) < 1e-12
        out32 = olref.wrp(g["x"], g["y"], g["z"], 24, 20, 8e-6, LAM, 0.15, square=sq)
        assert rel(out32, GOLD[key]) < 1e-6


def test_wrp_pixels_match_the_golden_plane():
    """The per-pixel WRP oracle (used at configuration size) gives the golden
    brute-force plane's values at every pixel, for both window shapes."""
    g = {k: GOLD[f"wrp_{k}"] for k in ("x", "y", "z")}
    hh, ww = np.meshgrid(np.arange(24), np.arange(20), indexing="ij")
    for sq, key in ((False, "wrp_circle"), (True, "wrp_square")):
        out = olref.wrp_pixels(g["x"], g["y"], g["z"], hh.ravel(), ww.ravel(), 8e-6, LAM, 0.15,
                               square=sq)
        assert rel(out.reshape(24, 20), GOLD[key]) < 1e-12


# ------------------------------------------------------------- closed forms
def test_asm_zero_distance_is_identity():
    u = synth.random_field(32, 48, seed=5, dtype=np.complex128)
    assert rel(olref.asm_prop(u, 8e-6, 8e-6, LAM, 0.0), u) < 1e-14


def _gauss_beam(n, dx, w0, z):
    """Paraxial Gaussian beam with the e^{+ikz} convention:
    U = q0/(z+q0) exp(ikz) exp(ik rho^2 / (2 (z+q0))), q0 = -i zR."""
    k = 2 * np.pi / LAM
    zr = np.pi * w0 ** 2 / LAM
    q0 = -1j * zr
    c = (np.arange(n) - n // 2) * dx
    r2 = c[:, None] ** 2 + c[None, :] ** 2
    return q0 / (z + q0) * np.exp(1j * k * z) * np.exp(1j * k * r2 / (2 * (z + q0)))


def test_asm_gaussian_beam_closed_form():
    n, dx, w0, z = 64, 8e-6, 6 * 8e-6, 0.005
    u0 = _gauss_beam(n, dx, w0, 0.0)
    out = olref.asm_prop(u0, dx, dx, LAM, z)
    # non-paraxial correction 2 pi z lambda^3 f^4 / 8 ~ 1e-6 rad at the beam's band
    assert rel(out, _gauss_beam(n, dx, w0, z)) < 2e-5
    # a sign or 2*pi slip in H breaks it
    assert rel(olref.asm_prop(u0, dx, dx, LAM, -z), _gauss_beam(n, dx, w0, z)) > 0.1


def test_fresnel_gaussian_beam_closed_form():
    n, dx, w0, z = 256, 8e-6, 4 * 8e-6, 0.05
    u0 = _gauss_beam(n, dx, w0, 0.0)
    out = olref.fresnel_prop(u0, dx, dx, LAM, z)
    M2 = (2 * n) ** 2
    k = 2 * np.pi / LAM
    # FresnelProp = circconv/M^2 of the bare chirp; restore dx dy e^{ikz}/(i lambda z)
    est = out * M2 * dx * dx * np.exp(1j * k * z) / (1j * LAM * z)
    assert rel(est, _gauss_beam(n, dx, w0, z)) < 1e-6


def test_arss_equals_scaled_fresnel_direct_sum():
    n, dx, z, s, oy, ox = 16, 8e-6, 0.01, 0.5, 1e-5, -2e-5
    u = synth.random_field(n, n, seed=11, dtype=np.complex128)
    out = olref.arss_prop(u, dx, dx, LAM, z, s, oy, ox)
    c = (np.arange(n) - n // 2) * dx
    # destination x2 + o, source s * x1; all windows are inactive at these values
    ex = np.exp(1j * np.pi * ((c[:, None] + ox) - s * c[None, :]) ** 2 / (LAM * z))
    ey = np.exp(1j * np.pi * ((c[:, None] + oy) - s * c[None, :]) ** 2 / (LAM * z))
    direct = ey @ u @ ex.T / (2 * n) ** 2
    assert rel(out, direct) < 1e-12


def test_shifted_fresnel_equals_scaled_direct_sum_over_m4():
    # CPU shiftedFresnelProp has no windows and net scale circconv / M^4
    n, dx, z, s, oy, ox = 16, 8e-6, 0.01, 0.5, 1e-5, -2e-5
    u = synth.random_field(n, n, seed=19, dtype=np.complex128)
    out = olref.shifted_fresnel_prop(u, dx, dx, LAM, z, s, oy, ox)
    c = (np.arange(n) - n // 2) * dx
    ex = np.exp(1j * np.pi * ((c[:, None] + ox) - s * c[None, :]) ** 2 / (LAM * z))
    ey = np.exp(1j * np.pi * ((c[:, None] + oy) - s * c[None, :]) ** 2 / (LAM * z))
    direct = ey @ u @ ex.T / (2 * n) ** 4
    assert rel(out, direct) < 1e-12
    # and at a distance where the GPU windows would cut, it keeps everything
    w = olref.shifted_fresnel_prop(u, dx, dx, LAM, 0.0005, s)
    a = olref.arss_prop(u, dx, dx, LAM, 0.0005, s) / (2 * n) ** 2
    assert rel(w, a) > 1e-3


def test_arss_unit_scale_equals_gpu_bandlimited_fresnel():
    u = synth.random_field(16, 16, seed=12, dtype=np.complex128)
    a = olref.arss_prop(u, 8e-6, 8e-6, LAM, 0.003, 1.0)
    f = olref.fresnel_prop(u, 8e-6, 8e-6, LAM, 0.003, olref.FRESNEL_BL_GPU)
    assert rel(a, f) < 1e-12


def test_cgh_gauss_sum_closed_form():
    # pi p^2/(lambda z) = pi/2  =>  U = i^((w-w0)^2 + (h-h0)^2) for one point on a pixel
    p = 8e-6
    z = 2 * p * p / LAM
    out = olref.cgh(np.array([5 * p]), np.array([3 * p]), np.array([z]), 10, 12, p, p, LAM,
                    dtype=np.complex128)
    hh, ww = np.meshgrid(np.arange(10), np.arange(12), indexing="ij")
    expect = 1j ** (((ww - 5) ** 2 + (hh - 3) ** 2) % 4)
    assert np.max(np.abs(out - expect)) < 1e-9


def test_cgh_exact_r_pythagorean():
    # p = 30 lambda, z = 40 lambda  =>  r = 40 lambda at the foot, 50 lambda one pixel away
    p, z = 30 * LAM, 40 * LAM
    A, phi = 0.7, 0.3
    out = olref.cgh(np.array([4 * p]), np.array([2 * p]), np.array([z]), 5, 9, p, p, LAM,
                    amp=np.array([A]), phase=np.array([phi]), flags=olref.CGH_EXACT_R,
                    dtype=np.complex128)
    want = A * np.exp(1j * phi)
    for h, w in [(2, 4), (2, 3), (2, 5), (1, 4), (3, 4)]:
        assert abs(out[h, w] - want) < 1e-9
    assert abs(out[1, 3] - want) > 1e-3


def test_cgh_normalized_is_the_cgh_of_the_scaled_cloud():
    """Cgh(..., normalized=true): pixel coordinates, x_j read for both axes.  A single
    point one pixel pitch squared over lambda away (theta = pi/2 per pixel^2 offset,
    as in the closed form above) gives the i^k pattern around (x_j, x_j) whatever y_j
    is; a random cloud equals the plain CGH of (p x_j, p x_j, p z_j)."""
    p = 8e-6
    zpix = 2 * p / LAM  # dz = p z = 2 p^2 / lambda
    ph = olref.cgh_normalized([5.0], [99.0], [zpix], 10, 12, p, LAM)
    hh, ww = np.meshgrid(np.arange(10), np.arange(12), indexing="ij")
    expect = np.angle(1j ** (((ww - 5) ** 2 + (hh - 5) ** 2) % 4))
    assert np.max(np.abs(np.angle(np.exp(1j * (ph - expect))))) < 1e-9
    rng = np.random.default_rng(21)
    x, y = rng.uniform(0, 16, 30), rng.uniform(0, 16, 30)
    z = rng.uniform(0.1, 0.2, 30) / p
    ph = olref.cgh_normalized(x, y, z, 16, 16, p, LAM)
    u = olref.cgh(p * x, p * x, p * z, 16, 16, p, p, LAM, dtype=np.complex128)
    assert np.max(np.abs(np.angle(np.exp(1j * (ph - np.angle(u)))))) < 1e-7


def test_cgh_linearity_and_permutation():
    pts = synth.random_points(20, seed=13, xmax=32 * 8e-6, ymax=32 * 8e-6, zmin=0.1, zmax=0.2)
    args = (16, 16, 8e-6, 8e-6, LAM)
    full = olref.cgh(pts["x"], pts["y"], pts["z"], *args, pts["amp"], pts["phase"],
                     dtype=np.complex128)
    a = olref.cgh(pts["x"][:9], pts["y"][:9], pts["z"][:9], *args, pts["amp"][:9],
                  pts["phase"][:9], dtype=np.complex128)
    b = olref.cgh(pts["x"][9:], pts["y"][9:], pts["z"][9:], *args, pts["amp"][9:],
                  pts["phase"][9:], dtype=np.complex128)
    assert rel(a + b, full) < 1e-13
    perm = np.random.default_rng(0).permutation(20)
    pf = olref.cgh(pts["x"][perm], pts["y"][perm], pts["z"][perm], *args, pts["amp"][perm],
                   pts["phase"][perm], dtype=np.complex128)
    assert rel(pf, full) < 1e-13


def test_wrp_wide_window_equals_cgh_over_dz():
    pts = synth.random_points(6, seed=14, xmax=20 * 8e-6, ymax=16 * 8e-6, zmin=0.151, zmax=0.155)
    plane = 0.15
    # p_limit just above lambda/2 makes W huge: every pixel is inside every window
    w = olref.wrp(pts["x"], pts["y"], pts["z"], 16, 20, 8e-6, LAM, plane,
                  p_limit=LAM * 0.5000001, dtype=np.complex128)
    dz = pts["z"] - plane
    c = olref.cgh(pts["x"], pts["y"], dz, 16, 20, 8e-6, 8e-6, LAM, amp=1 / np.abs(dz),
                  dtype=np.complex128)
    assert rel(w, c) < 1e-12


def test_wrp_window_lattice_count():
    p = 8e-6
    x0, y0, dz = 10.3 * p, 9.7 * p, 3e-3
    W = dz / np.sqrt(4 * (p / LAM) ** 2 - 1)
    out = olref.wrp(np.array([x0]), np.array([y0]), np.array([0.15 + dz]), 24, 24, p, LAM, 0.15,
                    dtype=np.complex128)
    count = sum(1 for h in range(24) for w in range(24) if (w * p - x0) ** 2 + (h * p - y0) ** 2 < W * W)
    assert count == int(np.sum(np.abs(out) > 0)) and count > 100
    outd = olref.wrp(np.array([x0]), np.array([y0]), np.array([0.15 + dz]), 24, 24, p, LAM, 0.15,
                     square=True, dtype=np.complex128)
    D = W * float(np.sqrt(np.float32(2)))
    cnt = sum(1 for h in range(24) for w in range(24)
              if abs(w * p - x0) * 2 <= D and abs(h * p - y0) * 2 <= D)
    assert cnt == int(np.sum(np.abs(outd) > 0))


def test_wrp_plane_rule():
    # plane_z = zmin * s with s a float, held in a float (golcgh.h:520-533)
    want = float(np.float32(0.15 * float(np.float32(0.9))))
    assert olref.wrp_plane_z(64, 64, 8e-6, LAM, 0.15, 0.9) == want
    pz = olref.wrp_plane_z(64, 64, 8e-6, LAM, 0.15, -1.0)
    wp2 = np.float32(64 * 64 * 8e-6 * 8e-6 / np.pi / 1024.0)
    assert abs(pz - (0.15 - np.sqrt((4 * (8e-6 / LAM) ** 2 - 1) * wp2))) < 1e-6


# ------------------------------------------------------------------ adjoints
@pytest.mark.parametrize("kind", ["asm", "fresnel"])
def test_adjoint_dot_product(kind):
    u = synth.random_field(16, 24, seed=15, dtype=np.complex128)
    v = synth.random_field(16, 24, seed=16, dtype=np.complex128)
    if kind == "asm":
        f = lambda a, fl: olref.asm_prop(a, 6e-6, 8e-6, LAM, 0.004, fl)
    else:
        f = lambda a, fl: olref.fresnel_prop(a, 6e-6, 8e-6, LAM, 0.004, fl)
    lhs = np.vdot(v, f(u, 0))
    rhs = np.vdot(f(v, olref.ADJOINT), u)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


# --------------------------------------------------------------- dispatcher
def test_prop_dispatch_choices():
    n, dx = 64, 8e-6
    u = synth.random_field(n, n, seed=17)
    du = 1 / (dx * 2 * n)
    umax = du * n
    zmax = np.sqrt(1 / LAM ** 2 - 2 * umax ** 2) / (2 * umax * du)
    zmin = 2 * (dx * n) * dx / LAM
    cases = [(0.5 * zmax, 0), (2 * zmin, 1), (0.5 * (zmax + zmin), 2)]
    assert zmax < zmin
    for z, want in cases:
        out, ch = olref.prop(u, dx, dx, LAM, z, mode=2)
        assert ch == want
        if want == 0:
            ref = olref.asm_prop(u, dx, dx, LAM, z)
        elif want == 1:
            ref = olref.fresnel_prop(u, dx, dx, LAM, z)
        else:
            ref = olref.fresnel_prop(u, dx, dx, LAM, z, olref.FRESNEL_BL_CPU)
        assert rel(out, ref) == 0.0


def test_prec_f32_variant_is_informational_only():
    # PREC_T=float rounds the 1e5-rad ASM phase: far from FP64 (reference default
    # tests/para.h:8), which is why parity is stated against PREC_T=double.
    u = synth.random_field(64, 64, seed=18)
    d = olref.asm_prop(u.astype(np.complex128), 8e-6, 8e-6, LAM, 0.1)
    f32 = olref.asm_prop(u, 8e-6, 8e-6, LAM, 0.1, olref.PREC_F32)
    f64 = olref.asm_prop(u, 8e-6, 8e-6, LAM, 0.1)
    assert rel(f64, d) < 1e-6
    assert rel(f32, d) > 1e-4


def test_cgh_fresnel_diff_recurrence():
    """_CghFresnelDiff's row recurrence: one point gives the closed form
    theta = p z / lambda + gamma / 2 ((w - x)^2 + (h - y)^2) cycles at every pixel, and a
    cloud equals the direct Fresnel sum of the cloud scaled to metres with the
    p z / lambda piston as a per-point phase (independent code paths)."""
    p, ny, nx = 8e-6, 24, 40
    x, y, z = np.array([13.3]), np.array([7.6]), np.array([0.125 / p])
    got = olref.cgh_fresnel_diff(x, y, z, ny, nx, p, LAM)
    hh, ww = np.mgrid[0:ny, 0:nx]
    gam = p / (LAM * z[0])
    theta = p * z[0] / LAM + gam / 2 * ((ww - x[0]) ** 2 + (hh - y[0]) ** 2)
    assert np.abs(got - np.exp(2j * np.pi * theta)).max() < 1e-8
    pts = synth.random_points(50, seed=17, xmax=nx, ymax=ny, zmin=0.10 / p, zmax=0.20 / p)
    amp = pts["amp"]
    got = olref.cgh_fresnel_diff(pts["x"], pts["y"], pts["z"], ny, nx, p, LAM, amp, pts["phase"])
    piston = 2 * np.pi * np.mod(p * pts["z"] / LAM, 1.0)
    ref = olref.cgh(p * pts["x"], p * pts["y"], p * pts["z"], ny, nx, p, p, LAM, amp,
                    pts["phase"] + piston, dtype=np.complex128)
    assert rel(got, ref) < 1e-8


# ------------------------------------------------------------- window edges
# The band-limit windows are written here as explicit half-widths worked out by hand
# from the parameters (a value v keeps offsets |d| < floor(v), i.e. |d| <= floor(v) - 1),
# at parameters where every window is active and v is far from an integer.  Each
# expected field is an FP64 direct sum over source pixels with those masks: a "<="
# for "<", a round for the floor, a y window from dx for one from dy (or the reverse),
# or a zeroed instead of unmodified outside region changes the result.
def _mask(n, half):
    d = np.arange(n) - n // 2
    return np.abs(d) <= half


def _kernel(n, d, lam, z, s, half):
    off = np.arange(n)[:, None] - np.arange(n)[None, :]
    k = np.exp(1j * np.pi * s * (off * d) ** 2 / (lam * z))
    return np.where(np.abs(off) <= half, k, 0)


@pytest.mark.parametrize("variant", ["per_axis", "x_both_axes"])
def test_fresnel_band_limit_window_edges(variant):
    ny, nx, dy, dx, z = 24, 20, 10e-6, 8e-6, 0.00115
    # lambda z / (2 dx^2) = 4.78 -> |dx offset| <= 3;  lambda z / (2 dy^2) = 3.06 -> <= 2
    # (the GPU form takes the x value for both axes: <= 3)
    hx, hy = 3, (2 if variant == "per_axis" else 3)
    u = synth.random_field(ny, nx, seed=31, dtype=np.complex128)
    flag = olref.FRESNEL_BL_CPU if variant == "per_axis" else olref.FRESNEL_BL_GPU
    out = olref.fresnel_prop(u, dy, dx, LAM, z, flag)
    direct = _kernel(ny, dy, LAM, z, 1.0, hy) @ u @ _kernel(nx, dx, LAM, z, 1.0, hx).T
    direct /= 4 * ny * nx
    assert rel(out, direct) < 1e-12
    # the same sum one step wider on either axis is far off
    wide = _kernel(ny, dy, LAM, z, 1.0, hy + 1) @ u @ _kernel(nx, dx, LAM, z, 1.0, hx).T
    assert rel(out, wide / (4 * ny * nx)) > 1e-2


def test_arss_window_edges():
    ny, nx, dy, dx, s, z, oy, ox = 30, 32, 10e-6, 8e-6, 0.25, 0.0006, 3e-6, -5e-6
    # source window, lambda |z| / (2 s |s-1| d^2): x 13.30 -> <= 12, y 8.51 -> <= 7
    # kernel window from the x value on both axes, lambda z / (2 s dx^2) = 9.975 -> <= 8
    # output chirp window, lambda |z| / (2 |1-s| d^2): x 3.325 -> <= 2, y 2.128 -> <= 1;
    # outside it the output is left as the convolution gave it
    tmp = np.pi / (LAM * z)
    u = synth.random_field(ny, nx, seed=32, dtype=np.complex128)
    yc, xc = (np.arange(ny) - ny // 2) * dy, (np.arange(nx) - nx // 2) * dx
    pre = np.exp(1j * tmp * (((s * s - s) * yc ** 2 - 2 * s * oy * yc)[:, None] +
                             ((s * s - s) * xc ** 2 - 2 * s * ox * xc)[None, :]))
    u1 = np.where(_mask(ny, 7)[:, None] & _mask(nx, 12)[None, :], u * pre, 0)
    conv = _kernel(ny, dy, LAM, z, s, 8) @ u1 @ _kernel(nx, dx, LAM, z, s, 8).T / (4 * ny * nx)
    post = np.exp(1j * tmp * (((1 - s) * yc ** 2 + 2 * oy * yc + oy * oy)[:, None] +
                              ((1 - s) * xc ** 2 + 2 * ox * xc + ox * ox)[None, :]))
    inside = _mask(ny, 1)[:, None] & _mask(nx, 2)[None, :]
    direct = np.where(inside, conv * post, conv)
    out = olref.arss_prop(u, dy, dx, LAM, z, s, oy, ox)
    # 1e-7: the reference scales by 1.0f / (float)(M^2) in float (golprop.h:965, 973),
    # inexact for M^2 = 60 x 64
    assert rel(out, direct) < 1e-6
    assert rel(out, conv * post) > 1e-2  # a chirp over the whole field is far off
