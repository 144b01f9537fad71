"""Pins for the oracle's tilted-plane propagation (CPU only).

* At angle 0 with bilinear sampling the resampling is the identity: both forms
  reduce to AsmProp / FresnelProp times the reference's extra 1/M^2.
* At angle 0 the reference's radial "bicubic" kernel is not the identity: at the
  integer sample points it weights the diagonal neighbours by k(sqrt 2) = 14 - 10
  sqrt 2; the oracle equals that 3 x 3 filter applied to the centred spectrum with
  numpy's FFT.
* A tilted plane wave: the output spectrum peaks where the rotated frequency
  equals the input's (a11 f + a13 sqrt(f0^2 - f^2) [- a13 f0] = fx0), which fixes
  the rotation's sign, the carrier term and the axis the rotation acts on.
"""
import numpy as np
import pytest
from scipy.optimize import brentq

import olref

LAM, P = 532e-9, 8e-6


def smooth_field(ny, nx, seed):
    """A field whose padded spectrum vanishes at the band edges: low-pass noise under
    a Gaussian envelope.  (Sample positions computed as ((m - c) du) / du + c may
    round just below an integer; at the first spectral column that selects the cell
    outside the array, in the reference as in the oracle, so the identity at angle 0
    holds only for spectra that are empty there.)"""
    rng = np.random.default_rng(seed)
    f = np.fft.fft2(rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx)))
    fy, fx = np.fft.fftfreq(ny)[:, None], np.fft.fftfreq(nx)[None, :]
    f *= np.exp(-(fx ** 2 + fy ** 2) / (2 * 0.06 ** 2))
    yy, xx = np.mgrid[0:ny, 0:nx]
    env = np.exp(-((xx - nx / 2) ** 2 / (2 * (nx / 6) ** 2) + (yy - ny / 2) ** 2 / (2 * (ny / 6) ** 2)))
    return np.fft.ifft2(f) * env


@pytest.mark.parametrize("kind", ["source", "reference"])
@pytest.mark.parametrize("mode", ["asm", "fresnel"])
def test_zero_angle_bilinear_is_plain_propagation(kind, mode):
    u = smooth_field(32, 48, 2)
    M2 = 64 * 96
    if mode == "asm":
        z = 0.01
        ref = olref.asm_prop(u, P, 6e-6, LAM, z)
    else:
        z = 0.08
        ref = olref.fresnel_prop(u, P, 6e-6, LAM, z)
    got = olref.tilted(u, P, 6e-6, LAM, z, 0.0, "bilinear", mode, kind)
    assert olref.rel_l2(got * M2, ref) < 1e-6


@pytest.mark.parametrize("kind", ["source", "reference"])
def test_zero_angle_bicubic_is_the_radial_kernel_filter(kind):
    n, z = 32, 0.01
    u = smooth_field(n, n, 4)
    M = 2 * n
    pad = np.zeros((M, M), complex)
    pad[n // 2:n // 2 + n, n // 2:n // 2 + n] = u
    S = np.fft.fftshift(np.fft.fft2(pad)) / M ** 2
    c = np.arange(M) - M // 2
    fu, fv = c[None, :] / (P * M), c[:, None] / (P * M)
    H = np.exp(2j * np.pi * np.sqrt(1 / LAM ** 2 - fu ** 2 - fv ** 2) * z)
    if kind == "reference":  # the transfer function comes before the resampling
        S = S * H
    k2 = 14 - 10 * np.sqrt(2)
    F = S.copy()
    for dy in (-1, 1):
        for dx in (-1, 1):
            sh = np.zeros_like(S)
            ys = slice(max(0, -dy), M - max(0, dy))
            yd = slice(max(0, dy), M - max(0, -dy))
            xs = slice(max(0, -dx), M - max(0, dx))
            xd = slice(max(0, dx), M - max(0, -dx))
            sh[yd, xd] = S[ys, xs]   # S[n - dy, m - dx]
            F += k2 * sh
    if kind == "source":
        F = F * H
    out = np.fft.ifft2(np.fft.ifftshift(F)) * M ** 2
    want = out[n // 2:n // 2 + n, n // 2:n // 2 + n] / M ** 2
    got = olref.tilted(u, P, P, LAM, z, 0.0, "bicubic", "asm", kind)
    assert olref.rel_l2(got, want) < 1e-6


def peak_bin(out, n):
    M = 2 * n
    pad = np.zeros((M, M), complex)
    pad[n // 2:n // 2 + n, n // 2:n // 2 + n] = out
    S = np.fft.fftshift(np.fft.fft2(pad))
    return np.unravel_index(np.argmax(np.abs(S)), S.shape)[1] - M // 2


@pytest.mark.parametrize("kind,deg,fx_bins", [("reference", 0.5, 6.4), ("reference", -0.5, 6.4),
                                              ("source", 20.0, 20.0), ("source", -20.0, 20.0)])
def test_tilted_plane_wave_peak(kind, deg, fx_bins):
    n = 64
    M = 2 * n
    du = 1 / (P * M)
    yy, xx = np.mgrid[0:n, 0:n]
    fx0 = fx_bins * du
    env = np.exp(-((xx - n / 2) ** 2 + (yy - n / 2) ** 2) / (2 * 12.0 ** 2))
    u = env * np.exp(2j * np.pi * fx0 * (xx - n / 2) * P)
    th = np.deg2rad(deg)
    out = olref.tilted(u, P, P, LAM, 1e-3, th, "bilinear", "asm", kind)
    rot = -th if kind == "source" else th
    a11, a13, f0 = np.cos(rot), -np.sin(rot), 1 / LAM
    g = lambda f: a11 * f + a13 * np.sqrt(f0 ** 2 - f ** 2) - (a13 * f0 if kind == "source" else 0) - fx0
    pred = brentq(g, -0.5 * f0, 0.5 * f0) / du
    assert abs(pred) < n / 2
    assert abs(peak_bin(out, n) - pred) <= 1.5
    if kind == "reference":
        assert abs(pred - fx_bins) > 10  # the tilt moves the peak: the sign is observable
