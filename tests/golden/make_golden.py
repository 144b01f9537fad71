"""Generate tests/golden/golden.npz: tiny FP64 brute-force vectors that pin the oracle.

The reference (somamake/openol) ships no golden outputs and cannot be built here
(FFTW and OpenCV are absent), so these vectors are computed from the definitions
of each reference routine by a route that shares no code with the oracle:

* transforms are explicit DFT matrices (O(M^4)), never an FFT;
* FFT-domain products are rewritten as direct circular convolutions
  (IFFT(FFT(a) FFT(b)) = M^2 circconv(a, b)), evaluated as direct sums;
* CGH / WRP are direct per-pixel sums.

Run:  python tests/golden/make_golden.py      (deterministic; seeds fixed below)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
from openol_amd import synth  # noqa: E402

LAM = 532e-9


def dft_mat(n, sign=-1):
    j = np.arange(n)
    return np.exp(sign * 2j * np.pi * np.outer(j, j) / n)


def pad_center(u, oh, ow):
    ih, iw = u.shape
    out = np.zeros((oh, ow), dtype=np.complex128)
    out[(oh - ih) // 2:(oh - ih) // 2 + ih, (ow - iw) // 2:(ow - iw) // 2 + iw] = u
    return out


def crop_center(u):
    h, w = u.shape
    return u[h // 4:h * 3 // 4, w // 4:w * 3 // 4]


def centred_to_natural(Hc):
    """Element at centred index c moves to natural index (c - M/2) mod M (fftshift)."""
    return np.roll(Hc, (-(Hc.shape[0] // 2), -(Hc.shape[1] // 2)), axis=(0, 1))


def asm_brute(u, dy, dx, z, bandlimit=False):
    """AsmProp (olprop.h:227-248) by DFT matrices."""
    ny, nx = u.shape
    My, Mx = 2 * ny, 2 * nx
    du, dv = 1 / (dx * Mx), 1 / (dy * My)
    uu = (np.arange(Mx) - Mx // 2) * du
    vv = (np.arange(My) - My // 2) * dv
    arg = 1 / LAM ** 2 - uu[None, :] ** 2 - vv[:, None] ** 2
    Hc = np.exp(1j * 2 * np.pi * np.sqrt(np.maximum(arg, 0)) * z) * (arg >= 0)
    if bandlimit:
        ulim = 1 / (LAM * np.sqrt((2 * du * z) ** 2 + 1))
        vlim = 1 / (LAM * np.sqrt((2 * dv * z) ** 2 + 1))
        Hc = Hc * ((np.abs(uu)[None, :] < ulim) & (np.abs(vv)[:, None] < vlim))
    H = centred_to_natural(Hc)
    Fy, Fx = dft_mat(My), dft_mat(Mx)
    U = Fy @ pad_center(u, My, Mx) @ Fx.T
    out = np.conj(Fy) @ (U * H) @ np.conj(Fx).T / (My * Mx)
    return crop_center(out)


def circconv_center(a, hc):
    """out[n] = sum_k a[k] * hc[((n - k) mod M + M/2) mod M]: direct circular
    convolution of a with the fftshift of the centred kernel hc."""
    My, Mx = a.shape
    out = np.zeros_like(a)
    ky, kx = np.nonzero(a)
    for n in range(My):
        iy = ((n - ky) % My + My // 2) % My
        for m in range(Mx):
            ix = ((m - kx) % Mx + Mx // 2) % Mx
            out[n, m] = np.sum(a[ky, kx] * hc[iy, ix])
    return out


def fresnel_brute(u, dy, dx, z, mode="plain"):
    """FresnelProp / FresnelPropBandLimit (olprop.h:125-170), GPU window golprop.h:223-238."""
    ny, nx = u.shape
    My, Mx = 2 * ny, 2 * nx
    x = (np.arange(Mx) - Mx // 2) * dx
    y = (np.arange(My) - My // 2) * dy
    hc = np.exp(1j * np.pi * (x[None, :] ** 2 + y[:, None] ** 2) / (LAM * z))
    mi = np.abs(np.arange(Mx) - Mx // 2)[None, :]
    ni = np.abs(np.arange(My) - My // 2)[:, None]
    if mode == "cpu":
        wl = int(np.floor(abs(LAM * z * 0.5 / dx) / dx))
        hl = int(np.floor(abs(LAM * z * 0.5 / dy) / dy))
        hc = hc * ((mi < wl) & (ni < hl))
    elif mode == "gpu":
        wl = int(np.floor(abs(LAM * z * 0.5 / dx) / dx))
        hc = hc * ((mi < wl) & (ni < wl))
    out = circconv_center(pad_center(u, My, Mx), hc) / (My * Mx)
    return crop_center(out)


def arss_brute(u, dy, dx, z, s, oy, ox):
    """ARSSFresnelProp (golprop.h:941-995) by direct circular convolution."""
    ny, nx = u.shape
    My, Mx = 2 * ny, 2 * nx
    t = np.pi / (LAM * z)
    # stage u
    wl = int(np.floor(LAM * abs(z) * 0.5 / (s * abs(s - 1) * dx) / dx))
    hl = int(np.floor(LAM * abs(z) * 0.5 / (s * abs(s - 1) * dy) / dy))
    x1 = (np.arange(nx) - nx // 2) * dx
    y1 = (np.arange(ny) - ny // 2) * dy
    ph = ((s * s - s) * x1[None, :] ** 2 - 2 * s * ox * x1[None, :]
          + (s * s - s) * y1[:, None] ** 2 - 2 * s * oy * y1[:, None]) * t
    win = ((np.abs(np.arange(nx) - nx // 2)[None, :] < wl)
           & (np.abs(np.arange(ny) - ny // 2)[:, None] < hl))
    pre = u * np.exp(1j * ph) * win
    # stage h (x-limit on both axes)
    lim = int(np.floor(abs(LAM * z * 0.5 / (dx * s)) / dx))
    xm = (np.arange(Mx) - Mx // 2) * dx
    ym = (np.arange(My) - My // 2) * dy
    hc = np.exp(1j * s * (xm[None, :] ** 2 + ym[:, None] ** 2) * t)
    hc = hc * ((np.abs(np.arange(Mx) - Mx // 2)[None, :] < lim)
               & (np.abs(np.arange(My) - My // 2)[:, None] < lim))
    out = crop_center(circconv_center(pad_center(pre, My, Mx), hc) / (My * Mx))
    # stage Cz: inside the window only
    wl = int(np.floor(LAM * abs(z) * 0.5 / (abs(1 - s) * dx) / dx))
    hl = int(np.floor(LAM * abs(z) * 0.5 / (abs(1 - s) * dy) / dy))
    phc = ((1 - s) * x1[None, :] ** 2 + 2 * ox * x1[None, :] + ox * ox
           + (1 - s) * y1[:, None] ** 2 + 2 * oy * y1[:, None] + oy * oy) * t
    win = ((np.abs(np.arange(nx) - nx // 2)[None, :] < wl)
           & (np.abs(np.arange(ny) - ny // 2)[:, None] < hl))
    return np.where(win, out * np.exp(1j * phc), out)


def cgh_brute(x, y, z, ny, nx, py, px, amp=None, phase=None, exact=False):
    """_CghFresnel (olcgh.h:296-313) + extension (A_j, phi_j, exact r): direct sum."""
    W = np.arange(nx)[None, :, None] * px
    Hh = np.arange(ny)[:, None, None] * py
    dxv = W - x[None, None, :]
    dyv = Hh - y[None, None, :]
    if exact:
        th = 2 * np.pi * np.sqrt(dxv ** 2 + dyv ** 2 + z[None, None, :] ** 2) / LAM
    else:
        th = 2 * np.pi * (dxv ** 2 + dyv ** 2) / (2 * z[None, None, :] * LAM)
    if phase is not None:
        th = th + phase[None, None, :]
    a = np.ones_like(x) if amp is None else amp
    return np.sum(a[None, None, :] * np.exp(1j * th), axis=2)


def wrp_brute(x, y, z, ny, nx, p, plane_z, square=False):
    """WRPStep1 (golcgh.h:303-367) / WRPStep1_D (:403-468): direct windowed sum."""
    W = np.abs(z - plane_z) / np.sqrt(4 * (p / LAM) ** 2 - 1)
    dxv = np.arange(nx)[None, :, None] * p - x[None, None, :]
    dyv = np.arange(ny)[:, None, None] * p - y[None, None, :]
    dz = (z - plane_z)[None, None, :]
    r2 = dxv ** 2 + dyv ** 2
    if square:
        D = W * float(np.sqrt(np.float32(2.0)))
        inside = (np.abs(dxv) * 2 <= D) & (np.abs(dyv) * 2 <= D)
    else:
        inside = r2 < W ** 2
    th = 2 * np.pi * r2 / (2 * dz * LAM)
    return np.sum(np.where(inside, np.exp(1j * th) / np.abs(dz), 0), axis=2)


def main():
    out = {}
    # propagation fields: square 8x8 and rectangular 8x4 (ny=8, nx=4), dy != dx
    for tag, (ny, nx) in {"sq": (8, 8), "rect": (8, 4)}.items():
        u = synth.random_field(ny, nx, seed=100 + ny * 10 + nx, dtype=np.complex128)
        out[f"field_{tag}"] = u
        dy, dx = 6e-6, 8e-6
        out[f"asm_{tag}"] = asm_brute(u, dy, dx, 0.01)
        out[f"asm_bl_{tag}"] = asm_brute(u, 1e-6, 1e-6, 1e-4, bandlimit=True)
        out[f"fresnel_{tag}"] = fresnel_brute(u, dy, dx, 0.002)
        out[f"fresnel_blcpu_{tag}"] = fresnel_brute(u, dy, dx, 0.001, mode="cpu")
        out[f"fresnel_blgpu_{tag}"] = fresnel_brute(u, dy, dx, 0.001, mode="gpu")
        out[f"arss_open_{tag}"] = arss_brute(u, dy, dx, 0.002, 0.5, 0.0, 0.0)
        out[f"arss_win_{tag}"] = arss_brute(u, 12e-6, 16e-6, -0.0005, 0.5, 3e-5, -2e-5)
    # CGH: 7 points over a 12 x 16 hologram (ny=12, nx=16)
    pts = synth.random_points(7, seed=7, xmax=16 * 8e-6, ymax=12 * 8e-6, zmin=0.01, zmax=0.02)
    out.update({f"cgh_{k}": v for k, v in pts.items()})
    out["cgh_plain"] = cgh_brute(pts["x"], pts["y"], pts["z"], 12, 16, 6e-6, 8e-6)
    out["cgh_ext"] = cgh_brute(pts["x"], pts["y"], pts["z"], 12, 16, 6e-6, 8e-6,
                               pts["amp"], pts["phase"], exact=False)
    out["cgh_exact"] = cgh_brute(pts["x"], pts["y"], pts["z"], 12, 16, 6e-6, 8e-6,
                                 pts["amp"], pts["phase"], exact=True)
    # WRP: 9 points 1-5 mm behind a plane at 0.15 m, 24 x 20 grid (ny=24, nx=20)
    wp = synth.random_points(9, seed=9, xmax=20 * 8e-6, ymax=24 * 8e-6, zmin=0.151, zmax=0.155)
    out.update({f"wrp_{k}": v for k, v in wp.items() if k in ("x", "y", "z")})
    out["wrp_circle"] = wrp_brute(wp["x"], wp["y"], wp["z"], 24, 20, 8e-6, 0.15)
    out["wrp_square"] = wrp_brute(wp["x"], wp["y"], wp["z"], 24, 20, 8e-6, 0.15, square=True)
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), sorted(out))


if __name__ == "__main__":
    main()
