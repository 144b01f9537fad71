"""CPU pins of the layered-Fresnel factorisation used by the MFMA kernel.

In exact arithmetic U = Y X^T with
  Y[h, j] = cy_l(h) e^{-2 pi i h beta_j} a_j,   X[w, j] = cx_l(w) e^{-2 pi i w alpha_j},
  cx_l(w) = e^{i pi (px w)^2 c_l},  alpha_j = px x_j c_l,  c_l = 1/(lambda z_l),
  a_j = A_j e^{i (phi_j + pi (x_j^2 + y_j^2) c_l)},
equals the direct Fresnel CGH of the cloud snapped to layer centres z_l.  These
tests evaluate both sides in FP64 (the direct side through the oracle), so a
dropped term, a sign slip or a transposed operand in the factor formulas fails.
"""
import numpy as np

import olref
from openol_amd import synth

LAM = 532e-9


def layer_depth(z, zmin, zmax, nl):
    l = np.clip(np.floor((z - zmin) / (zmax - zmin) * nl), 0, nl - 1)
    return zmin + (l + 0.5) * (zmax - zmin) / nl


def factors(pts, ny, nx, py, px, zl):
    c = 1.0 / (LAM * zl)
    h = np.arange(ny)[:, None]
    w = np.arange(nx)[:, None]
    cy = np.exp(1j * np.pi * (py * h) ** 2 * c[None, :])
    cx = np.exp(1j * np.pi * (px * w) ** 2 * c[None, :])
    alpha = px * pts["x"] * c
    beta = py * pts["y"] * c
    a = pts["amp"] * np.exp(1j * (pts["phase"] + np.pi * (pts["x"] ** 2 + pts["y"] ** 2) * c))
    Y = cy * np.exp(-2j * np.pi * h * beta[None, :]) * a[None, :]
    X = cx * np.exp(-2j * np.pi * w * alpha[None, :])
    return Y, X


def test_factorisation_equals_direct_sum_on_snapped_cloud():
    ny, nx, n, nl = 24, 40, 60, 8
    pts = synth.random_points(n, seed=50, xmax=nx * 8e-6, ymax=ny * 8e-6, zmin=0.10, zmax=0.20)
    zl = layer_depth(pts["z"], 0.10, 0.20, nl)
    Y, X = factors(pts, ny, nx, 8e-6, 8e-6, zl)
    U = Y @ X.T
    direct = olref.cgh(pts["x"], pts["y"], zl, ny, nx, 8e-6, 8e-6, LAM, pts["amp"], pts["phase"],
                       dtype=np.complex128)
    assert olref.rel_l2(U, direct) < 1e-10
    # the snapping itself moves the result (so the snapped reference is needed)
    raw = olref.cgh(pts["x"], pts["y"], pts["z"], ny, nx, 8e-6, 8e-6, LAM, pts["amp"], pts["phase"],
                    dtype=np.complex128)
    assert olref.rel_l2(raw, direct) > 1e-2


def test_real_gemm_interleaving():
    """Re U = A Br^T and Im U = A Bi^T with A = [Re Y, Im Y], Br = [Re X, -Im X],
    Bi = [Im X, Re X] interleaved per point; Bi is Br with each (re, im) pair
    swapped and the new first element negated."""
    rng = np.random.default_rng(1)
    Y = rng.normal(size=(5, 7)) + 1j * rng.normal(size=(5, 7))
    X = rng.normal(size=(6, 7)) + 1j * rng.normal(size=(6, 7))
    A = np.stack([Y.real, Y.imag], axis=-1).reshape(5, 14)
    Br = np.stack([X.real, -X.imag], axis=-1).reshape(6, 14)
    Bi = np.stack([-Br[:, 1::2], Br[:, 0::2]], axis=-1).reshape(6, 14)
    U = Y @ X.T
    assert np.allclose(A @ Br.T, U.real) and np.allclose(A @ Bi.T, U.imag)


def test_exact_phase_split():
    """w alpha mod 1 = frac(w hi) + w lo, with hi holding 24 - ceil(log2 P) fraction
    bits, is exact in FP32 for the w*hi product (the kernel's phase reduction)."""
    P = 4096
    nb = 24 - int(np.ceil(np.log2(P)))
    rng = np.random.default_rng(2)
    alpha = rng.uniform(0, 3000, 1000)
    fr = alpha - np.floor(alpha)
    hi = np.floor(fr * 2.0 ** nb) / 2.0 ** nb
    lo = fr - hi
    hi32, lo32 = hi.astype(np.float32), lo.astype(np.float32)
    assert np.all(hi32.astype(np.float64) == hi)
    w = np.float32(P - 1)
    prod = w * hi32  # FP32 product
    assert np.all(prod.astype(np.float64) == (P - 1) * hi)
    t = (prod - np.floor(prod)).astype(np.float64) + (w * lo32).astype(np.float64)
    exact = ((P - 1) * fr) % 1.0
    err = np.abs(((t - exact) + 0.5) % 1.0 - 0.5)
    assert err.max() < 1e-6
