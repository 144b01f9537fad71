"""Pins for the oracle's shifted (off-axis) ASM and split propagation (CPU only).

shifted_asm and split_prop are checked against what the physics fixes, not against
a retyping of their formulas:
* at zero shift they reduce to the pinned AsmProp (with and without the window);
* a destination window offset from the source equals the matching crop of one
  plain ASM over a canvas that holds both windows (the sign of the shift, the
  linear phase term and, far off axis, the window's outer branches);
* the adjoint flag satisfies the dot-product identity;
* the GPU split form sums to the plain ASM of the whole field;
* the reference's split offsets are (source - destination) block offsets, which
  the ASM form absorbs (the shift is a full period for adjacent blocks) and the
  Fresnel form does not: that quirk is kept and checked here.
"""
import numpy as np
import pytest

import olref
from openol_amd import synth

LAM, P = 532e-9, 8e-6


def gauss(n, cx, cy, sigma, fx=0.0, w=None):
    yy, xx = np.mgrid[0:n, 0:(w or n)]
    return (np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * sigma ** 2))
            * np.exp(2j * np.pi * fx * xx)).astype(np.complex64)


@pytest.mark.parametrize("shape", [(32, 32), (16, 48)])
def test_zero_shift_is_asm(shape):
    u = synth.random_field(*shape, seed=3)
    a = olref.asm_prop(u, P, 6e-6, LAM, 0.004)
    b = olref.shifted_asm(u, P, 6e-6, LAM, 0.004, 0.0, 0.0)
    assert olref.rel_l2(b, a) == 0.0
    a = olref.asm_prop(u, P, 6e-6, LAM, 0.004, olref.BANDLIMIT)
    b = olref.shifted_asm(u, P, 6e-6, LAM, 0.004, 0.0, 0.0, olref.SHIFT_WINDOW)
    assert olref.rel_l2(b, a) == 0.0
    # the CPU form's net scale is 1/M^2 of that
    c = olref.shifted_asm(u, P, 6e-6, LAM, 0.004, 0.0, 0.0, olref.SHIFT_WINDOW | olref.SHIFT_CPU_SCALE)
    assert olref.rel_l2(c * (4 * shape[0] * shape[1]), a) < 1e-6


@pytest.mark.parametrize("z", [2e-3, 1e-2])
def test_near_offset_window_matches_wide_canvas(z):
    """Destination window N/2 pixels to the right: equals the crop of a 2N-wide ASM."""
    N = 64
    src = gauss(N, 40, 32, 2.0, 0.15)
    big = np.zeros((N, 2 * N), np.complex64)
    big[:, :N] = src
    truth = olref.asm_prop(big, P, P, LAM, z)[:, N // 2:N // 2 + N]
    for flags in (0, olref.SHIFT_WINDOW):
        got = olref.shifted_asm(src, P, P, LAM, z, 0.0, (N // 2) * P, flags)
        assert olref.rel_l2(got, truth) < 2e-5
        wrong = olref.shifted_asm(src, P, P, LAM, z, 0.0, -(N // 2) * P, flags)
        assert olref.rel_l2(wrong, truth) > 0.9  # the sign is observable
    # the same along y
    srcy = np.ascontiguousarray(src.T)
    bigy = np.ascontiguousarray(big.T)
    truthy = olref.asm_prop(bigy, P, P, LAM, z)[N // 2:N // 2 + N, :]
    got = olref.shifted_asm(srcy, P, P, LAM, z, (N // 2) * P, 0.0, olref.SHIFT_WINDOW)
    assert olref.rel_l2(got, truthy) < 2e-5


@pytest.mark.parametrize("sign", [1, -1])
def test_far_offset_window_branches(sign):
    """|ox| beyond half the padded width (the outer branches of the window): a beam
    tilted towards a window 1.5 N away arrives there; the windowed result matches a
    4N-wide canvas."""
    N, off, z = 64, 96, 0.03
    src = gauss(N, 32, 32, 4.0, sign * 0.4)
    big = np.zeros((N, 4 * N), np.complex64)
    if sign > 0:
        big[:, :N] = src
        truth = olref.asm_prop(big, P, P, LAM, z)[:, off:off + N]
    else:
        big[:, 3 * N:] = src
        truth = olref.asm_prop(big, P, P, LAM, z)[:, 3 * N - off:4 * N - off]
    got = olref.shifted_asm(src, P, P, LAM, z, 0.0, sign * off * P, olref.SHIFT_WINDOW)
    assert olref.rel_l2(got, truth) < 1e-2
    wrong = olref.shifted_asm(src, P, P, LAM, z, 0.0, -sign * off * P, olref.SHIFT_WINDOW)
    assert olref.rel_l2(wrong, truth) > 0.9


def test_shifted_transfer_table():
    """The shifted table: at zero offset without the window, the plain transfer function
    exp(2 pi i w z); a shift multiplies it by exp(2 pi i (ox u + oy v)) (sign and axes);
    the window keeps exactly the band |u - u0| < uw / 2 (here centred, u0 = 0, at zero
    offset: the band limit |u| < 1 / (lambda sqrt((2 du z)^2 + 1)))."""
    ny, nx, du, dv, z = 32, 48, 1 / (64 * P), 1 / (96 * 6e-6), 0.01
    c_u = (np.arange(nx) - nx // 2) * du
    c_v = (np.arange(ny) - ny // 2) * dv
    w = np.sqrt(1 / LAM ** 2 - c_u[None, :] ** 2 - c_v[:, None] ** 2)
    H0 = olref.shifted_asm_transfer(ny, nx, dv, du, LAM, z, 0.0, 0.0)
    assert np.max(np.abs(H0 - np.exp(2j * np.pi * w * z))) < 1e-9
    oy, ox = 3e-4, -7e-4
    H1 = olref.shifted_asm_transfer(ny, nx, dv, du, LAM, z, oy, ox)
    ramp = np.exp(2j * np.pi * (ox * c_u[None, :] + oy * c_v[:, None]))
    assert np.max(np.abs(H1 - H0 * ramp)) < 1e-9
    zl = 0.2  # long enough for the band limit to cut
    Hw = olref.shifted_asm_transfer(ny, nx, dv, du, LAM, zl, 0.0, 0.0, olref.SHIFT_WINDOW)
    ulim = 1 / (LAM * np.sqrt((2 * du * zl) ** 2 + 1))
    vlim = 1 / (LAM * np.sqrt((2 * dv * zl) ** 2 + 1))
    keep = (np.abs(c_u)[None, :] < ulim) & (np.abs(c_v)[:, None] < vlim)
    assert 0 < keep.sum() < keep.size
    assert np.array_equal(Hw != 0, keep)


def test_shifted_adjoint_dot_product():
    u = synth.random_field(16, 24, seed=5, dtype=np.complex128)
    v = synth.random_field(16, 24, seed=6, dtype=np.complex128)
    for flags in (0, olref.SHIFT_WINDOW):
        f = lambda a, fl: olref.shifted_asm(a, 6e-6, 8e-6, LAM, 0.004, 3e-5, -7e-5, fl)
        lhs = np.vdot(v, f(u, flags))
        rhs = np.vdot(f(v, flags | olref.ADJOINT), u)
        assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def split_field(n=128):
    """Sources across the block borders, away from the outer edges (light that leaves
    the field would re-enter through the periodic block shifts)."""
    rng = np.random.default_rng(0)
    u = np.zeros((n, n), np.complex64)
    for cy, cx in [(30, 60), (70, 66), (100, 20), (64, 64)]:
        u += gauss(n, cx * n // 128, cy * n // 128, 2.5) * np.complex64(np.exp(1j * rng.uniform(0, 6)))
    return u


@pytest.mark.parametrize("z", [1e-3, 2e-3])
def test_gpu_split_sums_to_whole_field_asm(z):
    u = split_field()
    full = olref.asm_prop(u, P, P, LAM, z)
    got = olref.split_prop(u, P, P, LAM, z, 2, 2, 0, 0)
    assert olref.rel_l2(got, full) < 1e-5
    # one block is the plain shifted propagation at zero offset
    assert olref.rel_l2(olref.split_prop(u, P, P, LAM, z, 1, 1, 0, 0), full) < 1e-6


def test_split_fresnel_uses_source_minus_destination_offsets():
    """gsplitProp's FRESNEL branch passes (src - dst) offsets to ARSS (s = 1).  With
    (dst - src) offsets the sum would be 4 times the whole-field Fresnel propagation
    (the block transforms are normalised by a 4 times smaller M^2); the reference's
    sign gives something else, kept here."""
    N, z = 64, 0.05
    u = split_field(2 * N)

    def compose(sign):
        out = np.zeros_like(u)
        for gn in range(2):
            for gm in range(2):
                src = np.ascontiguousarray(u[gn * N:(gn + 1) * N, gm * N:(gm + 1) * N])
                for dn in range(2):
                    for dm in range(2):
                        oy, ox = sign * (gn - dn) * N * P, sign * (gm - dm) * N * P
                        out[dn * N:(dn + 1) * N, dm * N:(dm + 1) * N] += \
                            olref.arss_prop(src, P, P, LAM, z, 1.0, oy, ox)
        return out

    full = olref.fresnel_prop(u, P, P, LAM, z)
    assert olref.rel_l2(compose(-1) / 4, full) < 1e-5   # physics
    got = olref.split_prop(u, P, P, LAM, z, 2, 2, 1, 0)
    assert olref.rel_l2(got, compose(1)) < 1e-6          # the reference's offsets
    assert olref.rel_l2(got / 4, full) > 0.5


def test_cpu_split_is_blockwise_cpu_shifted():
    """splitAsmProp: every pair through the windowed, 1/M^2-scaled shiftedAsmProp."""
    N, z = 32, 3e-3
    u = split_field(2 * N)
    got = olref.split_prop(u, P, P, LAM, z, 2, 2, 0, olref.SHIFT_WINDOW | olref.SHIFT_CPU_SCALE)
    # the (0, 0) -> (0, 0) term dominates block (0, 0) when the sources sit inside it
    only = np.zeros_like(u)
    only[:N, :N] = u[:N, :N]
    got0 = olref.split_prop(only, P, P, LAM, z, 2, 2, 0, olref.SHIFT_WINDOW | olref.SHIFT_CPU_SCALE)
    self_term = olref.shifted_asm(np.ascontiguousarray(u[:N, :N]), P, P, LAM, z, 0.0, 0.0,
                                  olref.SHIFT_WINDOW | olref.SHIFT_CPU_SCALE)
    assert olref.rel_l2(got0[:N, :N], self_term) < 1e-6
    assert got.shape == u.shape and np.all(np.isfinite(got))
