"""Pins for the oracle's optical elements (CPU only): the MT19937 engine and the
53-bit draws of Diffuser::random_phase against the C++ standard's known answer and
numpy's independent MT19937 (RandomState draws the same res53 stream), the GPU
diffuser kernels' index maps, and the lens arrays against direct formulas, their
symmetries and the hexagonal layout."""
import numpy as np
import pytest

import olref

LAM = 532e-9


def test_mt19937_known_answer():
    # the C++ standard: the 10000th output of a default-constructed std::mt19937 (seed 5489)
    assert olref.mt19937_nth(5489, 10000) == 4123659995
    # numpy's legacy MT19937 with an integer seed uses the same initialisation
    rs = np.random.RandomState(1)
    assert [olref.mt19937_nth(1, k) for k in (1, 2, 3)] == [int(x) for x in rs.randint(0, 2 ** 32, 3, dtype=np.uint64)]


@pytest.mark.parametrize("step,oy,ox", [(olref.PLAY, 0, 0), (olref.RECORD, 2, -3), (olref.PLAY, -5, 7)])
def test_random_phase_is_the_res53_stream(step, oy, ox):
    ny, nx, seed, rng = 6, 8, 11, np.float32(1.5 * np.pi)
    u = np.ones((ny, nx), np.complex64)
    out = olref.diffuser_random_phase(u, step, seed, float(rng), oy, ox)
    v = np.random.RandomState(seed).random_sample(ny * nx).reshape(ny, nx)
    ph = ((1 if step == olref.PLAY else -1) * np.float64(rng) * (v - 0.5)).astype(np.float32)
    want = np.empty((ny, nx), np.float32)
    for h in range(ny):
        for w in range(nx):
            want[(h + oy) % ny, (w + ox) % nx] = ph[h, w]
    assert np.allclose(np.angle(out), want, atol=2e-7)
    assert np.allclose(np.abs(out), 1, atol=1e-6)


def test_diffuser_kernel_index_maps():
    ny, nx, bs = 8, 12, 4
    rng = np.random.default_rng(3)
    u = (rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))).astype(np.complex64)
    b = rng.random(ny * nx).astype(np.float32)
    sc = np.float32(2.0)
    yy, xx = np.mgrid[0:ny, 0:nx]
    blocks = b[(xx // bs) + (yy // bs) * (nx // bs)]
    cases = {
        "phase": (b.reshape(ny, nx) - 0.5) * sc,
        "xphase": np.broadcast_to((b[:nx] - 0.5) * sc, (ny, nx)),
        "yphase": np.broadcast_to(((b[:ny] - 0.5) * sc)[:, None], (ny, nx)),
        "bitrandom": np.where(b.reshape(ny, nx) > 0.5, np.float32(np.pi), 0),
        "block": (blocks - 0.5) * sc,
        "blockstep": (blocks + ((xx % bs) + (yy % bs)) / bs - 0.5) * sc,
        "bitphase": np.where((xx % 2) == (yy % 2), np.float32(np.pi), 0),
    }
    for kind, ph in cases.items():
        out = olref.diffuser_apply(u, b, float(sc), kind, bs)
        assert np.allclose(out, u * np.exp(1j * ph.astype(np.float64)), rtol=0, atol=2e-6), kind


def lens_direct(ny, nx, pitch, bw, f, lam, step, on, om):
    bp = int(np.round(np.float32(bw) / np.float32(pitch)))
    z = -f if step == olref.PLAY else f
    n, m = np.mgrid[0:ny, 0:nx]
    ns, ms = (n % bp - on) % bp, (m % bp - om) % bp
    x, y = (ms - bp // 2) * pitch, (ns - bp // 2) * pitch
    return np.exp(1j * np.pi * (x ** 2 + y ** 2) / (lam * z)), bp


@pytest.mark.parametrize("step,on,om", [(olref.PLAY, 0, 0), (olref.RECORD, 3, -2)])
def test_lens_array_direct_formula(step, on, om):
    ny, nx, pitch, bw, f = 40, 52, 8e-6, 8e-6 * 12, 0.02
    got = olref.lens_array(ny, nx, pitch, bw, f, LAM, step, on, om)
    want, bp = lens_direct(ny, nx, np.float32(pitch), np.float32(bw), np.float32(f), np.float32(LAM), step, on, om)
    assert bp == 12
    assert np.abs(got - want).max() < 2e-5      # float phase arithmetic
    # PLAY and RECORD are conjugate lenses; the array repeats with the lenslet pitch
    other = olref.lens_array(ny, nx, pitch, bw, f, LAM, 1 - step, on, om)
    assert np.array_equal(got, np.conj(other))
    assert np.array_equal(got[bp:2 * bp, :], got[:bp, :])


def test_lens_array_band_limit_window():
    ny, nx, pitch, bw, f, pl = 32, 32, 8e-6, 8e-6 * 16, 0.01, 60e-6
    got = olref.lens_array(ny, nx, pitch, bw, f, LAM, olref.PLAY, 0, 0, pl)
    full = olref.lens_array(ny, nx, pitch, bw, f, LAM, olref.PLAY, 0, 0)
    wlim = int(np.floor(np.float32(abs(np.float32(LAM) * np.float32(-f) * 0.5 / np.float32(pl))) / np.float32(pitch)))
    loc = np.arange(32) % 16 - 8
    inside = (np.abs(loc)[:, None] < wlim) & (np.abs(loc)[None, :] < wlim)
    assert 0 < wlim < 8
    assert np.array_equal(got[inside], full[inside]) and np.all(got[~inside] == 0)


def test_lens_circle_layout():
    ny, nx, pitch, bw, f = 64, 64, 8e-6, 8e-6 * 16, 0.02
    got = olref.lens_array(ny, nx, pitch, bw, f, LAM, olref.PLAY, 0, 0, 0.0, True)
    sq = olref.lens_array(ny, nx, pitch, bw, f, LAM, olref.PLAY, 0, 0)
    stepn = int(np.round(np.float32(bw) * np.sqrt(np.float32(3)) * np.float32(0.5) / np.float32(pitch)))
    assert stepn == 14
    # pixels in no circle are 1; row 0 of lenslets is the square lens inside the circle
    r = np.float32(bw) * np.float32(0.5)
    l = np.arange(16)
    inc = ((l[None, :] * np.float32(pitch) - r) ** 2 + (l[:, None] * np.float32(pitch) - r) ** 2) <= r ** 2
    assert np.array_equal(got[:14, :16][inc[:14]], sq[:14, :16][inc[:14]])
    assert np.all(got[:14, :16][~inc[:14]] == 1)
    # the second row (odd) is shifted by half a lenslet
    row1 = got[14 + 8, :]
    assert not np.array_equal(row1, got[8, :])


def _traversal(ny, nx, bp):
    """(n, m) of every pixel in the reference's duplicate traversal: lexicographic in
    (block row, block column, local row, local column), pixels past the field edge
    dropped (the loop's break skips the rest of a row; rows past the edge draw
    nothing)."""
    bn, bm, ln, lm = np.meshgrid(np.arange(0, ny, bp), np.arange(0, nx, bp), np.arange(bp),
                                 np.arange(bp), indexing="ij")
    n, m = (bn + ln).ravel(), (bm + lm).ravel()
    keep = (n < ny) & (m < nx)
    return n[keep], m[keep]


@pytest.mark.parametrize("shape,on,om", [((48, 64), 0, 0), ((50, 71), 3, -5)])
@pytest.mark.parametrize("kind", [0, 1, 2])
def test_lens_tile_forms(shape, on, om, kind):
    ny, nx = shape
    pitch, f, pl, bp = np.float32(8e-6), np.float32(0.01), np.float32(60e-6), 16
    block = olref.lens_block(bp, float(pitch), float(f), LAM, olref.PLAY, float(pl) if kind == 1 else 0.0)
    got = olref.lens_tile(block, ny, nx, on, om, float(pitch), float(f), LAM, float(pl), kind)
    n, m = np.mgrid[0:ny, 0:nx]
    ns, ms = (n % bp - on) % bp, (m % bp - om) % bp
    tiled = block[ns, ms]
    if kind == 0:
        # duplicate of the lenslet is gen_complex's field
        assert np.array_equal(got, tiled)
        assert np.array_equal(got, olref.lens_array(ny, nx, float(pitch), float(pitch) * bp, float(f),
                                                     LAM, olref.PLAY, on, om))
        return
    wlim = int(np.floor(np.float32(abs(np.float32(LAM) * f * 0.5 / pl)) / pitch))
    assert 0 < wlim < bp // 2
    rnd = (np.abs(ms - bp // 2) >= wlim) | (np.abs(ns - bp // 2) >= wlim)
    assert np.array_equal(got[~rnd], tiled[~rnd])
    # the random pixels, in traversal order, carry numpy's independent res53 stream
    tn, tm = _traversal(ny, nx, bp)
    sel = rnd[tn, tm]
    tn, tm = tn[sel], tm[sel]
    v = np.random.RandomState(0).random_sample(len(tn))
    span = 2 * np.float32(np.pi) if kind == 1 else 0.5 * np.float32(np.pi)
    ph = (span * (v - 0.5)).astype(np.float32)
    e = np.exp(1j * ph.astype(np.float64))
    want = e if kind == 1 else tiled[tn, tm] * e
    assert np.abs(got[tn, tm] - want).max() < 2e-6
