"""Point-cloud ingest and 8-bit encodings on the CPU: pins for the oracle's numpy
restatements (known answers decoded from the reference's own .3d fixtures with
od, hand-computed encodings, closed forms and invariants), and the library's
host-side entry points (file reader, BMP writer / reader) against the oracle,
byte for byte.  No GPU calls."""
import ctypes
import os

import numpy as np
import pytest

import olref
from openol_amd import _lib

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CUBE = os.path.join(GOLDEN, "cube284.3d")
HORN = os.path.join(GOLDEN, "horn02-f000.3d")
PT = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("amp", "<f4"), ("phase", "<f4")])


def lib_read(path, fmt=_lib.OBJ_AUTO, scale=1.0, offset=(0.0, 0.0, 0.0), cap=None):
    L = _lib.load()
    n = ctypes.c_int64()
    rc = L.olx_obj_read(os.fsencode(path), fmt, None, 0, ctypes.byref(n), scale, *offset)
    if rc != 0:
        return rc, None
    rec = np.zeros(n.value if cap is None else cap, dtype=PT)
    rc = L.olx_obj_read(os.fsencode(path), fmt, rec.ctypes.data, rec.size, ctypes.byref(n), scale,
                        *offset)
    return rc, rec


# --------------------------------------------------------------------------- reader
def test_oracle_reader_known_answers():
    """Values decoded with od -t d4 from the fixture files (reference 3ddata/)."""
    c = olref.obj_read(CUBE)
    assert c.shape == (284, 3) and os.path.getsize(CUBE) == 4 + 12 * 284
    assert c[0].tolist() == [-19, -6, 0] and c[1].tolist() == [-19, -6, 0]
    assert c[2].tolist() == [-18, -6, 0] and c[-1].tolist() == [19, 6, 0]
    h = olref.obj_read(HORN)
    assert h.shape == (5197, 3) and os.path.getsize(HORN) == 4 + 12 * 5197
    assert h[0].tolist() == [0, 2589, 0] and h[1].tolist() == [216, 2589, 125]
    assert h[2].tolist() == [175, 2736, 101] and h[-1].tolist() == [3856, -9727, -15306]
    hs = olref.obj_read(HORN, scale=0.5, offset=(1.0, 2.0, 3.0))
    assert hs[1].tolist() == [109.0, 1296.5, 65.5]


@pytest.mark.parametrize("path", [CUBE, HORN])
def test_library_reader_matches_oracle(path):
    rc, rec = lib_read(path)
    assert rc == 0
    ref = olref.obj_read(path)
    got = np.stack([rec["x"], rec["y"], rec["z"]], axis=1)
    assert np.array_equal(got, ref)
    assert np.all(rec["amp"] == 1) and np.all(rec["phase"] == 0)
    rc, rec = lib_read(path, scale=1e-3, offset=(0.25, -0.5, 0.1))
    ref = olref.obj_read(path, scale=1e-3, offset=(0.25, -0.5, 0.1))
    assert np.array_equal(np.stack([rec["x"], rec["y"], rec["z"]], axis=1), ref)


def test_library_reader_3df_and_capacity(tmp_path):
    rng = np.random.default_rng(5)
    v = rng.standard_normal((37, 3)).astype(np.float32) * 1e-3
    p = tmp_path / "c.3df"
    p.write_bytes(np.int32(37).tobytes() + v.tobytes())
    rc, rec = lib_read(str(p))
    assert rc == 0
    assert np.array_equal(np.stack([rec["x"], rec["y"], rec["z"]], axis=1), olref.obj_read(str(p)))
    assert np.array_equal(olref.obj_read(str(p)), v.astype(np.float64))
    # a .3d name forces int32 decoding of the same bytes
    q = tmp_path / "c.3d"
    q.write_bytes(p.read_bytes())
    rc, rec = lib_read(str(q))
    assert rc == 0 and np.array_equal(rec["x"], v.view(np.int32)[:, 0].astype(np.float64))
    # cap < N: the first cap points, *n still N
    L = _lib.load()
    n = ctypes.c_int64()
    part = np.zeros(5, dtype=PT)
    assert L.olx_obj_read(os.fsencode(str(p)), 0, part.ctypes.data, 5, ctypes.byref(n), 1.0, 0, 0, 0) == 0
    assert n.value == 37 and np.array_equal(part["x"], v[:5, 0].astype(np.float64))


def test_library_reader_errors(tmp_path):
    assert lib_read(str(tmp_path / "none.3d"))[0] == _lib.EIO
    assert "cannot open" in _lib.last_error()
    bad = tmp_path / "a.xyz"
    bad.write_bytes(b"\0" * 16)
    assert lib_read(str(bad))[0] == _lib.EINVAL
    short = tmp_path / "s.3d"
    short.write_bytes(np.int32(10).tobytes() + np.zeros(5, np.int32).tobytes())
    assert lib_read(str(short))[0] == _lib.EIO
    assert "short read" in _lib.last_error()
    empty = tmp_path / "e.3d"
    empty.write_bytes(b"")
    assert lib_read(str(empty))[0] == _lib.EIO
    zero = tmp_path / "z.3d"
    zero.write_bytes(np.int32(0).tobytes())
    rc, rec = lib_read(str(zero))
    assert rc == 0 and rec.size == 0
    with pytest.raises(ValueError):
        olref.obj_read(str(bad))


# --------------------------------------------------------------------------- objset / objsort
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("zmiddle", [False, True])
def test_oracle_objset_invariants(dtype, zmiddle):
    """After objset the x / y box is centred on the offset, the binding extent equals
    its range, and z's minimum (ZMIN) or mid-range (ZMIDDLE) sits at oz."""
    xyz = olref.obj_read(HORN)
    off = (3.2e-4, 2.1e-4, 0.1)
    W, H = 1.0e-3, 4.0e-4
    Wf, Hf = float(np.float32(W)), float(np.float32(H))  # the ranges are float parameters
    tol = 8 * np.finfo(dtype).eps
    for y_range in (None, H):
        out = olref.objset(xyz, W, y_range, off, zmiddle, dtype)
        assert out.dtype == dtype
        mn, mx = out.min(axis=0).astype(np.float64), out.max(axis=0).astype(np.float64)
        ext = mx - mn
        assert abs((mx[0] + mn[0]) / 2 - off[0]) < tol * 1e-2
        assert abs((mx[1] + mn[1]) / 2 - off[1]) < tol * 1e-2
        if y_range is None:
            assert abs(max(ext[0], ext[1]) - Wf) < tol * Wf * 4
        else:
            assert ext[0] <= Wf * (1 + 4 * tol) and ext[1] <= Hf * (1 + 4 * tol)
            assert min(abs(ext[0] - Wf) / Wf, abs(ext[1] - Hf) / Hf) < 4 * tol
        zref = (mx[2] + mn[2]) / 2 if zmiddle else mn[2]
        assert abs(zref - off[2]) < tol
        # the scale is uniform: every axis scaled by the same factor
        src_ext = xyz.max(axis=0) - xyz.min(axis=0)
        s = ext / src_ext
        assert np.allclose(s, s[0], rtol=16 * tol)


def test_oracle_objset_float_differs_from_double():
    """Object<float> arithmetic is really float: results are float32 values and
    differ from the double path at float resolution (a mistake that computed in
    double and rounded once at the end would agree with the double path after
    rounding)."""
    xyz = olref.obj_read(HORN)
    a = olref.objset(xyz, 1e-3, 5e-4, (1e-4, 2e-4, 0.1), True, np.float32)
    b = olref.objset(xyz, 1e-3, 5e-4, (1e-4, 2e-4, 0.1), True, np.float64)
    assert a.dtype == np.float32
    d = np.abs(a.astype(np.float64) - b)
    assert d.max() < 1e-6 and np.count_nonzero(a.astype(np.float64) != b.astype(np.float32).astype(np.float64)) > 0


def test_oracle_objsort_order():
    z = np.array([0.0, 2.0, -0.0, 2.0, 1.0, -1.0, 0.0])
    o = olref.objsort_order(z)
    assert o.tolist() == [1, 3, 4, 0, 2, 6, 5]  # descending, ties (2, 2), (0, -0, 0) in input order
    c = olref.obj_read(CUBE)
    o = olref.objsort_order(c[:, 2])
    zs = c[o, 2]
    assert np.all(np.diff(zs) <= 0) and sorted(o.tolist()) == list(range(len(c)))
    for v in np.unique(zs):  # stability
        idx = o[zs == v]
        assert np.all(np.diff(idx) > 0)


# --------------------------------------------------------------------------- encodings
def test_oracle_rounding_and_conversion():
    r = olref.c_round([-0.5, 0.5, 1.5, -1.5, 2.5, 0.49999999999999994, -127.5, -0.0])
    assert r.tolist() == [-1, 1, 2, -2, 3, 0, -128, 0]
    u = olref.x86_u8([-1.0, 256.0, -128.0, 255.0, 3e9, np.nan, -2.0 ** 31, 2.0 ** 31, np.inf, -300.0])
    assert u.tolist() == [255, 0, 128, 255, 0, 0, 0, 0, 0, 212]


def test_oracle_phase8bit_known_answers():
    # (phi - pi) / 2pi * 255: -pi -> -255 -> 1; 0 -> -127.5 -> -128 -> 128; pi -> 0
    got = olref.phase8bit(np.array([-np.pi, 0.0, np.pi, np.pi / 2], np.float64))
    assert got.tolist() == [1, 128, 0, 192]  # pi/2: -63.75 -> -64 -> 192
    got = olref.phase8bit(np.array([-np.pi, 0.0, np.pi, np.pi / 2]), offset=True)
    assert got.tolist() == [0, 128, 255, 191]  # 191.25 -> 191


def test_oracle_quant8bit_known_answers():
    assert olref.quant8bit(np.array([0.0, 1.0, 0.5, 0.25], np.float32)).tolist() == [0, 255, 128, 64]
    assert olref.quant8bit(np.array([2.0, 2.0, 2.0])).tolist() == [0, 0, 0]       # 0 / 0
    assert olref.quant8bit(np.array([np.nan, 0.0, 1.0])).tolist() == [0, 0, 0]   # NaN seed sticks
    assert olref.quant8bit(np.array([0.0, np.nan, 1.0])).tolist() == [0, 0, 255]


def test_oracle_complex2img_known_answers():
    u = np.array([-1 + 0j, 1 + 0j, complex(-1, -0.0), 1j], np.complex64)
    # atan2: pi -> 255, 0 -> 127.5 -> 128, -pi -> 0, pi/2 -> 191.25 -> 191
    assert olref.complex2img(u, olref.IMG_PHASE).tolist() == [255, 128, 0, 191]
    assert olref.complex2img(np.array([-2, 0, 2], np.complex64), olref.IMG_REAL).tolist() == [0, 128, 255]
    assert olref.complex2img(np.array([-4, -2], np.complex64), olref.IMG_REAL_LEGACY).tolist() == [0, 128]
    assert olref.complex2img(np.array([-4, -2], np.complex64), olref.IMG_REAL).tolist() == [0, 255]
    assert olref.complex2img(np.array([1j, 3j, 2j], np.complex64), olref.IMG_IMAG).tolist() == [0, 255, 128]
    assert olref.complex2img(np.array([3 + 4j, 0, 6 + 8j], np.complex64), olref.IMG_AMP).tolist() == [128, 0, 255]
    assert olref.complex2real(np.array([3 + 4j], np.complex64), olref.IMG_AMP).tolist() == [5.0]
    assert olref.complex2real(np.array([1j], np.complex64), olref.IMG_PHASE)[0] == np.float32(np.pi / 2)
    assert olref.complex2real(np.array([1j], np.complex128), olref.IMG_PHASE).dtype == np.float64
    # complex<float>: the products, their sum and the root are float (the reference's
    # include/olutils.h:5 includes <math.h>, so unqualified sqrt / atan2 on floats are
    # the float overloads); the angle is a float
    e = np.float32(1 + 2.0 ** -12)
    assert olref.complex_value(np.array([complex(e, 0)], np.complex64), olref.IMG_AMP)[0] == \
        np.float64(np.sqrt(np.float32(e * e)))
    assert olref.complex_value(np.array([1 + 1j], np.complex64), olref.IMG_AMP)[0] == \
        np.float64(np.float32(np.sqrt(2.0)))
    assert olref.complex_value(np.array([1 + 1j], np.complex64), olref.IMG_PHASE)[0] == \
        np.float64(np.float32(np.pi / 4))
    assert olref.complex_value(np.array([1 + 1j], np.complex128), olref.IMG_AMP)[0] == np.sqrt(2.0)


# --------------------------------------------------------------------------- BMP
@pytest.mark.parametrize("shape", [(3, 5), (4, 8), (2, 6), (5, 7), (1, 1)])
def test_oracle_bmp_layout(shape):
    ny, nx = shape
    img = (np.arange(ny * nx, dtype=np.int64) * 37 % 256).astype(np.uint8).reshape(ny, nx)
    b = olref.bmp_bytes(img)
    assert len(b) == 1078 + ny * (nx + nx % 4)
    assert b[:2] == b"BM"
    assert int.from_bytes(b[2:6], "little") == 1078 + nx * ny
    assert int.from_bytes(b[10:14], "little") == 1078
    assert int.from_bytes(b[14:18], "little") == 40
    assert int.from_bytes(b[18:22], "little") == nx and int.from_bytes(b[22:26], "little") == ny
    assert int.from_bytes(b[26:28], "little") == 1 and int.from_bytes(b[28:30], "little") == 8
    assert all(b[54 + 4 * k: 58 + 4 * k] == bytes([k, k, k, 0]) for k in range(256))
    assert b[1078:1078 + nx] == img[0].tobytes()
    assert b[1078 + nx: 1078 + nx + nx % 4] == bytes(nx % 4)
    assert np.array_equal(olref.bmp_pixels(b, ny, nx), img)


@pytest.mark.parametrize("shape", [(3, 5), (4, 8), (2, 6), (5, 7), (64, 33)])
def test_library_bmp_matches_oracle(tmp_path, shape):
    ny, nx = shape
    img = np.random.default_rng(ny * nx).integers(0, 256, size=(ny, nx), dtype=np.uint8)
    p = str(tmp_path / "a.bmp")
    L = _lib.load()
    assert L.olx_bmp_write(os.fsencode(p), img.ctypes.data, ny, nx) == 0
    assert open(p, "rb").read() == olref.bmp_bytes(img)
    back = np.zeros_like(img)
    assert L.olx_bmp_read(os.fsencode(p), back.ctypes.data, ny, nx) == 0
    assert np.array_equal(back, img)
    # errors: missing file, file shorter than asked, bad sizes
    assert L.olx_bmp_read(os.fsencode(str(tmp_path / "none.bmp")), back.ctypes.data, ny, nx) == _lib.EIO
    big = np.zeros((ny + 1, nx), np.uint8)
    assert L.olx_bmp_read(os.fsencode(p), big.ctypes.data, ny + 1, nx) == _lib.EIO
    assert L.olx_bmp_write(os.fsencode(p), img.ctypes.data, 0, nx) == _lib.EINVAL
    assert L.olx_bmp_write(os.fsencode(str(tmp_path / "no" / "dir.bmp")), img.ctypes.data, ny, nx) == _lib.EIO
