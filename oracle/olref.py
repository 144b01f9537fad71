"""ctypes wrapper around the oracle (liboref.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker.  The product path (openol_amd) never imports it.

Each wrapper restates the reference function named in its docstring (see
oracle/olref.c for the line-by-line citations).  Arrays are numpy complex64 /
complex128, row-major (idx = h * nx + w, include/olutils.h:127 in the reference).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "libolref.so")

BANDLIMIT = 0x1
ADJOINT = 0x2
PREC_F32 = 0x4
FRESNEL_BL_CPU = 0x10
FRESNEL_BL_GPU = 0x20
CGH_EXACT_R = 0x100
WRP_SQUARE = 0x200
SHIFT_WINDOW = 0x1000
SHIFT_CPU_SCALE = 0x2000

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        L = ctypes.CDLL(_LIBPATH)
        i64, d, u32, p = ctypes.c_int64, ctypes.c_double, ctypes.c_uint, ctypes.c_void_p
        sig = {
            "olref_fft2d_c64": [p, i64, i64, ctypes.c_int],
            "olref_fft2d_c128": [p, i64, i64, ctypes.c_int],
            "olref_asm_c64": [p, i64, i64, d, d, d, d, u32],
            "olref_asm_c128": [p, i64, i64, d, d, d, d, u32],
            "olref_fresnel_c64": [p, i64, i64, d, d, d, d, u32],
            "olref_fresnel_c128": [p, i64, i64, d, d, d, d, u32],
            "olref_arss_c64": [p, p, i64, i64, d, d, d, d, d, d, d, u32],
            "olref_arss_c128": [p, p, i64, i64, d, d, d, d, d, d, d, u32],
            "olref_prop_c64": [p, i64, i64, d, d, d, d, ctypes.c_int, p],
            "olref_shifted_fresnel_c64": [p, i64, i64, d, d, d, d, d, d, d],
            "olref_shifted_fresnel_c128": [p, i64, i64, d, d, d, d, d, d, d],
            "olref_shifted_asm_c64": [p, p, i64, i64, d, d, d, d, d, d, u32],
            "olref_shifted_asm_c128": [p, p, i64, i64, d, d, d, d, d, d, u32],
            "olref_shifted_asm_transfer_c128": [p, i64, i64, d, d, d, d, d, d, u32],
            "olref_split_prop_c64": [p, i64, i64, d, d, d, d, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, u32],
            "olref_tilted_c128": [p, i64, i64, d, d, d, d, d, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int],
            "olref_diffuser_random_phase": [p, i64, i64, ctypes.c_int, u32, ctypes.c_float,
                                            ctypes.c_int, ctypes.c_int],
            "olref_diffuser_apply": [p, i64, i64, p, ctypes.c_float, ctypes.c_int, ctypes.c_int],
            "olref_lens_tile": [p, ctypes.c_int, p, i64, i64, ctypes.c_int, ctypes.c_int,
                                ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                ctypes.c_int],
            "olref_lens_block": [p, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_int, ctypes.c_float],
            "olref_lens_array": [p, i64, i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                 ctypes.c_float, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_float, ctypes.c_int],
            "olref_asm_prop_check": [i64, i64, d, d, d, d],
            "olref_fresnel_prop_check": [i64, i64, d, d, d, d],
            "olref_cgh_c64": [p, p, p, p, p, i64, p, i64, i64, d, d, d, u32],
            "olref_cgh_c128": [p, p, p, p, p, i64, p, i64, i64, d, d, d, u32],
            "olref_cgh_pixels_c128": [p, p, p, p, p, i64, p, p, i64, p, d, d, d, u32],
            "olref_wrp_c64": [p, p, p, i64, p, i64, i64, d, d, d, d, u32],
            "olref_cgh_fresnel_diff": [p, p, p, p, p, i64, i64, i64, d, d, p],
            "olref_wrp_c128": [p, p, p, i64, p, i64, i64, d, d, d, d, u32],
            "olref_wrp_pixels_c128": [p, p, p, i64, p, p, i64, p, d, d, d, d, u32],
            "olref_wrp_plane_z": [i64, i64, d, d, d, d, ctypes.c_float, ctypes.c_int],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.olref_wrp_plane_z.restype = ctypes.c_double
        L.olref_mt19937_nth.argtypes = [u32, i64]
        L.olref_mt19937_nth.restype = ctypes.c_uint32
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _cplx(u, dtype):
    return np.ascontiguousarray(u, dtype=dtype)


def _check(rc, name):
    if rc != 0:
        raise ValueError(f"{name}: invalid arguments (rc={rc})")


def fft2(u, sign=-1):
    """Unnormalised 2-D FFT (olutils.h:19-66 semantics), returns a new array."""
    u = np.array(u, copy=True)
    fn = lib().olref_fft2d_c64 if u.dtype == np.complex64 else lib().olref_fft2d_c128
    if u.dtype not in (np.complex64, np.complex128):
        u = u.astype(np.complex128)
        fn = lib().olref_fft2d_c128
    _check(fn(_ptr(u), u.shape[0], u.shape[1], sign), "fft2")
    return u


def asm_prop(u, dy, dx, lam, z, flags=0):
    """AsmProp, include/olprop.h:227-248.  complex64 -> fftw3f path, complex128 ->
    fftw3 path; PREC_T = double unless flags has PREC_F32."""
    u = np.array(u, copy=True)
    fn = lib().olref_asm_c64 if u.dtype == np.complex64 else lib().olref_asm_c128
    _check(fn(_ptr(u), u.shape[0], u.shape[1], dy, dx, lam, z, flags), "asm_prop")
    return u


def fresnel_prop(u, dy, dx, lam, z, flags=0):
    """FresnelProp / FresnelPropBandLimit, include/olprop.h:125-170 (flags pick the
    CPU or GPU band-limit window)."""
    u = np.array(u, copy=True)
    fn = lib().olref_fresnel_c64 if u.dtype == np.complex64 else lib().olref_fresnel_c128
    _check(fn(_ptr(u), u.shape[0], u.shape[1], dy, dx, lam, z, flags), "fresnel_prop")
    return u


def arss_prop(u, dy, dx, lam, z, s, oy=0.0, ox=0.0, flags=0):
    """ARSSFresnelProp, include/golprop.h:941-995."""
    src = np.ascontiguousarray(u)
    dst = np.empty_like(src)
    fn = lib().olref_arss_c64 if src.dtype == np.complex64 else lib().olref_arss_c128
    _check(fn(_ptr(src), _ptr(dst), src.shape[0], src.shape[1], dy, dx, lam, z, s, oy, ox,
              flags), "arss_prop")
    return dst


def shifted_fresnel_prop(u, dy, dx, lam, z, s, oy=0.0, ox=0.0):
    """shiftedFresnelProp (CPU ARSS analogue), include/olprop.h:338-408."""
    u = np.array(u, copy=True)
    fn = lib().olref_shifted_fresnel_c64 if u.dtype == np.complex64 else lib().olref_shifted_fresnel_c128
    _check(fn(_ptr(u), u.shape[0], u.shape[1], dy, dx, lam, z, s, oy, ox), "shifted_fresnel_prop")
    return u


def shifted_asm(u, dy, dx, lam, z, oy, ox, flags=0):
    """shiftedAsmProp (include/olprop.h:307-330, flags SHIFT_WINDOW | SHIFT_CPU_SCALE)
    or gshiftedAsmProp (include/golprop.h:536-576, flags 0)."""
    src = _cplx(u, u.dtype if u.dtype == np.complex128 else np.complex64)
    out = np.empty_like(src)
    fn = lib().olref_shifted_asm_c128 if src.dtype == np.complex128 else lib().olref_shifted_asm_c64
    _check(fn(_ptr(src), _ptr(out), src.shape[0], src.shape[1], dy, dx, lam, z, oy, ox, flags),
           "shifted_asm")
    return out


def shifted_asm_transfer(ny, nx, dv, du, lam, z, oy, ox, flags=0):
    """The centred table of shiftedAsmTransferF (include/olprop.h:250-305, flags
    SHIFT_WINDOW) or gshiftedAsmTransferF (include/golprop.h:336-419, flags 0),
    complex128; evanescent entries 0."""
    H = np.empty((ny, nx), np.complex128)
    _check(lib().olref_shifted_asm_transfer_c128(_ptr(H), ny, nx, dv, du, lam, z, oy, ox, flags),
           "shifted_asm_transfer")
    return H


def split_prop(u, dy, dx, lam, z, by=2, bx=2, mode=0, flags=0):
    """splitAsmProp (include/olprop.h:614-651; mode 0 with SHIFT_WINDOW |
    SHIFT_CPU_SCALE) or gsplitProp (include/golprop.h:773-813; mode 0 ASM, 1 FRESNEL)."""
    v = _cplx(u, np.complex64).copy()
    _check(lib().olref_split_prop_c64(_ptr(v), v.shape[0], v.shape[1], dy, dx, lam, z, mode, by,
                                      bx, flags), "split_prop")
    return v


def tilted(u, dy, dx, lam, z, angle, interp="bilinear", mode="asm", kind="source"):
    """TiltedSourceProp (include/olprop.h:414-526) or TiltedReferenceProp (:528-594)
    on complex128 data; returns the propagated field."""
    v = _cplx(u, np.complex128).copy()
    _check(lib().olref_tilted_c128(_ptr(v), v.shape[0], v.shape[1], dy, dx, lam, z, angle,
                                   {"bilinear": 0, "bicubic": 1}[interp],
                                   {"asm": 0, "fresnel": 1}[mode],
                                   {"source": 0, "reference": 1}[kind]), "tilted")
    return v


PLAY, RECORD = 1, 0
DIFFUSER_KINDS = {"phase": 0, "xphase": 1, "yphase": 2, "bitrandom": 3, "block": 4, "blockstep": 5,
                  "bitphase": 6}


def mt19937_nth(seed, n):
    """n-th output of MT19937 seeded with seed (std::mt19937)."""
    return int(lib().olref_mt19937_nth(seed, n))


def diffuser_random_phase(u, step, seed=1, rng=2 * np.pi, oy=0, ox=0):
    """Diffuser::random_phase, CPU form (include/Diffuser.h:191-215)."""
    v = _cplx(u, np.complex64).copy()
    _check(lib().olref_diffuser_random_phase(_ptr(v), v.shape[0], v.shape[1], step, seed, rng, oy,
                                             ox), "diffuser_random_phase")
    return v


def diffuser_apply(u, b, scalar, kind, blocksize=1):
    """The GPU diffuser kernels (include/Diffuser.h:14-129) given uniforms b."""
    v = _cplx(u, np.complex64).copy()
    b = np.ascontiguousarray(b, dtype=np.float32)
    _check(lib().olref_diffuser_apply(_ptr(v), v.shape[0], v.shape[1], _ptr(b), scalar,
                                      DIFFUSER_KINDS[kind], blocksize), "diffuser_apply")
    return v


def lens_array(height, width, pitch, blockwidth, f, lam, step=PLAY, on=0, om=0, p_limit=0.0,
               circle=False):
    """Lensarry::gen_complex(_limit) / LensarryCircle lens field (include/Lensarry.h)."""
    out = np.empty((height, width), dtype=np.complex64)
    _check(lib().olref_lens_array(_ptr(out), height, width, pitch, blockwidth, f, lam, step, on, om,
                                  p_limit, 1 if circle else 0), "lens_array")
    return out


def lens_block(bp, pitch, f, lam, step=PLAY, p_limit=0.0):
    """The bp x bp lenslet of Lensarry::gen_complex(_limit) (FresnelResponse at -+f)."""
    out = np.empty((bp, bp), dtype=np.complex64)
    _check(lib().olref_lens_block(_ptr(out), bp, pitch, f, lam, step, p_limit), "lens_block")
    return out


def lens_tile(block, height, width, on=0, om=0, pitch=0.0, f=0.0, lam=0.0, p_limit=0.0, kind=0):
    """Lensarry::duplicate / duplicate_random / duplicate_multrandom (kind 0 / 1 / 2),
    include/Lensarry.h:137-252."""
    block = np.ascontiguousarray(block, dtype=np.complex64)
    out = np.empty((height, width), dtype=np.complex64)
    _check(lib().olref_lens_tile(_ptr(block), block.shape[0], _ptr(out), height, width, on, om,
                                 pitch, f, lam, p_limit, kind), "lens_tile")
    return out


def prop(u, dy, dx, lam, z, mode=2):
    """Prop dispatcher, include/olprop.h:656-682 (mode 0 ASM, 1 FRESNEL, 2 AUTO).
    Returns (field, chosen) with chosen 0 ASM, 1 Fresnel, 2 Fresnel band-limited."""
    u = _cplx(u, np.complex64).copy()
    ch = ctypes.c_int(-1)
    _check(lib().olref_prop_c64(_ptr(u), u.shape[0], u.shape[1], dy, dx, lam, z, mode,
                                ctypes.byref(ch)), "prop")
    return u, ch.value


def asm_prop_check(ny, nx, dy, dx, lam, z):
    """AsmPropCheck, include/olprop.h:209-225."""
    return bool(lib().olref_asm_prop_check(ny, nx, dy, dx, lam, z))


def fresnel_prop_check(ny, nx, dy, dx, lam, z):
    """FresnelPropCheck, include/olprop.h:109-123."""
    return bool(lib().olref_fresnel_prop_check(ny, nx, dy, dx, lam, z))


def _pts(x, y, z, amp=None, phase=None):
    f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
    return f(x), f(y), f(z), f(amp), f(phase)


def cgh(x, y, z, ny, nx, py, px, lam, amp=None, phase=None, flags=0, dtype=np.complex64):
    """_CghFresnel (complex out), include/olcgh.h:296-313; amp/phase/exact-r are the
    documented extension (A_j, phi_j, spherical r)."""
    x, y, z, amp, phase = _pts(x, y, z, amp, phase)
    U = np.empty((ny, nx), dtype=dtype)
    fn = lib().olref_cgh_c64 if dtype == np.complex64 else lib().olref_cgh_c128
    _check(fn(_ptr(x), _ptr(y), _ptr(z), _ptr(amp), _ptr(phase), len(x), _ptr(U), ny, nx,
              py, px, lam, flags), "cgh")
    return U


def cgh_normalized(x, y, z, ny, nx, p, lam):
    """_CghFresnelNormalized, include/olcgh.h:213-231 (what Cgh(obj, phase, ny, nx, p,
    lambda, true) runs): point coordinates in pixels, dx = p (w - x_j), dy = p (h - x_j)
    (the reference reads x_j for both), dz = p z_j, theta = 2 pi (dx^2 + dy^2) / (2 dz
    lambda), unit amplitude; returns the phase atan2(Im U, Re U).  FP64, numpy, for
    small sizes."""
    x = np.asarray(x, np.float64)
    z = np.asarray(z, np.float64)
    w = np.arange(nx, dtype=np.float64)[None, :]
    h = np.arange(ny, dtype=np.float64)[:, None]
    U = np.zeros((ny, nx), np.complex128)
    for xj, zj in zip(x, z):
        dx = p * (w - xj)
        dy = p * (h - xj)
        dz = p * zj
        U += np.exp(1j * (2 * np.pi * (dx * dx + dy * dy) / (2 * dz * lam)))
    return np.arctan2(U.imag, U.real)


def cgh_pixels(x, y, z, ph, pw, py, px, lam, amp=None, phase=None, flags=0):
    """FP64 CGH sum at the listed pixels only (sampled check at configuration size)."""
    x, y, z, amp, phase = _pts(x, y, z, amp, phase)
    ph = np.ascontiguousarray(ph, dtype=np.int64)
    pw = np.ascontiguousarray(pw, dtype=np.int64)
    out = np.empty(len(ph), dtype=np.complex128)
    _check(lib().olref_cgh_pixels_c128(_ptr(x), _ptr(y), _ptr(z), _ptr(amp), _ptr(phase),
                                       len(x), _ptr(ph), _ptr(pw), len(ph), _ptr(out), py, px,
                                       lam, flags), "cgh_pixels")
    return out


def cgh_fresnel_diff(x, y, z, ny, nx, p, lam, amp=None, phase=None):
    """_CghFresnelDiff (include/olcgh.h:315-349), the recurrence carried along each row
    in FP64; complex128 field (the reference outputs its atan2)."""
    x, y, z, amp, phase = _pts(x, y, z, amp, phase)
    out = np.empty((ny, nx), dtype=np.complex128)
    _check(lib().olref_cgh_fresnel_diff(_ptr(x), _ptr(y), _ptr(z), _ptr(amp), _ptr(phase), len(x),
                                        ny, nx, p, lam, _ptr(out)), "cgh_fresnel_diff")
    return out


def wrp(x, y, z, ny, nx, p, lam, plane_z, p_limit=None, square=False, dtype=np.complex64):
    """WRPStep1 (circle, golcgh.h:291-378) or WRPStep1_D (square, :391-491)."""
    x, y, z, _, _ = _pts(x, y, z)
    U = np.empty((ny, nx), dtype=dtype)
    fn = lib().olref_wrp_c64 if dtype == np.complex64 else lib().olref_wrp_c128
    _check(fn(_ptr(x), _ptr(y), _ptr(z), len(x), _ptr(U), ny, nx, p, lam, plane_z,
              p if p_limit is None else p_limit, WRP_SQUARE if square else 0), "wrp")
    return U


def wrp_pixels(x, y, z, ph, pw, p, lam, plane_z, p_limit=None, square=False):
    """The FP64 WRP sum (WRPStep1 / WRPStep1_D windows) at the listed pixels only."""
    x, y, z, _, _ = _pts(x, y, z)
    ph = np.ascontiguousarray(ph, dtype=np.int64)
    pw = np.ascontiguousarray(pw, dtype=np.int64)
    out = np.empty(len(ph), dtype=np.complex128)
    _check(lib().olref_wrp_pixels_c128(_ptr(x), _ptr(y), _ptr(z), len(x), _ptr(ph), _ptr(pw),
                                       len(ph), _ptr(out), p, lam, plane_z,
                                       p if p_limit is None else p_limit,
                                       WRP_SQUARE if square else 0), "wrp_pixels")
    return out


def wrp_plane_z(ny, nx, p, lam, zmin, s=0.9, p_limit=None, square=False):
    """WRPMethod / WRPMethod_D plane position, include/golcgh.h:519-581."""
    return lib().olref_wrp_plane_z(ny, nx, p, lam, zmin, p if p_limit is None else p_limit,
                                   s, 1 if square else 0)


def rel_l2(a, b):
    """Relative L2 error ||a - b|| / ||b|| in float64."""
    a = np.asarray(a, dtype=np.complex128)
    b = np.asarray(b, dtype=np.complex128)
    nb = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0))


# ---------------------------------------------------------------------------
# Point-cloud ingest and 8-bit encodings: numpy restatements (byte and integer
# work, plus scalar float arithmetic done one IEEE operation at a time).
# ---------------------------------------------------------------------------
import struct as _struct


def obj_read(path, scale=1.0, offset=(0.0, 0.0, 0.0)):
    """objread (include/olobject.h:79-91): dispatch on the extension to objread_3d
    (:9-42, int32 N then N x int32 x, y, z) or objread_3df (:44-77, float triples),
    read native (little-endian) order.  With scale / offset: x = x_int scale + o in
    double (:93-126 with SCALE_T = OFFSET_T = double).  Returns an (N, 3) float64
    array."""
    ext = os.path.splitext(str(path))[1]
    if ext not in (".3d", ".3df"):
        raise ValueError("extension is neither .3d nor .3df")
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], dtype="<i4")[0])
    vals = np.frombuffer(raw[4:4 + 12 * n], dtype="<i4" if ext == ".3d" else "<f4")
    if vals.size != 3 * n:
        raise ValueError("short file")
    xyz = vals.reshape(n, 3).astype(np.float64)
    if scale != 1.0 or any(o != 0.0 for o in offset):
        xyz = xyz * np.float64(scale) + np.asarray(offset, dtype=np.float64)[None, :]
    return xyz


def objset(xyz, x_range, y_range=None, offset=(0.0, 0.0, 0.0), zmiddle=False, dtype=np.float64):
    """objset in the object's element type (dtype float32 = Object<float>, float64 =
    Object<double>).  One range: include/olobject.h:196-228 (scale = range /
    max(range_x, range_y)); two ranges: :232-264 (scale = min(x_range / range_x,
    y_range / range_y)).  The ranges are float parameters.  Offsets are shifted by
    the centre of x and y and by z's minimum (ZMIN, :215-216) or mid-range
    (ZMIDDLE, :218-219); then every point becomes p scale + offset (:223-227)."""
    T = dtype
    p = np.asarray(xyz, dtype=T)
    mn = p.min(axis=0)
    mx = p.max(axis=0)
    f32 = np.float32
    rx, ry = T(mx[0] - mn[0]), T(mx[1] - mn[1])
    if y_range is None:
        scale = rx if rx > ry else ry
        scale = T(T(f32(x_range)) / scale)
    else:
        sx = T(T(f32(x_range)) / rx)
        sy = T(T(f32(y_range)) / ry)
        scale = sx if sx < sy else sy
    o = [T(v) for v in offset]
    o[0] = T(o[0] - T(T(T(mx[0] + mn[0]) / T(2)) * scale))
    o[1] = T(o[1] - T(T(T(mx[1] + mn[1]) / T(2)) * scale))
    if zmiddle:
        o[2] = T(o[2] - T(T(T(mx[2] + mn[2]) / T(2)) * scale))
    else:
        o[2] = T(o[2] - T(mn[2] * scale))
    out = np.empty_like(p)
    for c in range(3):
        out[:, c] = (p[:, c] * scale).astype(T) + o[c]
    return out


def objsort_order(z):
    """objsort (include/olobject.h:176-181): descending z.  std::sort is not stable;
    the order returned keeps equal depths in input order (one of the orders it may
    produce).  Returns the permutation."""
    z = np.asarray(z, dtype=np.float64)
    return np.argsort(-z, kind="stable")


def c_round(t):
    """C round(): half away from zero (numpy's round is half to even)."""
    t = np.asarray(t, dtype=np.float64)
    r = np.trunc(t)
    with np.errstate(invalid="ignore"):
        bump = np.abs(t - r) >= 0.5
    return r + np.where(bump, np.sign(t), 0.0)


def x86_u8(v):
    """The reference's x86-64 conversion of a double to uint8 (cvttsd2si to int32, then
    the low byte): modulo 256 inside the int32 range, 0 for NaN and outside it."""
    v = np.asarray(v, dtype=np.float64)
    out = np.zeros(v.shape, dtype=np.uint8)
    with np.errstate(invalid="ignore"):
        ok = (v >= -2147483648.0) & (v < 2147483648.0)
    out[ok] = (v[ok].astype(np.int64) & 0xFF).astype(np.uint8)
    return out


def phase8bit(phase, offset=False):
    """phase8bit (include/olimg.h:206-213): round((phi - pi) / (2 pi) * 255) stored
    to uint8; offset=True uses phi + pi (no reference counterpart)."""
    t = np.asarray(phase).astype(np.float64)
    t = t + np.pi if offset else t - np.pi
    return x86_u8(c_round((t / (2 * np.pi)) * 255.0))


def _ref_minmax(v, max_start=None):
    """The reference's running min / max seeded with element 0, so a NaN there sticks
    (every output is then 0) and later NaNs are skipped.  max_start replaces the max
    seed (complex2img's bool overload starts it at 0, include/olimg.h:331)."""
    v = np.asarray(v, dtype=np.float64).ravel()
    if np.isnan(v[0]):
        return np.nan, np.nan
    mn, mx = np.nanmin(v), np.nanmax(v)
    if max_start is not None:
        mx = max(mx, max_start)
    return mn, mx


def quant8bit(real):
    """quant8bit (include/olimg.h:178-198): round((v - min) / (max - min) * 255)."""
    v = np.asarray(real).astype(np.float64)
    mn, mx = _ref_minmax(v)
    with np.errstate(invalid="ignore", divide="ignore"):
        return x86_u8(c_round(((v - mn) / (mx - mn)) * 255.0))


IMG_PHASE, IMG_REAL, IMG_IMAG, IMG_AMP, IMG_REAL_LEGACY = 0, 1, 2, 3, 4


def complex_value(u, mode):
    """complex2img / complex2real per-element value (include/olimg.h:355-366,
    399-410): real, imag, sqrt(re re + im im), atan2(im, re).  The reference's headers
    include <math.h> (include/olutils.h:5) before these templates, so for
    complex<float> the unqualified sqrt and atan2 are the float overloads: the
    products, sum and root are float and the angle is a float (taken here as the
    correctly rounded one: the FP64 angle rounded to float); complex<double> is FP64
    throughout."""
    u = np.asarray(u)
    if mode in (IMG_REAL, IMG_REAL_LEGACY):
        return u.real.astype(np.float64)
    if mode == IMG_IMAG:
        return u.imag.astype(np.float64)
    if u.dtype == np.complex64:
        re, im = u.real, u.imag
        if mode == IMG_AMP:
            return np.sqrt(re * re + im * im).astype(np.float64)   # float32 throughout
        return np.arctan2(im.astype(np.float64), re.astype(np.float64)).astype(np.float32).astype(np.float64)
    if mode == IMG_AMP:
        return np.sqrt(u.real * u.real + u.imag * u.imag)
    return np.arctan2(u.imag, u.real)


def complex2img(u, mode):
    """complex2img: MODE overload (include/olimg.h:348-391; PHASE range fixed to
    [-pi, pi], others min / max seeded with element 0) or the bool overload's real
    mode (IMG_REAL_LEGACY, :314-346: max seeded with 0).  img = round(255 (v - min) /
    (max - min))."""
    v = complex_value(u, mode)
    if mode == IMG_PHASE:
        mn, mx = -np.pi, np.pi
    else:
        mn, mx = _ref_minmax(v, 0.0 if mode == IMG_REAL_LEGACY else None)
    with np.errstate(invalid="ignore", divide="ignore"):
        return x86_u8(c_round((255.0 * (v - mn)) / (mx - mn)))


def complex2real(u, mode):
    """complex2real (include/olimg.h:393-413) in the field's real type."""
    u = np.asarray(u)
    rt = np.float32 if u.dtype == np.complex64 else np.float64
    return complex_value(u, mode).astype(rt)


def bmp_bytes(img):
    """bmpwrite (include/olimg.h:136-165): 14-byte file header (type 19778, size
    1078 + nx ny, offset 1078), 40-byte info header (nx, ny, 1 plane, 8 bits), a grey
    256-entry palette, then each row followed by nx % 4 zero bytes."""
    img = np.ascontiguousarray(img, dtype=np.uint8)
    ny, nx = img.shape
    out = bytearray()
    out += _struct.pack("<HIHHI", 19778, (14 + 40 + 1024 + nx * ny) & 0xFFFFFFFF, 0, 0, 14 + 40 + 1024)
    out += _struct.pack("<IiiHHIIiiII", 40, nx, ny, 1, 8, 0, 0, 0, 0, 0, 0)
    out += bytes(b for i in range(256) for b in (i, i, i, 0))
    pad = bytes(nx % 4)
    for r in range(ny):
        out += img[r].tobytes() + pad
    return bytes(out)


def bmp_pixels(data, ny, nx):
    """bmpread (include/olimg.h:56-99): skip 14 + 40 + 1024 bytes, then ny rows of nx
    bytes, each followed by nx % 4 skipped bytes."""
    off, pad = 14 + 40 + 1024, nx % 4
    rows = [np.frombuffer(data, dtype=np.uint8, count=nx, offset=off + r * (nx + pad)) for r in range(ny)]
    return np.stack(rows)
