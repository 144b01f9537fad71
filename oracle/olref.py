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
            "olref_asm_prop_check": [i64, i64, d, d, d, d],
            "olref_fresnel_prop_check": [i64, i64, d, d, d, d],
            "olref_cgh_c64": [p, p, p, p, p, i64, p, i64, i64, d, d, d, u32],
            "olref_cgh_c128": [p, p, p, p, p, i64, p, i64, i64, d, d, d, u32],
            "olref_cgh_pixels_c128": [p, p, p, p, p, i64, p, p, i64, p, d, d, d, u32],
            "olref_wrp_c64": [p, p, p, i64, p, i64, i64, d, d, d, d, u32],
            "olref_wrp_c128": [p, p, p, i64, p, i64, i64, d, d, d, d, u32],
            "olref_wrp_plane_z": [i64, i64, d, d, d, d, ctypes.c_float, ctypes.c_int],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.olref_wrp_plane_z.restype = ctypes.c_double
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


def wrp(x, y, z, ny, nx, p, lam, plane_z, p_limit=None, square=False, dtype=np.complex64):
    """WRPStep1 (circle, golcgh.h:291-378) or WRPStep1_D (square, :391-491)."""
    x, y, z, _, _ = _pts(x, y, z)
    U = np.empty((ny, nx), dtype=dtype)
    fn = lib().olref_wrp_c64 if dtype == np.complex64 else lib().olref_wrp_c128
    _check(fn(_ptr(x), _ptr(y), _ptr(z), len(x), _ptr(U), ny, nx, p, lam, plane_z,
              p if p_limit is None else p_limit, WRP_SQUARE if square else 0), "wrp")
    return U


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
