"""Python mirror of OpenOL's diffraction / CGH operators on MI355X.

Same names, argument order and meaning as the reference's C++ templates
(include/olprop.h, include/olcgh.h, include/golprop.h, include/golcgh.h,
include/olobject.h, include/olimg.h), on
torch tensors that live on the GPU.  Fields are complex64 tensors of shape
(height, width), row-major (idx = h * width + w), updated in place like the
reference's unique_ptr buffers.  Every call goes through the C ABI of libolx.so;
errors raise (OlxInvalid for bad arguments) instead of exiting the process.
"""
import ctypes
import os

import numpy as np
import torch

from . import _lib
from ._lib import OlxInvalid, check

ASM, FRESNEL, AUTO = _lib.MODE_ASM, _lib.MODE_FRESNEL, _lib.MODE_AUTO
FRESNEL_BL = _lib.MODE_FRESNEL_BL

_p = ctypes.c_void_p


def _ptr(t):
    return _p(t.data_ptr())


def _field(u, height, width, name="u", c128=False):
    """Checks a field tensor; with c128, complex128 is accepted too and the entry-point
    suffix ("c64" / "c128") is returned."""
    if not isinstance(u, torch.Tensor) or not u.is_cuda:
        raise _lib.OlxInvalid(_lib.EINVAL, f"{name} must be a CUDA (HIP) tensor")
    ok = (torch.complex64, torch.complex128) if c128 else (torch.complex64,)
    if u.dtype not in ok:
        raise _lib.OlxInvalid(_lib.EINVAL, f"{name} must be {' or '.join(map(str, ok))}, got {u.dtype}")
    if not u.is_contiguous():
        raise _lib.OlxInvalid(_lib.EINVAL, f"{name} must be contiguous")
    if u.dim() != 2 or tuple(u.shape) != (int(height), int(width)):
        raise _lib.OlxInvalid(_lib.EINVAL, f"{name} shape {tuple(u.shape)} != ({height}, {width})")
    return ("c128" if u.dtype == torch.complex128 else "c64") if c128 else u


# ------------------------------------------------------------------ propagation
def AsmProp(u, height, width, dy, dx, lam, z, bandlimit=False, adjoint=False, prec_f32=False):
    """ol::AsmProp (include/olprop.h:227-248): in place, net scale 1.  prec_f32 computes
    the transfer function with the reference's PREC_T = float arithmetic (complex64).
    A complex128 field runs the FP64 passes (the reference's fftw3 path)."""
    suf = _field(u, height, width, c128=True)
    if suf == "c128" and prec_f32:
        raise _lib.OlxInvalid(_lib.EINVAL, "AsmProp: prec_f32 takes complex64 fields")
    flags = (_lib.BANDLIMIT if bandlimit else 0) | (_lib.ADJOINT if adjoint else 0) | \
        (_lib.PREC_F32 if prec_f32 else 0)
    check(_lib.call(f"olx_asm_{suf}", u.device.index, _ptr(u), height, width, dy, dx, lam, z,
                    flags), "AsmProp")
    return u


def multiplane_mse(phi, targets, zs, dy, dx, lam, bandlimit=True, grad=True):
    """Mean-normalised multi-plane intensity loss of the phase-only field exp(i phi) and,
    with grad, its gradient dL/dphi, in one native pass (no reference counterpart):
    L = sum_p mean((I_p / mean I_p - T_p / mean T_p)^2), I_p = |AsmProp_{z_p}(exp(i phi))|^2.
    Returns (loss as a float64 0-d tensor on the device, gradient or None)."""
    if phi.dtype != torch.float32 or phi.dim() != 2 or not phi.is_cuda or not phi.is_contiguous():
        raise OlxInvalid(_lib.EINVAL, "multiplane_mse: phi must be a contiguous 2-D float32 CUDA tensor")
    if len(targets) != len(zs) or not zs:
        raise OlxInvalid(_lib.EINVAL, "multiplane_mse: one target per plane")
    for t in targets:
        if t.dtype != torch.float32 or t.shape != phi.shape or t.device != phi.device \
                or not t.is_contiguous():
            raise OlxInvalid(_lib.EINVAL, "multiplane_mse: targets must be contiguous float32 "
                             "tensors shaped like phi on its device")
    ny, nx = phi.shape
    loss = torch.empty((), dtype=torch.float64, device=phi.device)
    g = torch.empty_like(phi) if grad else None
    zarr = (ctypes.c_double * len(zs))(*[float(z) for z in zs])
    tarr = (ctypes.c_void_p * len(targets))(*[t.data_ptr() for t in targets])
    check(_lib.call("olx_multiplane_mse_c64", phi.device.index, _ptr(phi), ny, nx, len(zs), ctypes.cast(zarr, ctypes.c_void_p),
        ctypes.cast(tarr, ctypes.c_void_p), dy, dx, lam, _lib.BANDLIMIT if bandlimit else 0,
        _ptr(loss), _ptr(g) if grad else None), "multiplane_mse")
    return loss, g


def mAsmProp(u, height, width, dy, dx, lam, z):
    """ol::mAsmProp (include/molprop.h:80-101): the angular spectrum propagation whose
    transforms the reference stages through host memory, net scale 1/M^2 (it scales
    before and after the inverse transform).  The field stays in HBM here."""
    suf = _field(u, height, width, c128=True)
    check(_lib.call(f"olx_asm_{suf}", u.device.index, _ptr(u), height, width, dy, dx, lam, z,
                    _lib.EXTRA_INV_M2), "mAsmProp")
    return u


def FresnelProp(u, height, width, dy, dx, lam, z, adjoint=False):
    """ol::FresnelProp (include/olprop.h:125-148): circconv(pad u, h) / M^2, in place."""
    suf = _field(u, height, width, c128=True)
    check(_lib.call(f"olx_fresnel_{suf}", u.device.index, _ptr(u), height, width, dy, dx, lam, z,
                    _lib.ADJOINT if adjoint else 0), "FresnelProp")
    return u


def FresnelPropBandLimit(u, height, width, dy, dx, lam, z, variant="cpu"):
    """ol::FresnelPropBandLimit (include/olprop.h:150-170, per-axis window) or, with
    variant="gpu", gFresnelPropBandLimit (include/golprop.h:258-302, x-limit on both axes)."""
    suf = _field(u, height, width, c128=True)
    flags = {"cpu": _lib.FRESNEL_BL_CPU, "gpu": _lib.FRESNEL_BL_GPU}[variant]
    check(_lib.call(f"olx_fresnel_{suf}", u.device.index, _ptr(u), height, width, dy, dx, lam, z,
                    flags), "FresnelPropBandLimit")
    return u


def gFresnelPropBandLimit(u, height, width, dy, dx, lam, z):
    return FresnelPropBandLimit(u, height, width, dy, dx, lam, z, variant="gpu")


def prop_batch(u, dy, dx, lam, z, propmode=ASM, flags=0, pre=None, post=None, img=None,
               img_offset=False):
    """Every field of the (nb, ny, nx) complex64 tensor u, in place, replaced by
    post * Prop(pre * u[k]): Prop is AsmProp (propmode ASM, flags BANDLIMIT / ADJOINT /
    PREC_F32) or FresnelProp (FRESNEL, flags FRESNEL_BL_CPU / FRESNEL_BL_GPU / ADJOINT).
    pre / post are (ny, nx) complex64 factor fields shared by every field (a lens array,
    a diffuser's phase, an aperture; None: 1), multiplied in as the first pass loads
    and the last pass stores, where the reference multiplies in passes of their own
    (Lensarry::mult_complex, include/Lensarry.h:262-270).  img, a (nb, ny, nx) uint8
    tensor, also receives the 8-bit phase of every output value (phase8bit,
    include/olimg.h:205-213; img_offset: the +pi form) from the last pass."""
    if not isinstance(u, torch.Tensor) or u.dim() != 3:
        raise _lib.OlxInvalid(_lib.EINVAL, "prop_batch: u must be a (nb, ny, nx) tensor")
    nb, ny, nx = u.shape
    _field(u.view(nb * ny, nx), nb * ny, nx)
    for f, name in ((pre, "pre"), (post, "post")):
        if f is not None:
            _field(f, ny, nx, name)
            if f.device != u.device:
                raise _lib.OlxInvalid(_lib.EINVAL, f"prop_batch: {name} on {f.device}, u on {u.device}")
    if img is not None:
        _out2d(img.view(nb * ny, nx) if img.dim() == 3 and tuple(img.shape) == (nb, ny, nx)
               else img, nb * ny, nx, torch.uint8, "img")
    if propmode not in (ASM, FRESNEL):
        raise _lib.OlxInvalid(_lib.EINVAL, "prop_batch: propmode must be ASM or FRESNEL")
    nul = lambda t: None if t is None else _ptr(t)
    check(_lib.call("olx_prop_batch_c64", u.device.index, propmode, _ptr(u), nb, ny, nx, dy, dx,
                    lam, z, flags, nul(pre), nul(post), nul(img),
                    _lib.PHASE8_OFFSET if img_offset else 0), "prop_batch")
    return u


def ARSSFresnelProp(usrc, udst, ny, nx, dy, dx, lam, z, s, oy=0.0, ox=0.0):
    """ol::ARSSFresnelProp (include/golprop.h:941-995); usrc and udst may alias."""
    _field(usrc, ny, nx, "usrc")
    _field(udst, ny, nx, "udst")
    check(_lib.call("olx_arss_c64", usrc.device.index, _ptr(usrc), _ptr(udst), ny, nx, dy, dx,
                         lam, z, s, oy, ox, 0), "ARSSFresnelProp")
    return udst


def _shifted(usrc, udst, ny, nx, dy, dx, lam, z, oy, ox, flags, name):
    _field(usrc, ny, nx, "usrc")
    _field(udst, ny, nx, "udst")
    check(_lib.call("olx_shifted_asm_c64", usrc.device.index, _ptr(usrc), _ptr(udst), ny,
                                          nx, dy, dx, lam, z, oy, ox, flags), name)
    return udst


def shiftedAsmProp(usrc, udst, ny, nx, dy, dx, lam, z, oy, ox, adjoint=False):
    """ol::shiftedAsmProp (include/olprop.h:307-330): off-axis ASM to the window offset
    by (oy, ox), with the shifted band-limit window of shiftedAsmTransferF (:250-305)
    and the reference's net scale 1/M^2.  usrc and udst may alias."""
    return _shifted(usrc, udst, ny, nx, dy, dx, lam, z, oy, ox,
                    _lib.SHIFTED_CPU | (_lib.ADJOINT if adjoint else 0), "shiftedAsmProp")


def gshiftedAsmProp(usrc, udst, ny, nx, dy, dx, lam, z, oy, ox, bandlimit=False, adjoint=False):
    """ol::gshiftedAsmProp (include/golprop.h:536-576): no window (its window is
    commented out there; bandlimit=True adds it), net scale 1."""
    flags = (_lib.BANDLIMIT if bandlimit else 0) | (_lib.ADJOINT if adjoint else 0)
    return _shifted(usrc, udst, ny, nx, dy, dx, lam, z, oy, ox, flags, "gshiftedAsmProp")


def shiftedProp(usrc, udst, ny, nx, dy, dx, lam, d, oy, ox, propmode):
    """ol::shiftedProp (include/golprop.h:756-770): gshiftedAsmProp for ASM, ARSS with
    s = 1 for FRESNEL."""
    if propmode == ASM:
        return gshiftedAsmProp(usrc, udst, ny, nx, dy, dx, lam, d, oy, ox)
    if propmode == FRESNEL:
        return ARSSFresnelProp(usrc, udst, ny, nx, dy, dx, lam, d, 1.0, oy, ox)
    raise OlxInvalid(_lib.EINVAL, "shiftedProp: propmode must be ASM or FRESNEL")


def splitAsmProp(u, height, width, dy, dx, lam, z, blocknum_y=2, blocknum_x=2):
    """ol::splitAsmProp (include/olprop.h:614-651): every block to every block through
    shiftedAsmProp, summed, in place."""
    _field(u, height, width)
    check(_lib.call("olx_split_prop_c64", u.device.index, _ptr(u), height, width, dy, dx,
                                         lam, z, _lib.MODE_ASM, blocknum_y, blocknum_x,
                                         _lib.SHIFTED_CPU), "splitAsmProp")
    return u


def gsplitProp(u, height, width, dy, dx, lam, z, propmode, blocknum_y=2, blocknum_x=2):
    """ol::gsplitProp (include/golprop.h:773-813): blocks through shiftedProp, in place."""
    _field(u, height, width)
    if propmode not in (ASM, FRESNEL):
        raise OlxInvalid(_lib.EINVAL, "gsplitProp: propmode must be ASM or FRESNEL")
    check(_lib.call("olx_split_prop_c64", u.device.index, _ptr(u), height, width, dy, dx,
                                         lam, z, propmode, blocknum_y, blocknum_x, 0), "gsplitProp")
    return u


BILINEAR, BICUBIC = _lib.INTERP_BILINEAR, _lib.INTERP_BICUBIC


def _tilted(u, ny, nx, dy, dx, lam, z, angle, inp, propmode, kind, name):
    _field(u, ny, nx)
    if propmode not in (ASM, FRESNEL):
        raise OlxInvalid(_lib.EINVAL, f"{name}: propmode must be ASM or FRESNEL")
    check(_lib.call("olx_tilted_c64", u.device.index, _ptr(u), ny, nx, dy, dx, lam, z,
                                     angle, inp, propmode, kind), name)
    return u


def TiltedSourceProp(u, ny, nx, dy, dx, lam, z, angle_src2dst, inp=BILINEAR, propmode=ASM):
    """ol::TiltedSourceProp (include/olprop.h:414-526): from a source plane tilted by
    angle_src2dst to a parallel destination plane; in place, net scale 1/M^2."""
    return _tilted(u, ny, nx, dy, dx, lam, z, angle_src2dst, inp, propmode, _lib.TILT_SOURCE,
                   "TiltedSourceProp")


def TiltedReferenceProp(u, ny, nx, dy, dx, lam, z, rotate, inp=BILINEAR, propmode=ASM):
    """ol::TiltedReferenceProp (include/olprop.h:528-594): propagation then rotation of
    the reference plane by rotate; in place, net scale 1/M^2."""
    return _tilted(u, ny, nx, dy, dx, lam, z, rotate, inp, propmode, _lib.TILT_REFERENCE,
                   "TiltedReferenceProp")


def Prop(u, ny, nx, dy, dx, lam, d, propmode=AUTO):
    """ol::Prop (include/olprop.h:656-682).  Returns the mode used (ASM, FRESNEL or
    FRESNEL_BL when AUTO falls back to the band-limited Fresnel)."""
    suf = _field(u, ny, nx, c128=True)
    ch = ctypes.c_int(-1)
    check(_lib.call(f"olx_prop_{suf}", u.device.index, _ptr(u), ny, nx, dy, dx, lam, d, propmode,
                    ctypes.byref(ch)), "Prop")
    return ch.value


def AsmPropCheck(ny, nx, dy, dx, lam, z):
    """ol::AsmPropCheck (include/olprop.h:209-225)."""
    return bool(_lib.load().olx_asm_prop_check(ny, nx, dy, dx, lam, z))


def FresnelPropCheck(ny, nx, dy, dx, lam, z):
    """ol::FresnelPropCheck (include/olprop.h:109-123)."""
    return bool(_lib.load().olx_fresnel_prop_check(ny, nx, dy, dx, lam, z))


def fft(u):
    """Unnormalised forward 2-D FFT in place (ol::fft, include/olutils.h:19-41);
    complex128 tensors in FP64."""
    suf = _field(u, u.shape[-2], u.shape[-1], c128=True)
    check(_lib.call(f"olx_fft2d_{suf}", u.device.index, _ptr(u), u.shape[-2], u.shape[-1], -1),
          "fft")
    return u


def ifft(u):
    """Unnormalised backward 2-D FFT in place (ol::ifft, include/olutils.h:44-66);
    complex128 tensors in FP64."""
    suf = _field(u, u.shape[-2], u.shape[-1], c128=True)
    check(_lib.call(f"olx_fft2d_{suf}", u.device.index, _ptr(u), u.shape[-2], u.shape[-1], 1),
          "ifft")
    return u


# ------------------------------------------------------------------ array helpers
_ELEM_OK = (1, 2, 4, 8, 16)


def _arr2d(t, name):
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dim() != 2 or not t.is_contiguous():
        raise OlxInvalid(_lib.EINVAL, f"{name} must be a contiguous 2-D CUDA tensor")
    if t.element_size() not in _ELEM_OK:
        raise OlxInvalid(_lib.EINVAL, f"{name}: element size {t.element_size()} not in {_ELEM_OK}")
    return t


def _window(src, dny, dnx, oy, ox, name):
    _arr2d(src, "src")
    dst = torch.empty((dny, dnx), dtype=src.dtype, device=src.device)
    sny, snx = src.shape
    check(_lib.call("olx_window_copy", src.device.index, _ptr(src), sny, snx, _ptr(dst), dny, dnx,
                    oy, ox, src.element_size()), name)
    return dst


def zeropadding(src, out_height=None, out_width=None):
    """ol::zeropadding / gzeropadding (include/olutils.h:133-178, include/golutils.h:
    189-244): a new (out_height, out_width) array (default twice the size) holding src
    centred, left pads (out - in) // 2, zeros elsewhere."""
    ny, nx = src.shape
    oh = 2 * ny if out_height is None else int(out_height)
    ow = 2 * nx if out_width is None else int(out_width)
    return _window(src, oh, ow, (oh - ny) // 2, (ow - nx) // 2, "zeropadding")


def del_zero(src):
    """ol::del_zero / gdel_zero (include/olutils.h:181-198, include/golutils.h:247-278):
    the centre [h/4, 3h/4) x [w/4, 3w/4) of src as a new (h/2, w/2) array."""
    ny, nx = src.shape
    return _window(src, ny // 2, nx // 2, -(ny // 4), -(nx // 4), "del_zero")


def cut(src, y, x, ysize, xsize):
    """ol::cut / gcut (include/olutils.h:200-211, include/golutils.h:281-308): the
    (ysize, xsize) window at row y, column x as a new array."""
    ny, nx = _arr2d(src, "src").shape
    if y < 0 or x < 0 or ysize < 0 or xsize < 0 or y + ysize > ny or x + xsize > nx:
        raise OlxInvalid(_lib.EINVAL, "cut: window outside the array")
    return _window(src, ysize, xsize, -y, -x, "cut")


def fftshift(u):
    """ol::fftshift / gfftshift (include/olutils.h:95-117, include/golutils.h:80-110):
    the quadrant swap, in place."""
    _arr2d(u, "u")
    if u.element_size() not in (4, 8, 16):
        raise OlxInvalid(_lib.EINVAL, "fftshift: element size must be 4, 8 or 16")
    check(_lib.call("olx_fftshift", u.device.index, _ptr(u), u.shape[0], u.shape[1],
                    u.element_size()), "fftshift")
    return u


def _same(a, b, c):
    for t, nm in ((a, "a"), (b, "b"), (c, "c")):
        if not isinstance(t, torch.Tensor) or not t.is_cuda or not t.is_contiguous():
            raise OlxInvalid(_lib.EINVAL, f"{nm} must be a contiguous CUDA tensor")
    if not (a.shape == b.shape == c.shape and a.dtype == b.dtype == c.dtype):
        raise OlxInvalid(_lib.EINVAL, "operands must share shape and dtype")
    if a.dtype not in (torch.complex64, torch.complex128):
        raise OlxInvalid(_lib.EINVAL, "operands must be complex64 or complex128")


def mult(a, b, c=None):
    """ol::mult / mul_complex (include/golutils.h:123-143, include/olutils.h:121-130):
    c = a * b elementwise (c may be a or b; a new tensor when omitted)."""
    c = torch.empty_like(a) if c is None else c
    _same(a, b, c)
    name = "olx_mul_c64" if a.dtype == torch.complex64 else "olx_mul_c128"
    check(_lib.call(name, a.device.index, _ptr(a), _ptr(b), _ptr(c), a.numel()), "mult")
    return c


mul_complex = mult


def add(a, b, c=None):
    """ol::add (include/golutils.h:145-164): c = a + b elementwise (complex64)."""
    c = torch.empty_like(a) if c is None else c
    _same(a, b, c)
    if a.dtype != torch.complex64:
        raise OlxInvalid(_lib.EINVAL, "add: complex64 operands")
    check(_lib.call("olx_add_c64", a.device.index, _ptr(a), _ptr(b), _ptr(c), a.numel()), "add")
    return c


def mul_scalar(u, c):
    """ol::mul_scalar (include/olutils.h:81-89, include/golutils.h:166-187): u *= c in
    place (c rounded to float for complex64 data)."""
    _same(u, u, u)
    if u.dtype == torch.complex64:
        check(_lib.call("olx_scale_c64", u.device.index, _ptr(u), u.numel(), float(c)), "mul_scalar")
    else:
        check(_lib.call("olx_scale_c128", u.device.index, _ptr(u), u.numel(), float(c)), "mul_scalar")
    return u


def standardization(u):
    """ol::standardization (include/olutils.h:70-79): u *= 1 / (ny nx), the factor
    rounded to float as there."""
    ny, nx = u.shape[-2], u.shape[-1]
    return mul_scalar(u, float(np.float32(1.0 / (ny * nx))))


# ------------------------------------------------------------------ point clouds
POINT_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("amp", "<f4"),
                        ("phase", "<f4")])


class Object:
    """Point cloud on the GPU (ol::Object / ol::gObject, include/olclass.h:54-70,
    include/golclass.h:58-71) packed as olx_point records (x, y, z, amp, phase)."""

    def __init__(self, x, y, z, amp=None, phase=None, device=None):
        x = np.asarray(x, dtype=np.float64).ravel()
        n = x.size
        rec = np.zeros(n, dtype=POINT_DTYPE)
        rec["x"] = x
        rec["y"] = np.asarray(y, dtype=np.float64).ravel()
        rec["z"] = np.asarray(z, dtype=np.float64).ravel()
        rec["amp"] = 1.0 if amp is None else np.asarray(amp, dtype=np.float32).ravel()
        rec["phase"] = 0.0 if phase is None else np.asarray(phase, dtype=np.float32).ravel()
        self._upload(rec, device)

    def _upload(self, rec, device):
        self.size = rec.size
        self._host = rec
        dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        buf = torch.from_numpy(rec.view(np.uint8).copy())
        self.data = buf.to(dev) if rec.size else torch.empty(32, dtype=torch.uint8, device=dev)

    @classmethod
    def from_records(cls, rec, device=None):
        """A cloud from a POINT_DTYPE record array (as returned by objread_host)."""
        obj = cls.__new__(cls)
        obj._upload(np.ascontiguousarray(rec, dtype=POINT_DTYPE), device)
        return obj

    @property
    def host(self):
        """The records, read back from the device after an in-place device update."""
        if self._host is None:
            self._host = self.data[: self.size * POINT_DTYPE.itemsize].cpu().numpy().view(POINT_DTYPE).copy()
        return self._host

    def _touched(self):
        self._host = None

    @property
    def device(self):
        return self.data.device

    def ptr(self):
        return _ptr(self.data)


def Object2cuda(obj_or_xyz):
    """ol::Object2cuda (include/golclass.h:73-78): accepts an Object or an (n, 3) array."""
    if isinstance(obj_or_xyz, Object):
        return obj_or_xyz
    a = np.asarray(obj_or_xyz, dtype=np.float64)
    return Object(a[:, 0], a[:, 1], a[:, 2])


def _out2d(t, ny, nx, dtype, name):
    if t is None:
        return None
    if not t.is_cuda or t.dtype != dtype or not t.is_contiguous() or t.dim() != 2 \
            or tuple(t.shape) != (ny, nx):
        raise _lib.OlxInvalid(_lib.EINVAL, f"{name} must be a contiguous CUDA {dtype} ({ny}, {nx})")
    return t


def Cgh(obj, out, ny, nx, py, px, lam, exact_r=False, phase_out=None, amp_out=None,
        accumulate=False, prec_f32=False):
    """ol::Cgh / gCgh (include/olcgh.h:351-379, include/golcgh.h:127-145).

    out is a complex64 tensor (complex hologram, _CghFresnel include/olcgh.h:296-313)
    or a float32 tensor (phase-only hologram via atan2, include/olcgh.h:275-295).
    Optional phase_out / amp_out receive the encodings in the same pass.  prec_f32
    evaluates the Fresnel sum with the reference's PREC_T = float arithmetic."""
    if out is not None and out.dtype == torch.float32:
        phase_out, out = out, None
    _out2d(out, ny, nx, torch.complex64, "out")
    _out2d(phase_out, ny, nx, torch.float32, "phase_out")
    _out2d(amp_out, ny, nx, torch.float32, "amp_out")
    flags = (_lib.CGH_EXACT_R if exact_r else 0) | (_lib.CGH_ACCUMULATE if accumulate else 0) | \
        (_lib.PREC_F32 if prec_f32 else 0)
    nul = lambda t: None if t is None else _ptr(t)
    check(_lib.call("olx_cgh_c64", obj.device.index, obj.ptr(), obj.size, nul(out),
                                  nul(phase_out), nul(amp_out), ny, nx, py, px, lam, flags), "Cgh")
    return out if out is not None else phase_out


def CghPhase8(obj, img, ny, nx, py, px, lam, out=None, exact_r=False, offset=False,
              accumulate=False, prec_f32=False):
    """Cgh -> phase8bit in one kernel (include/olcgh.h:275-295, include/olimg.h:205-213):
    img (uint8, ny x nx) receives the 8-bit phase of the hologram, the byte phase8bit
    gives for the phase Cgh writes, without the phase array in between.  out
    (complex64) optionally receives the field as well; offset selects the +pi mapping."""
    if img is None:
        raise _lib.OlxInvalid(_lib.EINVAL, "CghPhase8: img is required")
    _out2d(img, ny, nx, torch.uint8, "img")
    _out2d(out, ny, nx, torch.complex64, "out")
    flags = (_lib.CGH_EXACT_R if exact_r else 0) | (_lib.CGH_ACCUMULATE if accumulate else 0) | \
        (_lib.PREC_F32 if prec_f32 else 0) | (_lib.PHASE8_OFFSET if offset else 0)
    check(_lib.call("olx_cgh_img8_c64", obj.device.index, obj.ptr(), obj.size,
                    None if out is None else _ptr(out), _ptr(img), ny, nx, py, px, lam, flags),
          "CghPhase8")
    return img


def CghFP64(obj, out, ny, nx, py, px, lam, exact_r=False):
    """Same sum in FP64 into a complex128 tensor (the FP64 column)."""
    _out2d(out, ny, nx, torch.complex128, "out")
    check(_lib.call("olx_cgh_c128", obj.device.index, obj.ptr(), obj.size, _ptr(out), ny,
                                   nx, py, px, lam, _lib.CGH_EXACT_R if exact_r else 0), "CghFP64")
    return out


def encode(U, phase_out=None, amp_out=None):
    """Phase-only (atan2) and amplitude encodings of a complex64 field."""
    if phase_out is None:
        phase_out = torch.empty(U.shape, dtype=torch.float32, device=U.device)
    if amp_out is None:
        amp_out = torch.empty(U.shape, dtype=torch.float32, device=U.device)
    check(_lib.call("olx_encode_c64", U.device.index, _ptr(U), _ptr(phase_out),
                                     _ptr(amp_out), U.numel()), "encode")
    return phase_out, amp_out


# ------------------------------------------------------------------ WRP
def WRPStep1(obj, U, ny, nx, p, lam, d, plimit=None, square=False):
    """WRPStep1 (include/golcgh.h:369-389); square=True is WRPStep1_D (:471-491).
    d is the WRP plane position; plimit the pitch that sets the window."""
    _out2d(U, ny, nx, torch.complex64, "U")
    check(_lib.call("olx_wrp_c64", U.device.index, obj.ptr(), obj.size, _ptr(U), ny, nx,
                                  p, lam, d, p if plimit is None else plimit,
                                  _lib.WRP_SQUARE if square else 0), "WRPStep1")
    return U


def WRPStep1_D(obj, U, ny, nx, p, lam, d, plimit=None):
    return WRPStep1(obj, U, ny, nx, p, lam, d, plimit, square=True)


def WRPStep2(U, ny, nx, p, lam, d, propmode=AUTO):
    """WRPStep2 (include/golcgh.h:494-517): AUTO uses ASM when AsmPropCheck holds,
    else Fresnel when FresnelPropCheck holds, else ASM.  Returns the mode used."""
    if propmode == AUTO:
        if AsmPropCheck(ny, nx, p, p, lam, d) or not FresnelPropCheck(ny, nx, p, p, lam, d):
            propmode = ASM
        else:
            propmode = FRESNEL
    if propmode == ASM:
        AsmProp(U, ny, nx, p, p, lam, d)
    elif propmode == FRESNEL:
        FresnelProp(U, ny, nx, p, p, lam, d)
    else:
        raise _lib.OlxInvalid(_lib.EINVAL, f"WRPStep2: unknown mode {propmode}")
    return propmode


def wrp_plane_z(ny, nx, p, lam, zmin, s=0.9, plimit=None, square=False):
    """Plane rule of WRPMethod / WRPMethod_D (include/golcgh.h:524-533, :545-554)."""
    return _lib.load().olx_wrp_plane_z(ny, nx, p, lam, zmin, p if plimit is None else plimit,
                                       s, 1 if square else 0)


def WRPMethod(obj, U, ny, nx, p, lam, zmin, s=0.9, propmode=AUTO, square=False, plimit=None):
    """WRPMethod (include/golcgh.h:561-581); square=True is WRPMethod_D (:519-559).
    Returns the WRP plane position used."""
    _out2d(U, ny, nx, torch.complex64, "U")
    pz = ctypes.c_double(0.0)
    check(_lib.call("olx_wrp_method_c64", U.device.index, obj.ptr(), obj.size, _ptr(U),
                                         ny, nx, p, lam, zmin, p if plimit is None else plimit, s,
                                         propmode, _lib.WRP_SQUARE if square else 0,
                                         ctypes.byref(pz)), "WRPMethod")
    return pz.value


def WRPMethod_D(obj, U, ny, nx, p, lam, zmin, s=0.9, propmode=AUTO, plimit=None):
    return WRPMethod(obj, U, ny, nx, p, lam, zmin, s, propmode, square=True, plimit=plimit)


# ------------------------------------------------------------------ layered CGH
_PREC = {"f16": _lib.LAYERED_F16, "bf16": _lib.LAYERED_BF16, "bf16x3": _lib.LAYERED_BF16X3}


def CghFresnelDiff(obj, ny, nx, p, lam, U=None, phase=None):
    """ol::_CghFresnelDiff (include/olcgh.h:315-349): the recurrence-formula CGH of a cloud
    in pixel units, theta_j = p z_j / lambda + gamma_j / 2 ((w - x_j)^2 + (h - y_j)^2)
    cycles, gamma_j = p / (lambda z_j), every pixel at the recurrence's exact value.
    Writes U (complex64) and / or phase (float32, atan2); returns them."""
    if U is None and phase is None:
        phase = torch.empty((ny, nx), dtype=torch.float32, device=obj.device)
    if U is not None:
        _out2d(U, ny, nx, torch.complex64, "U")
    if phase is not None:
        _out2d(phase, ny, nx, torch.float32, "phase")
    check(_lib.call("olx_cgh_fresnel_diff_c64", obj.device.index, obj.ptr(), obj.size,
                    _ptr(U) if U is not None else None,
                    _ptr(phase) if phase is not None else None, ny, nx, p, lam),
          "CghFresnelDiff")
    return U, phase


def CghLayered(obj, U, ny, nx, py, px, lam, nlayers, zmin, zmax, precision="f16",
               accumulate=False, tile128=False):
    """Layered Fresnel CGH as complex GEMMs on MFMA (no reference counterpart).
    Points are snapped to the centres of nlayers uniform layers over [zmin, zmax).
    tile128 forces the 128 x 128 work-group tile where 256 x 128 would be used."""
    _out2d(U, ny, nx, torch.complex64, "U")
    flags = (_lib.CGH_ACCUMULATE if accumulate else 0) | (_lib.LAYERED_TILE128 if tile128 else 0)
    check(_lib.call("olx_cgh_layered_c64", U.device.index, obj.ptr(), obj.size, _ptr(U),
                                          ny, nx, py, px, lam, nlayers, zmin, zmax,
                                          _PREC[precision], flags), "CghLayered")
    return U


def layer_depth(z, zmin, zmax, nlayers):
    """Depth a point is snapped to by CghLayered."""
    return _lib.load().olx_layer_depth(z, zmin, zmax, nlayers)


# ------------------------------------------------------------------ point-cloud ingest
ZMIN, ZMIDDLE = _lib.OBJSET_ZMIN, _lib.OBJSET_ZMIDDLE
_OBJFMT = {"auto": _lib.OBJ_AUTO, "3d": _lib.OBJ_3D, "3df": _lib.OBJ_3DF}


def objread_host(path, scale=1.0, offset=(0.0, 0.0, 0.0), fmt="auto"):
    """ol::objread (include/olobject.h:9-162) into host POINT_DTYPE records:
    .3d (int32) or .3df (float) triples, x = v scale + offset, amp 1, phase 0."""
    L = _lib.load()
    n = ctypes.c_int64()
    bpath = os.fsencode(path)
    check(L.olx_obj_read(bpath, _OBJFMT[fmt], None, 0, ctypes.byref(n), scale, *offset), "objread")
    rec = np.zeros(n.value, dtype=POINT_DTYPE)
    if n.value:
        check(L.olx_obj_read(bpath, _OBJFMT[fmt], rec.ctypes.data, n.value, ctypes.byref(n), scale,
                             *offset), "objread")
    return rec


def objread(path, scale=1.0, offset=(0.0, 0.0, 0.0), fmt="auto", device=None):
    """ol::objread followed by Object2cuda: the cloud in device memory."""
    return Object.from_records(objread_host(path, scale, offset, fmt), device)


def objinfo(obj):
    """Bounding box of a device cloud, ((xmin, ymin, zmin), (xmax, ymax, zmax))
    (ol::objinfo, include/olobject.h:164-175, without printing)."""
    mn, mx = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
    check(_lib.call("olx_obj_bounds", obj.device.index, obj.ptr(), obj.size, mn, mx),
          "objinfo")
    return tuple(mn), tuple(mx)


def objset(obj, x_range, y_range=None, offset=(0.0, 0.0, 0.0), mode=ZMIN, f32=False):
    """ol::objset (include/olobject.h:195-264) in place on the device: one range
    (the larger extent fits it) or x and y ranges (the tighter scale wins).
    f32=True reproduces the arithmetic of an Object<float> (float coordinates,
    scale and offsets); the default that of an Object<double>."""
    m = mode | (_lib.OBJSET_F32 if f32 else 0)
    check(_lib.call("olx_obj_set", obj.device.index, obj.ptr(), obj.size, x_range,
                                  -1.0 if y_range is None else y_range, *offset, m), "objset")
    obj._touched()
    return obj


def objsort(obj):
    """ol::objsort (include/olobject.h:177-182) in place on the device: descending z,
    equal depths in input order."""
    check(_lib.call("olx_obj_sort_z", obj.device.index, obj.ptr(), obj.size), "objsort")
    obj._touched()
    return obj


# ------------------------------------------------------------------ 8-bit encodings
PHASE, REAL, IMAG, AMP = _lib.IMG_PHASE, _lib.IMG_REAL, _lib.IMG_IMAG, _lib.IMG_AMP
_IMGMODE = {"phase": _lib.IMG_PHASE, "real": _lib.IMG_REAL, "imag": _lib.IMG_IMAG,
            "amp": _lib.IMG_AMP, "real_legacy": _lib.IMG_REAL_LEGACY}


def _img_like(t, img):
    if img is None:
        img = torch.empty(t.shape, dtype=torch.uint8, device=t.device)
    if img.dtype != torch.uint8 or img.numel() != t.numel() or not img.is_contiguous():
        raise OlxInvalid(_lib.EINVAL, "img must be a contiguous uint8 tensor of the input's size")
    return img


_REAL_T = {torch.float32: 0, torch.float64: _lib.IN_F64}
_COMPLEX_T = {torch.complex64: 0, torch.complex128: _lib.IN_F64}


def _typed(t, table, name):
    if t.dtype not in table or not t.is_contiguous() or not t.is_cuda:
        raise OlxInvalid(_lib.EINVAL, f"{name} must be a contiguous CUDA tensor of type "
                         + " or ".join(str(d) for d in table))
    return table[t.dtype]


def phase8bit(phase, img=None, offset=False):
    """ol::phase8bit (include/olimg.h:205-213) on a float32 / float64 device phase:
    the reference's mapping round((phi - pi) / 2pi * 255) taken modulo 256 as its
    x86 build converts it; offset=True maps round((phi + pi) / 2pi * 255)."""
    f = _typed(phase, _REAL_T, "phase")
    img = _img_like(phase, img)
    check(_lib.call("olx_phase8bit", phase.device.index, _ptr(phase), _ptr(img),
                                    phase.numel(), f | (_lib.PHASE8_OFFSET if offset else 0)),
          "phase8bit")
    return img


def quant8bit(real, img=None):
    """ol::quant8bit (include/olimg.h:178-199): min/max scaled to 0..255."""
    f = _typed(real, _REAL_T, "real")
    img = _img_like(real, img)
    check(_lib.call("olx_quant8bit", real.device.index, _ptr(real), _ptr(img), real.numel(),
                                    f), "quant8bit")
    return img


def complex2img(u, img=None, mode="phase"):
    """ol::complex2img (include/olimg.h:314-391) of a complex64 / complex128 device
    field; mode "phase", "real", "imag", "amp" or "real_legacy" (the bool overload)."""
    f = _typed(u, _COMPLEX_T, "u")
    img = _img_like(u, img)
    check(_lib.call("olx_complex2img", u.device.index, _ptr(u), _ptr(img), u.numel(),
                                      _IMGMODE[mode] | f), "complex2img")
    return img


def complex2real(u, out=None, mode="phase"):
    """ol::complex2real (include/olimg.h:393-420) into the field's real type."""
    f = _typed(u, _COMPLEX_T, "u")
    rt = torch.float64 if f else torch.float32
    if out is None:
        out = torch.empty(u.shape, dtype=rt, device=u.device)
    if out.dtype != rt or out.numel() != u.numel() or not out.is_contiguous():
        raise OlxInvalid(_lib.EINVAL, f"out must be a contiguous {rt} tensor of u's size")
    check(_lib.call("olx_complex2real", u.device.index, _ptr(u), _ptr(out), u.numel(),
                                       _IMGMODE[mode] | f), "complex2real")
    return out


def Save(path, u, mode="phase"):
    """ol::Save (include/olimg.h:430-454) for .bmp paths: a complex field through
    complex2img, or a real array through quant8bit, then bmpwrite.  Other image
    formats need OpenCV in the reference and are refused."""
    if not str(path).endswith(".bmp"):
        raise OlxInvalid(_lib.EINVAL, "Save writes .bmp files only")
    img = complex2img(u, mode=mode) if u.is_complex() else quant8bit(u)
    bmpwrite(path, img.reshape(u.shape[-2], u.shape[-1]))


def bmpwrite(path, img):
    """ol::bmpwrite (include/olimg.h:136-165): 8-bit grey BMP from a 2-D uint8 array."""
    a = np.ascontiguousarray(img.cpu().numpy() if isinstance(img, torch.Tensor) else img, dtype=np.uint8)
    if a.ndim != 2:
        raise OlxInvalid(_lib.EINVAL, "bmpwrite needs a 2-D image")
    check(_lib.load().olx_bmp_write(os.fsencode(path), a.ctypes.data, a.shape[0], a.shape[1]),
          "bmpwrite")


def bmpread(path, ny, nx):
    """ol::bmpread (include/olimg.h:56-104) of a file bmpwrite made: (ny, nx) uint8."""
    a = np.zeros((ny, nx), dtype=np.uint8)
    check(_lib.load().olx_bmp_read(os.fsencode(path), a.ctypes.data, ny, nx), "bmpread")
    return a


# ------------------------------------------------------------------ optical elements
RECORD, PLAY = _lib.STEP_RECORD, _lib.STEP_PLAY
_DIFFUSER_KIND = {"phase": 0, "xphase": 1, "yphase": 2, "bitrandom": 3, "block": 4, "blockstep": 5,
                  "bitphase": 6}


def uniform(n, seed=1, device=None):
    """n uniform floats in (0, 1] on the device (Philox4x32-10 stream keyed by seed)."""
    dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    check(_lib.call("olx_uniform_f32", dev.index, _ptr(out), n, seed), "uniform")
    return out


class Diffuser:
    """ol::Diffuser (include/Diffuser.h:172-326) on complex64 device fields."""

    def __init__(self, height, width, pitch=None):
        self.height, self.width, self.pitch = height, width, pitch

    def random_phase(self, u, step, seed=1, range=2 * np.pi, oy=0, ox=0):
        """The CPU form (include/Diffuser.h:191-215): the std::mt19937(seed) stream in
        pixel order, phases range (v - 0.5) (PLAY) or their negative (RECORD)."""
        _field(u, self.height, self.width)
        check(_lib.call("olx_diffuser_random_phase_c64", u.device.index, _ptr(u), self.height,
                                                        self.width, step, seed, range, oy, ox),
              "random_phase")
        return u

    def _apply(self, u, kind, b, scalar, blocksize=1):
        _field(u, self.height, self.width)
        if b.dtype != torch.float32 or not b.is_contiguous():
            raise OlxInvalid(_lib.EINVAL, "uniforms must be a contiguous float32 tensor")
        check(_lib.call("olx_diffuser_apply_c64", u.device.index, _ptr(u), self.height,
                                                 self.width, _ptr(b), scalar, _DIFFUSER_KIND[kind],
                                                 blocksize), "diffuser")
        return u

    def _uniforms(self, n, seed, u):
        return uniform(n, seed, u.device.index)

    def grandom_phase(self, u, step, seed=1, range=2 * np.pi, b=None):
        """The GPU form (include/Diffuser.h:218-238, KernelMultPhase) with uniforms b
        (default: the Philox stream for seed).  The reference's shift of the uniforms
        is applied only to a quarter of the array (its grid is halved on both axes,
        include/Diffuser.h:151-158); no shift is offered here."""
        b = self._uniforms(self.height * self.width, seed, u) if b is None else b
        return self._apply(u, "phase", b, range if step == PLAY else -range)

    def random_xphase(self, u, step, seed=1, range=2 * np.pi, b=None):
        """include/Diffuser.h:240-259 (one value per column)."""
        b = self._uniforms(self.width, seed, u) if b is None else b
        return self._apply(u, "xphase", b, range if step == PLAY else -range)

    def random_blockphase(self, u, step, seed=1, range=2 * np.pi, blocksize=1, b=None):
        """include/Diffuser.h:261-279 (one value per blocksize^2 block)."""
        b = self._uniforms(self.height * self.width // (blocksize * blocksize), seed, u) if b is None else b
        return self._apply(u, "block", b, range if step == PLAY else -range, blocksize)

    def random_blockstepphase(self, u, step, seed=1, range=2 * np.pi, blocksize=4, b=None):
        """include/Diffuser.h:281-299 (per block, plus a (x + y) / blocksize ramp)."""
        b = self._uniforms(self.height * self.width // (blocksize * blocksize), seed, u) if b is None else b
        return self._apply(u, "blockstep", b, range if step == PLAY else -range, blocksize)

    def bitphase(self, u):
        """include/Diffuser.h:303-308: exp(i F32_PI) on the even checkerboard."""
        dummy = torch.empty(1, dtype=torch.float32, device=u.device)
        return self._apply(u, "bitphase", dummy, 0.0)


class Lensarry:
    """ol::Lensarry (include/Lensarry.h:84-311): square lenslets of blockwidth over a
    height x width field of pitch; the lens field is computed per pixel on the device."""
    circle = False

    def __init__(self, height, width, pitch, blockwidth, f, lam):
        self.height, self.width = height, width
        self.pitch, self.blockwidth, self.f, self.lam = pitch, blockwidth, f, lam
        self.blockpx = int(np.round(np.float32(blockwidth) / np.float32(pitch)))

    def _run(self, u, step, oy, ox, p_limit, write):
        flags = (_lib.LENS_CIRCLE if self.circle else 0) | (_lib.LENS_WRITE if write else 0)
        check(_lib.call("olx_lens_array_c64", u.device.index, _ptr(u), self.height, self.width,
                                             self.pitch, self.blockwidth, self.f, self.lam, step, oy,
                                             ox, p_limit, flags), "Lensarry")
        return u

    def gen_complex(self, step=PLAY, oy=0, ox=0, device=None):
        """The lens field (gen_complex, include/Lensarry.h:254-265) as a new tensor."""
        dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        u = torch.empty((self.height, self.width), dtype=torch.complex64, device=dev)
        return self._run(u, step, oy, ox, 0.0, True)

    def gen_complex_limit(self, p_limit, step=PLAY, oy=0, ox=0, device=None):
        """include/Lensarry.h:300-311: band-limited lenslets."""
        dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        u = torch.empty((self.height, self.width), dtype=torch.complex64, device=dev)
        return self._run(u, step, oy, ox, p_limit, True)

    def mult_complex(self, u, step=PLAY, oy=0, ox=0):
        """include/Lensarry.h:266-271: u times the lens field, in place."""
        _field(u, self.height, self.width)
        return self._run(u, step, oy, ox, 0.0, False)

    def mult_complex_limit(self, u, p_limit, step=PLAY, oy=0, ox=0):
        """include/Lensarry.h:312-317."""
        _field(u, self.height, self.width)
        return self._run(u, step, oy, ox, p_limit, False)

    def lenslet(self, step=PLAY, p_limit=0.0, device=None):
        """The blockpx x blockpx lenslet (FresnelResponse at -+f, its band-limited form
        with p_limit > 0) that the array tiles."""
        dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        b = torch.empty((self.blockpx, self.blockpx), dtype=torch.complex64, device=dev)
        bp = self.blockpx
        check(_lib.call("olx_lens_array_c64", dev.index, _ptr(b), bp, bp, self.pitch,
                        self.blockwidth, self.f, self.lam, step, 0, 0, p_limit, _lib.LENS_WRITE),
              "Lensarry")
        return b

    def _tile(self, block, on, om, p_limit, kind):
        if block.dtype != torch.complex64 or tuple(block.shape) != (self.blockpx, self.blockpx) \
                or not block.is_cuda or not block.is_contiguous():
            raise OlxInvalid(_lib.EINVAL, "duplicate: block must be a contiguous blockpx x blockpx "
                             "complex64 CUDA tensor")
        out = torch.empty((self.height, self.width), dtype=torch.complex64, device=block.device)
        check(_lib.call("olx_lens_tile_c64", block.device.index, _ptr(block), self.blockpx,
                        _ptr(out), self.height, self.width, on, om, self.pitch, self.f, self.lam,
                        p_limit, kind), "duplicate")
        return out

    def duplicate(self, block, on=0, om=0):
        """include/Lensarry.h:136-161: the block tiled over the field, shifted by (on, om)."""
        return self._tile(block, on, om, 0.0, 0)

    def duplicate_random(self, block, p_limit, on=0, om=0):
        """include/Lensarry.h:163-206: outside the p_limit window each pixel takes
        exp(i 2 pi (v - 0.5)), v from std::mt19937(0) in the reference's traversal."""
        return self._tile(block, on, om, p_limit, 1)

    def duplicate_multrandom(self, block, p_limit, on=0, om=0):
        """include/Lensarry.h:207-250: outside the window the block value times
        exp(i pi/2 (v - 0.5))."""
        return self._tile(block, on, om, p_limit, 2)

    def gen_complex_random(self, p_limit, step=PLAY, oy=0, ox=0, device=None):
        """include/Lensarry.h:270-281: the unwindowed lenslet tiled with a random phase
        outside the p_limit window (duplicate_multrandom)."""
        return self.duplicate_multrandom(self.lenslet(step, 0.0, device), p_limit, oy, ox)

    def gen_complex_limit_random(self, p_limit, step=PLAY, oy=0, ox=0, device=None):
        """include/Lensarry.h:310-321: the band-limited lenslet, random phases outside its
        window (duplicate_random)."""
        return self.duplicate_random(self.lenslet(step, p_limit, device), p_limit, oy, ox)

    def mult_complex_random(self, u, p_limit, step=PLAY, oy=0, ox=0):
        """include/Lensarry.h:282-288: u times gen_complex_random's field, in place."""
        _field(u, self.height, self.width)
        return mult(u, self.gen_complex_random(p_limit, step, oy, ox, u.device.index), u)

    def mult_complex_limit_random(self, u, p_limit, step=PLAY, oy=0, ox=0):
        """include/Lensarry.h:322-328."""
        _field(u, self.height, self.width)
        return mult(u, self.gen_complex_limit_random(p_limit, step, oy, ox, u.device.index), u)


class LensarryCircle(Lensarry):
    """ol::LensarryCircle (include/Lensarry.h:383-447): hexagonal rows of circular
    lenslets, ones between them."""
    circle = True
