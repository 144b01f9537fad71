"""ctypes binding of libolx.so (the C ABI declared in include/olx.h).

The library must be built in-tree (openol_amd/libolx.so).  There is no CPU
fallback: if the library or a HIP device is missing, every entry point raises.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libolx.so")

OK = 0
EINVAL, ENOMEM, EDEVICE, EUNSUPPORTED, EIO = -1, -2, -3, -4, -5

BANDLIMIT = 0x1
ADJOINT = 0x2
FRESNEL_BL_CPU = 0x10
FRESNEL_BL_GPU = 0x20
CGH_EXACT_R = 0x100
CGH_ACCUMULATE = 0x400
WRP_SQUARE = 0x200
ARSS_CPU = 0x800
SHIFTED_CPU = 0x1000
PREC_F32 = 0x8000
EXTRA_INV_M2 = 0x2000
TILT_SOURCE, TILT_REFERENCE = 0, 1
STEP_RECORD, STEP_PLAY = 0, 1
LENS_CIRCLE, LENS_WRITE = 0x1, 0x2
INTERP_BILINEAR, INTERP_BICUBIC = 0, 1

MODE_ASM, MODE_FRESNEL, MODE_AUTO, MODE_FRESNEL_BL = 0, 1, 2, 3
LAYERED_F16, LAYERED_BF16, LAYERED_BF16X3 = 0, 1, 2
LAYERED_TILE128 = 0x4000
OBJ_AUTO, OBJ_3D, OBJ_3DF = 0, 1, 2
OBJSET_ZMIN, OBJSET_ZMIDDLE, OBJSET_F32 = 0, 1, 0x100
IN_F64 = 0x100
PHASE8_OFFSET = 0x1
IMG_PHASE, IMG_REAL, IMG_IMAG, IMG_AMP, IMG_REAL_LEGACY = 0, 1, 2, 3, 4


class OlxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"olx error {code}: {msg}")
        self.code = code


class OlxInvalid(OlxError, ValueError):
    pass


_lib = None
_lock = threading.Lock()

_i64, _d, _u32, _p, _i = ctypes.c_int64, ctypes.c_double, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_int
SIGNATURES = {
    "olx_ctx_create": [ctypes.POINTER(_p), _i],
    "olx_ctx_destroy": [_p],
    "olx_ctx_set_stream": [_p, _p],
    "olx_ctx_clear_tables": [_p],
    "olx_ctx_synchronize": [_p],
    "olx_version": [],
    "olx_dev_alloc": [_p, ctypes.POINTER(_p), ctypes.c_size_t],
    "olx_dev_free": [_p, _p],
    "olx_copy_h2d": [_p, _p, _p, ctypes.c_size_t],
    "olx_copy_d2h": [_p, _p, _p, ctypes.c_size_t],
    "olx_dev_memset": [_p, _p, _i, ctypes.c_size_t],
    "olx_asm_c64": [_p, _p, _i64, _i64, _d, _d, _d, _d, _u32],
    "olx_asm_c64_noalloc": [_p, _p, _i64, _i64, _d, _d, _d, _d, _u32],
    "olx_fresnel_c64": [_p, _p, _i64, _i64, _d, _d, _d, _d, _u32],
    "olx_asm_c128": [_p, _p, _i64, _i64, _d, _d, _d, _d, _u32],
    "olx_fresnel_c128": [_p, _p, _i64, _i64, _d, _d, _d, _d, _u32],
    "olx_prop_c128": [_p, _p, _i64, _i64, _d, _d, _d, _d, _i, ctypes.POINTER(_i)],
    "olx_fft2d_c128": [_p, _p, _i64, _i64, _i],
    "olx_prop_batch_c64": [_p, _i, _p, _i64, _i64, _i64, _d, _d, _d, _d, _u32, _p, _p, _p, _u32],
    "olx_arss_c64": [_p, _p, _p, _i64, _i64, _d, _d, _d, _d, _d, _d, _d, _u32],
    "olx_shifted_asm_c64": [_p, _p, _p, _i64, _i64, _d, _d, _d, _d, _d, _d, _u32],
    "olx_split_prop_c64": [_p, _p, _i64, _i64, _d, _d, _d, _d, _i, _i, _i, _u32],
    "olx_tilted_c64": [_p, _p, _i64, _i64, _d, _d, _d, _d, _d, _i, _i, _i],
    "olx_diffuser_random_phase_c64": [_p, _p, _i64, _i64, _i, _u32, ctypes.c_float, _i, _i],
    "olx_diffuser_apply_c64": [_p, _p, _i64, _i64, _p, ctypes.c_float, _i, _i],
    "olx_uniform_f32": [_p, _p, _i64, ctypes.c_uint64],
    "olx_lens_array_c64": [_p, _p, _i64, _i64, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                           ctypes.c_float, _i, _i, _i, ctypes.c_float, _u32],
    "olx_lens_tile_c64": [_p, _p, _i, _p, _i64, _i64, _i, _i, ctypes.c_float, ctypes.c_float,
                          ctypes.c_float, ctypes.c_float, _i],
    "olx_prop_c64": [_p, _p, _i64, _i64, _d, _d, _d, _d, _i, ctypes.POINTER(_i)],
    "olx_asm_prop_check": [_i64, _i64, _d, _d, _d, _d],
    "olx_fresnel_prop_check": [_i64, _i64, _d, _d, _d, _d],
    "olx_fft2d_c64": [_p, _p, _i64, _i64, _i],
    "olx_window_copy": [_p, _p, _i64, _i64, _p, _i64, _i64, _i64, _i64, ctypes.c_size_t],
    "olx_fftshift": [_p, _p, _i64, _i64, ctypes.c_size_t],
    "olx_mul_c64": [_p, _p, _p, _p, _i64],
    "olx_add_c64": [_p, _p, _p, _p, _i64],
    "olx_scale_c64": [_p, _p, _i64, ctypes.c_float],
    "olx_mul_c128": [_p, _p, _p, _p, _i64],
    "olx_scale_c128": [_p, _p, _i64, _d],
    "olx_asm_transfer_c64": [_p, _p, _i64, _i64, _d, _d, _d, _d, _u32],
    "olx_fresnel_response_c64": [_p, _p, _i64, _i64, _d, _d, _d, _d, _i64, _i64],
    "olx_cgh_c64": [_p, _p, _i64, _p, _p, _p, _i64, _i64, _d, _d, _d, _u32],
    "olx_cgh_img8_c64": [_p, _p, _i64, _p, _p, _i64, _i64, _d, _d, _d, _u32],
    "olx_cgh_c128": [_p, _p, _i64, _p, _i64, _i64, _d, _d, _d, _u32],
    "olx_cgh_fresnel_diff_c64": [_p, _p, _i64, _p, _p, _i64, _i64, _d, _d],
    "olx_encode_c64": [_p, _p, _p, _p, _i64],
    "olx_wrp_c64": [_p, _p, _i64, _p, _i64, _i64, _d, _d, _d, _d, _u32],
    "olx_wrp_method_c64": [_p, _p, _i64, _p, _i64, _i64, _d, _d, _d, _d, ctypes.c_float, _i,
                           _u32, ctypes.POINTER(_d)],
    "olx_wrp_plane_z": [_i64, _i64, _d, _d, _d, _d, ctypes.c_float, _i],
    "olx_cgh_layered_c64": [_p, _p, _i64, _p, _i64, _i64, _d, _d, _d, _i, _d, _d, _i, _u32],
    "olx_layer_depth": [_d, _d, _d, _i],
    "olx_ctx_enable_timing": [_p, _i],
    "olx_ctx_reset_timing": [_p],
    "olx_ctx_timing_report": [_p, ctypes.c_char_p, ctypes.c_size_t],
    "olx_stream_copy": [_p, _p, _p, ctypes.c_size_t],
    "olx_peak_probe": [_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
    "olx_obj_read": [ctypes.c_char_p, _i, _p, _i64, ctypes.POINTER(_i64), _d, _d, _d, _d],
    "olx_obj_bounds": [_p, _p, _i64, ctypes.POINTER(_d), ctypes.POINTER(_d)],
    "olx_obj_set": [_p, _p, _i64, _d, _d, _d, _d, _d, _i],
    "olx_obj_sort_z": [_p, _p, _i64],
    "olx_phase8bit": [_p, _p, _p, _i64, _u32],
    "olx_quant8bit": [_p, _p, _p, _i64, _u32],
    "olx_complex2img": [_p, _p, _p, _i64, _i],
    "olx_complex2real": [_p, _p, _p, _i64, _i],
    "olx_bmp_write": [ctypes.c_char_p, _p, _i, _i],
    "olx_bmp_read": [ctypes.c_char_p, _p, _i, _i],
    "olx_multiplane_mse_c64": [_p, _p, _i64, _i64, _i, _p, _p, _d, _d, _d, _u32, _p, _p],
    "olx_shifted_asm_transfer_c64": [_p, _p, _i64, _i64, _d, _d, _d, _d, _d, _d, _u32],
}
DOUBLE_RESULT = {"olx_wrp_plane_z", "olx_layer_depth"}


def load():
    """Load libolx.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIBPATH):
            raise OlxError(EDEVICE, f"{LIBPATH} not built: run python -m openol_amd.build")
        L = ctypes.CDLL(LIBPATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_double if name in DOUBLE_RESULT else ctypes.c_int
        L.olx_last_error.argtypes = []
        L.olx_last_error.restype = ctypes.c_char_p
        _lib = L
        return L


def last_error():
    return load().olx_last_error().decode(errors="replace")


def check(rc, what=""):
    if rc != OK:
        msg = f"{what}: {last_error()}" if what else last_error()
        if rc == EINVAL:
            raise OlxInvalid(rc, msg)
        raise OlxError(rc, msg)
    return rc


_ctx = {}
# Held from binding a context to the caller's stream through the launch, so that
# two threads cannot interleave a bind and another thread's launch on one context.
_call_lock = threading.RLock()


def ctx(device=None):
    """Per-device library context bound to torch's current stream.

    One context (and so one set of workspaces) serves every stream of a device.
    When the bound stream changes, the library makes the new stream wait for the
    work already queued on the previous one (olx_ctx_set_stream), so calls issued
    on different streams are ordered, never overlapped, on the shared workspaces."""
    import torch
    if not torch.cuda.is_available():
        raise OlxError(EDEVICE, "no HIP device visible: the MI355X path has no CPU fallback")
    if device is None:
        device = torch.cuda.current_device()
    L = load()
    with _lock:
        c = _ctx.get(device)
        if c is None:
            h = _p()
            with torch.cuda.device(device):
                check(L.olx_ctx_create(ctypes.byref(h), device), "olx_ctx_create")
            c = h
            _ctx[device] = c
    check(L.olx_ctx_set_stream(c, _p(torch.cuda.current_stream(device).cuda_stream)))
    return c


def call(name, device, *args):
    """Bind the device's context to torch's current stream and call entry point
    `name` with the context as its first argument, both under one lock."""
    L = load()
    with _call_lock:
        return getattr(L, name)(ctx(device), *args)


def timing_report(device=None):
    """{kernel: (launches, total_ms)} of the context's instrumented launches."""
    L = load()
    buf = ctypes.create_string_buffer(1 << 16)
    check(L.olx_ctx_timing_report(ctx(device), buf, len(buf)))
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.split()
        out[name] = (int(cnt), float(ms))
    return out
