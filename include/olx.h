/*
 * olx.h -- C ABI of the MI355X-native OpenOL diffraction / CGH library (libolx.so).
 *
 * Plain C: pointers, sizes and scalars only.  Every field pointer named d_* is a
 * device (HBM) pointer owned by the caller; complex fields are interleaved
 * float pairs (re, im) -- the layout of std::complex<float> and of
 * thrust::complex<float> in the reference -- row-major with idx = h * nx + w
 * (include/olutils.h:127, include/golutils.h:117 in the reference).
 *
 * Errors: every entry point returns OLX_OK (0) or a negative OLX_E* code and
 * never exits the process (the reference calls exit(1) on a bad mode,
 * include/olprop.h:611-613).  olx_last_error() returns a thread-local message.
 *
 * Threading: a context is not thread safe; independent contexts are.  All work
 * is enqueued on the context's stream (olx_ctx_set_stream) and is asynchronous
 * with respect to the host unless an entry point says otherwise.
 *
 * Sizes: propagation needs ny, nx to be powers of two with 4 <= n <= 8192
 * (padded transform length 2n <= 16384).
 */
#ifndef OLX_H
#define OLX_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define OLX_OK 0
#define OLX_EINVAL (-1)       /* invalid argument */
#define OLX_ENOMEM (-2)       /* device allocation failed */
#define OLX_EDEVICE (-3)      /* HIP runtime / launch error */
#define OLX_EUNSUPPORTED (-4) /* valid but unsupported size or mode */

/* propagation flags */
#define OLX_BANDLIMIT 0x1u    /* ASM band limit |u| < 1/(lambda sqrt((2 du z)^2 + 1)), include/olprop.h:250-305 */
#define OLX_ADJOINT 0x2u      /* apply the adjoint (conjugate transfer function) */
#define OLX_FRESNEL_BL_CPU 0x10u /* Fresnel window, per-axis limits (include/olprop.h:26-54) */
#define OLX_FRESNEL_BL_GPU 0x20u /* Fresnel window, x-limit on both axes (include/golprop.h:223-238) */
#define OLX_ARSS_CPU 0x800u      /* ARSS with the CPU shiftedFresnelProp semantics: no windows and
                                    net scale circconv / M^4 (include/olprop.h:338-408) */

/* CGH flags */
#define OLX_CGH_EXACT_R 0x100u    /* exact spherical distance r instead of the Fresnel quadratic */
#define OLX_CGH_ACCUMULATE 0x400u /* add into d_U instead of overwriting it */
/* WRP flags */
#define OLX_WRP_SQUARE 0x200u     /* square window D = sqrt(2) W, include/golcgh.h:391-468 */

/* propagation modes (ol::PROPMODE, include/oldefine.h:7-9) */
#define OLX_MODE_ASM 0
#define OLX_MODE_FRESNEL 1
#define OLX_MODE_AUTO 2
#define OLX_MODE_FRESNEL_BL 3     /* reported by olx_prop_c64 when AUTO picked the band-limited Fresnel */

/* layered CGH factor precisions */
#define OLX_LAYERED_F16 0
#define OLX_LAYERED_BF16 1
#define OLX_LAYERED_BF16X3 2      /* split bf16 hi+lo, three products */

typedef struct olx_ctx olx_ctx;

/* One object point (32 bytes).  The reference's ol::Vect3<T> holds x, y, z only
 * (include/olclass.h:44-52); amp and phase carry the per-point A_j and phi_j of
 * the extended sum (amp = 1, phase = 0 reproduce the reference). */
typedef struct olx_point {
    double x, y, z;
    float amp, phase;
} olx_point;

/* ---- context, stream, memory ------------------------------------------- */
int olx_ctx_create(olx_ctx** out, int device);
int olx_ctx_destroy(olx_ctx* ctx);
int olx_ctx_set_stream(olx_ctx* ctx, void* hip_stream); /* NULL = default stream */
int olx_ctx_synchronize(olx_ctx* ctx);
const char* olx_last_error(void);
int olx_version(void);

/* device memory helpers for hosts without their own allocator
 * (the roles of cuda::make_unique / cpu2cuda / cuda2cpu, include/golclass.h:36-121) */
int olx_dev_alloc(olx_ctx* ctx, void** d_ptr, size_t bytes);
int olx_dev_free(olx_ctx* ctx, void* d_ptr);
int olx_copy_h2d(olx_ctx* ctx, void* d_dst, const void* h_src, size_t bytes); /* synchronous */
int olx_copy_d2h(olx_ctx* ctx, void* h_dst, const void* d_src, size_t bytes); /* synchronous */
int olx_dev_memset(olx_ctx* ctx, void* d_ptr, int value, size_t bytes);

/* ---- propagation (include/olprop.h, include/golprop.h) ------------------ */
/* AsmProp: in place on an ny x nx field, pad 2x, IFFT(FFT(u) H)/M^2, crop.
 * Replaces ol::AsmProp (include/olprop.h:227-248) and gAsmProp (include/golprop.h:506-534). */
int olx_asm_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                double lambda, double z, uint32_t flags);
/* FresnelProp / FresnelPropBandLimit: circconv(pad u, h)/M^2, crop.  Replaces
 * ol::FresnelProp (include/olprop.h:125-148), FresnelPropBandLimit (:150-170),
 * gFresnelProp (include/golprop.h:167-198), gFresnelPropBandLimit (:258-302). */
int olx_fresnel_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                    double lambda, double z, uint32_t flags);
/* ARSSFresnelProp: scaled/shifted Fresnel, src -> dst (may alias).  Replaces
 * ol::ARSSFresnelProp (include/golprop.h:941-995). */
int olx_arss_c64(olx_ctx* ctx, const void* d_src, void* d_dst, int64_t ny, int64_t nx,
                 double dy, double dx, double lambda, double z, double s, double oy,
                 double ox, uint32_t flags);
/* Prop dispatcher with AUTO selection; *chosen receives OLX_MODE_*.  Replaces
 * ol::Prop (include/olprop.h:656-682, include/golprop.h:578-604). */
int olx_prop_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                 double lambda, double z, int mode, int* chosen);
/* host-only checks: 1 true, 0 false.  AsmPropCheck include/olprop.h:209-225,
 * FresnelPropCheck :109-123 */
int olx_asm_prop_check(int64_t ny, int64_t nx, double dy, double dx, double lambda, double z);
int olx_fresnel_prop_check(int64_t ny, int64_t nx, double dy, double dx, double lambda,
                           double z);
/* materialise the centred ASM transfer function H[n][m], u = (m - nx/2) du,
 * v = (n - ny/2) dv (AsmTransferF, include/olprop.h:172-189; OLX_BANDLIMIT adds the
 * window of :250-305; evanescent entries are 0) into a complex64 ny x nx array */
int olx_asm_transfer_c64(olx_ctx* ctx, void* d_H, int64_t ny, int64_t nx, double dv, double du,
                         double lambda, double z, uint32_t flags);
/* materialise the centred Fresnel impulse response h = exp(i pi (x^2 + y^2)/(lambda z)),
 * x = (m - nx/2) dx, zero where |m - nx/2| >= wlim or |n - ny/2| >= hlim
 * (FresnelResponse / FresnelResponseBandLimit, include/olprop.h:8-84; pass
 * INT32_MAX limits for no window) */
int olx_fresnel_response_c64(olx_ctx* ctx, void* d_h, int64_t ny, int64_t nx, double dy,
                             double dx, double lambda, double z, int64_t wlim, int64_t hlim);
/* plain unnormalised 2-D FFT in place (sign -1 forward, +1 backward);
 * the role of ol::fft / ol::ifft (include/olutils.h:19-66, include/golutils.h:55-75) */
int olx_fft2d_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, int sign);

/* ---- point-cloud CGH (include/olcgh.h, include/golcgh.h) ---------------- */
/* U[h,w] = sum_j A_j exp(i (theta_j(h,w) + phi_j)), theta = pi ((px w - x)^2 +
 * (py h - y)^2)/(lambda z) (Fresnel, _CghFresnel include/olcgh.h:296-313, gCgh
 * include/golcgh.h:108-145) or 2 pi r/lambda (OLX_CGH_EXACT_R).  Outputs are
 * optional: d_U (complex64), d_phase (atan2, include/olcgh.h:291), d_amp (|U|). */
int olx_cgh_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U, float* d_phase,
                float* d_amp, int64_t ny, int64_t nx, double py, double px, double lambda,
                uint32_t flags);
/* same sum with FP64 arithmetic and a complex128 output (the FP64 column) */
int olx_cgh_c128(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U, int64_t ny,
                 int64_t nx, double py, double px, double lambda, uint32_t flags);
/* phase-only / amplitude encodings of a complex64 field (either output may be NULL) */
int olx_encode_c64(olx_ctx* ctx, const void* d_U, float* d_phase, float* d_amp, int64_t count);

/* ---- WRP (include/golcgh.h:291-581) ------------------------------------- */
/* WRPStep1 / WRPStep1_D: windowed accumulation of exp(i pi rho^2/(lambda dz))/|dz|
 * into the ny x nx plane at plane_z; p_limit is the pitch that sets the window. */
int olx_wrp_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U, int64_t ny,
                int64_t nx, double p, double lambda, double plane_z, double p_limit,
                uint32_t flags);
/* WRPMethod / WRPMethod_D: plane rule, step 1, then WRPStep2 propagation by mode
 * (ASM / FRESNEL / AUTO as include/golcgh.h:494-517).  *plane_z_out may be NULL. */
int olx_wrp_method_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                       int64_t ny, int64_t nx, double p, double lambda, double zmin,
                       double p_limit, float s, int mode, uint32_t flags,
                       double* plane_z_out);
/* plane position rule of WRPMethod(_D), include/golcgh.h:524-533 / :545-554 */
double olx_wrp_plane_z(int64_t ny, int64_t nx, double p, double lambda, double zmin,
                       double p_limit, float s, int square);

/* ---- layered Fresnel CGH on MFMA (no reference counterpart) -------------- */
/* Points are binned into nlayers uniform depth layers over [zmin, zmax) and
 * snapped to the layer centres; each layer is the complex GEMM
 * U += Y_l diag(a_l) X_l^T of unit-modulus exponential factors, stored in
 * F16 / BF16 / split BF16 with FP32 accumulation. */
int olx_cgh_layered_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                        int64_t ny, int64_t nx, double py, double px, double lambda,
                        int nlayers, double zmin, double zmax, int precision,
                        uint32_t flags);
/* the snapping rule, on the host: layer l = clamp(floor((z - zmin)/(zmax - zmin) * L), 0, L-1),
 * z_l = zmin + (l + 0.5) (zmax - zmin)/L */
double olx_layer_depth(double z, double zmin, double zmax, int nlayers);

/* ---- instrumentation ---------------------------------------------------- */
/* per-kernel device time measured with HIP events on the context stream */
int olx_ctx_enable_timing(olx_ctx* ctx, int on);
int olx_ctx_reset_timing(olx_ctx* ctx);
/* writes "name count total_ms\n" lines into buf (NUL terminated) */
int olx_ctx_timing_report(olx_ctx* ctx, char* buf, size_t buflen);
/* plain device copy kernel (calibrates the HBM byte counters) */
int olx_stream_copy(olx_ctx* ctx, void* d_dst, const void* d_src, size_t bytes);
/* On-device ceilings (no reference counterpart; measurement only): runs a kernel
 * that does nothing but the named operation on every CU and returns its rate:
 * FP32 / FP64 FMA in flop/s (2 per FMA), hardware sine in evaluations/s, FP16 /
 * BF16 32x32x16 MFMA in flop/s.  Synchronises the context stream. */
#define OLX_PROBE_FMA_F32 0
#define OLX_PROBE_FMA_F64 1
#define OLX_PROBE_SIN_F32 2
#define OLX_PROBE_MFMA_F16 3
#define OLX_PROBE_MFMA_BF16 4
int olx_peak_probe(olx_ctx* ctx, int kind, double* rate);

#ifdef __cplusplus
}
#endif
#endif
