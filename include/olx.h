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
 * with respect to the host unless an entry point says otherwise.  A context's
 * workspaces and cached tables serve every stream it is bound to: rebinding it to
 * another stream makes that stream wait for the work queued on the previous one,
 * so calls on different streams through one context are ordered, not overlapped.
 *
 * Cached tables (Fresnel / ARSS chirp spectra, pre- and post-chirps, twiddles of
 * non-power-of-two lengths) are kept per context, keyed by their parameters; the
 * least recently used is freed beyond 64 entries, and olx_ctx_clear_tables frees
 * them all.
 *
 * Sizes: propagation takes 2 <= ny, nx <= 8192 whose padded lengths 2n have no
 * prime factor above 13 (the reference's FFTW / cuFFT take any length; 1920 x
 * 1080 and 3840 x 2160, its demo sizes, qualify).  Power-of-two sides from 4 run
 * the fused three-pass kernels; other sides the mixed-radix passes.  Other sizes
 * fail with OLX_EUNSUPPORTED.
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
#define OLX_EIO (-5)          /* file cannot be opened, read or written */

/* propagation flags */
#define OLX_BANDLIMIT 0x1u    /* ASM band limit |u| < 1/(lambda sqrt((2 du z)^2 + 1)), include/olprop.h:250-305 */
#define OLX_ADJOINT 0x2u      /* apply the adjoint (conjugate transfer function) */
#define OLX_FRESNEL_BL_CPU 0x10u /* Fresnel window, per-axis limits (include/olprop.h:26-54) */
#define OLX_FRESNEL_BL_GPU 0x20u /* Fresnel window, x-limit on both axes (include/golprop.h:223-238) */
#define OLX_ARSS_CPU 0x800u      /* ARSS with the CPU shiftedFresnelProp semantics: no windows and
                                    net scale circconv / M^4 (include/olprop.h:338-408) */
#define OLX_EXTRA_INV_M2 0x2000u /* ASM: multiply the result by one more 1/M^2, the net scale of
                                    mAsmProp (include/molprop.h:80-101) */
#define OLX_PREC_F32 0x8000u      /* ASM and Fresnel CGH with the reference's PREC_T = float
                                    arithmetic (its tests' default, tests/para.h:8): float
                                    frequencies / coordinates and phases, cos / sin of the float
                                    phase (AsmTransferF include/olprop.h:172-189, _CghFresnel
                                    include/olcgh.h:296-313).  Far less accurate than the default
                                    (FP64 phases); for reproducing float-PREC_T results */
#define OLX_SHIFTED_CPU 0x1000u  /* shifted ASM with the CPU shiftedAsmProp semantics: the shifted
                                    band-limit window and net scale 1/M^2 (include/olprop.h:250-330) */

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
#define OLX_LAYERED_TILE128 0x4000u /* layered: use the 128 x 128 work-group tile even where the
                                      sides allow 256 x 128 (same sums, same results) */

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
int olx_ctx_clear_tables(olx_ctx* ctx); /* synchronises the stream, frees every cached table */
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

/* ---- array helpers (include/olutils.h, include/golutils.h) -------------- */
/* Window copy: d_dst is dny x dnx, dst[h][w] = src[h - oy][w - ox] inside the
 * sny x snx source and 0 elsewhere (elements of 1, 2, 4, 8 or 16 bytes; source and
 * destination distinct).  With oy = (dny - sny) / 2, ox = (dnx - snx) / 2 it is
 * zeropadding (include/olutils.h:133-178) / gzeropadding (include/golutils.h:
 * 189-244); with oy = -sny / 4, ox = -snx / 4 into sny/2 x snx/2 del_zero
 * (include/olutils.h:181-198) / gdel_zero (include/golutils.h:247-278); with
 * oy = -y, ox = -x cut (include/olutils.h:200-211) / gcut (include/golutils.h:
 * 281-308). */
int olx_window_copy(olx_ctx* ctx, const void* d_src, int64_t sny, int64_t snx, void* d_dst,
                    int64_t dny, int64_t dnx, int64_t oy, int64_t ox, size_t elem_bytes);
/* fftshift (include/olutils.h:95-117) / gfftshift (include/golutils.h:80-110): the
 * reference's quadrant swap in place, elements of 4, 8 or 16 bytes. */
int olx_fftshift(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, size_t elem_bytes);
/* c = a * b elementwise over n complex64 values (c may alias a or b): mul_complex
 * (include/olutils.h:121-130) / mult (include/golutils.h:123-143). */
int olx_mul_c64(olx_ctx* ctx, const void* d_a, const void* d_b, void* d_c, int64_t n);
/* c = a + b elementwise (add, include/golutils.h:145-164). */
int olx_add_c64(olx_ctx* ctx, const void* d_a, const void* d_b, void* d_c, int64_t n);
/* u *= s for n complex64 values: mul_scalar (include/olutils.h:81-89,
 * include/golutils.h:166-187), standardization with s = 1 / (ny nx) (:70-79). */
int olx_scale_c64(olx_ctx* ctx, void* d_u, int64_t n, float s);
/* complex128 twins of olx_mul_c64 and olx_scale_c64 (host std::complex<double>
 * arrays of mul_complex / mul_scalar / standardization stay in FP64) */
int olx_mul_c128(olx_ctx* ctx, const void* d_a, const void* d_b, void* d_c, int64_t n);
int olx_scale_c128(olx_ctx* ctx, void* d_u, int64_t n, double s);

/* ---- propagation (include/olprop.h, include/golprop.h) ------------------ */
/* AsmProp: in place on an ny x nx field, pad 2x, IFFT(FFT(u) H)/M^2, crop.
 * Replaces ol::AsmProp (include/olprop.h:227-248) and gAsmProp (include/golprop.h:506-534). */
int olx_asm_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                double lambda, double z, uint32_t flags);
/* olx_asm_c64 without any device allocation: no counterpart in the reference.  For
 * callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, flags and results as
 * olx_asm_c64; the padded intermediate (ny x 2nx complex64) lives in a static
 * 64 MiB device array per GPU, set up by olx_ctx_create, so ny * nx <= 2048 * 2048
 * (larger fields fail with OLX_EINVAL before any work is queued).  Every call on one
 * GPU shares that array, across contexts and streams: calls that overlap on one
 * GPU race and must be ordered by the caller. */
int olx_asm_c64_noalloc(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                        double lambda, double z, uint32_t flags);
/* FresnelProp / FresnelPropBandLimit: circconv(pad u, h)/M^2, crop.  Replaces
 * ol::FresnelProp (include/olprop.h:125-148), FresnelPropBandLimit (:150-170),
 * gFresnelProp (include/golprop.h:167-198), gFresnelPropBandLimit (:258-302). */
int olx_fresnel_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                    double lambda, double z, uint32_t flags);
/* complex128 twins of olx_asm_c64 / olx_fresnel_c64 / olx_prop_c64 / olx_fft2d_c64:
 * FP64 throughout (transforms, twiddles, transfer function and spectra), the
 * reference's fftw3 (double) path for std::complex<double> fields (include/olprop.h:
 * 125-170, 227-248, 656-682; include/olutils.h:19-66).  d_u holds interleaved double
 * pairs.  Propagation sides in [2, 4096], the FFT's in [2, 8192], padded / transform
 * lengths with prime factors 2, 3, 5, 7, 11, 13 only; flags: ASM BANDLIMIT, ADJOINT,
 * EXTRA_INV_M2; Fresnel FRESNEL_BL_CPU / FRESNEL_BL_GPU, ADJOINT.  Evanescent waves
 * are 0 (the reference: NaN). */
int olx_asm_c128(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                 double lambda, double z, uint32_t flags);
int olx_fresnel_c128(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                     double lambda, double z, uint32_t flags);
int olx_prop_c128(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                  double lambda, double z, int mode, int* chosen);
int olx_fft2d_c128(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, int sign);
/* Batched propagation with fused factor fields.  d_u holds nb fields of ny x nx,
 * field k at d_u + k ny nx; each is replaced in place by post * Prop(pre * u_k),
 * where Prop is olx_asm_c64 (mode OLX_MODE_ASM, its flags) or olx_fresnel_c64
 * (OLX_MODE_FRESNEL, its flags) and pre / post are ny x nx complex64 factor fields
 * shared by every field (null: 1).  The factors are the optical-element multiplies
 * that the reference runs as passes of their own before or after a propagation
 * (Lensarry::mult_complex, include/Lensarry.h:262-270; Diffuser's random phase,
 * include/Diffuser.h:173-326; mult, include/golutils.h:123-143): here pass A
 * multiplies as it loads and pass C as it stores.  nb in [1, 65535] with the
 * padded intermediate (nb ny 2nx complex64) at most 2 GiB; fields follow the
 * propagation size rule.  No counterpart for nb > 1 in the reference, which
 * propagates one field per call.  d_img (nb ny nx bytes, optional) also receives
 * phase8bit(atan2) of every output value, the byte olx_phase8bit gives for the
 * output's phase (include/olimg.h:205-213), from the last pass's epilogue; img_flags
 * 0 or OLX_PHASE8_OFFSET. */
int olx_prop_batch_c64(olx_ctx* ctx, int mode, void* d_u, int64_t nb, int64_t ny, int64_t nx,
                       double dy, double dx, double lambda, double z, uint32_t flags,
                       const void* d_pre, const void* d_post, uint8_t* d_img,
                       uint32_t img_flags);
/* ARSSFresnelProp: scaled/shifted Fresnel, src -> dst (may alias).  Replaces
 * ol::ARSSFresnelProp (include/golprop.h:941-995). */
int olx_arss_c64(olx_ctx* ctx, const void* d_src, void* d_dst, int64_t ny, int64_t nx,
                 double dy, double dx, double lambda, double z, double s, double oy,
                 double ox, uint32_t flags);
/* shifted (off-axis) ASM, src -> dst (may alias): H = exp(2 pi i (w z + ox u + oy v))
 * maps an ny x nx source window to the destination window offset by (oy, ox) (in
 * length units, source minus destination).  No window and net scale 1 is
 * gshiftedAsmProp (include/golprop.h:536-576, its window is commented out);
 * OLX_BANDLIMIT adds the shifted band-limit window of shiftedAsmTransferF
 * (include/olprop.h:250-305); OLX_SHIFTED_CPU is ol::shiftedAsmProp (include/
 * olprop.h:307-330: window and net scale 1/M^2).  OLX_ADJOINT conjugates H. */
int olx_shifted_asm_c64(olx_ctx* ctx, const void* d_src, void* d_dst, int64_t ny, int64_t nx,
                        double dy, double dx, double lambda, double z, double oy, double ox,
                        uint32_t flags);
/* split propagation of an ny x nx field in place as by x bx blocks: every block is
 * propagated to every block with the shift of their offsets (in float, as the
 * reference computes them) and the contributions are summed.  mode OLX_MODE_ASM:
 * gshiftedAsmProp per pair (gsplitProp, include/golprop.h:773-813), or with
 * OLX_SHIFTED_CPU the CPU shiftedAsmProp (splitAsmProp, include/olprop.h:614-651);
 * OLX_MODE_FRESNEL: ARSS with s = 1 per pair (gsplitProp's FRESNEL branch through
 * shiftedProp, include/golprop.h:756-770).  Blocks follow the propagation size rule. */
int olx_split_prop_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                       double lambda, double z, int mode, int by, int bx, uint32_t flags);
/* tilted-plane propagation in place: kind OLX_TILT_SOURCE is TiltedSourceProp
 * (include/olprop.h:414-526), OLX_TILT_REFERENCE TiltedReferenceProp (:528-594).
 * The padded spectrum is resampled at frequencies rotated by angle about the y
 * axis (interp OLX_INTERP_BILINEAR / OLX_INTERP_BICUBIC: include/olutils.h:213-234,
 * 290-320), weighted by the rotation's Jacobian and an ASM (mode OLX_MODE_ASM) or
 * Fresnel (OLX_MODE_FRESNEL) transfer function.  Net scale 1/M^2, as the
 * reference's (it scales before and after the inverse transform).  Evanescent
 * samples are 0 (the reference: NaN).  Sides: powers of two in [4, 8192]. */
#define OLX_TILT_SOURCE 0
#define OLX_TILT_REFERENCE 1
#define OLX_INTERP_BILINEAR 0
#define OLX_INTERP_BICUBIC 1
int olx_tilted_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                   double lambda, double z, double angle, int interp, int mode, int kind);
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
/* materialise the centred shifted transfer function H = exp(2 pi i (w z + ox u + oy v)),
 * u = (m - nx/2) du, v = (n - ny/2) dv, evanescent entries 0.  OLX_BANDLIMIT adds the
 * off-axis window of the CPU form.  Replaces shiftedAsmTransferF (include/olprop.h:250-305,
 * with OLX_BANDLIMIT) and gshiftedAsmTransferF (include/golprop.h:336-419, window
 * commented out there: flags 0). */
int olx_shifted_asm_transfer_c64(olx_ctx* ctx, void* d_H, int64_t ny, int64_t nx, double dv,
                                 double du, double lambda, double z, double oy, double ox,
                                 uint32_t flags);
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
/* olx_cgh_c64 with the hologram written as 8-bit phase: d_img[h nx + w] =
 * phase8bit(atan2(U)) (include/olimg.h:205-213, the byte olx_phase8bit gives for
 * the phase olx_cgh_c64 writes), computed in the CGH kernel's epilogue instead of a
 * phase array and a second pass: the reference's Cgh -> phase8bit -> Save chain
 * (include/olcgh.h:275-295, include/olimg.h:205-213).  d_U optional (also stored);
 * flags as olx_cgh_c64 plus OLX_PHASE8_OFFSET (defined with olx_phase8bit below). */
int olx_cgh_img8_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                     uint8_t* d_img, int64_t ny, int64_t nx, double py, double px, double lambda,
                     uint32_t flags);
/* same sum with FP64 arithmetic and a complex128 output (the FP64 column) */
int olx_cgh_c128(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U, int64_t ny,
                 int64_t nx, double py, double px, double lambda, uint32_t flags);
/* _CghFresnelDiff (include/olcgh.h:315-349): point coordinates x, y, z in pixel
 * units, the phase in cycles theta_j(h, w) = p z_j / lambda + gamma_j / 2 ((w - x_j)^2 +
 * (h - y_j)^2), gamma_j = p / (lambda z_j), which the reference carries along each row
 * by theta <- theta + delta, delta <- delta + gamma (its theta and delta are declared
 * inside the column loop, so only w = 0 is defined there; here every w takes the
 * recurrence's exact value).  U = sum_j A_j exp(i (2 pi theta_j + phi_j)) (A = 1,
 * phi = 0 reproduce the reference); outputs d_U (complex64) and / or d_phase (atan2). */
int olx_cgh_fresnel_diff_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                             float* d_phase, int64_t ny, int64_t nx, double p, double lambda);
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
 * F16 / BF16 / split BF16 with FP32 accumulation.  Sides in [1, 8192]; the GEMM
 * runs on whole 128 x 128 or 256 x 128 tiles and stores only the pixels inside
 * the hologram. */
int olx_cgh_layered_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                        int64_t ny, int64_t nx, double py, double px, double lambda,
                        int nlayers, double zmin, double zmax, int precision,
                        uint32_t flags);
/* the snapping rule, on the host: layer l = clamp(floor((z - zmin)/(zmax - zmin) * L), 0, L-1),
 * z_l = zmin + (l + 0.5) (zmax - zmin)/L */
double olx_layer_depth(double z, double zmin, double zmax, int nlayers);

/* ---- multi-plane phase-only hologram loss (no reference counterpart) ------ */
/* u = exp(i phi) (phi: float, ny x nx); f_p = AsmProp_{zs[p]}(u) (2x pad, OLX_BANDLIMIT
 * honoured); L = sum_p mean((I_p / mean I_p - T_p / mean T_p)^2), I_p = |f_p|^2.  Writes L
 * to *d_loss (one device double) and, when d_grad is not NULL, dL/dphi to d_grad (float,
 * ny x nx), using the adjoint ASM.  zs and d_targets are host arrays of nplanes entries;
 * the targets are device float arrays of ny x nx.  Each plane's field is formed,
 * reduced, turned into its gradient and propagated back before the next plane. */
int olx_multiplane_mse_c64(olx_ctx* ctx, const float* d_phi, int64_t ny, int64_t nx, int nplanes,
                           const double* zs, const float* const* d_targets, double dy, double dx,
                           double lambda, uint32_t flags, double* d_loss, float* d_grad);

/* ---- point-cloud ingest (include/olobject.h) ----------------------------- */
/* .3d: int32 N, then N x (int32 x, y, z); .3df: int32 N, then N x (float x, y, z),
 * little-endian.  OLX_OBJ_AUTO picks the format by extension as objread does
 * (include/olobject.h:88-100).  Each point becomes x = v_x scale + ox (likewise y, z;
 * objread with offset and scale, include/olobject.h:102-162; scale 1 and a zero
 * offset give objread_3d / objread_3df, :9-86), amp = 1, phase = 0.  With pts ==
 * NULL only *n is set (the file's N); otherwise min(N, cap) points are written to
 * host memory and *n is set to N. */
#define OLX_OBJ_AUTO 0
#define OLX_OBJ_3D 1
#define OLX_OBJ_3DF 2
int olx_obj_read(const char* path, int format, olx_point* pts, int64_t cap, int64_t* n,
                 double scale, double ox, double oy, double oz);
/* device bounding box (objinfo's minmax_element, include/olobject.h:164-175):
 * mn[3] / mx[3] receive (x, y, z) minima / maxima (host pointers) */
int olx_obj_bounds(olx_ctx* ctx, const olx_point* d_pts, int64_t n, double* mn, double* mx);
/* objset on the device (include/olobject.h:195-264): scale the cloud so that its
 * x/y extent fits the range(s), centre x and y on the offset and put z's minimum
 * (OLX_OBJSET_ZMIN) or mid-range (OLX_OBJSET_ZMIDDLE) at oz.  y_range <= 0 selects
 * the single-range overload (scale = xy_range / max(range_x, range_y)), else
 * scale = min(x_range / range_x, y_range / range_y).  The ranges are rounded to
 * float first, as the reference's parameters are.  The arithmetic is that of the
 * reference's Object<double> (FP64, one rounding per operation); with
 * OLX_OBJSET_F32 or'ed into mode it is that of Object<float> (ranges, scale,
 * offsets and every point in FP32, coordinates and offsets rounded to float first). */
#define OLX_OBJSET_ZMIN 0
#define OLX_OBJSET_ZMIDDLE 1
#define OLX_OBJSET_F32 0x100
int olx_obj_set(olx_ctx* ctx, olx_point* d_pts, int64_t n, double x_range, double y_range,
                double ox, double oy, double oz, int mode);
/* objsort on the device (include/olobject.h:177-182): descending z.  Stable: points
 * with equal z (+0 and -0 are equal) keep their input order (the reference's
 * std::sort leaves that order unspecified). */
int olx_obj_sort_z(olx_ctx* ctx, olx_point* d_pts, int64_t n);

/* ---- 8-bit hologram encodings and BMP files (include/olimg.h) ------------- */
/* The device encodings take float (complex<float>) input, or double (complex<double>)
 * with OLX_IN_F64 or'ed into flags / mode: the reference's templates over REAL_T /
 * PREC_T.  Arithmetic is FP64 with one rounding per operation, rounding is C round()
 * (half away from zero), and the conversion of the rounded value to uint8 is the one
 * the reference's x86-64 build performs (cvttsd2si: modulo 256 inside the int32
 * range, 0 for NaN and outside it). */
#define OLX_IN_F64 0x100
/* phase8bit (include/olimg.h:205-213) on n device values: round((phi - pi) / 2pi
 * * 255), which is negative for phi in (-pi, pi) and so wraps modulo 256.
 * OLX_PHASE8_OFFSET maps round((phi + pi) / 2pi * 255) instead (no wrap for phi in
 * [-pi, pi]; no reference counterpart). */
#define OLX_PHASE8_OFFSET 0x1u
int olx_phase8bit(olx_ctx* ctx, const void* d_phase, uint8_t* d_img, int64_t n, uint32_t flags);
/* quant8bit (include/olimg.h:178-199): round((v - min) / (max - min) * 255) with the
 * extremes of the n values; all-equal input gives 0 (the reference's 0 / 0 NaN,
 * converted as above).  flags: 0 or OLX_IN_F64. */
int olx_quant8bit(olx_ctx* ctx, const void* d_real, uint8_t* d_img, int64_t n, uint32_t flags);
/* complex2img (include/olimg.h:314-391) on n complex device values, fused (two
 * reads of u, no intermediate array; one read for PHASE): REAL, IMAG, AMP
 * (sqrt(re*re + im*im), products and sum in the element type, square root in FP64)
 * or PHASE (atan2 in FP64, fixed range [-pi, pi]), mapped by
 * round(255 (v - min) / (max - min)).  OLX_IMG_REAL_LEGACY is the bool overload's
 * real mode, whose maximum starts at 0 (:314-346); its amplitude mode equals
 * OLX_IMG_AMP.  mode may carry OLX_IN_F64. */
#define OLX_IMG_PHASE 0
#define OLX_IMG_REAL 1
#define OLX_IMG_IMAG 2
#define OLX_IMG_AMP 3
#define OLX_IMG_REAL_LEGACY 4
int olx_complex2img(olx_ctx* ctx, const void* d_u, uint8_t* d_img, int64_t n, int mode);
/* complex2real (include/olimg.h:393-420): REAL / IMAG / AMP / PHASE into the element
 * type (float, or double with OLX_IN_F64 in mode) */
int olx_complex2real(olx_ctx* ctx, const void* d_u, void* d_real, int64_t n, int mode);
/* bmpwrite (include/olimg.h:136-165): 8-bit grey BMP with a 256-entry palette from
 * a host image of ny rows of nx bytes, written in order, each row followed by
 * nx % 4 zero bytes (the reference's padding) */
int olx_bmp_write(const char* path, const uint8_t* img, int ny, int nx);
/* bmpread (include/olimg.h:56-99): skips the 14 + 40 + 1024 header bytes, then reads
 * ny rows of nx bytes, each followed by nx % 4 skipped bytes */
int olx_bmp_read(const char* path, uint8_t* img, int ny, int nx);

/* ---- optical elements (include/Diffuser.h, include/Lensarry.h) ---------- */
/* hologram step (ol::HOLOGRAMSTEP, include/oldefine.h:11-13) */
#define OLX_STEP_RECORD 0
#define OLX_STEP_PLAY 1
/* Diffuser::random_phase, CPU form (include/Diffuser.h:191-215): pixel (h, w), in
 * row-major order, takes two std::mt19937(seed) draws (a 53-bit value v) and
 * multiplies u at ((h + oy) mod ny, (w + ox) mod nx) by exp(i phase), phase =
 * range (v - 0.5) (PLAY) or its negative (RECORD), rounded to float.  The stream is
 * generated on the host (it is sequential) and applied on the device; synchronises
 * the context stream. */
int olx_diffuser_random_phase_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, int step,
                                  uint32_t seed, float range, int oy, int ox);
/* the GPU diffuser kernels (include/Diffuser.h:14-129) on a field, given the uniform
 * array d_b the reference takes from cuRAND: exp(i (b - 0.5) scalar) per pixel
 * (PHASE, KernelMultPhase), per column (XPHASE) or row (YPHASE); pi where b > 0.5
 * (BITRANDOM); per blocksize^2 block (BLOCK) plus a (x + y) / blocksize ramp
 * (BLOCKSTEP); exp(i F32_PI) on the even checkerboard (BITPHASE, b unused). */
#define OLX_DIFFUSER_PHASE 0
#define OLX_DIFFUSER_XPHASE 1
#define OLX_DIFFUSER_YPHASE 2
#define OLX_DIFFUSER_BITRANDOM 3
#define OLX_DIFFUSER_BLOCK 4
#define OLX_DIFFUSER_BLOCKSTEP 5
#define OLX_DIFFUSER_BITPHASE 6
int olx_diffuser_apply_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, const float* d_b,
                           float scalar, int kind, int blocksize);
/* n uniform floats in (0, 1] from a Philox4x32-10 counter stream keyed by seed (no
 * reference counterpart: cuRAND's MT19937 ordering is its own) */
int olx_uniform_f32(olx_ctx* ctx, float* d_out, int64_t n, uint64_t seed);
/* lens array of Lensarry (include/Lensarry.h:84-311): square lenslets of
 * round(blockwidth / pitch) pixels tiling the field, each the Fresnel response at
 * z = -f (PLAY) or +f (RECORD) computed in float, shifted by (oy, ox) pixels;
 * p_limit > 0 windows each lenslet (FresnelResponseBandLimit, include/olprop.h:56-84);
 * OLX_LENS_CIRCLE: LensarryCircle's hexagonal rows of circular lenslets on a field
 * of ones (:383-447).  The field is multiplied by the lens array (mult_complex), or
 * overwritten by it with OLX_LENS_WRITE (gen_complex). */
#define OLX_LENS_CIRCLE 0x1u
#define OLX_LENS_WRITE 0x2u
int olx_lens_array_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, float pitch,
                       float blockwidth, float f, float lambda, int step, int oy, int ox,
                       float p_limit, uint32_t flags);
/* Lensarry::duplicate (kind 0), duplicate_random (kind 1) and duplicate_multrandom
 * (kind 2) (include/Lensarry.h:137-252): the bp x bp block d_block tiles the ny x nx
 * field d_out with local shift (on, om); for kinds 1 and 2, pixels whose local
 * position lies outside the window |. - bp/2| < floor(|lambda f / (2 p_limit)| / pitch)
 * take exp(i phase) (kind 1, phase 2 pi (v - 0.5)) or the block value times
 * exp(i phase) (kind 2, phase pi/2 (v - 0.5)), v the 53-bit value of two draws of
 * std::mt19937(0) taken in the reference's block traversal order (generated on the
 * host; the call then waits for its upload). */
int olx_lens_tile_c64(olx_ctx* ctx, const void* d_block, int bp, void* d_out, int64_t ny,
                      int64_t nx, int on, int om, float pitch, float f, float lambda,
                      float p_limit, int kind);

/* ---- instrumentation ---------------------------------------------------- */
/* per-kernel device time measured with HIP events on the context stream.  on = 0
 * stops (and resolves the pending events), on = 1 times every launch, on = k > 1
 * times every k-th launch (the events' own cost, about 20 us per three-kernel
 * propagation at 4096^2, then falls on one launch in k). */
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
