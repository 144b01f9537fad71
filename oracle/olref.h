/*
 * olref.h -- CPU restatement of OpenOL's diffraction / CGH path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product path (openol_amd/, include/) never links or calls it.
 *
 * Each function restates the reference algorithm named in its comment, keeping
 * the reference's expression order with PREC_T = double (coordinates/phases)
 * and COMPLEX_T = float ("c64") or double ("c128").  The FFT is this file's own
 * radix-2 FFT (FFTW, which the reference links, is not installed here): forward
 * uses e^{-i}, backward e^{+i}, both unnormalised (FFTW convention,
 * include/olutils.h:27,52 in the reference).
 *
 * Deliberate divergences from the reference (shared with the HIP library):
 *   - evanescent angular-spectrum components (1/lambda^2 - u^2 - v^2 < 0) are set
 *     to zero; the reference takes sqrt of a negative number and produces NaN
 *     (include/olprop.h:184).
 *   - window limits that overflow int (e.g. ARSS with s == 1) mean "no window";
 *     the reference converts +inf to int (undefined behaviour).
 *   - CGH / WRP phases use the double-precision pi, not F32_PI.
 */
#ifndef OLREF_H
#define OLREF_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* flags shared by the propagation routines */
#define OLREF_BANDLIMIT   0x1u   /* Matsushima band limit (olprop.h:250-305 at ox=oy=0) */
#define OLREF_ADJOINT     0x2u   /* conjugate transfer function (adjoint operator) */
#define OLREF_PREC_F32    0x4u   /* emulate PREC_T=float coordinate/phase arithmetic */
/* Fresnel band-limit variants */
#define OLREF_FRESNEL_BL_CPU 0x10u /* per-axis limits, olprop.h:26-54 */
#define OLREF_FRESNEL_BL_GPU 0x20u /* x-limit on both axes, golprop.h:223-238 */
/* CGH flags */
#define OLREF_CGH_EXACT_R 0x100u  /* exact spherical r instead of the Fresnel quadratic */
/* WRP flags */
#define OLREF_WRP_SQUARE  0x200u  /* square window D = sqrt(2) W (golcgh.h:391-468) */
#define OLREF_SHIFT_WINDOW 0x1000u /* shifted ASM: shiftedAsmTransferF window (olprop.h:250-305) */
#define OLREF_SHIFT_CPU_SCALE 0x2000u /* shifted ASM: the CPU's second 1/M^2 (olprop.h:325-329) */

/* 2-D FFT, in place, row-major ny x nx, sign -1 forward / +1 backward, unnormalised */
int olref_fft2d_c64(float* data, int64_t ny, int64_t nx, int sign);
int olref_fft2d_c128(double* data, int64_t ny, int64_t nx, int sign);

/* AsmProp, include/olprop.h:227-248 (+ AsmTransferF :172-189, band limit :250-305) */
int olref_asm_c64(float* u, int64_t ny, int64_t nx, double dy, double dx,
                  double lambda, double z, unsigned flags);
int olref_asm_c128(double* u, int64_t ny, int64_t nx, double dy, double dx,
                   double lambda, double z, unsigned flags);

/* FresnelProp / FresnelPropBandLimit, include/olprop.h:125-170 (+ :8-54);
 * GPU band-limit variant golprop.h:258-302 */
int olref_fresnel_c64(float* u, int64_t ny, int64_t nx, double dy, double dx,
                      double lambda, double z, unsigned flags);
int olref_fresnel_c128(double* u, int64_t ny, int64_t nx, double dy, double dx,
                       double lambda, double z, unsigned flags);

/* ARSSFresnelProp, include/golprop.h:816-995 (src and dst may alias) */
int olref_arss_c64(const float* src, float* dst, int64_t ny, int64_t nx, double dy,
                   double dx, double lambda, double z, double s, double oy,
                   double ox, unsigned flags);
int olref_arss_c128(const double* src, double* dst, int64_t ny, int64_t nx, double dy,
                    double dx, double lambda, double z, double s, double oy,
                    double ox, unsigned flags);

/* shiftedFresnelProp (CPU ARSS analogue), include/olprop.h:338-408: no windows,
 * u scaled by 1/M^2 twice and the output once more (net circconv / M^4); in place */
int olref_shifted_fresnel_c64(float* u, int64_t ny, int64_t nx, double dy, double dx,
                              double lambda, double z, double s, double oy, double ox);
int olref_shifted_fresnel_c128(double* u, int64_t ny, int64_t nx, double dy, double dx,
                               double lambda, double z, double s, double oy, double ox);

/* shifted ASM: shiftedAsmProp include/olprop.h:307-330 with shiftedAsmTransferF
 * :250-305 (flags SHIFT_WINDOW | SHIFT_CPU_SCALE), or gshiftedAsmProp
 * include/golprop.h:536-576 (flags 0: no window, net scale 1); ADJOINT conjugates
 * H.  src and dst may alias. */
int olref_shifted_asm_c64(const float* src, float* dst, int64_t ny, int64_t nx, double dy,
                          double dx, double lambda, double z, double oy, double ox,
                          unsigned flags);
int olref_shifted_asm_c128(const double* src, double* dst, int64_t ny, int64_t nx, double dy,
                           double dx, double lambda, double z, double oy, double ox,
                           unsigned flags);
/* splitAsmProp include/olprop.h:614-651 / gsplitProp include/golprop.h:773-813, in
 * place: mode 0 ASM (shifted ASM per block pair with the given flags), 1 FRESNEL
 * (ARSS with s = 1 per block pair, golprop.h:756-770) */
int olref_split_prop_c64(float* u, int64_t ny, int64_t nx, double dy, double dx,
                         double lambda, double z, int mode, int by, int bx, unsigned flags);

/* tilted-plane propagation in place on complex<double> data: kind 0 TiltedSourceProp
 * (include/olprop.h:414-526), 1 TiltedReferenceProp (:528-594); interp 0 bilinear,
 * 1 bicubic (include/olutils.h:213-234, 290-320); mode 0 ASM, 1 FRESNEL */
int olref_tilted_c128(double* u, int64_t ny, int64_t nx, double dy, double dx, double lambda,
                      double z, double angle, int interp, int mode, int kind);

/* MT19937 (the std::mt19937 engine): the n-th output for a seed (n >= 1) */
uint32_t olref_mt19937_nth(uint32_t seed, int64_t n);
/* Diffuser::random_phase, CPU form (include/Diffuser.h:191-215), complex<float> in place */
int olref_diffuser_random_phase(float* u, int64_t height, int64_t width, int step,
                                uint32_t seed, float range, int oy, int ox);
/* the GPU diffuser kernels (include/Diffuser.h:14-129) given the uniform array b */
int olref_diffuser_apply(float* u, int64_t height, int64_t width, const float* b, float scalar,
                         int kind, int blocksize);
/* Lensarry / LensarryCircle lens field (include/Lensarry.h:137-311, 406-447) */
int olref_lens_tile(const float* block, int bp, float* out, int64_t height, int64_t width,
                    int on, int om, float pitch, float f, float lambda, float p_limit, int kind);
int olref_lens_block(float* out, int bp, float pitch, float f, float lambda, int step,
                     float p_limit);
int olref_lens_array(float* out, int64_t height, int64_t width, float pitch, float blockwidth,
                     float f, float lambda, int step, int on, int om, float p_limit, int circle);

/* Prop dispatcher, include/olprop.h:656-682; returns chosen mode in *chosen
 * (0 ASM, 1 FRESNEL, 2 FRESNEL band-limited) */
int olref_prop_c64(float* u, int64_t ny, int64_t nx, double dy, double dx,
                   double lambda, double z, int mode, int* chosen);
int olref_asm_prop_check(int64_t ny, int64_t nx, double dy, double dx,
                         double lambda, double z);       /* olprop.h:209-225 */
int olref_fresnel_prop_check(int64_t ny, int64_t nx, double dy, double dx,
                             double lambda, double z);   /* olprop.h:109-123 */

/* Point-cloud CGH, include/olcgh.h:296-313 (complex out).  Points are SoA;
 * amp / phase may be NULL (reference behaviour: A = 1, phi = 0).
 * c64: float accumulation as complex<float> += complex<double>; c128: double. */
int olref_cgh_c64(const double* x, const double* y, const double* z,
                  const double* amp, const double* phase, int64_t n,
                  float* U, int64_t ny, int64_t nx, double py, double px,
                  double lambda, unsigned flags);
int olref_cgh_c128(const double* x, const double* y, const double* z,
                   const double* amp, const double* phase, int64_t n,
                   double* U, int64_t ny, int64_t nx, double py, double px,
                   double lambda, unsigned flags);
/* same sum evaluated only at a list of pixels (h[i], w[i]) -- for sampled checks
 * at configuration size; FP64 throughout */
int olref_cgh_pixels_c128(const double* x, const double* y, const double* z,
                          const double* amp, const double* phase, int64_t n,
                          const int64_t* ph, const int64_t* pw, int64_t npix,
                          double* out, double py, double px, double lambda,
                          unsigned flags);

/* WRP step 1, include/golcgh.h:291-378 (circle) and :391-491 (square, flag);
 * p_limit is the pitch used for the window (p for the plain overloads). */
/* _CghFresnelDiff (include/olcgh.h:315-349): the recurrence CGH of a cloud in pixel
 * units, complex128 output (amp / phase may be NULL) */
int olref_cgh_fresnel_diff(const double* x, const double* y, const double* z, const double* amp,
                           const double* phase, int64_t n, int64_t ny, int64_t nx, double p,
                           double lambda, double* out);
int olref_wrp_c64(const double* x, const double* y, const double* z, int64_t n,
                  float* U, int64_t ny, int64_t nx, double p, double lambda,
                  double plane_z, double p_limit, unsigned flags);
int olref_wrp_c128(const double* x, const double* y, const double* z, int64_t n,
                   double* U, int64_t ny, int64_t nx, double p, double lambda,
                   double plane_z, double p_limit, unsigned flags);
/* WRPMethod(_D) plane position rule, golcgh.h:519-581 (returns float value) */
/* the c128 WRP sum at the listed pixels (ph[i], pw[i]) only, into out (complex128) */
int olref_wrp_pixels_c128(const double* x, const double* y, const double* z, int64_t n,
                          const int64_t* ph, const int64_t* pw, int64_t npix, double* out,
                          double p, double lambda, double plane_z, double p_limit,
                          unsigned flags);
double olref_wrp_plane_z(int64_t ny, int64_t nx, double p, double lambda,
                         double zmin, double p_limit, float s, int square);

#ifdef __cplusplus
}
#endif
int olref_shifted_asm_transfer_c128(double* Hout, int64_t ny, int64_t nx, double dv, double du,
                                    double lambda, double z, double oy, double ox,
                                    unsigned flags);
#endif
