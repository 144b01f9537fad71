/*
 * olref.c -- CPU restatement of OpenOL's diffraction / CGH path.
 * TEST INFRASTRUCTURE ONLY (the checker); see olref.h for scope and the list of
 * deliberate divergences.  Every routine cites the reference lines it follows.
 */
#include "olref.h"
#include <complex.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef float complex cf;
typedef double complex cd;

static int is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

/* int window limit from floor(x); "no window" when it does not fit an int */
static int64_t win_lim(double v) {
    double f = floor(v);
    if (!(f == f) || f >= (double)INT_MAX) return (int64_t)INT_MAX;
    if (f <= (double)INT_MIN) return (int64_t)INT_MIN;
    return (int64_t)f;
}

/* ------------------------------------------------------------------------- */
/* FFT: radix-2 iterative (power-of-two lengths) or recursive mixed radix by    */
/* the smallest prime factor (other lengths; prime lengths by a direct DFT).    */
/* Twiddles are computed in double and rounded to the working precision.       */
/* ------------------------------------------------------------------------- */
static int64_t small_factor(int64_t n) {
    for (int64_t p = 2; p * p <= n && p <= 256; p++)
        if (n % p == 0) return p;
    return n;
}
#define DEFINE_FFT(SUF, CT, RT, COSF, SINF)                                         \
static CT* twiddles_##SUF(int64_t n, int sign) {                                   \
    CT* tw = (CT*)malloc(sizeof(CT) * (size_t)(n > 1 ? n : 1));                     \
    for (int64_t k = 0; k < n; k++) {                                               \
        double a = 2.0 * M_PI * (double)k / (double)n;                              \
        tw[k] = (RT)cos(a) + I * (RT)(sign * sin(a));                               \
    }                                                                               \
    return tw;                                                                      \
}                                                                                   \
/* out[k] = sum_j in[j s] W_n^{jk}, W_n^x = tw[x st] (tw holds W_N, st = N / n).  \
   Decimation in time by p, the smallest prime factor of n = p m: the p          \
   subsequences in[(r + p j) s] transform into out[r m .. r m + m), then          \
   X[k + m q] = sum_r W_p^{rq} (W_n^{rk} Y_r[k]) reads and writes the same set    \
   {k + m r} for each k.  A prime n is a direct DFT. */                          \
static void mrfft_##SUF(const CT* in, int64_t s, CT* out, int64_t n, const CT* tw,  \
                        int64_t st) {                                               \
    if (n == 1) { out[0] = in[0]; return; }                                         \
    const int64_t p = small_factor(n), m = n / p;                                   \
    if (p == n) {                                                                   \
        for (int64_t k = 0; k < n; k++) {                                           \
            CT acc = 0;                                                             \
            for (int64_t j = 0; j < n; j++) acc += in[j * s] * tw[((j * k) % n) * st]; \
            out[k] = acc;                                                           \
        }                                                                           \
        return;                                                                     \
    }                                                                               \
    for (int64_t r = 0; r < p; r++) mrfft_##SUF(in + r * s, s * p, out + r * m, m, tw, st * p); \
    CT t[256];                                                                      \
    for (int64_t k = 0; k < m; k++) {                                               \
        for (int64_t r = 0; r < p; r++) t[r] = out[r * m + k] * tw[r * k * st];     \
        for (int64_t q = 0; q < p; q++) {                                           \
            CT acc = 0;                                                             \
            for (int64_t r = 0; r < p; r++) acc += t[r] * tw[((r * q) % p) * m * st]; \
            out[k + m * q] = acc;                                                   \
        }                                                                           \
    }                                                                               \
}                                                                                   \
static void fft1d_##SUF(CT* a, int64_t n, const CT* tw, CT* tmp) {                 \
    if (n <= 1) return;                                                             \
    if (!is_pow2(n)) {                                                              \
        mrfft_##SUF(a, 1, tmp, n, tw, 1);                                           \
        memcpy(a, tmp, sizeof(CT) * (size_t)n);                                     \
        return;                                                                     \
    }                                                                               \
    for (int64_t i = 1, j = 0; i < n; i++) { /* bit reversal */                     \
        int64_t bit = n >> 1;                                                       \
        for (; j & bit; bit >>= 1) j ^= bit;                                        \
        j ^= bit;                                                                   \
        if (i < j) { CT t = a[i]; a[i] = a[j]; a[j] = t; }                          \
    }                                                                               \
    for (int64_t len = 2; len <= n; len <<= 1) {                                    \
        int64_t half = len >> 1, step = n / len;                                    \
        for (int64_t i = 0; i < n; i += len) {                                      \
            for (int64_t k = 0; k < half; k++) {                                    \
                CT w = tw[k * step];                                                \
                CT u = a[i + k], v = a[i + k + half] * w;                           \
                a[i + k] = u + v;                                                   \
                a[i + k + half] = u - v;                                            \
            }                                                                       \
        }                                                                           \
    }                                                                               \
}                                                                                   \
static int fft2d_##SUF(CT* d, int64_t ny, int64_t nx, int sign) {                  \
    CT* twx = twiddles_##SUF(nx, sign);                                             \
    CT* twy = twiddles_##SUF(ny, sign);                                             \
    _Pragma("omp parallel")                                                         \
    {                                                                               \
        CT* tmp = (CT*)malloc(sizeof(CT) * (size_t)(nx > ny ? nx : ny) * 17);       \
        _Pragma("omp for schedule(static)")                                         \
        for (int64_t r = 0; r < ny; r++) fft1d_##SUF(d + r * nx, nx, twx, tmp);      \
        /* columns in blocks of 16 to stay cache friendly */                        \
        _Pragma("omp for schedule(static)")                                         \
        for (int64_t c0 = 0; c0 < nx; c0 += 16) {                                   \
            int64_t cb = nx - c0 < 16 ? nx - c0 : 16;                               \
            CT* col = tmp + ny;                                                     \
            for (int64_t r = 0; r < ny; r++)                                        \
                for (int64_t c = 0; c < cb; c++) col[c * ny + r] = d[r * nx + c0 + c];\
            for (int64_t c = 0; c < cb; c++) fft1d_##SUF(col + c * ny, ny, twy, tmp); \
            for (int64_t r = 0; r < ny; r++)                                        \
                for (int64_t c = 0; c < cb; c++) d[r * nx + c0 + c] = col[c * ny + r];\
        }                                                                           \
        free(tmp);                                                                  \
    }                                                                               \
    free(twx); free(twy);                                                           \
    return 0;                                                                       \
}

DEFINE_FFT(c64, cf, float, cosf, sinf)
DEFINE_FFT(c128, cd, double, cos, sin)

int olref_fft2d_c64(float* data, int64_t ny, int64_t nx, int sign) {
    if (!data || ny <= 0 || nx <= 0) return -1;
    return fft2d_c64((cf*)data, ny, nx, sign < 0 ? -1 : 1);
}
int olref_fft2d_c128(double* data, int64_t ny, int64_t nx, int sign) {
    if (!data || ny <= 0 || nx <= 0) return -1;
    return fft2d_c128((cd*)data, ny, nx, sign < 0 ? -1 : 1);
}

/* ------------------------------------------------------------------------- */
/* Array helpers: include/olutils.h:95-198                                    */
/* ------------------------------------------------------------------------- */
#define DEFINE_HELPERS(SUF, CT)                                                     \
/* zeropadding(src,dst,in_h,in_w,out_h,out_w), olutils.h:155-178 */                 \
static CT* pad_##SUF(const CT* in, int64_t ih, int64_t iw, int64_t oh, int64_t ow) {\
    int64_t yl = (oh - ih) / 2, xl = (ow - iw) / 2;                                 \
    int64_t yr = (oh - ih + 1) / 2, xr = (ow - iw + 1) / 2;                         \
    CT* out = (CT*)calloc((size_t)(oh * ow), sizeof(CT));                           \
    for (int64_t h = 0; h < oh; h++)                                                \
        for (int64_t w = 0; w < ow; w++)                                            \
            if (yl <= h && h < oh - yr && xl <= w && w < ow - xr)                   \
                out[h * ow + w] = in[(h - yl) * iw + w - xl];                       \
    return out;                                                                     \
}                                                                                   \
/* del_zero(in,out,in_h,in_w), olutils.h:181-198: keep [h/4,3h/4) x [w/4,3w/4) */    \
static void crop_##SUF(const CT* in, CT* out, int64_t ih, int64_t iw) {            \
    for (int64_t n = ih / 4; n < ih * 3 / 4; n++)                                   \
        for (int64_t m = iw / 4; m < iw * 3 / 4; m++)                               \
            out[m - iw / 4 + (n - ih / 4) * (iw / 2)] = in[m + n * iw];             \
}                                                                                   \
/* fftshift, olutils.h:95-117 (quadrant swap) */                                    \
static void fftshift_##SUF(CT* u, int64_t height, int64_t width) {                 \
    int64_t hh = height / 2, hw = width / 2;                                        \
    for (int64_t h = 0; h < hh; h++)                                                \
        for (int64_t w = 0; w < hw; w++) {                                          \
            int64_t i1 = h * width + w, i2 = (h + hh) * width + (w + hw);           \
            CT t = u[i1]; u[i1] = u[i2]; u[i2] = t;                                 \
            i1 = h * width + (w + hw); i2 = (h + hh) * width + w;                   \
            t = u[i1]; u[i1] = u[i2]; u[i2] = t;                                    \
        }                                                                           \
}

DEFINE_HELPERS(c64, cf)
DEFINE_HELPERS(c128, cd)

/* ------------------------------------------------------------------------- */
/* Angular spectrum: AsmProp olprop.h:227-248, AsmTransferF :172-189,        */
/* band-limit window shiftedAsmTransferF :250-305 evaluated at ox = oy = 0.   */
/* ------------------------------------------------------------------------- */
static void asm_transfer(cd* H, int64_t ny, int64_t nx, double dv, double du,
                         double lambda, double z, unsigned flags) {
    double ulim = INFINITY, vlim = INFINITY;
    if (flags & OLREF_BANDLIMIT) {
        /* ulm = 1/sqrt(pow(ox - 0.5/du,-2) z^2 + 1)/lambda, ulp likewise; at
         * ox = 0 the branch -sx <= ox < sx gives u0 = 0, uw = ulp + ulm */
        double ulm = 1.0 / sqrt(pow(0.0 - 0.5f / du, -2) * z * z + 1.0f) / lambda;
        double ulp = 1.0 / sqrt(pow(0.0 + 0.5f / du, -2) * z * z + 1.0f) / lambda;
        double vlm = 1.0 / sqrt(pow(0.0 - 0.5f / dv, -2) * z * z + 1.0f) / lambda;
        double vlp = 1.0 / sqrt(pow(0.0 + 0.5f / dv, -2) * z * z + 1.0f) / lambda;
        ulim = ulp + ulm; /* condition |u - u0| * 2 < uw */
        vlim = vlp + vlm;
    }
    double tmp = 1.0 / (lambda * lambda);
    #pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < ny; n++) {
        double v = (double)(n - ny / 2) * dv;
        for (int64_t m = 0; m < nx; m++) {
            int64_t idx = m + n * nx;
            double u = (double)(m - nx / 2) * du;
            double arg = tmp - u * u - v * v;
            if (arg < 0.0 || !(fabs(u) * 2 < ulim && fabs(v) * 2 < vlim)) {
                H[idx] = 0;
                continue;
            }
            double w = sqrt(arg);
            double phase = (flags & OLREF_BANDLIMIT) ? 2 * M_PI * (w * z + 0.0 * u + 0.0 * v)
                                                     : 2 * M_PI * w * z;
            double s = sin(phase);
            if (flags & OLREF_ADJOINT) s = -s;
            H[idx] = cos(phase) + I * s;
        }
    }
}

/* PREC_T = float emulation of AsmTransferF (informational variant) */
static void asm_transfer_f32(cd* H, int64_t ny, int64_t nx, float dv, float du,
                             float lambda, float z, unsigned flags) {
    float tmp = 1 / (lambda * lambda);
    #pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < ny; n++) {
        float v = (float)(n - ny / 2) * dv;
        for (int64_t m = 0; m < nx; m++) {
            float u = (float)(m - nx / 2) * du;
            float arg = tmp - u * u - v * v;
            if (arg < 0) { H[m + n * nx] = 0; continue; }
            float w = sqrtf(arg);
            float phase = (float)(2 * M_PI * w * z);
            float s = sinf(phase);
            if (flags & OLREF_ADJOINT) s = -s;
            H[m + n * nx] = cosf(phase) + I * s;
        }
    }
}

#define DEFINE_ASM(SUF, CT, RT)                                                     \
int olref_asm_##SUF(RT* uin, int64_t ny, int64_t nx, double dy, double dx,          \
                    double lambda, double z, unsigned flags) {                      \
    if (!uin || ny <= 0 || nx <= 0 || !(lambda > 0) || !(dx > 0) || !(dy > 0))      \
        return -1;                                                                  \
    CT* u = (CT*)uin;                                                               \
    int64_t ny2 = ny * 2, nx2 = nx * 2;                                             \
    CT* buf = pad_##SUF(u, ny, nx, ny2, nx2);                                       \
    cd* Hd = (cd*)malloc(sizeof(cd) * (size_t)(ny2 * nx2));                         \
    if (flags & OLREF_PREC_F32) {                                                   \
        float du = 1 / ((float)dx * nx2), dv = 1 / ((float)dy * ny2);               \
        asm_transfer_f32(Hd, ny2, nx2, dv, du, (float)lambda, (float)z, flags);     \
    } else {                                                                        \
        double du = 1 / (dx * nx2), dv = 1 / (dy * ny2);                            \
        asm_transfer(Hd, ny2, nx2, dv, du, lambda, z, flags);                       \
    }                                                                               \
    CT* H = (CT*)malloc(sizeof(CT) * (size_t)(ny2 * nx2));                          \
    for (int64_t i = 0; i < ny2 * nx2; i++) H[i] = (CT)Hd[i];                       \
    free(Hd);                                                                       \
    fftshift_##SUF(H, ny2, nx2);                                                    \
    fft2d_##SUF(buf, ny2, nx2, -1);                                                 \
    RT sc = (RT)(1.0 / (double)(ny2 * nx2));                                        \
    for (int64_t i = 0; i < ny2 * nx2; i++) { buf[i] = buf[i] * H[i]; buf[i] *= sc; }\
    fft2d_##SUF(buf, ny2, nx2, +1);                                                 \
    crop_##SUF(buf, u, ny2, nx2);                                                   \
    free(H); free(buf);                                                             \
    return 0;                                                                       \
}

DEFINE_ASM(c64, cf, float)
DEFINE_ASM(c128, cd, double)

/* ------------------------------------------------------------------------- */
/* Shifted ASM: shiftedAsmTransferF olprop.h:250-305 (window; the GPU twin,  */
/* golprop.h:333-356, has it commented out), shiftedAsmProp :307-330 and      */
/* gshiftedAsmProp golprop.h:536-576.                                         */
/* ------------------------------------------------------------------------- */
/* window centre and width on one axis, olprop.h:255-285 */
static void shift_window(double o, double d, double z, double lambda, double* c, double* w) {
    double sx = 0.5f / d;
    double lm = 1 / sqrt(pow((o - 0.5f / d), -2) * z * z + 1.0f) / lambda;
    double lp = 1 / sqrt(pow((o + 0.5f / d), -2) * z * z + 1.0f) / lambda;
    if (sx < o) {
        *c = 0.5f * (lp + lm);
        *w = lp - lm;
    } else if (-sx <= o && o < sx) {
        *c = 0.5f * (lp - lm);
        *w = lp + lm;
    } else {
        *c = -0.5f * (lp + lm);
        *w = lm - lp;
    }
}

static void shifted_asm_transfer(cd* H, int64_t ny, int64_t nx, double dv, double du,
                                 double lambda, double z, double oy, double ox, unsigned flags) {
    int win = (flags & OLREF_SHIFT_WINDOW) != 0;
    double u0 = 0, uw = 0, v0 = 0, vw = 0;
    if (win) {
        shift_window(ox, du, z, lambda, &u0, &uw);
        shift_window(oy, dv, z, lambda, &v0, &vw);
    }
    double tmp = 1 / (lambda * lambda);
    #pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < ny; n++) {
        double v = (n - ny / 2) * dv;
        for (int64_t m = 0; m < nx; m++) {
            int64_t idx = m + n * nx;
            double u = (m - nx / 2) * du;
            double arg = tmp - u * u - v * v;
            if ((win && !((fabs(u - u0) * 2 < uw) && (fabs(v - v0) * 2 < vw))) || arg < 0) {
                H[idx] = 0;
                continue;
            }
            double w = sqrt(arg);
            double phase = 2 * M_PI * (w * z + ox * u + oy * v);
            double s = sin(phase);
            if (flags & OLREF_ADJOINT) s = -s;
            H[idx] = cos(phase) + I * s;
        }
    }
}

/* the table itself, complex128 (shiftedAsmTransferF with OLREF_SHIFT_WINDOW,
 * gshiftedAsmTransferF without) */
int olref_shifted_asm_transfer_c128(double* Hout, int64_t ny, int64_t nx, double dv, double du,
                                    double lambda, double z, double oy, double ox,
                                    unsigned flags) {
    if (!Hout || ny <= 0 || nx <= 0) return -1;
    shifted_asm_transfer((cd*)Hout, ny, nx, dv, du, lambda, z, oy, ox, flags);
    return 0;
}

#define DEFINE_SHIFTED_ASM(SUF, CT, RT)                                             \
int olref_shifted_asm_##SUF(const RT* srcin, RT* dstin, int64_t ny, int64_t nx,     \
                            double dy, double dx, double lambda, double z,          \
                            double oy, double ox, unsigned flags) {                 \
    if (!srcin || !dstin || ny <= 0 || nx <= 0 || !(lambda > 0) || !(dx > 0) ||     \
        !(dy > 0))                                                                  \
        return -1;                                                                  \
    int64_t ny2 = ny * 2, nx2 = nx * 2;                                             \
    double du = 1 / (dx * nx2), dv = 1 / (dy * ny2);                                \
    CT* buf = pad_##SUF((const CT*)srcin, ny, nx, ny2, nx2);                        \
    cd* Hd = (cd*)malloc(sizeof(cd) * (size_t)(ny2 * nx2));                         \
    shifted_asm_transfer(Hd, ny2, nx2, dv, du, lambda, z, oy, ox, flags);           \
    CT* H = (CT*)malloc(sizeof(CT) * (size_t)(ny2 * nx2));                          \
    for (int64_t i = 0; i < ny2 * nx2; i++) H[i] = (CT)Hd[i];                       \
    free(Hd);                                                                       \
    fftshift_##SUF(H, ny2, nx2);                                                    \
    fft2d_##SUF(buf, ny2, nx2, -1);                                                 \
    RT sc = (RT)(1.0 / (double)(ny2 * nx2));                                        \
    for (int64_t i = 0; i < ny2 * nx2; i++) { buf[i] = buf[i] * H[i]; buf[i] *= sc; }\
    fft2d_##SUF(buf, ny2, nx2, +1);                                                 \
    CT* dst = (CT*)dstin;                                                           \
    crop_##SUF(buf, dst, ny2, nx2);                                                 \
    if (flags & OLREF_SHIFT_CPU_SCALE)                                              \
        for (int64_t i = 0; i < ny * nx; i++) dst[i] *= sc;                         \
    free(H); free(buf);                                                             \
    return 0;                                                                       \
}

DEFINE_SHIFTED_ASM(c64, cf, float)
DEFINE_SHIFTED_ASM(c128, cd, double)

/* ------------------------------------------------------------------------- */
/* Fresnel: FresnelProp olprop.h:125-148 with FresnelResponse :8-24;         */
/* FresnelPropBandLimit :150-170 with FresnelResponseBandLimit :26-54;        */
/* GPU band limit (x-limit on both axes) golprop.h:201-238.                   */
/* ------------------------------------------------------------------------- */
#define DEFINE_FRESNEL(SUF, CT, RT)                                                 \
int olref_fresnel_##SUF(RT* uin, int64_t ny, int64_t nx, double dy, double dx,      \
                        double lambda, double z, unsigned flags) {                  \
    if (!uin || ny <= 0 || nx <= 0 || !(lambda > 0) || z == 0) return -1;            \
    CT* u = (CT*)uin;                                                               \
    CT* buf = pad_##SUF(u, ny, nx, ny * 2, nx * 2);                                 \
    int64_t height = ny * 2, width = nx * 2;                                        \
    CT* h = (CT*)malloc(sizeof(CT) * (size_t)(height * width));                     \
    double tmp = 1 / (lambda * z);                                                  \
    int64_t hh = height / 2, hw = width / 2;                                        \
    int64_t wlim = INT_MAX, hlim = INT_MAX;                                         \
    if (flags & OLREF_FRESNEL_BL_CPU) {                                             \
        double xlim = fabs(lambda * z * 0.5 / dx), ylim = fabs(lambda * z * 0.5 / dy);\
        wlim = win_lim(xlim / dx); hlim = win_lim(ylim / dy);                       \
    } else if (flags & OLREF_FRESNEL_BL_GPU) {                                      \
        double xlim = fabs(lambda * z * 0.5 / dx);                                  \
        wlim = hlim = win_lim(xlim / dx);                                           \
    }                                                                               \
    for (int64_t n = 0; n < height; n++) {                                          \
        double y = ((double)(n - hh)) * dy;                                         \
        for (int64_t m = 0; m < width; m++) {                                       \
            int64_t idx = n * width + m;                                            \
            if (llabs(m - hw) < wlim && llabs(n - hh) < hlim) {                     \
                double x = ((double)(m - hw)) * dx;                                 \
                double phase = M_PI * (x * x + y * y) * tmp;                        \
                double s = sin(phase);                                              \
                if (flags & OLREF_ADJOINT) s = -s;                                  \
                h[idx] = (RT)cos(phase) + I * (RT)s;                                \
            } else h[idx] = 0;                                                      \
        }                                                                           \
    }                                                                               \
    fftshift_##SUF(h, height, width);                                               \
    fft2d_##SUF(buf, height, width, -1);                                            \
    fft2d_##SUF(h, height, width, -1);                                              \
    RT sc = (RT)(1.0 / (double)(height * width));                                   \
    for (int64_t i = 0; i < height * width; i++) {                                  \
        buf[i] *= sc; h[i] *= sc; buf[i] = buf[i] * h[i];                           \
    }                                                                               \
    fft2d_##SUF(buf, height, width, +1);                                            \
    crop_##SUF(buf, u, height, width);                                              \
    free(h); free(buf);                                                             \
    return 0;                                                                       \
}

DEFINE_FRESNEL(c64, cf, float)
DEFINE_FRESNEL(c128, cd, double)

/* ------------------------------------------------------------------------- */
/* ARSS: ARSSFresnelProp golprop.h:941-989 with stages _u :816-859,          */
/* _h :861-898 and _Cz :900-938.                                             */
/* ------------------------------------------------------------------------- */
#define DEFINE_ARSS(SUF, CT, RT)                                                    \
int olref_arss_##SUF(const RT* srcin, RT* dstin, int64_t ny, int64_t nx, double dy, \
                     double dx, double lambda, double z, double s, double oy,       \
                     double ox, unsigned flags) {                                   \
    (void)flags;                                                                    \
    if (!srcin || !dstin || ny <= 0 || nx <= 0 || !(lambda > 0) || z == 0 || s == 0)\
        return -1;                                                                  \
    const CT* src = (const CT*)srcin;                                               \
    CT* dst = (CT*)dstin;                                                           \
    int64_t ny2 = ny * 2, nx2 = nx * 2;                                             \
    /* stage u: pre-chirp inside window, zero outside */                            \
    {                                                                               \
        double xlim = lambda * fabs(z) * 0.5 / (s * fabs(s - 1) * dx);              \
        int64_t wlim = win_lim(xlim / dx);                                          \
        double ylim = lambda * fabs(z) * 0.5 / (s * fabs(s - 1) * dy);              \
        int64_t hlim = win_lim(ylim / dy);                                          \
        int64_t hny = ny / 2, hnx = nx / 2;                                         \
        for (int64_t h = 0; h < ny; h++)                                            \
            for (int64_t w = 0; w < nx; w++) {                                      \
                int64_t idx = h * nx + w;                                           \
                if (llabs(w - hnx) >= wlim || llabs(h - hny) >= hlim) { dst[idx] = 0; continue; }\
                double tmp = M_PI / (lambda * z);                                   \
                double y1 = (double)(h - hny) * dy;                                 \
                double phiu_y = (s * s - s) * y1 * y1 - 2 * s * oy * y1;            \
                double x1 = (double)(w - hnx) * dx;                                 \
                double phiu = (s * s - s) * x1 * x1 - 2 * s * ox * x1;              \
                phiu += phiu_y;                                                     \
                phiu *= tmp;                                                        \
                dst[idx] = src[idx] * ((RT)cos(phiu) + I * (RT)sin(phiu));          \
            }                                                                       \
    }                                                                               \
    CT* buf1 = pad_##SUF(dst, ny, nx, ny2, nx2);                                    \
    fft2d_##SUF(buf1, ny2, nx2, -1);                                                \
    RT sc = (RT)(1.0f / (float)(ny2 * nx2));                                        \
    for (int64_t i = 0; i < ny2 * nx2; i++) buf1[i] *= sc;                          \
    /* stage h: kernel chirp inside a square window of the x-limit */               \
    CT* buf2 = (CT*)malloc(sizeof(CT) * (size_t)(ny2 * nx2));                       \
    {                                                                               \
        double xlim = fabs(lambda * z * 0.5 / (dx * s));                            \
        int64_t lim = win_lim(xlim / dx);                                           \
        int64_t hny = ny2 / 2, hnx = nx2 / 2;                                       \
        for (int64_t h = 0; h < ny2; h++)                                           \
            for (int64_t w = 0; w < nx2; w++) {                                     \
                int64_t idx = h * nx2 + w;                                          \
                if (llabs(w - hnx) >= lim || llabs(h - hny) >= lim) { buf2[idx] = 0; continue; }\
                double tmp = M_PI / (lambda * z);                                   \
                double y1 = (double)(h - hny) * dy;                                 \
                double x1 = (double)(w - hnx) * dx;                                 \
                double phih_y = s * y1 * y1;                                        \
                double phih = s * x1 * x1 + phih_y;                                 \
                phih *= tmp;                                                        \
                buf2[idx] = (RT)cos(phih) + I * (RT)sin(phih);                      \
            }                                                                       \
    }                                                                               \
    fftshift_##SUF(buf2, ny2, nx2);                                                 \
    fft2d_##SUF(buf2, ny2, nx2, -1);                                                \
    for (int64_t i = 0; i < ny2 * nx2; i++) { buf2[i] *= sc; buf1[i] = buf1[i] * buf2[i]; }\
    free(buf2);                                                                     \
    fft2d_##SUF(buf1, ny2, nx2, +1);                                                \
    crop_##SUF(buf1, dst, ny2, nx2);                                                \
    free(buf1);                                                                     \
    /* stage Cz: post-chirp only inside its window, outside left unmodified */      \
    {                                                                               \
        double xlim = lambda * fabs(z) * 0.5 / (fabs(1 - s) * dx);                  \
        int64_t wlim = win_lim(xlim / dx);                                          \
        double ylim = lambda * fabs(z) * 0.5 / (fabs(1 - s) * dy);                  \
        int64_t hlim = win_lim(ylim / dy);                                          \
        int64_t hny = ny / 2, hnx = nx / 2;                                         \
        for (int64_t h = 0; h < ny; h++)                                            \
            for (int64_t w = 0; w < nx; w++) {                                      \
                if (!(llabs(w - hnx) < wlim && llabs(h - hny) < hlim)) continue;    \
                int64_t idx = h * nx + w;                                           \
                double tmp = M_PI / (lambda * z);                                   \
                double y2 = (double)(h - hny) * dy;                                 \
                double phic_y = ((1 - s) * y2 * y2 + 2 * oy * y2 + oy * oy);        \
                double x2 = (double)(w - hnx) * dx;                                 \
                double phic = ((1 - s) * x2 * x2 + 2 * ox * x2 + ox * ox);          \
                phic += phic_y;                                                     \
                phic *= tmp;                                                        \
                dst[idx] = dst[idx] * ((RT)cos(phic) + I * (RT)sin(phic));          \
            }                                                                       \
    }                                                                               \
    return 0;                                                                       \
}

DEFINE_ARSS(c64, cf, float)
DEFINE_ARSS(c128, cd, double)

/* ------------------------------------------------------------------------- */
/* shiftedFresnelProp, include/olprop.h:338-408 (CPU; no windows)            */
/* ------------------------------------------------------------------------- */
#define DEFINE_SHIFTED(SUF, CT, RT)                                                 \
int olref_shifted_fresnel_##SUF(RT* uin, int64_t height, int64_t width, double dy,  \
                                double dx, double lambda, double z, double s,       \
                                double oy, double ox) {                             \
    if (!uin || height <= 0 || width <= 0 || !(lambda > 0) || z == 0) return -1;     \
    CT* u = (CT*)uin;                                                               \
    double tmp = M_PI / (lambda * z);                                               \
    int64_t hh = height / 2, hw = width / 2;                                        \
    for (int64_t h = 0; h < height; h++) {                                          \
        double y1 = (double)(h - hh) * dy;                                          \
        double phiu_y = (s * s - s) * y1 * y1 - 2 * s * oy * y1;                    \
        for (int64_t w = 0; w < width; w++) {                                       \
            int64_t idx = h * width + w;                                            \
            double x1 = (double)(w - hw) * dx;                                      \
            double phiu = (s * s - s) * x1 * x1 - 2 * s * ox * x1;                  \
            phiu += phiu_y;                                                         \
            phiu *= tmp;                                                            \
            u[idx] = u[idx] * ((RT)cos(phiu) + I * (RT)sin(phiu));                  \
        }                                                                           \
    }                                                                               \
    int64_t H = height * 2, Wd = width * 2;                                         \
    CT* buf = pad_##SUF(u, height, width, H, Wd);                                   \
    hh *= 2; hw *= 2;                                                               \
    fft2d_##SUF(buf, H, Wd, -1);                                                    \
    RT sc = (RT)(1.0 / (double)(H * Wd));                                           \
    for (int64_t i = 0; i < H * Wd; i++) buf[i] *= sc;                              \
    CT* hk = (CT*)malloc(sizeof(CT) * (size_t)(H * Wd));                            \
    for (int64_t n = 0; n < H; n++) {                                               \
        double y1 = (double)(n - hh) * dy;                                          \
        double phih_y = s * y1 * y1;                                                \
        for (int64_t m = 0; m < Wd; m++) {                                          \
            double x1 = (double)(m - hw) * dx;                                      \
            double phih = s * x1 * x1 + phih_y;                                     \
            phih *= tmp;                                                            \
            hk[n * Wd + m] = (RT)cos(phih) + I * (RT)sin(phih);                     \
        }                                                                           \
    }                                                                               \
    fftshift_##SUF(hk, H, Wd);                                                      \
    fft2d_##SUF(hk, H, Wd, -1);                                                     \
    for (int64_t i = 0; i < H * Wd; i++) { buf[i] *= sc; buf[i] = buf[i] * hk[i]; }  \
    free(hk);                                                                       \
    fft2d_##SUF(buf, H, Wd, +1);                                                    \
    crop_##SUF(buf, u, H, Wd);                                                      \
    free(buf);                                                                      \
    for (int64_t i = 0; i < height * width; i++) u[i] *= sc;                        \
    hh /= 2; hw /= 2;                                                               \
    for (int64_t h = 0; h < height; h++) {                                          \
        double y2 = (double)(h - hh) * dy;                                          \
        double phic_y = ((1 - s) * y2 * y2 + 2 * oy * y2 + oy * oy);                \
        for (int64_t w = 0; w < width; w++) {                                       \
            int64_t idx = h * width + w;                                            \
            double x2 = (double)(w - hw) * dx;                                      \
            double phic = ((1 - s) * x2 * x2 + 2 * ox * x2 + ox * ox);              \
            phic += phic_y;                                                         \
            phic *= tmp;                                                            \
            u[idx] = u[idx] * ((RT)cos(phic) + I * (RT)sin(phic));                  \
        }                                                                           \
    }                                                                               \
    return 0;                                                                       \
}

DEFINE_SHIFTED(c64, cf, float)
DEFINE_SHIFTED(c128, cd, double)

/* ------------------------------------------------------------------------- */
/* Dispatcher and checks: olprop.h:109-123, :209-225, :656-682                */
/* ------------------------------------------------------------------------- */
/* ------------------------------------------------------------------------- */
/* Tilted-plane propagation: TiltedSourceProp olprop.h:414-526 and            */
/* TiltedReferenceProp :528-594, with bilinear olutils.h:213-234 and bicubic  */
/* :290-320 sampling the centred spectrum.  complex<double> data (the         */
/* reference compiles them with PREC_T = COMPLEX_T); evanescent samples are 0 */
/* (the reference: NaN) and a NaN sample position reads 0 (the reference: UB). */
/* ------------------------------------------------------------------------- */
static cd tilt_bilinear(const cd* src, double dsty, double dstx, int64_t ny, int64_t nx) {
    if (!(dstx == dstx) || !(dsty == dsty)) return 0;
    int64_t ms = (int64_t)floor(dstx), ns = (int64_t)floor(dsty);
    if (0 <= ms && ms < (nx - 1) && 0 <= ns && ns < (ny - 1)) {
        double s = dstx - ms, t = dsty - ns;
        cd U1 = (1 - s) * src[ms + ns * nx] + s * src[(ms + 1) + ns * nx];
        cd U2 = (1 - s) * src[ms + (ns + 1) * nx] + s * src[(ms + 1) + (ns + 1) * nx];
        return (1 - t) * U1 + t * U2;
    } else if (0 <= ms && ms < nx && 0 <= ns && ns < ny) {
        return src[ms + ns * nx];
    }
    return 0;
}

static cd tilt_bicubic(const cd* src, double dsty, double dstx, int64_t ny, int64_t nx) {
    const double a = -1;
    if (!(dstx == dstx) || !(dsty == dsty)) return 0;
    int64_t ms = (int64_t)floor(dstx), ns = (int64_t)floor(dsty);
    if (!(0 <= ms && ms < nx && 0 <= ns && ns < ny)) return 0;
    cd avg = 0;
    for (int64_t j = 0; j < 4; j++)
        for (int64_t i = 0; i < 4; i++) {
            int64_t m = ms - 1 + i, n = ns - 1 + j;
            if (0 <= m && m < nx && 0 <= n && n < ny) {
                double d = sqrt((m - dstx) * (m - dstx) + (n - dsty) * (n - dsty));
                double k;
                if (0 <= d && d < 1) k = 1 - (a + 3) * d * d + (a + 2) * d * d * d;
                else if (1 <= d && d < 2) k = -4 * a + 8 * a * d - 5 * a * d * d + a * d * d * d;
                else k = 0;
                avg += k * src[m + n * nx];
            }
        }
    return avg;
}

/* centred ASM transfer function with the carrier u shift of the source form (ush = 0
 * for the reference form's AsmTransferF) */
static void tilt_asm_h(cd* H, int64_t ny2, int64_t nx2, double dv, double du, double lambda,
                       double z, double ush) {
    double tmp = 1 / (lambda * lambda);
    for (int64_t n = 0; n < ny2; n++) {
        double v = (n - ny2 / 2) * dv;
        for (int64_t m = 0; m < nx2; m++) {
            double u = (m - nx2 / 2) * du + ush;
            double arg = tmp - u * u - v * v;
            if (arg < 0) { H[m + n * nx2] = 0; continue; }
            double w = sqrt(arg);
            double phase = 2 * M_PI * w * z;
            H[m + n * nx2] = cos(phase) + I * sin(phase);
        }
    }
}

/* centred spectrum of FresnelResponse (olprop.h:8-24) times the source form's tilt
 * ramp exp(-2 pi i x a13 f0) (olprop.h:480-489); ramp = 0 for the reference form */
static void tilt_fresnel_h(cd* H, int64_t ny2, int64_t nx2, double dy, double dx, double lambda,
                           double z, double ramp) {
    double tmp = 1 / (lambda * z);
    int64_t hh = ny2 / 2, hw = nx2 / 2;
    for (int64_t n = 0; n < ny2; n++) {
        double y = ((double)(n - hh)) * dy;
        for (int64_t m = 0; m < nx2; m++) {
            double x = ((double)(m - hw)) * dx;
            double phase = M_PI * (x * x + y * y) * tmp;
            cd h = cos(phase) + I * sin(phase);
            double w = (m - nx2 / 2) * dx;
            double ph2 = -2 * M_PI * w * ramp;
            H[m + n * nx2] = h * (cos(ph2) + I * sin(ph2));
        }
    }
    fftshift_c128(H, ny2, nx2);
    fft2d_c128(H, ny2, nx2, -1);
    double sc = 1.0 / (double)(ny2 * nx2);
    for (int64_t i = 0; i < ny2 * nx2; i++) H[i] *= sc;
    fftshift_c128(H, ny2, nx2);
}

int olref_tilted_c128(double* uin, int64_t ny, int64_t nx, double dy, double dx, double lambda,
                      double z, double angle, int interp, int mode, int kind) {
    if (!uin || ny <= 0 || nx <= 0 || !(lambda > 0) || !(dx > 0) || !(dy > 0)) return -1;
    cd* u = (cd*)uin;
    int64_t ny2 = ny * 2, nx2 = nx * 2;
    double du = 1 / (dx * nx2), dv = 1 / (dy * ny2);
    double fsqr = 1 / (lambda * lambda), f0 = 1 / lambda;
    double sc = 1.0 / (double)(ny2 * nx2);
    cd* U = pad_c128(u, ny, nx, ny2, nx2);
    fft2d_c128(U, ny2, nx2, -1);
    fftshift_c128(U, ny2, nx2);
    for (int64_t i = 0; i < ny2 * nx2; i++) U[i] *= sc;
    /* source form: rotation dst -> src by -angle, carrier term -a13 f0;
     * reference form: rotation by angle, no carrier term */
    double rot = kind == 0 ? -angle : angle;
    double a11 = cos(rot), a12 = 0, a13 = -sin(rot), a21 = 0, a22 = 1, a23 = 0;
    cd* H = (cd*)malloc(sizeof(cd) * (size_t)(ny2 * nx2));
    if (mode == 0) tilt_asm_h(H, ny2, nx2, dv, du, lambda, z, kind == 0 ? a13 * f0 : 0.0);
    else tilt_fresnel_h(H, ny2, nx2, dy, dx, lambda, z, kind == 0 ? a13 * f0 : 0.0);
    if (kind == 1)
        for (int64_t i = 0; i < ny2 * nx2; i++) U[i] = U[i] * H[i];
    cd* Up = (cd*)malloc(sizeof(cd) * (size_t)(ny2 * nx2));
    #pragma omp parallel for schedule(static)
    for (int64_t n = 0; n < ny2; n++)
        for (int64_t m = 0; m < nx2; m++) {
            double fpx = (m - nx2 / 2) * du;
            double fpy = (n - ny2 / 2) * dv;
            double arg = fsqr - fpx * fpx - fpy * fpy;
            if (arg < 0) { Up[m + n * nx2] = 0; continue; }
            double fpz = sqrt(arg);
            double fsx = kind == 0 ? a11 * fpx + a12 * fpy + a13 * fpz - a13 * f0
                                   : a11 * fpx + a12 * fpy + a13 * fpz;
            double fsy = a21 * fpx + a22 * fpy + a23 * fpz;
            double py = fsy / dv + ny2 / 2, px = fsx / du + nx2 / 2;
            cd v = interp == 0 ? tilt_bilinear(U, py, px, ny2, nx2) : tilt_bicubic(U, py, px, ny2, nx2);
            double J = (a11 * a22 - a12 * a21) + (a12 * a23 - a13 * a22) * fpx / fpz +
                       (a13 * a21 - a11 * a23) * fpy / fpz;
            Up[m + n * nx2] = v * J;
        }
    if (kind == 0)
        for (int64_t i = 0; i < ny2 * nx2; i++) Up[i] = Up[i] * H[i];
    fftshift_c128(Up, ny2, nx2);
    fft2d_c128(Up, ny2, nx2, +1);
    crop_c128(Up, u, ny2, nx2);
    for (int64_t i = 0; i < ny * nx; i++) u[i] *= sc;
    free(U); free(H); free(Up);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Optical elements: Diffuser include/Diffuser.h, Lensarry include/Lensarry.h */
/* ------------------------------------------------------------------------- */
/* MT19937 (Matsumoto & Nishimura 1998), the engine std::mt19937 specifies */
typedef struct { uint32_t mt[624]; int i; } mt19937;
static void mt_seed(mt19937* g, uint32_t s) {
    g->mt[0] = s;
    for (int i = 1; i < 624; i++)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->i = 624;
}
static uint32_t mt_next(mt19937* g) {
    if (g->i >= 624) {
        for (int k = 0; k < 624; k++) {
            uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7fffffffu);
            g->mt[k] = g->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        g->i = 0;
    }
    uint32_t y = g->mt[g->i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}
uint32_t olref_mt19937_nth(uint32_t seed, int64_t n) {
    mt19937 g;
    mt_seed(&g, seed);
    uint32_t v = 0;
    for (int64_t k = 0; k < n; k++) v = mt_next(&g);
    return v;
}
/* two draws -> a double in [0, 1) with 53 random bits (Diffuser.h:205-207) */
static double mt_res53(mt19937* g) {
    int a = (int)(mt_next(g) >> 5);
    int b = (int)(mt_next(g) >> 6);
    return (a * 67108864.0 + b) / 9007199254740992.0;
}

/* Diffuser::random_phase (CPU), include/Diffuser.h:191-215: pixel (h, w) in row-major
 * order draws a value and multiplies u at ((h + oy) mod H, (w + ox) mod W) by
 * exp(i phase), phase = +-range (value - 0.5) rounded to float, cos / sin in float;
 * step 1 = PLAY (+), 0 = RECORD (-).  complex<float> data. */
int olref_diffuser_random_phase(float* uin, int64_t height, int64_t width, int step,
                                uint32_t seed, float range, int oy, int ox) {
    cf* u = (cf*)uin;
    mt19937 g;
    mt_seed(&g, seed);
    for (int h = 0; h < height; h++)
        for (int w = 0; w < width; w++) {
            int _h = (int)((h + oy) % height), _w = (int)((w + ox) % width);
            _h = (_h >= 0) ? _h : _h + (int)height;
            _w = (_w >= 0) ? _w : _w + (int)width;
            double value = mt_res53(&g);
            float phase = step == 1 ? (float)(range * (value - 0.5f))
                                    : (float)(-range * (value - 0.5f));
            /* cos / sin of a float: the float overloads (<math.h> comes first in the
             * reference's includes) */
            cf c = cosf(phase) + I * sinf(phase);
            cf a = u[_h * width + _w];
            float re = crealf(a) * crealf(c) - cimagf(a) * cimagf(c);
            float im = crealf(a) * cimagf(c) + cimagf(a) * crealf(c);
            u[_h * width + _w] = re + I * im;
        }
    return 0;
}

/* the GPU diffuser kernels (include/Diffuser.h:14-129) given the uniform array b:
 * kind 0 KernelMultPhase (b[adr]), 1 KernelMultLinePhase (b[x]), 2 KernelMultLineYPhase
 * (b[y]), 3 KernelbitRandomPhase (pi if b > 0.5), 4 KernelBlockPhase, 5
 * KernelBlockStepPhase, 6 Kernelbitphase (checkerboard of exp(i F32_PI)); float math */
int olref_diffuser_apply(float* uin, int64_t height, int64_t width, const float* b, float scalar,
                         int kind, int blocksize) {
    cf* u = (cf*)uin;
    const float F32PI = 3.14159265f;
    for (int64_t y = 0; y < height; y++)
        for (int64_t x = 0; x < width; x++) {
            float phase;
            if (kind == 0) phase = (b[x + y * width] - 0.5f) * scalar;
            else if (kind == 1) phase = (b[x] - 0.5f) * scalar;
            else if (kind == 2) phase = (b[y] - 0.5f) * scalar;
            else if (kind == 3) phase = (b[x + y * width] > 0.5f) ? F32PI : 0;
            else if (kind == 4 || kind == 5) {
                int64_t bw = x / blocksize, bh = y / blocksize;
                int64_t root = bw + bh * width / blocksize;
                float stp = kind == 5 ? ((float)(x % blocksize) + (float)(y % blocksize)) / (float)blocksize : 0.0f;
                phase = kind == 5 ? (b[root] + stp - 0.5f) * scalar : (b[root] - 0.5f) * scalar;
            } else {
                if ((x % 2) != (y % 2)) continue;
                phase = F32PI;
            }
            cf c = cosf(phase) + I * sinf(phase);
            cf a = u[x + y * width];
            float re = crealf(a) * crealf(c) - cimagf(a) * cimagf(c);
            float im = crealf(a) * cimagf(c) + cimagf(a) * crealf(c);
            u[x + y * width] = re + I * im;
        }
    return 0;
}

/* FresnelResponse on a blockpx^2 lens (include/olprop.h:8-24 with PREC_T = float:
 * float coordinates, the phase M_PI (x^2 + y^2) tmp formed in double and stored to
 * float, cos / sin in float) or its p_limit window (:56-84, p_limit > 0) */
static void lens_block(cf* ub, int bp, float pitch, float lambda, float z, float p_limit) {
    float tmp = 1 / (lambda * z);
    int64_t hh = bp / 2, hw = bp / 2;
    int64_t wlim = INT_MAX, hlim = INT_MAX;
    if (p_limit > 0) {
        float xlim = fabsf((float)(lambda * z * 0.5 / p_limit));
        wlim = (int64_t)floorf(xlim / pitch);
        hlim = wlim;
    }
    for (int64_t n = 0; n < bp; n++) {
        float y = ((float)(n - hh)) * pitch;
        for (int64_t m = 0; m < bp; m++) {
            float x = ((float)(m - hw)) * pitch;
            if (llabs(m - hw) < wlim && llabs(n - hh) < hlim) {
                float phase = (float)(M_PI * (x * x + y * y) * tmp);
                ub[n * bp + m] = cosf(phase) + I * sinf(phase);
            } else ub[n * bp + m] = 0;
        }
    }
}

/* Lensarry::gen_complex (include/Lensarry.h:254-265) / gen_complex_limit (:300-311)
 * with duplicate (:137-160), or LensarryCircle::duplicate (:406-447) when circle;
 * step 1 = PLAY, 0 = RECORD (focal length negated).  Writes the lens field. */
int olref_lens_array(float* out, int64_t height, int64_t width, float pitch, float blockwidth,
                     float f, float lambda, int step, int on, int om, float p_limit, int circle) {
    cf* u = (cf*)out;
    int bp = (int)roundf(blockwidth / pitch);
    float tmpf = step == 0 ? -f : f;
    cf* ub = (cf*)malloc(sizeof(cf) * (size_t)bp * bp);
    lens_block(ub, bp, pitch, lambda, -tmpf, p_limit);
    if (!circle) {
        for (int64_t i = 0; i < height * width; i++) u[i] = 0;
        for (int bn = 0; bn < height; bn += bp)
            for (int bm = 0; bm < width; bm += bp)
                for (int ln = 0; ln < bp; ln++)
                    for (int lm = 0; lm < bp; lm++) {
                        int ms = (lm - om) % bp, ns = (ln - on) % bp;
                        ms = ms >= 0 ? ms : ms + bp;
                        ns = ns >= 0 ? ns : ns + bp;
                        int m = bm + lm, n = bn + ln;
                        if (m >= width || n >= height) break;
                        u[m + n * width] = ub[ms + ns * bp];
                    }
    } else {
        float stepy = blockwidth * sqrtf(3.0f) * 0.5f, stepx = blockwidth, r = blockwidth * 0.5f;
        int stepn = (int)roundf(stepy / pitch), stepm = (int)roundf(stepx / pitch);
        for (int64_t i = 0; i < height * width; i++) u[i] = 1.0f;
        for (int bn = 0; bn < height; bn += stepn)
            for (int bm = 0; bm < width; bm += stepm)
                for (int ln = 0; ln < bp; ln++)
                    for (int lm = 0; lm < bp; lm++) {
                        int ms = (lm - om + (bn / stepn % 2) * bp / 2) % bp;
                        int ns = (ln - on) % bp;
                        ms = ms >= 0 ? ms : ms + bp;
                        ns = ns >= 0 ? ns : ns + bp;
                        int m = bm + lm, n = bn + ln;
                        if (m >= width || n >= height) break;
                        float dxr = ms * pitch - blockwidth * 0.5f, dyr = ns * pitch - blockwidth * 0.5f;
                        if (powf(dxr, 2) + powf(dyr, 2) <= powf(r, 2.0f)) u[m + n * width] = ub[ms + ns * bp];
                    }
    }
    free(ub);
    return 0;
}

/* Lensarry::duplicate (kind 0), duplicate_random (kind 1), duplicate_multrandom
 * (kind 2), include/Lensarry.h:137-252: the bp x bp block tiles the field in the
 * reference's traversal (blocks row-major, then the block's rows and columns, a row
 * cut at the field edge); outside the window |local - bp/2| < wlim, wlim =
 * floor(|lambda f 0.5 / p_limit| / pitch) (float), each pixel draws a 53-bit value
 * from std::mt19937(0) and takes exp(i phase) (kind 1, phase = 2 F32_PI (v - 0.5f))
 * or the block value times it (kind 2, phase = 0.5 F32_PI (v - 0.5f)); cos / sin of
 * the float phase in float. */
int olref_lens_tile(const float* blockin, int bp, float* out, int64_t height, int64_t width,
                    int on, int om, float pitch, float f, float lambda, float p_limit, int kind) {
    if (!blockin || !out || bp <= 0 || kind < 0 || kind > 2) return -1;
    const cf* ub = (const cf*)blockin;
    cf* u = (cf*)out;
    const float F32PI = 3.14159265f;
    int wlim = INT_MAX;
    if (kind) {
        float xlim = (float)fabs(lambda * f * 0.5 / p_limit);
        wlim = (int)floorf(xlim / pitch);
    }
    mt19937 g;
    mt_seed(&g, 0);
    for (int64_t i = 0; i < height * width; i++) u[i] = 0;
    for (int bn = 0; bn < height; bn += bp)
        for (int bm = 0; bm < width; bm += bp)
            for (int ln = 0; ln < bp; ln++)
                for (int lm = 0; lm < bp; lm++) {
                    int ms = (lm - om) % bp, ns = (ln - on) % bp;
                    ms = ms >= 0 ? ms : ms + bp;
                    ns = ns >= 0 ? ns : ns + bp;
                    int m = bm + lm, n = bn + ln;
                    if (m >= width || n >= height) break;
                    cf v = ub[ms + ns * bp];
                    if (kind && (abs(ms - bp / 2) >= wlim || abs(ns - bp / 2) >= wlim)) {
                        double value = mt_res53(&g);
                        float phase = kind == 1 ? (float)(2 * F32PI * (value - 0.5f))
                                                : (float)(0.5 * F32PI * (value - 0.5f));
                        cf e = cosf(phase) + I * sinf(phase);
                        if (kind == 1) v = e;
                        else {
                            float re = crealf(v) * crealf(e) - cimagf(v) * cimagf(e);
                            float im = crealf(v) * cimagf(e) + cimagf(v) * crealf(e);
                            v = re + I * im;
                        }
                    }
                    u[m + n * width] = v;
                }
    return 0;
}

/* the bp x bp lenslet of Lensarry (FresnelResponse at z = -f for PLAY, +f for
 * RECORD, PREC_T = float; p_limit > 0: its band-limited form) */
int olref_lens_block(float* out, int bp, float pitch, float f, float lambda, int step,
                     float p_limit) {
    if (!out || bp <= 0) return -1;
    float tmpf = step == 0 ? -f : f;
    lens_block((cf*)out, bp, pitch, lambda, -tmpf, p_limit);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Split propagation: splitAsmProp olprop.h:614-651, gsplitProp golprop.h:    */
/* 773-813 (shiftedProp golprop.h:756-770 picks ASM or ARSS with s = 1).      */
/* ------------------------------------------------------------------------- */
int olref_split_prop_c64(float* uin, int64_t height, int64_t width, double dy, double dx,
                         double lambda, double z, int mode, int by, int bx, unsigned flags) {
    if (!uin || by <= 0 || bx <= 0 || height % by || width % bx) return -1;
    cf* u = (cf*)uin;
    int64_t bw = width / bx, bh = height / by;
    cf* udst = (cf*)calloc((size_t)(height * width), sizeof(cf));
    cf* src = (cf*)malloc(sizeof(cf) * (size_t)(bh * bw));
    cf* dst = (cf*)malloc(sizeof(cf) * (size_t)(bh * bw));
    for (int gn = 0; gn < by; gn++)
        for (int gm = 0; gm < bx; gm++) {
            for (int64_t ln = 0; ln < bh; ln++)
                for (int64_t lm = 0; lm < bw; lm++)
                    src[ln * bw + lm] = u[(gn * bh + ln) * width + gm * bw + lm];
            for (int dn = 0; dn < by; dn++)
                for (int dm = 0; dm < bx; dm++) {
                    /* float oy = (global_n - dst_n) * blockheight * dy (PREC_T float) */
                    float oy = (float)((gn - dn) * (int)bh) * (float)dy;
                    float ox = (float)((gm - dm) * (int)bw) * (float)dx;
                    if (mode == 0)
                        olref_shifted_asm_c64((const float*)src, (float*)dst, bh, bw, dy, dx,
                                              lambda, z, oy, ox, flags);
                    else
                        olref_arss_c64((const float*)src, (float*)dst, bh, bw, dy, dx, lambda,
                                       z, 1.0, oy, ox, 0);
                    for (int64_t ln = 0; ln < bh; ln++)
                        for (int64_t lm = 0; lm < bw; lm++)
                            udst[(dn * bh + ln) * width + dm * bw + lm] += dst[ln * bw + lm];
                }
        }
    memcpy(u, udst, sizeof(cf) * (size_t)(height * width));
    free(udst); free(src); free(dst);
    return 0;
}

int olref_asm_prop_check(int64_t ny, int64_t nx, double dy, double dx, double lambda,
                         double z) {
    int64_t nx2 = 2 * nx, ny2 = 2 * ny;
    double du = 1 / (dx * nx2), dv = 1 / (dy * ny2);
    double umax = du * nx2 / 2, vmax = dv * ny2 / 2;
    double tmpsqrt = sqrt(1 / (lambda * lambda) - umax * umax - vmax * vmax);
    double zmax_u = tmpsqrt / (2 * umax * du), zmax_v = tmpsqrt / (2 * vmax * dv);
    if (zmax_u < fabs(z) || zmax_v < fabs(z)) return 0;
    return 1;
}

int olref_fresnel_prop_check(int64_t ny, int64_t nx, double dy, double dx, double lambda,
                             double z) {
    int64_t nx2 = 2 * nx, ny2 = 2 * ny;
    double xmax = dx * nx2 / 2, ymax = dy * ny2 / 2;
    double zmin_x = 2 * xmax * dx / lambda, zmin_y = 2 * ymax * dy / lambda;
    if (zmin_x > fabs(z) || zmin_y > fabs(z)) return 0;
    return 1;
}

int olref_prop_c64(float* u, int64_t ny, int64_t nx, double dy, double dx, double lambda,
                   double z, int mode, int* chosen) {
    int c;
    if (mode == 1) c = 1;
    else if (mode == 0) c = 0;
    else if (mode == 2) {
        if (olref_asm_prop_check(ny, nx, dy, dx, lambda, z)) c = 0;
        else if (!olref_fresnel_prop_check(ny, nx, dy, dx, lambda, z)) c = 2;
        else c = 1;
    } else return -1;
    if (chosen) *chosen = c;
    if (c == 0) return olref_asm_c64(u, ny, nx, dy, dx, lambda, z, 0);
    if (c == 1) return olref_fresnel_c64(u, ny, nx, dy, dx, lambda, z, 0);
    return olref_fresnel_c64(u, ny, nx, dy, dx, lambda, z, OLREF_FRESNEL_BL_CPU);
}

/* ------------------------------------------------------------------------- */
/* CGH: _CghFresnel olcgh.h:296-313 (complex out), :275-295 (phase out).     */
/* Extension (flag): exact spherical r; amp/phase per point.                  */
/* ------------------------------------------------------------------------- */
static inline double cgh_theta(double dxv, double dyv, double dz, double lambda,
                               unsigned flags) {
    if (flags & OLREF_CGH_EXACT_R)
        return 2 * M_PI * sqrt(dxv * dxv + dyv * dyv + dz * dz) / lambda;
    return 2 * M_PI * (dxv * dxv + dyv * dyv) / (2 * dz * lambda);
}

int olref_cgh_c64(const double* x, const double* y, const double* z, const double* amp,
                  const double* phase, int64_t n, float* Uin, int64_t ny, int64_t nx,
                  double py, double px, double lambda, unsigned flags) {
    if (!x || !y || !z || !Uin || n < 0 || ny <= 0 || nx <= 0) return -1;
    cf* U = (cf*)Uin;
    for (int64_t h = 0; h < ny; h++) {
        #pragma omp parallel for schedule(static)
        for (int64_t w = 0; w < nx; w++) {
            cf acc = 0;
            for (int64_t j = 0; j < n; j++) {
                double dxv, dyv, theta;
                if (flags & OLREF_PREC_F32) {
                    float fdx = (float)px * (float)w - (float)x[j];
                    float fdy = (float)py * (float)h - (float)y[j];
                    float fdz = (float)z[j];
                    float th = (float)(2 * M_PI * (fdx * fdx + fdy * fdy) / (2 * fdz * (float)lambda));
                    acc += (cf)(cosf(th) + I * sinf(th));
                    continue;
                }
                dxv = px * w - x[j];
                dyv = py * h - y[j];
                theta = cgh_theta(dxv, dyv, z[j], lambda, flags);
                if (phase) theta += phase[j];
                double a = amp ? amp[j] : 1.0;
                cd t = a * cos(theta) + I * (a * sin(theta));
                acc += (cf)t; /* complex<float> += complex<double> */
            }
            U[h * nx + w] = acc;
        }
    }
    return 0;
}

int olref_cgh_c128(const double* x, const double* y, const double* z, const double* amp,
                   const double* phase, int64_t n, double* Uin, int64_t ny, int64_t nx,
                   double py, double px, double lambda, unsigned flags) {
    if (!x || !y || !z || !Uin || n < 0 || ny <= 0 || nx <= 0) return -1;
    cd* U = (cd*)Uin;
    #pragma omp parallel for schedule(static) collapse(2)
    for (int64_t h = 0; h < ny; h++)
        for (int64_t w = 0; w < nx; w++) {
            cd acc = 0;
            for (int64_t j = 0; j < n; j++) {
                double theta = cgh_theta(px * w - x[j], py * h - y[j], z[j], lambda, flags);
                if (phase) theta += phase[j];
                double a = amp ? amp[j] : 1.0;
                acc += a * cos(theta) + I * (a * sin(theta));
            }
            U[h * nx + w] = acc;
        }
    return 0;
}

int olref_cgh_pixels_c128(const double* x, const double* y, const double* z,
                          const double* amp, const double* phase, int64_t n,
                          const int64_t* ph, const int64_t* pw, int64_t npix, double* out,
                          double py, double px, double lambda, unsigned flags) {
    if (!x || !y || !z || !ph || !pw || !out || n < 0 || npix < 0) return -1;
    cd* o = (cd*)out;
    #pragma omp parallel for schedule(dynamic, 4)
    for (int64_t i = 0; i < npix; i++) {
        cd acc = 0;
        for (int64_t j = 0; j < n; j++) {
            double theta = cgh_theta(px * pw[i] - x[j], py * ph[i] - y[j], z[j], lambda, flags);
            if (phase) theta += phase[j];
            double a = amp ? amp[j] : 1.0;
            acc += a * cos(theta) + I * (a * sin(theta));
        }
        o[i] = acc;
    }
    return 0;
}

/* _CghFresnelDiff, include/olcgh.h:315-349, complex sum in FP64: coordinates in pixel
 * units; per point and row, theta (cycles) and delta start at w = 0 from
 * theta = p z / lambda + gamma / 2 (dx0^2 + dy0^2), delta = gamma / 2 (2 dx0 + 1),
 * gamma = p / (lambda z), and are carried along the row (theta += delta, delta +=
 * gamma), the reference's evident intent (it declares them inside the column loop).
 * amp / phase (may be NULL) weight each point as in the direct sum. */
int olref_cgh_fresnel_diff(const double* x, const double* y, const double* z, const double* amp,
                           const double* phase, int64_t n, int64_t ny, int64_t nx, double p,
                           double lambda, double* out) {
    if (!x || !y || !z || !out || n < 0 || ny <= 0 || nx <= 0) return -1;
    cd* U = (cd*)out;
    #pragma omp parallel for schedule(static)
    for (int64_t h = 0; h < ny; h++) {
        for (int64_t w = 0; w < nx; w++) U[h * nx + w] = 0;
        for (int64_t j = 0; j < n; j++) {
            double gam = p / (lambda * z[j]);
            double theta = 0, delta = 0;
            double a = amp ? amp[j] : 1.0, ph = phase ? phase[j] : 0.0;
            for (int64_t w = 0; w < nx; w++) {
                if (w == 0) {
                    double dx0 = w - x[j], dy0 = h - y[j];
                    delta = gam / 2 * (2 * dx0 + 1);
                    theta = p * z[j] / lambda + gam / 2 * (dx0 * dx0 + dy0 * dy0);
                } else {
                    theta = theta + delta;
                    delta = delta + gam;
                }
                double t = 2 * M_PI * theta + ph;
                U[h * nx + w] += a * cos(t) + I * (a * sin(t));
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* WRP step 1: WRPStep1Pre / WRPStep1 golcgh.h:291-367 (circle, float        */
/* accumulation into U) and WRPStep1Pre_D / WRPStep1_D :391-468 (square,     */
/* PREC_T accumulation).                                                     */
/* ------------------------------------------------------------------------- */
/* per-point window of WRPStep1Pre (golcgh.h:291-299): W_j = |z_j - z_wrp| /          */
/* sqrt(4 (p / lambda)^2 - 1), times sqrt(2) for the square window (WRPStep1Pre_D)    */
static double* wrp_windows(const double* z, int64_t n, double lambda, double plane_z,
                           double p_limit, unsigned flags) {
    double* W = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    double rt2 = (double)sqrtf(2.0f);
    for (int64_t j = 0; j < n; j++) {
        double Wj = 1 / (sqrt(4 * pow((p_limit / lambda), 2) - 1)) * fabs(z[j] - plane_z);
        W[j] = (flags & OLREF_WRP_SQUARE) ? Wj * rt2 : Wj;
    }
    return W;
}

/* the WRP sum at the listed pixels only, FP64 accumulation (sampled check at
 * configuration size); the window tests and terms are those of olref_wrp_* */
int olref_wrp_pixels_c128(const double* x, const double* y, const double* z, int64_t n,
                          const int64_t* ph, const int64_t* pw, int64_t npix, double* out,
                          double p, double lambda, double plane_z, double p_limit,
                          unsigned flags) {
    if (!x || !y || !z || !ph || !pw || !out || n < 0 || npix < 0) return -1;
    cd* o = (cd*)out;
    double* W = wrp_windows(z, n, lambda, plane_z, p_limit, flags);
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < npix; i++) {
        cd acc = 0;
        const int64_t h = ph[i], w = pw[i];
        for (int64_t j = 0; j < n; j++) {
            double dxv = p * w - x[j], dyv = p * h - y[j], dz = z[j] - plane_z;
            double tmp = dxv * dxv + dyv * dyv;
            int in = (flags & OLREF_WRP_SQUARE) ? (fabs(dxv) * 2 <= W[j] && fabs(dyv) * 2 <= W[j])
                                                : (tmp < W[j] * W[j]);
            if (in) {
                double th = 2 * M_PI * tmp / (2 * dz * lambda);
                acc += (cos(th) + I * sin(th)) / fabs(dz);
            }
        }
        o[i] = acc;
    }
    free(W);
    return 0;
}

#define DEFINE_WRP(SUF, CT, RT)                                                     \
int olref_wrp_##SUF(const double* x, const double* y, const double* z, int64_t n,   \
                    RT* Uin, int64_t ny, int64_t nx, double p, double lambda,       \
                    double plane_z, double p_limit, unsigned flags) {               \
    if (!x || !y || !z || !Uin || n < 0 || ny <= 0 || nx <= 0) return -1;           \
    CT* U = (CT*)Uin;                                                               \
    double* W = wrp_windows(z, n, lambda, plane_z, p_limit, flags);                 \
    _Pragma("omp parallel for schedule(static) collapse(2)")                        \
    for (int64_t h = 0; h < ny; h++)                                                \
        for (int64_t w = 0; w < nx; w++) {                                          \
            CT acc = 0;                                                             \
            cd accd = 0;                                                            \
            for (int64_t j = 0; j < n; j++) {                                       \
                double dxv = p * w - x[j], dyv = p * h - y[j], dz = z[j] - plane_z; \
                double tmp = dxv * dxv + dyv * dyv;                                 \
                if (flags & OLREF_WRP_SQUARE) {                                     \
                    if (fabs(dxv) * 2 <= W[j] && fabs(dyv) * 2 <= W[j]) {           \
                        double th = 2 * M_PI * tmp / (2 * dz * lambda);             \
                        accd += (cos(th) + I * sin(th)) / fabs(dz);                 \
                    }                                                               \
                } else if (tmp < W[j] * W[j]) {                                     \
                    double th = 2 * M_PI * tmp / (2 * dz * lambda);                 \
                    acc += (CT)((cos(th) + I * sin(th)) / fabs(dz));                \
                }                                                                   \
            }                                                                       \
            U[h * nx + w] = (flags & OLREF_WRP_SQUARE) ? (CT)accd : acc;            \
        }                                                                           \
    free(W);                                                                        \
    return 0;                                                                       \
}

DEFINE_WRP(c64, cf, float)
DEFINE_WRP(c128, cd, double)

double olref_wrp_plane_z(int64_t ny, int64_t nx, double p, double lambda, double zmin,
                         double p_limit, float s, int square) {
    float plane_z;
    if (s < 0) {
        if (square) { /* WRPMethod_D, golcgh.h:546-550 */
            float D = (float)(nx * p * 0.05);
            float d = (float)(sqrt(4 * pow((p_limit / lambda), 2) - 1) / sqrt(2) * D);
            plane_z = (float)(zmin - d);
        } else { /* WRPMethod, golcgh.h:567-571 */
            float Wpow2 = (float)(nx * ny * p * p / M_PI * 1 / 1024.f);
            float d = (float)sqrt((4 * pow((p / lambda), 2) - 1) * Wpow2);
            plane_z = (float)(zmin - d);
        }
    } else {
        plane_z = (float)(zmin * s);
    }
    return (double)plane_z;
}
