// olx_tilt.hip -- tilted-plane propagation: TiltedSourceProp (include/olprop.h:414-526)
// and TiltedReferenceProp (:528-594).
//
// Both rotate the angular spectrum about the y axis: the spectrum of the padded
// field is resampled at the rotated spatial frequencies (bilinear or the
// reference's radial "bicubic" kernel, include/olutils.h:213-234, 290-320) with the
// Jacobian of the rotation, combined with an ASM or Fresnel transfer function, and
// transformed back.  On the device:
//   pad (centred, 2x) -> 2-D FFT -> [reference form: x H / M^2] -> resampling
//   gather (centred coordinates read through the fftshift index map, so no shift
//   pass; source form: x H / M^2 here) -> 2-D IFFT -> crop x 1/M^2.
// Sample positions are FP64 with one rounding per operation (this file is built
// without FMA contraction) so that floor() picks the same cells as the reference's
// double arithmetic; data are complex64, bilinear weights and sums FP64, bicubic
// weights and sums FP32.
#include <climits>
#include <cmath>

#include "olx_common.hpp"

extern "C" int olx_fft2d_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, int sign);

namespace olx {

struct TiltArgs {
    int ny2, nx2;
    double du, dv, dx, dy, fsqr, f0, z, sc;
    double a11, a12, a13, a21, a22, a23;
    int kind;  // 0 source, 1 reference
    int mode;  // 0 ASM, 1 Fresnel (tables)
    const float2* sx;  // Fresnel: natural-order 1-D spectra (x with the source form's ramp)
    const float2* sy;
};

// centred index c of a length-M axis stored in natural (FFT) order: fftshift
__device__ __forceinline__ int nat(int c, int M) { return (c + M / 2) & (M - 1); }

__device__ __forceinline__ double2 ld(const float2* P, int n, int m, const TiltArgs& a) {
    const float2 v = P[(size_t)nat(n, a.ny2) * a.nx2 + nat(m, a.nx2)];
    return make_double2((double)v.x, (double)v.y);
}

// include/olutils.h:213-234 on the centred spectrum
__device__ double2 bilinear(const float2* P, double dsty, double dstx, const TiltArgs& a) {
    if (!(dstx == dstx) || !(dsty == dsty)) return make_double2(0.0, 0.0);
    const long long ms = (long long)floor(dstx), ns = (long long)floor(dsty);
    const int nx = a.nx2, ny = a.ny2;
    if (0 <= ms && ms < (nx - 1) && 0 <= ns && ns < (ny - 1)) {
        const double s = dstx - (double)ms, t = dsty - (double)ns;
        const double2 p00 = ld(P, (int)ns, (int)ms, a), p01 = ld(P, (int)ns, (int)ms + 1, a);
        const double2 p10 = ld(P, (int)ns + 1, (int)ms, a), p11 = ld(P, (int)ns + 1, (int)ms + 1, a);
        const double2 U1 = make_double2((1 - s) * p00.x + s * p01.x, (1 - s) * p00.y + s * p01.y);
        const double2 U2 = make_double2((1 - s) * p10.x + s * p11.x, (1 - s) * p10.y + s * p11.y);
        return make_double2((1 - t) * U1.x + t * U2.x, (1 - t) * U1.y + t * U2.y);
    }
    if (0 <= ms && ms < nx && 0 <= ns && ns < ny) return ld(P, (int)ns, (int)ms, a);
    return make_double2(0.0, 0.0);
}

// include/olutils.h:290-320 (a = -1): a 4 x 4 window weighted by the cubic kernel of
// the radial distance, k(d) = (1 - d)(1 + d - d^2) for d < 1, (2 - d)^2 (1 - d) for
// 1 <= d < 2, 0 beyond (taps at d >= 2 are skipped).  The cell (ms, ns) comes from
// floor() of the FP64 position, as in the reference; the tap offsets are formed from
// the exact FP64 fraction dst - floor(dst) rounded once to FP32, and the weights and
// the sum are FP32 (relative error ~1e-7).
__device__ double2 bicubic(const float2* P, double dsty, double dstx, const TiltArgs& a) {
    if (!(dstx == dstx) || !(dsty == dsty)) return make_double2(0.0, 0.0);
    const long long ms = (long long)floor(dstx), ns = (long long)floor(dsty);
    const int nx = a.nx2, ny = a.ny2;
    if (!(0 <= ms && ms < nx && 0 <= ns && ns < ny)) return make_double2(0.0, 0.0);
    const float fx = (float)(dstx - (double)ms), fy = (float)(dsty - (double)ns);
    float sx = 0.f, sy = 0.f;
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const long long m = ms - 1 + i, n = ns - 1 + j;
            if (0 <= m && m < nx && 0 <= n && n < ny) {
                const float dx = (float)(i - 1) - fx, dy = (float)(j - 1) - fy;
                const float d2 = __builtin_fmaf(dx, dx, dy * dy);
                if (d2 >= 4.f) continue;
                const float d = __builtin_amdgcn_sqrtf(d2);
                const float e = 1.f - d;
                const float k = d < 1.f ? e * __builtin_fmaf(-d, d, 1.f + d)
                                        : (2.f - d) * (2.f - d) * e;
                const float2 p = P[(size_t)nat((int)n, ny) * nx + nat((int)m, nx)];
                sx = __builtin_fmaf(k, p.x, sx);
                sy = __builtin_fmaf(k, p.y, sy);
            }
        }
    return make_double2((double)sx, (double)sy);
}

// centred ASM transfer function exp(2 pi i w z), w = sqrt(1/lambda^2 - u^2 - v^2), the
// u axis shifted by ush (source form: a13 f0, include/olprop.h:461-470); 0 where
// evanescent
__device__ double2 tilt_asm_h(const TiltArgs& a, int n, int m, double ush) {
    const double v = (double)(n - a.ny2 / 2) * a.dv;
    const double u = (double)(m - a.nx2 / 2) * a.du + ush;
    const double arg = a.fsqr - u * u - v * v;
    if (arg < 0) return make_double2(0.0, 0.0);
    double cyc = sqrt(arg) * a.z;
    cyc -= rint(cyc);
    double s, c;
    sincospi(2.0 * cyc, &s, &c);
    return make_double2(c, s);
}

// H at centred (n, m) for the Fresnel form: the product of the 1-D spectra
__device__ __forceinline__ double2 tilt_fresnel_h(const TiltArgs& a, int n, int m) {
    const float2 x = a.sx[nat(m, a.nx2)], y = a.sy[nat(n, a.ny2)];
    return make_double2((double)x.x * y.x - (double)x.y * y.y, (double)x.x * y.y + (double)x.y * y.x);
}

__device__ __forceinline__ double2 cmul2(double2 p, double2 q) {
    return make_double2(p.x * q.x - p.y * q.y, p.x * q.y + p.y * q.x);
}

__global__ void k_tilt_pad(const float2* __restrict__ u, float2* __restrict__ P, int ny, int nx) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y;
    const int nx2 = 2 * nx;
    if (m >= nx2) return;
    const int r = n - ny / 2, c = m - nx / 2;
    float2 v = make_float2(0.f, 0.f);
    if (r >= 0 && r < ny && c >= 0 && c < nx) v = u[(size_t)r * nx + c];
    P[(size_t)n * nx2 + m] = v;
}

// reference form: the source spectrum times H (centred) / M^2, in natural order
__global__ void k_tilt_hmul(float2* __restrict__ P, TiltArgs a) {
    const int mn = blockIdx.x * blockDim.x + threadIdx.x, nn = blockIdx.y;
    if (mn >= a.nx2) return;
    const int m = nat(mn, a.nx2), n = nat(nn, a.ny2);  // centred indices (fftshift is an involution)
    const double2 h = a.mode == 0 ? tilt_asm_h(a, n, m, 0.0) : tilt_fresnel_h(a, n, m);
    float2* p = &P[(size_t)nn * a.nx2 + mn];
    const double2 v = cmul2(make_double2((double)p->x * a.sc, (double)p->y * a.sc), h);
    *p = make_float2((float)v.x, (float)v.y);
}

// the resampled spectrum at centred output (n, m), written to natural order
template <int INTERP>
__global__ void k_tilt_gather(const float2* __restrict__ P, float2* __restrict__ Q, TiltArgs a) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y;
    if (m >= a.nx2) return;
    const double fpx = (double)(m - a.nx2 / 2) * a.du;
    const double fpy = (double)(n - a.ny2 / 2) * a.dv;
    const double arg = a.fsqr - fpx * fpx - fpy * fpy;
    double2 r = make_double2(0.0, 0.0);
    if (arg >= 0) {
        const double fpz = sqrt(arg);
        // include/olprop.h:447-449 (source, with the carrier term) and :569-571
        const double fsx = a.kind == 0 ? a.a11 * fpx + a.a12 * fpy + a.a13 * fpz - a.a13 * a.f0
                                       : a.a11 * fpx + a.a12 * fpy + a.a13 * fpz;
        const double fsy = a.a21 * fpx + a.a22 * fpy + a.a23 * fpz;
        const double py = fsy / a.dv + a.ny2 / 2, px = fsx / a.du + a.nx2 / 2;
        const double2 v = INTERP == 0 ? bilinear(P, py, px, a) : bicubic(P, py, px, a);
        const double J = (a.a11 * a.a22 - a.a12 * a.a21) + (a.a12 * a.a23 - a.a13 * a.a22) * fpx / fpz +
                         (a.a13 * a.a21 - a.a11 * a.a23) * fpy / fpz;
        r = make_double2(v.x * J, v.y * J);
        if (a.kind == 0) {
            const double2 h = a.mode == 0 ? tilt_asm_h(a, n, m, a.a13 * a.f0) : tilt_fresnel_h(a, n, m);
            r = cmul2(make_double2(r.x * a.sc, r.y * a.sc), h);
        }
    }
    Q[(size_t)nat(n, a.ny2) * a.nx2 + nat(m, a.nx2)] = make_float2((float)r.x, (float)r.y);
}

__global__ void k_tilt_crop(const float2* __restrict__ Q, float2* __restrict__ u, int ny, int nx,
                            float sc) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y;
    if (m >= nx) return;
    const float2 v = Q[(size_t)(n + ny / 2) * (2 * nx) + m + nx / 2];
    u[(size_t)n * nx + m] = make_float2(v.x * sc, v.y * sc);
}

}  // namespace olx

using namespace olx;

extern "C" int olx_tilted_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy,
                              double dx, double lambda, double z, double angle, int interp,
                              int mode, int kind) {
    if (!ctx || !d_u) OLX_FAIL(OLX_EINVAL, "tilted: null argument");
    if (!is_pow2(ny) || !is_pow2(nx) || ny < 4 || nx < 4 || ny > 8192 || nx > 8192)
        OLX_FAIL(OLX_EUNSUPPORTED, "tilted: sides must be powers of two in [4, 8192]");
    if (!(lambda > 0) || !(dx > 0) || !(dy > 0) || !std::isfinite(z) || !std::isfinite(angle))
        OLX_FAIL(OLX_EINVAL, "tilted: lambda, dx, dy must be > 0, z and angle finite");
    if ((interp != 0 && interp != 1) || (mode != 0 && mode != 1) || (kind != 0 && kind != 1))
        OLX_FAIL(OLX_EINVAL, "tilted: interp, mode and kind must be 0 or 1");
    if (mode == 1 && !(z != 0)) OLX_FAIL(OLX_EINVAL, "tilted: Fresnel needs z != 0");
    TiltArgs a;
    a.ny2 = (int)(2 * ny);
    a.nx2 = (int)(2 * nx);
    a.du = 1 / (dx * a.nx2);
    a.dv = 1 / (dy * a.ny2);
    a.dx = dx;
    a.dy = dy;
    a.fsqr = 1 / (lambda * lambda);
    a.f0 = 1 / lambda;
    a.z = z;
    a.sc = 1.0 / ((double)a.ny2 * (double)a.nx2);
    // source form: rotation dst -> src by -angle (include/olprop.h:429-436);
    // reference form: by angle (:553-559)
    const double rot = kind == 0 ? -angle : angle;
    a.a11 = std::cos(rot);
    a.a12 = 0;
    a.a13 = -std::sin(rot);
    a.a21 = 0;
    a.a22 = 1;
    a.a23 = 0;
    a.kind = kind;
    a.mode = mode;
    a.sx = a.sy = nullptr;
    int rc = OLX_OK;
    if (mode == 1) {
        // FresnelResponse (include/olprop.h:8-24) times the source form's ramp
        // exp(-2 pi i x a13 f0) (:480-489): K (x^2 + qb x), K = pi / (lambda z)
        const double K = M_PI * (1 / (lambda * z));
        const double qb = kind == 0 ? -2 * a.a13 * a.f0 * lambda * z : 0.0;
        a.sx = spectrum_table(ctx, a.nx2, dx, K, 1.0, INT_MAX, 0, &rc, qb);
        if (!a.sx) return rc;
        a.sy = spectrum_table(ctx, a.ny2, dy, K, 1.0, INT_MAX, 0, &rc);
        if (!a.sy) return rc;
    }
    const size_t M2 = (size_t)a.ny2 * a.nx2;
    OLX_TRY(ctx->work.reserve(sizeof(float2) * M2));
    OLX_TRY(ctx->scratch2.reserve(sizeof(float2) * M2));
    float2* P = (float2*)ctx->work.p;
    float2* Q = (float2*)ctx->scratch2.p;
    hipStream_t s = ctx->stream;
    const dim3 gM((unsigned)((a.nx2 + 255) / 256), (unsigned)a.ny2);
    k_tilt_pad<<<gM, 256, 0, s>>>((const float2*)d_u, P, (int)ny, (int)nx);
    OLX_TRY(check_launch("tilted pad"));
    OLX_TRY(olx_fft2d_c64(ctx, P, a.ny2, a.nx2, -1));
    if (kind == 1) {
        k_tilt_hmul<<<gM, 256, 0, s>>>(P, a);
        OLX_TRY(check_launch("tilted transfer function"));
    }
    {
        TimedLaunch tl(ctx, "tilt_gather");
        if (interp == 0) k_tilt_gather<0><<<gM, 256, 0, s>>>(P, Q, a);
        else k_tilt_gather<1><<<gM, 256, 0, s>>>(P, Q, a);
    }
    OLX_TRY(check_launch("tilted gather"));
    OLX_TRY(olx_fft2d_c64(ctx, Q, a.ny2, a.nx2, +1));
    k_tilt_crop<<<dim3((unsigned)((nx + 255) / 256), (unsigned)ny), 256, 0, s>>>(
        Q, (float2*)d_u, (int)ny, (int)nx, (float)a.sc);
    return check_launch("tilted crop");
}
