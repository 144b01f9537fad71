// olx_optics.hip -- optical elements applied to a field in HBM: the random-phase
// diffuser (include/Diffuser.h) and the lens arrays (include/Lensarry.h), each one
// elementwise pass over the field (read + write 8 bytes per pixel).
//
// Diffuser::random_phase (CPU form) draws its phases from std::mt19937 in pixel
// order; that stream is sequential, so the host generates it (the same engine the
// reference uses, bit for bit) and the device applies it.  The GPU forms draw from
// cuRAND, whose MT19937 ordering is its own; here they take a uniform array, which
// olx_uniform_f32 fills from a Philox4x32-10 counter stream.  Lens fields are
// computed per pixel from the lenslet's local coordinates (nothing is
// materialised), with the reference's float arithmetic: this file is built
// without FMA contraction so the phases round as the reference's do.
#include <climits>
#include <cmath>
#include <random>
#include <vector>

#include "olx_common.hpp"

namespace olx {

// ------------------------------------------------------------------ Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3" (SC'11)
__device__ inline uint4 philox4x32(uint4 c, uint2 k) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += 0x9E3779B9u;
        k.y += 0xBB67AE85u;
    }
    return c;
}

// uniform floats in (0, 1]: 24 random bits, (v + 1) 2^-24
__global__ void k_uniform(float* __restrict__ out, int64_t n, uint64_t seed) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;  // one counter per 4 values
    if (q * 4 >= n) return;
    const uint4 r = philox4x32(make_uint4((uint32_t)q, (uint32_t)(q >> 32), 0u, 0u),
                               make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
    const uint32_t v[4] = {r.x, r.y, r.z, r.w};
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (q * 4 + i < n) out[q * 4 + i] = (float)((v[i] >> 8) + 1u) * (1.0f / 16777216.0f);
}

// a * c in float, one rounding per operation (std::complex<float> multiplication)
__device__ __forceinline__ float2 cmulf(float2 a, float2 c) {
    return make_float2(a.x * c.x - a.y * c.y, a.x * c.y + a.y * c.x);
}

// ------------------------------------------------------------------ diffuser
// Diffuser::random_phase, CPU form: pixel j (row-major generation order) multiplies
// u at its shifted position by exp(i phase[j]), cos / sin in double rounded to float
__global__ void k_diffuser_cpu(float2* __restrict__ u, const float* __restrict__ phase, int ny,
                               int nx, int oy, int ox) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x, h = blockIdx.y;
    if (w >= nx) return;
    int hh = (h + oy) % ny, ww = (w + ox) % nx;
    hh = hh >= 0 ? hh : hh + ny;
    ww = ww >= 0 ? ww : ww + nx;
    const double p = (double)phase[(size_t)h * nx + w];
    double s, c;
    sincos(p, &s, &c);
    float2* q = &u[(size_t)hh * nx + ww];
    *q = cmulf(*q, make_float2((float)c, (float)s));
}

// the GPU diffuser kernels (include/Diffuser.h:14-129) given the uniforms b
__global__ void k_diffuser_apply(float2* __restrict__ u, const float* __restrict__ b, int ny,
                                 int nx, float scalar, int kind, int bs) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= nx) return;
    const float F32PI = 3.14159265f;
    float phase;
    switch (kind) {
        case OLX_DIFFUSER_PHASE: phase = (b[(size_t)y * nx + x] - 0.5f) * scalar; break;
        case OLX_DIFFUSER_XPHASE: phase = (b[x] - 0.5f) * scalar; break;
        case OLX_DIFFUSER_YPHASE: phase = (b[y] - 0.5f) * scalar; break;
        case OLX_DIFFUSER_BITRANDOM: phase = (b[(size_t)y * nx + x] > 0.5f) ? F32PI : 0.f; break;
        case OLX_DIFFUSER_BLOCK:
        case OLX_DIFFUSER_BLOCKSTEP: {
            const int64_t root = x / bs + (int64_t)(y / bs) * nx / bs;
            const float stp = ((float)(x % bs) + (float)(y % bs)) / (float)bs;
            phase = kind == OLX_DIFFUSER_BLOCKSTEP ? (b[root] + stp - 0.5f) * scalar
                                                   : (b[root] - 0.5f) * scalar;
            break;
        }
        default:  // OLX_DIFFUSER_BITPHASE: exp(i F32_PI) on the (x + y) even checkerboard
            if ((x % 2) != (y % 2)) return;
            phase = F32PI;
    }
    float s, c;
    sincosf(phase, &s, &c);
    float2* q = &u[(size_t)y * nx + x];
    *q = cmulf(*q, make_float2(c, s));
}

// ------------------------------------------------------------------ lens arrays
struct LensArgs {
    int ny, nx, bp, on, om;
    float pitch, tmp, blockwidth, r;
    int wlim;         // window of the band-limited lenslet (INT_MAX: none)
    int circle, stepn, stepm;
    int write;        // write the lens field instead of multiplying by it
};

// lenslet value at local (ns, ms): FresnelResponse with PREC_T = float
// (include/olprop.h:8-24; band limit :56-84)
__device__ inline float2 lenslet(const LensArgs& a, int ns, int ms) {
    const int hh = a.bp / 2, hw = a.bp / 2;
    if (!(abs(ms - hw) < a.wlim && abs(ns - hh) < a.wlim)) return make_float2(0.f, 0.f);
    const float y = ((float)(ns - hh)) * a.pitch, x = ((float)(ms - hw)) * a.pitch;
    const float phase = (float)(M_PI * (double)(x * x + y * y) * (double)a.tmp);
    double s, c;
    sincos((double)phase, &s, &c);
    return make_float2((float)c, (float)s);
}

__global__ void k_lens(float2* __restrict__ u, LensArgs a) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y;
    if (m >= a.nx) return;
    float2 v = make_float2(1.f, 0.f);
    if (!a.circle) {
        // duplicate (include/Lensarry.h:137-160): the block tiles the field
        int ms = (m % a.bp - a.om) % a.bp, ns = (n % a.bp - a.on) % a.bp;
        ms = ms >= 0 ? ms : ms + a.bp;
        ns = ns >= 0 ? ns : ns + a.bp;
        v = lenslet(a, ns, ms);
    } else {
        // LensarryCircle::duplicate (:406-447): rows of circular lenslets stepn apart,
        // odd rows shifted by half a lenslet; a later row overwrites an earlier one
        // inside its circle, pixels in no circle stay 1
        for (int k = n / a.stepn; k >= 0 && k * a.stepn + a.bp > n; k--) {
            const int bn = k * a.stepn, bm = (m / a.stepm) * a.stepm;
            const int ln = n - bn, lm = m - bm;
            if (lm >= a.bp) continue;
            int ms = (lm - a.om + (k % 2) * a.bp / 2) % a.bp, ns = (ln - a.on) % a.bp;
            ms = ms >= 0 ? ms : ms + a.bp;
            ns = ns >= 0 ? ns : ns + a.bp;
            const float dxr = ms * a.pitch - a.blockwidth * 0.5f, dyr = ns * a.pitch - a.blockwidth * 0.5f;
            if (powf(dxr, 2.f) + powf(dyr, 2.f) <= powf(a.r, 2.0f)) {
                v = lenslet(a, ns, ms);
                break;
            }
        }
    }
    float2* q = &u[(size_t)n * a.nx + m];
    *q = a.write ? v : cmulf(*q, v);
}

}  // namespace olx

using namespace olx;

extern "C" int olx_uniform_f32(olx_ctx* ctx, float* d_out, int64_t n, uint64_t seed) {
    if (!ctx || (n > 0 && !d_out)) OLX_FAIL(OLX_EINVAL, "uniform: null argument");
    if (n <= 0) return OLX_OK;
    const int64_t q = (n + 3) / 4;
    k_uniform<<<(unsigned)((q + 255) / 256), 256, 0, ctx->stream>>>(d_out, n, seed);
    return check_launch("uniform");
}

extern "C" int olx_diffuser_apply_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx,
                                      const float* d_b, float scalar, int kind, int blocksize) {
    if (!ctx || !d_u || !d_b) OLX_FAIL(OLX_EINVAL, "diffuser: null argument");
    if (ny <= 0 || nx <= 0 || ny > 65535 || nx > (1 << 24)) OLX_FAIL(OLX_EINVAL, "diffuser: bad size");
    if (kind < OLX_DIFFUSER_PHASE || kind > OLX_DIFFUSER_BITPHASE)
        OLX_FAIL(OLX_EINVAL, "diffuser: unknown kind %d", kind);
    if ((kind == OLX_DIFFUSER_BLOCK || kind == OLX_DIFFUSER_BLOCKSTEP) &&
        (blocksize <= 0 || nx % blocksize || ny % blocksize))
        OLX_FAIL(OLX_EINVAL, "diffuser: block size %d must divide the field", blocksize);
    k_diffuser_apply<<<dim3((unsigned)((nx + 255) / 256), (unsigned)ny), 256, 0, ctx->stream>>>(
        (float2*)d_u, d_b, (int)ny, (int)nx, scalar, kind, blocksize);
    return check_launch("diffuser");
}

extern "C" int olx_diffuser_random_phase_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx,
                                             int step, uint32_t seed, float range, int oy, int ox) {
    if (!ctx || !d_u) OLX_FAIL(OLX_EINVAL, "random_phase: null argument");
    if (ny <= 0 || nx <= 0 || ny > 65535 || nx > (1 << 24)) OLX_FAIL(OLX_EINVAL, "random_phase: bad size");
    if (step != OLX_STEP_PLAY && step != OLX_STEP_RECORD)
        OLX_FAIL(OLX_EINVAL, "random_phase: step must be PLAY or RECORD");
    // include/Diffuser.h:191-215: two draws per pixel, 53-bit value, float phase
    const size_t n = (size_t)ny * nx;
    std::vector<float> ph(n);
    std::mt19937 e2(seed);
    for (size_t j = 0; j < n; j++) {
        const int a = (int)(e2() >> 5);
        const int b = (int)(e2() >> 6);
        const double value = (a * 67108864.0 + b) / 9007199254740992.0;
        ph[j] = step == OLX_STEP_PLAY ? (float)(range * (value - 0.5f)) : (float)(-range * (value - 0.5f));
    }
    OLX_TRY(ctx->obj_tmp.reserve(sizeof(float) * n));
    OLX_HIP(hipMemcpyAsync(ctx->obj_tmp.p, ph.data(), sizeof(float) * n, hipMemcpyHostToDevice,
                           ctx->stream));
    k_diffuser_cpu<<<dim3((unsigned)((nx + 255) / 256), (unsigned)ny), 256, 0, ctx->stream>>>(
        (float2*)d_u, (const float*)ctx->obj_tmp.p, (int)ny, (int)nx, oy, ox);
    OLX_TRY(check_launch("random_phase"));
    // the host vector is freed on return: wait for the copy
    OLX_HIP(hipStreamSynchronize(ctx->stream));
    return OLX_OK;
}

// Lensarry::duplicate / duplicate_random / duplicate_multrandom (include/Lensarry.h:
// 137-252): the bp x bp block tiles the field, each pixel taking the block value at
// its shifted local position; with a phase array, pixels outside the +-wlim window
// of the block centre take exp(i phase) (kind 1) or the block value times it (kind
// 2), cos / sin in double rounded to float (std::cos / std::sin of a float)
__global__ void k_lens_tile(const float2* __restrict__ block, int bp, float2* __restrict__ out,
                            int ny, int nx, int on, int om, const float* __restrict__ phase,
                            int kind, int wlim) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y;
    if (m >= nx) return;
    int ms = (m % bp - om) % bp, ns = (n % bp - on) % bp;
    ms = ms >= 0 ? ms : ms + bp;
    ns = ns >= 0 ? ns : ns + bp;
    float2 v = block[(size_t)ns * bp + ms];
    if (kind != 0 && (abs(ms - bp / 2) >= wlim || abs(ns - bp / 2) >= wlim)) {
        double s, c;
        sincos((double)phase[(size_t)n * nx + m], &s, &c);
        const float2 e = make_float2((float)c, (float)s);
        v = kind == 1 ? e : cmulf(v, e);
    }
    out[(size_t)n * nx + m] = v;
}

extern "C" int olx_lens_array_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, float pitch,
                                  float blockwidth, float f, float lambda, int step, int oy, int ox,
                                  float p_limit, uint32_t flags) {
    if (!ctx || !d_u) OLX_FAIL(OLX_EINVAL, "lens array: null argument");
    if (ny <= 0 || nx <= 0 || ny > 65535 || nx > (1 << 24)) OLX_FAIL(OLX_EINVAL, "lens array: bad size");
    if (!(pitch > 0) || !(blockwidth > 0) || !(lambda > 0) || !(f != 0) || !std::isfinite(f))
        OLX_FAIL(OLX_EINVAL, "lens array: pitch, block width, lambda > 0 and f finite, nonzero");
    if (step != OLX_STEP_PLAY && step != OLX_STEP_RECORD)
        OLX_FAIL(OLX_EINVAL, "lens array: step must be PLAY or RECORD");
    LensArgs a;
    a.ny = (int)ny;
    a.nx = (int)nx;
    a.bp = (int)std::round(blockwidth / pitch);  // include/Lensarry.h:104
    if (a.bp <= 0) OLX_FAIL(OLX_EINVAL, "lens array: the lenslet is narrower than a pixel");
    a.on = oy;
    a.om = ox;
    a.pitch = pitch;
    // gen_complex: FresnelResponse at z = -f (PLAY) or +f (RECORD), include/Lensarry.h:254-265
    const float tmpf = step == OLX_STEP_RECORD ? -f : f;
    const float z = -tmpf;
    a.tmp = 1 / (lambda * z);
    a.blockwidth = blockwidth;
    a.r = blockwidth * 0.5f;
    a.wlim = INT_MAX;
    if (p_limit > 0) {  // include/olprop.h:63-68 with PREC_T = float
        const float xlim = (float)std::fabs(lambda * z * 0.5 / p_limit);
        a.wlim = (int)std::floor(xlim / pitch);
    }
    a.circle = (flags & OLX_LENS_CIRCLE) ? 1 : 0;
    a.stepn = (int)std::round(blockwidth * std::sqrt(3.0f) * 0.5f / pitch);  // :409-413
    a.stepm = (int)std::round(blockwidth / pitch);
    if (a.circle && (a.stepn <= 0 || a.stepm <= 0))
        OLX_FAIL(OLX_EINVAL, "lens array: lenslet rows narrower than a pixel");
    a.write = (flags & OLX_LENS_WRITE) ? 1 : 0;
    k_lens<<<dim3((unsigned)((nx + 255) / 256), (unsigned)ny), 256, 0, ctx->stream>>>((float2*)d_u, a);
    return check_launch("lens array");
}

extern "C" int olx_lens_tile_c64(olx_ctx* ctx, const void* d_block, int bp, void* d_out,
                                 int64_t ny, int64_t nx, int on, int om, float pitch, float f,
                                 float lambda, float p_limit, int kind) {
    if (!ctx || !d_block || !d_out) OLX_FAIL(OLX_EINVAL, "lens tile: null argument");
    if (ny <= 0 || nx <= 0 || ny > 65535 || nx > (1 << 24) || bp <= 0)
        OLX_FAIL(OLX_EINVAL, "lens tile: bad size");
    if (kind < 0 || kind > 2) OLX_FAIL(OLX_EINVAL, "lens tile: kind must be 0, 1 or 2");
    int wlim = INT_MAX;
    const float* dph = nullptr;
    if (kind != 0) {
        if (!(p_limit > 0) || !(pitch > 0)) OLX_FAIL(OLX_EINVAL, "lens tile: random forms need p_limit, pitch > 0");
        // include/Lensarry.h:167-169: xlim = |lambda f 0.5 / p_limit| (float), floor(xlim / pitch)
        const float xlim = (float)std::fabs(lambda * f * 0.5 / p_limit);
        wlim = (int)std::floor(xlim / pitch);
        // the std::mt19937(0) stream in the reference's traversal: blocks row-major, then
        // the block's rows and columns, a row cut at the field edge (its break), two draws
        // per pixel outside the window
        const size_t npx = (size_t)ny * nx;
        std::vector<float> ph(npx, 0.f);
        std::mt19937 e2(0);
        const float span = kind == 1 ? 2 * 3.14159265f : 0.5f * 3.14159265f;
        for (int bn = 0; bn < ny; bn += bp)
            for (int bm = 0; bm < nx; bm += bp)
                for (int ln = 0; ln < bp; ln++)
                    for (int lm = 0; lm < bp; lm++) {
                        int lms = (lm - om) % bp, lns = (ln - on) % bp;
                        lms = lms >= 0 ? lms : lms + bp;
                        lns = lns >= 0 ? lns : lns + bp;
                        const int m = bm + lm, n = bn + ln;
                        if (m >= nx || n >= ny) break;
                        if (std::abs(lms - bp / 2) >= wlim || std::abs(lns - bp / 2) >= wlim) {
                            const int a = (int)(e2() >> 5);
                            const int b = (int)(e2() >> 6);
                            const double value = (a * 67108864.0 + b) / 9007199254740992.0;
                            ph[(size_t)n * nx + m] = (float)(span * (value - 0.5f));
                        }
                    }
        OLX_TRY(ctx->obj_tmp.reserve(sizeof(float) * npx));
        OLX_HIP(hipMemcpyAsync(ctx->obj_tmp.p, ph.data(), sizeof(float) * npx, hipMemcpyHostToDevice,
                               ctx->stream));
        dph = (const float*)ctx->obj_tmp.p;
        k_lens_tile<<<dim3((unsigned)((nx + 255) / 256), (unsigned)ny), 256, 0, ctx->stream>>>(
            (const float2*)d_block, bp, (float2*)d_out, (int)ny, (int)nx, on, om, dph, kind, wlim);
        OLX_TRY(check_launch("lens tile"));
        // the host vector is freed on return: wait for the copy
        OLX_HIP(hipStreamSynchronize(ctx->stream));
        return OLX_OK;
    }
    k_lens_tile<<<dim3((unsigned)((nx + 255) / 256), (unsigned)ny), 256, 0, ctx->stream>>>(
        (const float2*)d_block, bp, (float2*)d_out, (int)ny, (int)nx, on, om, dph, 0, wlim);
    return check_launch("lens tile");
}
