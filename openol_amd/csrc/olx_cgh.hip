// olx_cgh.hip -- direct point-cloud CGH.
//
// U[h,w] = sum_j A_j exp(2 pi i theta_j(h,w)),  theta in cycles:
//   Fresnel:  theta = ((px w - x_j)^2 + (py h - y_j)^2) / (2 lambda z_j) + phi_j / 2 pi
//             (the reference's 2 pi (dx^2+dy^2)/(2 dz lambda), include/olcgh.h:304-308)
//   exact r:  theta = sqrt(dx^2 + dy^2 + z_j^2) / lambda + phi_j / 2 pi
//
// Points are staged through LDS 256 at a time: each thread converts one point into
// the tile's FP64 anchors (phase at an anchor pixel reduced mod 1, and the linear
// and quadratic phase slopes around it), rounded to FP32 once.  The inner loop then
// evaluates every (pixel, point) pair in FP32 from an anchor with the pixel's small
// offset (at most 16 pixels), so the 1e4..1e5-cycle phases never meet FP32
// rounding.  sin/cos use the hardware transcendental unit in revolutions.  Each
// chunk of 256 points is summed on its own and then added to the pixel's total, so
// the roundings of the running sum happen once per chunk.
//
// Exact r (k_cgh_exact): a work-group owns a 32 x 32 pixel tile (256 threads, 4
// pixels each, one row of 32 contiguous pixels per 32 lanes, so every store is a
// 256-byte line) with one anchor at its centre.
//
// Exact r, per (pixel, point): with dq = r^2 - ra^2 and e = dq / ra^2,
// r - ra = ra (sqrt(1 + e) - 1).  When e stays small over the whole tile, the
// series ra (e/2 - e^2/8 + e^3/16) is used (3 FMA); its truncation error is below
// (5/128) e^4 ra, kept under 4e-8 cycles by the per-point test.  Other points
// (near field, wide tiles) use dq / (ra + sqrt(ra^2 + dq)).  Each 256-point chunk
// is partitioned so both groups run as uniform loops.

#include "olx_common.hpp"
#include "img_common.hpp"

namespace olx {

constexpr int CGH_TX = 32, CGH_TY = 32, CGH_NT = 256, CGH_RPT = CGH_TY / (CGH_NT / CGH_TX);

// The outputs of a CGH kernel, any of them null: the field (added to what U holds
// with accumulate), its phase atan2(im, re), its amplitude, and the phase8bit byte of
// that phase (include/olimg.h:205-213; img_offset: the +pi form) fused into the
// epilogue in place of a phase array and a second pass.
struct CghOut {
    float2* U;
    float* phase;
    float* amp;
    uint8_t* img;
    int img_offset;
    int accumulate;
};

__device__ __forceinline__ void cgh_store(const CghOut& o, size_t idx, float2 v) {
    if (o.accumulate && o.U) {
        const float2 q = o.U[idx];
        v = make_float2(v.x + q.x, v.y + q.y);
    }
    if (o.U) o.U[idx] = v;
    if (o.phase || o.img) {
        const float ph = atan2f(v.y, v.x);
        if (o.phase) o.phase[idx] = ph;
        if (o.img) o.img[idx] = phase8_one<float>(ph, o.img_offset);
    }
    if (o.amp) o.amp[idx] = sqrtf(v.x * v.x + v.y * v.y);
}

// A hologram of few tiles splits its points into gridDim.z slices of `slice` points
// (a multiple of 256, so chunking is unchanged); slice z writes its partial plane to
// part + z ny nx, and k_cgh_sum_slices adds the planes in slice order.
__device__ __forceinline__ void cgh_slice(const olx_point* __restrict__& pts, int& n, CghOut& o,
                                          float2* part, int slice, int ny, int nx) {
    if (!part) return;
    const int start = (int)blockIdx.z * slice;
    pts += start;
    n = max(0, min(slice, n - start));
    o = CghOut{part + (size_t)blockIdx.z * ny * nx, nullptr, nullptr, nullptr, 0, 0};
}

// sin / cos of t revolutions: the hardware units take revolutions; t is first
// reduced to [0, 1) with one v_fract (exact for a float)
__device__ __forceinline__ void sincos_rev(float t, float* s, float* c) {
    t = __builtin_amdgcn_fractf(t);
    *s = __builtin_amdgcn_sinf(t);
    *c = __builtin_amdgcn_cosf(t);
}

struct PtAnchorR {  // per-(tile, point) constants of the exact-r kernel, FP32
    float th;       // phase at the anchor pixel, cycles mod 1 (phi folded in)
    float gx, gy;   // series: 2 dxa px / ra^2, 2 dya py / ra^2; sqrt: 2 dxa px, 2 dya py
    float qx, qy;   // series: px^2 / ra^2, py^2 / ra^2; sqrt: ra, ra^2
    float amp;
    float c1, c2, c3;  // series coefficients (ra / lambda) (1/2, -1/8, 1/16)
};

// series bound: (5/128) e^4 ra / lambda <= 4e-8 cycles
constexpr double CGH_SERIES_E4K = 1.0e-6;

__global__ void __launch_bounds__(CGH_NT)
k_cgh_exact(const olx_point* __restrict__ pts, int n, CghOut o, int ny, int nx, double py,
            double px, double lambda, float2* __restrict__ part, int slice) {
    cgh_slice(pts, n, o, part, slice, ny, nx);
    __shared__ PtAnchorR sp[CGH_NT];
    __shared__ int wcount[2][CGH_NT / 64];
    const int tid = threadIdx.x, tx = tid % CGH_TX, ty = tid / CGH_TX;
    const int w0 = blockIdx.x * CGH_TX, h0 = blockIdx.y * CGH_TY;
    const int wa = w0 + CGH_TX / 2, ha = h0 + CGH_TY / 2;  // anchor pixel
    const float a = (float)(tx - CGH_TX / 2);               // this thread's x offset
    float bk[CGH_RPT];
#pragma unroll
    for (int k = 0; k < CGH_RPT; k++) bk[k] = (float)(ty + k * (CGH_NT / CGH_TX) - CGH_TY / 2);
    const float px2 = (float)(px * px), py2 = (float)(py * py);
    const float inv_lambda = (float)(1.0 / lambda);
    // largest pixel offsets from the anchor, metres
    const double hx = (CGH_TX / 2) * px, hy = (CGH_TY / 2) * py;
    float re[CGH_RPT], im[CGH_RPT];
#pragma unroll
    for (int k = 0; k < CGH_RPT; k++) re[k] = im[k] = 0.f;

    for (int base = 0; base < n; base += CGH_NT) {
        const int j = base + tid;
        PtAnchorR pa;
        bool series = false;
        if (j < n) {
            const olx_point p = pts[j];
            const double dxa = px * wa - p.x, dya = py * ha - p.y;
            const double ra2 = dxa * dxa + dya * dya + p.z * p.z;
            const double ra = sqrt(ra2);
            double th = ra / lambda;
            // |e| over the tile
            const double em = (2.0 * (fabs(dxa) * hx + fabs(dya) * hy) + hx * hx + hy * hy) / ra2;
            series = em <= 0.01 && em * em * em * em * th <= CGH_SERIES_E4K;
            if (series) {
                const double ir2 = 1.0 / ra2;
                pa.gx = (float)(2.0 * dxa * px * ir2);
                pa.gy = (float)(2.0 * dya * py * ir2);
                pa.qx = (float)(px * px * ir2);
                pa.qy = (float)(py * py * ir2);
                pa.c1 = (float)(0.5 * th);
                pa.c2 = (float)(-0.125 * th);
                pa.c3 = (float)(0.0625 * th);
            } else {
                pa.gx = (float)(2.0 * dxa * px);
                pa.gy = (float)(2.0 * dya * py);
                pa.qx = (float)ra;
                pa.qy = (float)ra2;
                pa.c1 = pa.c2 = pa.c3 = 0.f;
            }
            th += (double)p.phase * (0.5 / M_PI);
            pa.th = (float)(th - rint(th));
            pa.amp = p.amp;
        }
        const int cnt = min(CGH_NT, n - base);
        int nser = 0;  // slots [0, nser) hold series points, [nser, cnt) the others
        {
            const int lane = tid & 63, wv = tid >> 6;
            const unsigned long long ms = __ballot(j < n && series);
            const unsigned long long mo = __ballot(j < n && !series);
            if (lane == 0) {
                wcount[0][wv] = __popcll(ms);
                wcount[1][wv] = __popcll(mo);
            }
            __syncthreads();
            int sb = 0, ob = 0;
#pragma unroll
            for (int v = 0; v < CGH_NT / 64; v++) {
                nser += wcount[0][v];
                if (v < wv) {
                    sb += wcount[0][v];
                    ob += wcount[1][v];
                }
            }
            const unsigned long long below = (1ull << lane) - 1ull;
            if (j < n)
                sp[series ? sb + __popcll(ms & below) : nser + ob + __popcll(mo & below)] = pa;
        }
        __syncthreads();
        float cre[CGH_RPT], cim[CGH_RPT];  // this chunk's sums
#pragma unroll
        for (int k = 0; k < CGH_RPT; k++) cre[k] = cim[k] = 0.f;
        for (int q = 0; q < nser; q++) {
            const PtAnchorR P = sp[q];
            const float ex = a * (P.gx + a * P.qx);
#pragma unroll
            for (int k = 0; k < CGH_RPT; k++) {
                const float e = ex + bk[k] * (P.gy + bk[k] * P.qy);  // dq / ra^2
                const float poly = fmaf(e, fmaf(e, P.c3, P.c2), P.c1);
                const float t = fmaf(e, poly, P.th);  // + (r - ra) / lambda
                float s, c;
                sincos_rev(t, &s, &c);
                cre[k] += P.amp * c;
                cim[k] += P.amp * s;
            }
        }
        for (int q = nser; q < cnt; q++) {
            const PtAnchorR P = sp[q];
            const float dqx = a * (P.gx + a * px2);
#pragma unroll
            for (int k = 0; k < CGH_RPT; k++) {
                const float dq = dqx + bk[k] * (P.gy + bk[k] * py2);
                const float dr = dq / (P.qx + sqrtf(P.qy + dq));  // r - ra
                const float t = P.th + dr * inv_lambda;
                float s, c;
                sincos_rev(t, &s, &c);
                cre[k] += P.amp * c;
                cim[k] += P.amp * s;
            }
        }
#pragma unroll
        for (int k = 0; k < CGH_RPT; k++) {
            re[k] += cre[k];
            im[k] += cim[k];
        }
        __syncthreads();
    }
    const int w = w0 + tx;
    if (w >= nx) return;
#pragma unroll
    for (int k = 0; k < CGH_RPT; k++) {
        const int h = h0 + ty + k * (CGH_NT / CGH_TX);
        if (h >= ny) continue;
        cgh_store(o, (size_t)h * nx + w, make_float2(re[k], im[k]));
    }
}

// Fresnel, separable.  theta = X(w) + Y(h) with X = c (px w - x_j)^2 + phi_j / 2 pi and
// Y = c (py h - y_j)^2, c = 1 / (2 lambda z_j), so exp(2 pi i theta) = exp(2 pi i X)
// exp(2 pi i Y): a thread holding 4 columns x 8 rows takes 4 + 8 sine / cosine pairs
// per point for its 32 pixels, and each pixel is then the complex product accumulated
// (4 FMA, no transcendental) instead of a phase, a reduction and a sine / cosine pair.
// A work-group owns a 128 x 64 tile (256 threads): thread (tx, ty) holds columns
// w0 + 32 i + tx and rows h0 + ty + 8 k, so every store is a 256-byte line.  Each
// column group i and each 32-row half m has its own FP64 anchor (its phase mod 1 and
// slope at the group's centre), so the FP32 offsets stay within 16 pixels, as in the
// exact-r kernel.
constexpr int SEP_TX = 128, SEP_TY = 64, SEP_NT = 256, SEP_NI = 4, SEP_NK = 8;

struct SepAnchor {  // per (tile, point), FP32
    float thx[SEP_NI];  // X at column anchor i, cycles mod 1 (phi folded in)
    float gx[SEP_NI];   // 2 c dx_i px (cycles per pixel)
    float thy[2];       // Y at row anchor m, cycles mod 1
    float gy[2];        // 2 c dy_m py
    float qx, qy;       // c px^2, c py^2
    float amp, pad;
};

__global__ void __launch_bounds__(SEP_NT)
k_cgh_fresnel_sep(const olx_point* __restrict__ pts, int n, CghOut o, int ny, int nx,
                  double py, double px, double lambda, float2* __restrict__ part, int slice) {
    cgh_slice(pts, n, o, part, slice, ny, nx);
    __shared__ SepAnchor sp[SEP_NT];
    const int tid = threadIdx.x, tx = tid % 32, ty = tid / 32;
    const int w0 = blockIdx.x * SEP_TX, h0 = blockIdx.y * SEP_TY;
    const float a = (float)(tx - 16);  // offset from every column anchor
    float bk[4];                       // row offsets from the half's anchor
#pragma unroll
    for (int k = 0; k < 4; k++) bk[k] = (float)(ty + 8 * k - 16);
    float re[SEP_NI][SEP_NK], im[SEP_NI][SEP_NK];
#pragma unroll
    for (int i = 0; i < SEP_NI; i++)
#pragma unroll
        for (int k = 0; k < SEP_NK; k++) re[i][k] = im[i][k] = 0.f;

    for (int base = 0; base < n; base += SEP_NT) {
        const int j = base + tid;
        if (j < n) {
            const olx_point p = pts[j];
            const double c = 1.0 / (2.0 * lambda * p.z);
            const double ph = (double)p.phase * (0.5 / M_PI);
            SepAnchor pa;
#pragma unroll
            for (int i = 0; i < SEP_NI; i++) {
                const double dx = px * (w0 + 32 * i + 16) - p.x;
                const double th = c * dx * dx + ph;
                pa.thx[i] = (float)(th - rint(th));
                pa.gx[i] = (float)(2.0 * c * dx * px);
            }
#pragma unroll
            for (int m = 0; m < 2; m++) {
                const double dy = py * (h0 + 32 * m + 16) - p.y;
                const double th = c * dy * dy;
                pa.thy[m] = (float)(th - rint(th));
                pa.gy[m] = (float)(2.0 * c * dy * py);
            }
            pa.qx = (float)(c * px * px);
            pa.qy = (float)(c * py * py);
            pa.amp = p.amp;
            pa.pad = 0.f;
            sp[tid] = pa;
        }
        __syncthreads();
        const int cnt = min(SEP_NT, n - base);
        // the chunk's sum first: its roundings happen at the chunk's magnitude (about
        // 16 amp) rather than the whole sum's (sqrt(n) amp), once per chunk there
        float cre[SEP_NI][SEP_NK], cim[SEP_NI][SEP_NK];
#pragma unroll
        for (int i = 0; i < SEP_NI; i++)
#pragma unroll
            for (int k = 0; k < SEP_NK; k++) cre[i][k] = cim[i][k] = 0.f;
        for (int q = 0; q < cnt; q++) {
            const SepAnchor P = sp[q];
            float cx[SEP_NI], sx[SEP_NI], cy[SEP_NK], sy[SEP_NK];
#pragma unroll
            for (int i = 0; i < SEP_NI; i++) {
                sincos_rev(P.thx[i] + a * (P.gx[i] + a * P.qx), &sx[i], &cx[i]);
                cx[i] *= P.amp;
                sx[i] *= P.amp;
            }
#pragma unroll
            for (int k = 0; k < SEP_NK; k++) {
                const int m = k / 4;
                const float b = bk[k % 4];
                sincos_rev(P.thy[m] + b * (P.gy[m] + b * P.qy), &sy[k], &cy[k]);
            }
#pragma unroll
            for (int i = 0; i < SEP_NI; i++)
#pragma unroll
                for (int k = 0; k < SEP_NK; k++) {
                    cre[i][k] = fmaf(cx[i], cy[k], fmaf(-sx[i], sy[k], cre[i][k]));
                    cim[i][k] = fmaf(sx[i], cy[k], fmaf(cx[i], sy[k], cim[i][k]));
                }
        }
#pragma unroll
        for (int i = 0; i < SEP_NI; i++)
#pragma unroll
            for (int k = 0; k < SEP_NK; k++) {
                re[i][k] += cre[i][k];
                im[i][k] += cim[i][k];
            }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < SEP_NI; i++) {
        const int w = w0 + 32 * i + tx;
        if (w >= nx) continue;
#pragma unroll
        for (int k = 0; k < SEP_NK; k++) {
            const int h = h0 + ty + 8 * k;
            if (h >= ny) continue;
            cgh_store(o, (size_t)h * nx + w, make_float2(re[i][k], im[i][k]));
        }
    }
}

// FP64 reference kernel: one pixel per thread, FP64 throughout
template <bool EXACT>
__global__ void __launch_bounds__(256)
k_cgh_f64(const olx_point* __restrict__ pts, int n, double2* __restrict__ U, int ny, int nx,
          double py, double px, double lambda) {
    __shared__ olx_point sp[256];
    const int w = blockIdx.x * 16 + threadIdx.x % 16, h = blockIdx.y * 16 + threadIdx.x / 16;
    const double xw = px * w, yh = py * h;
    double re = 0.0, im = 0.0;
    for (int base = 0; base < n; base += 256) {
        const int j = base + threadIdx.x;
        if (j < n) sp[threadIdx.x] = pts[j];
        __syncthreads();
        const int cnt = min(256, n - base);
        for (int q = 0; q < cnt; q++) {
            const olx_point p = sp[q];
            const double dx = xw - p.x, dy = yh - p.y;
            double th = EXACT ? sqrt(dx * dx + dy * dy + p.z * p.z) / lambda
                              : (dx * dx + dy * dy) / (2.0 * p.z * lambda);
            th += (double)p.phase * (0.5 / M_PI);
            th -= rint(th);
            double s, c;
            sincospi(2.0 * th, &s, &c);
            re += (double)p.amp * c;
            im += (double)p.amp * s;
        }
        __syncthreads();
    }
    if (w < nx && h < ny) U[(size_t)h * nx + w] = make_double2(re, im);
}

__global__ void k_cgh_sum_slices(const float2* __restrict__ part, int S, int64_t np, CghOut o) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np;
         i += (int64_t)gridDim.x * blockDim.x) {
        float2 v = part[i];
        for (int s = 1; s < S; s++) {
            const float2 w = part[(size_t)s * np + i];
            v = make_float2(v.x + w.x, v.y + w.y);
        }
        cgh_store(o, (size_t)i, v);
    }
}

__global__ void k_encode(const float2* __restrict__ U, float* __restrict__ ph,
                         float* __restrict__ am, int64_t count) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float2 v = U[i];
        if (ph) ph[i] = atan2f(v.y, v.x);
        if (am) am[i] = sqrtf(v.x * v.x + v.y * v.y);
    }
}

// _CghFresnel with PREC_T = float (include/olcgh.h:296-313; the reference tests'
// default): dx = px w - x_j, dy = py h - y_j, dz = z_j in float, theta = 2 pi (dx^2 +
// dy^2) / (2 dz lambda) formed in double from the float terms and stored to float,
// cos / sin of that float angle accumulated in float.  Each float operation rounds
// (no contraction).  One thread per pixel; points staged through LDS in tiles.
constexpr int F32C_NT = 256, F32C_TILE = 256;
__global__ void __launch_bounds__(F32C_NT)
k_cgh_f32compat(const olx_point* __restrict__ pts, int n, CghOut o, int ny, int nx,
                float pyf, float pxf, float lf) {
#pragma clang fp contract(off)
    __shared__ float4 sp[F32C_TILE];
    const int64_t pix = blockIdx.x * (int64_t)F32C_NT + threadIdx.x;
    const bool valid = pix < (int64_t)ny * nx;
    const int h = valid ? (int)(pix / nx) : 0, w = valid ? (int)(pix - (int64_t)h * nx) : 0;
    const float xw = pxf * (float)w, yh = pyf * (float)h;
    float re = 0.f, im = 0.f;
    for (int j0 = 0; j0 < n; j0 += F32C_TILE) {
        __syncthreads();
        for (int k = threadIdx.x; k < F32C_TILE && j0 + k < n; k += F32C_NT) {
            const olx_point q = pts[j0 + k];
            sp[k] = make_float4((float)q.x, (float)q.y, (float)q.z, q.amp);
        }
        __syncthreads();
        const int m = min(F32C_TILE, n - j0);
        for (int k = 0; k < m; k++) {
            const float4 q = sp[k];
            const float dx = xw - q.x, dy = yh - q.y, dz = q.z;
            const float th = (float)(2.0 * M_PI * (double)(dx * dx + dy * dy) / (double)(2.0f * dz * lf));
            float sn, cs;
            sincosf(th, &sn, &cs);
            re += cs;
            im += sn;
        }
    }
    if (!valid) return;
    cgh_store(o, (size_t)pix, make_float2(re, im));
}

static int check_cgh(olx_ctx* ctx, const olx_point* pts, int64_t n, int64_t ny, int64_t nx,
                     double py, double px, double lambda) {
    if (!ctx) OLX_FAIL(OLX_EINVAL, "null context");
    if (n < 0 || (n > 0 && !pts)) OLX_FAIL(OLX_EINVAL, "cgh: bad point array");
    if (n > INT32_MAX) OLX_FAIL(OLX_EUNSUPPORTED, "cgh: more than 2^31-1 points");
    if (ny <= 0 || nx <= 0 || ny > (1 << 20) || nx > (1 << 20))
        OLX_FAIL(OLX_EINVAL, "cgh: bad hologram size %lld x %lld", (long long)ny, (long long)nx);
    if (!(py > 0) || !(px > 0) || !(lambda > 0)) OLX_FAIL(OLX_EINVAL, "cgh: pitch and lambda must be > 0");
    return OLX_OK;
}

}  // namespace olx

using namespace olx;

static int cgh_run(olx_ctx* ctx, const olx_point* d_pts, int64_t n, CghOut o, int64_t ny,
                   int64_t nx, double py, double px, double lambda, uint32_t flags) {
    OLX_TRY(check_cgh(ctx, d_pts, n, ny, nx, py, px, lambda));
    if (!o.U && !o.phase && !o.amp && !o.img) OLX_FAIL(OLX_EINVAL, "cgh: no output requested");
    if ((flags & OLX_CGH_ACCUMULATE) && !o.U) OLX_FAIL(OLX_EINVAL, "cgh: accumulate needs d_U");
    o.accumulate = (flags & OLX_CGH_ACCUMULATE) ? 1 : 0;
    const bool exact = (flags & OLX_CGH_EXACT_R) != 0;
    if (flags & OLX_PREC_F32) {
        if (exact) OLX_FAIL(OLX_EINVAL, "cgh: the float-PREC_T form is the Fresnel sum only");
        const int64_t np = ny * nx;
        if (np / F32C_NT >= INT32_MAX) OLX_FAIL(OLX_EUNSUPPORTED, "cgh: hologram too large");
        TimedLaunch tl(ctx, "cgh_f32compat");
        k_cgh_f32compat<<<(unsigned)((np + F32C_NT - 1) / F32C_NT), F32C_NT, 0, ctx->stream>>>(
            d_pts, (int)n, o, (int)ny, (int)nx, (float)py, (float)px, (float)lambda);
        return check_launch("cgh_f32compat");
    }
    const int tx = exact ? CGH_TX : SEP_TX, ty = exact ? CGH_TY : SEP_TY;
    dim3 grid((unsigned)((nx + tx - 1) / tx), (unsigned)((ny + ty - 1) / ty));
    // fewer tiles than CUs: split the points into slices of at least 2048 (8 chunks),
    // up to about 1024 work-groups, and sum the partial planes in slice order
    const int64_t tiles = (int64_t)grid.x * grid.y, np = ny * nx;
    int64_t S = 1, slice = n;
    if (tiles < 256 && n >= 4096) {
        S = std::min<int64_t>((1024 + tiles - 1) / tiles, n / 2048);
        slice = ((n + S - 1) / S + 255) / 256 * 256;
        S = (n + slice - 1) / slice;
    }
    float2* part = nullptr;
    if (S > 1) {
        OLX_TRY(ctx->cgh_part.reserve(sizeof(float2) * (size_t)(S * np)));
        part = (float2*)ctx->cgh_part.p;
        grid.z = (unsigned)S;
    }
    {
        TimedLaunch tl(ctx, exact ? "cgh_f32_exact" : "cgh_f32_fresnel");
        if (exact)
            k_cgh_exact<<<grid, CGH_NT, 0, ctx->stream>>>(d_pts, (int)n, o, (int)ny, (int)nx, py,
                                                          px, lambda, part, (int)slice);
        else
            k_cgh_fresnel_sep<<<grid, SEP_NT, 0, ctx->stream>>>(d_pts, (int)n, o, (int)ny, (int)nx,
                                                                py, px, lambda, part, (int)slice);
        if (part) {
            const int64_t blocks = std::min<int64_t>((np + 255) / 256, 65536);
            k_cgh_sum_slices<<<(unsigned)blocks, 256, 0, ctx->stream>>>(part, (int)S, np, o);
        }
    }
    return check_launch("cgh_f32");
}

extern "C" int olx_cgh_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                           float* d_phase, float* d_amp, int64_t ny, int64_t nx, double py,
                           double px, double lambda, uint32_t flags) {
    return cgh_run(ctx, d_pts, n, CghOut{(float2*)d_U, d_phase, d_amp, nullptr, 0, 0}, ny, nx, py,
                   px, lambda, flags);
}

extern "C" int olx_cgh_img8_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                                uint8_t* d_img, int64_t ny, int64_t nx, double py, double px,
                                double lambda, uint32_t flags) {
    if (!d_img) OLX_FAIL(OLX_EINVAL, "cgh img8: null image");
    const int off = (flags & OLX_PHASE8_OFFSET) ? 1 : 0;
    return cgh_run(ctx, d_pts, n, CghOut{(float2*)d_U, nullptr, nullptr, d_img, off, 0}, ny, nx,
                   py, px, lambda, flags & ~OLX_PHASE8_OFFSET);
}

extern "C" int olx_cgh_c128(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                            int64_t ny, int64_t nx, double py, double px, double lambda,
                            uint32_t flags) {
    OLX_TRY(check_cgh(ctx, d_pts, n, ny, nx, py, px, lambda));
    if (!d_U) OLX_FAIL(OLX_EINVAL, "cgh_c128: null output");
    dim3 grid((unsigned)((nx + 15) / 16), (unsigned)((ny + 15) / 16));
    {
        TimedLaunch tl(ctx, "cgh_f64");
        if (flags & OLX_CGH_EXACT_R)
            k_cgh_f64<true><<<grid, 256, 0, ctx->stream>>>(d_pts, (int)n, (double2*)d_U, (int)ny,
                                                           (int)nx, py, px, lambda);
        else
            k_cgh_f64<false><<<grid, 256, 0, ctx->stream>>>(d_pts, (int)n, (double2*)d_U, (int)ny,
                                                            (int)nx, py, px, lambda);
    }
    return check_launch("cgh_f64");
}

extern "C" int olx_encode_c64(olx_ctx* ctx, const void* d_U, float* d_phase, float* d_amp,
                              int64_t count) {
    if (!ctx || !d_U || count < 0) OLX_FAIL(OLX_EINVAL, "encode: bad argument");
    if (count == 0 || (!d_phase && !d_amp)) return OLX_OK;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, 65536);
    k_encode<<<(unsigned)blocks, 256, 0, ctx->stream>>>((const float2*)d_U, d_phase, d_amp, count);
    return check_launch("encode");
}

// _CghFresnelDiff (include/olcgh.h:315-349): coordinates in pixel units, the phase in
// cycles carried along each row by theta <- theta + delta, delta <- delta + gamma
// from theta(0) = p z / lambda + gamma / 2 (dx0^2 + dy0^2), delta(0) = gamma / 2
// (2 dx0 + 1), gamma = p / (lambda z).  In exact arithmetic that is theta(w) =
// p z / lambda + p ((w - x)^2 + (h - y)^2) / (2 lambda z): the Fresnel sum of the
// cloud scaled to metres (p x, p y, p z) plus a per-point piston 2 pi frac(p z /
// lambda).  Each point is converted once (FP64 piston reduced mod 1, added to the
// point's phase) and the direct kernel evaluates the sum.
__global__ void k_diff_points(const olx_point* __restrict__ in, olx_point* __restrict__ out,
                              int n, double p, double lambda) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    olx_point q = in[j];
    const double piston = p * q.z / lambda;
    const double cyc = piston - floor(piston);
    q.phase = (float)((double)q.phase + 2.0 * M_PI * cyc);
    q.x *= p;
    q.y *= p;
    q.z *= p;
    out[j] = q;
}

extern "C" int olx_cgh_fresnel_diff_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n,
                                        void* d_U, float* d_phase, int64_t ny, int64_t nx,
                                        double p, double lambda) {
    if (!ctx) OLX_FAIL(OLX_EINVAL, "cgh diff: null context");
    if (!(p > 0)) OLX_FAIL(OLX_EINVAL, "cgh diff: pitch must be > 0");
    if (n < 0 || (n > 0 && !d_pts)) OLX_FAIL(OLX_EINVAL, "cgh diff: bad point array");
    if (n > INT32_MAX) OLX_FAIL(OLX_EUNSUPPORTED, "cgh diff: more than 2^31-1 points");
    if (!d_U && !d_phase) OLX_FAIL(OLX_EINVAL, "cgh diff: no output requested");
    const olx_point* pts = d_pts;
    if (n > 0) {
        OLX_TRY(ctx->lay_pts.reserve(sizeof(olx_point) * (size_t)n));
        k_diff_points<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(
            d_pts, (olx_point*)ctx->lay_pts.p, (int)n, p, lambda);
        OLX_TRY(check_launch("cgh diff points"));
        pts = (const olx_point*)ctx->lay_pts.p;
    }
    // unit pixel pitch in the scaled cloud: p w - p x = p (w - x)
    return olx_cgh_c64(ctx, pts, n, d_U, d_phase, nullptr, ny, nx, p, p, lambda, 0);
}
