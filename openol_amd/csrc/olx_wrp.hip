// olx_wrp.hip -- wavefront-recording-plane accumulation and the WRP method.
//
// WRPStep1 (include/golcgh.h:303-367 circle, :403-468 square) loops every point
// for every pixel.  Here points are binned instead:
//   1. per point: window radius W_j = |z_j - z_wrp| / sqrt(4 (p_lim/lambda)^2 - 1)
//      (square: D_j = sqrt(2) W_j) and the 16 x 16 pixel tiles its window touches;
//   2. counting sort of (tile, point) pairs: count, exclusive scan, fill;
//   3. one work-group per tile: load the tile's point list, sort it by point index
//      (fixed summation order), and accumulate exp(i pi rho^2/(lambda dz))/|dz|
//      for the pixels inside each window.
// Window tests and rho^2 use FP64 in the reference's expression order, so pixel
// membership is decided exactly as the reference does with PREC_T = double.
#include <climits>

#include "olx_common.hpp"

namespace olx {

constexpr int WT = 16;  // tile side (pixels)

__device__ __forceinline__ double wrp_radius(double z, double plane_z, double plim,
                                             double lambda, int square) {
    // WRPStep1Pre(_D), include/golcgh.h:298, :399-400
    const double W = 1.0 / sqrt(4.0 * (plim / lambda) * (plim / lambda) - 1.0) * fabs(z - plane_z);
    return square ? W * (double)1.41421353816986083984375 /* sqrtf(2.0f) */ : W;
}

// tile range [t0, t1] touched by a point's window along one axis
__device__ __forceinline__ void tile_span(double c, double R, double p, int n, int* t0, int* t1) {
    double lo = floor((c - R) / p) - 1.0, hi = ceil((c + R) / p) + 1.0;
    lo = fmax(lo, 0.0);
    hi = fmin(hi, (double)(n - 1));
    if (hi < lo) {
        *t0 = 1;
        *t1 = 0;
        return;
    }
    *t0 = (int)lo / WT;
    *t1 = (int)hi / WT;
}

__device__ __forceinline__ double window_extent(double Wd, int square) {
    return square ? 0.5 * Wd : Wd;
}

__global__ void k_wrp_count(const olx_point* __restrict__ pts, int n, int* __restrict__ counts,
                            int ny, int nx, double p, double lambda, double plane_z, double plim,
                            int square) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const olx_point q = pts[j];
    const double R = window_extent(wrp_radius(q.z, plane_z, plim, lambda, square), square);
    int x0, x1, y0, y1;
    tile_span(q.x, R, p, nx, &x0, &x1);
    tile_span(q.y, R, p, ny, &y0, &y1);
    const int tiles_x = (nx + WT - 1) / WT;
    for (int ty = y0; ty <= y1; ty++)
        for (int tx = x0; tx <= x1; tx++) atomicAdd(&counts[ty * tiles_x + tx], 1);
}

__global__ void k_wrp_fill(const olx_point* __restrict__ pts, int n, const int* __restrict__ offs,
                           int* __restrict__ cursor, int* __restrict__ list, int ny, int nx,
                           double p, double lambda, double plane_z, double plim, int square) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const olx_point q = pts[j];
    const double R = window_extent(wrp_radius(q.z, plane_z, plim, lambda, square), square);
    int x0, x1, y0, y1;
    tile_span(q.x, R, p, nx, &x0, &x1);
    tile_span(q.y, R, p, ny, &y0, &y1);
    const int tiles_x = (nx + WT - 1) / WT;
    for (int ty = y0; ty <= y1; ty++)
        for (int tx = x0; tx <= x1; tx++) {
            const int t = ty * tiles_x + tx;
            list[offs[t] + atomicAdd(&cursor[t], 1)] = j;
        }
}

// single work-group exclusive scan; offs[ntiles] = total
__global__ void __launch_bounds__(1024) k_scan(const int* __restrict__ counts, int* __restrict__ offs,
                                               int ntiles) {
    __shared__ long long part[1024];
    const int tid = threadIdx.x;
    const int per = (ntiles + 1023) / 1024;
    const int b = tid * per, e = min(ntiles, b + per);
    long long s = 0;
    for (int i = b; i < e; i++) s += counts[i];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        long long v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    long long run = tid ? part[tid - 1] : 0;
    for (int i = b; i < e; i++) {
        offs[i] = (int)run;
        run += counts[i];
    }
    if (tid == 1023) offs[ntiles] = part[1023] > INT_MAX ? -1 : (int)part[1023];
}

constexpr int WRP_SORT_CAP = 4096;

__global__ void __launch_bounds__(256)
k_wrp_accum(const olx_point* __restrict__ pts, const int* __restrict__ offs,
            const int* __restrict__ list, float2* __restrict__ U, int ny, int nx, double p,
            double lambda, double plane_z, double plim, int square) {
    __shared__ int ids[WRP_SORT_CAP];
    const int tiles_x = (nx + WT - 1) / WT;
    const int tile = blockIdx.x;
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int w = tx * WT + threadIdx.x % WT, h = ty * WT + threadIdx.x / WT;
    const int beg = offs[tile], cnt = offs[tile + 1] - beg;
    const double xw = p * w, yh = p * h;
    float re = 0.f, im = 0.f;
    const bool sorted = cnt <= WRP_SORT_CAP;
    if (sorted) {
        int np2 = 1;
        while (np2 < cnt) np2 <<= 1;
        for (int i = threadIdx.x; i < np2; i += 256) ids[i] = i < cnt ? list[beg + i] : INT_MAX;
        __syncthreads();
        // bitonic sort, ascending point index
        for (int k = 2; k <= np2; k <<= 1)
            for (int jj = k >> 1; jj > 0; jj >>= 1) {
                for (int i = threadIdx.x; i < np2; i += 256) {
                    const int l = i ^ jj;
                    if (l > i) {
                        const int a = ids[i], b = ids[l];
                        const bool up = (i & k) == 0;
                        if ((a > b) == up) {
                            ids[i] = b;
                            ids[l] = a;
                        }
                    }
                }
                __syncthreads();
            }
    }
    for (int q = 0; q < cnt; q++) {
        const int j = sorted ? ids[q] : list[beg + q];
        const olx_point P = pts[j];
        const double Wd = wrp_radius(P.z, plane_z, plim, lambda, square);
        const double dx = xw - P.x, dy = yh - P.y, dz = P.z - plane_z;
        const double tmp = dx * dx + dy * dy;
        const bool inside = square ? (fabs(dx) * 2 <= Wd && fabs(dy) * 2 <= Wd) : (tmp < Wd * Wd);
        if (inside) {
            double cyc = tmp / (2.0 * dz * lambda);  // 2 pi tmp/(2 dz lambda) rad
            cyc -= rint(cyc);
            const float t = (float)cyc;
            const float inv = (float)(1.0 / fabs(dz));
            re += __builtin_amdgcn_cosf(t) * inv;
            im += __builtin_amdgcn_sinf(t) * inv;
        }
    }
    if (w < nx && h < ny) U[(size_t)h * nx + w] = make_float2(re, im);
}

}  // namespace olx

using namespace olx;

extern "C" double olx_wrp_plane_z(int64_t ny, int64_t nx, double p, double lambda, double zmin,
                                  double p_limit, float s, int square) {
    float plane_z;
    if (s < 0) {
        if (square) {  // WRPMethod_D, include/golcgh.h:546-550
            const float D = (float)(nx * p * 0.05);
            const float d = (float)(std::sqrt(4 * std::pow((p_limit / lambda), 2) - 1) /
                                    std::sqrt(2) * D);
            plane_z = (float)(zmin - d);
        } else {  // WRPMethod, include/golcgh.h:567-571
            const float Wpow2 = (float)(nx * ny * p * p / M_PI * 1 / 1024.f);
            const float d = (float)std::sqrt((4 * std::pow((p / lambda), 2) - 1) * Wpow2);
            plane_z = (float)(zmin - d);
        }
    } else {
        plane_z = (float)(zmin * s);
    }
    return (double)plane_z;
}

extern "C" int olx_wrp_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                           int64_t ny, int64_t nx, double p, double lambda, double plane_z,
                           double p_limit, uint32_t flags) {
    if (!ctx || !d_U) OLX_FAIL(OLX_EINVAL, "wrp: null argument");
    if (n < 0 || (n > 0 && !d_pts)) OLX_FAIL(OLX_EINVAL, "wrp: bad point array");
    if (n > INT32_MAX / 4) OLX_FAIL(OLX_EUNSUPPORTED, "wrp: too many points");
    if (ny <= 0 || nx <= 0 || ny > 65536 || nx > 65536) OLX_FAIL(OLX_EINVAL, "wrp: bad plane size");
    if (!(p > 0) || !(lambda > 0) || !(p_limit > 0.5 * lambda))
        OLX_FAIL(OLX_EINVAL, "wrp: need p > 0, lambda > 0 and p_limit > lambda/2");
    const int square = (flags & OLX_WRP_SQUARE) ? 1 : 0;
    const int tiles_x = (int)((nx + WT - 1) / WT), tiles_y = (int)((ny + WT - 1) / WT);
    const int ntiles = tiles_x * tiles_y;
    OLX_TRY(ctx->wrp_counts.reserve(sizeof(int) * (size_t)(3 * ntiles + 2)));
    int* counts = (int*)ctx->wrp_counts.p;
    int* offs = counts + ntiles;       // ntiles + 1
    int* cursor = offs + ntiles + 1;   // ntiles
    hipStream_t s = ctx->stream;
    OLX_HIP(hipMemsetAsync(counts, 0, sizeof(int) * (size_t)ntiles, s));
    OLX_HIP(hipMemsetAsync(cursor, 0, sizeof(int) * (size_t)ntiles, s));
    const int nb = (int)((n + 255) / 256);
    {
        TimedLaunch tl(ctx, "wrp_bin");
        if (n > 0)
            k_wrp_count<<<nb, 256, 0, s>>>(d_pts, (int)n, counts, (int)ny, (int)nx, p, lambda,
                                           plane_z, p_limit, square);
        k_scan<<<1, 1024, 0, s>>>(counts, offs, ntiles);
    }
    OLX_TRY(check_launch("wrp count/scan"));
    int total = 0;
    OLX_HIP(hipMemcpyAsync(&total, offs + ntiles, sizeof(int), hipMemcpyDeviceToHost, s));
    OLX_HIP(hipStreamSynchronize(s));
    if (total < 0) OLX_FAIL(OLX_EUNSUPPORTED, "wrp: more than 2^31 (tile, point) pairs");
    OLX_TRY(ctx->wrp_lists.reserve(sizeof(int) * (size_t)std::max(total, 1)));
    int* list = (int*)ctx->wrp_lists.p;
    {
        TimedLaunch tl(ctx, "wrp_fill");
        if (n > 0)
            k_wrp_fill<<<nb, 256, 0, s>>>(d_pts, (int)n, offs, cursor, list, (int)ny, (int)nx, p,
                                          lambda, plane_z, p_limit, square);
    }
    OLX_TRY(check_launch("wrp fill"));
    {
        TimedLaunch tl(ctx, "wrp_accum");
        k_wrp_accum<<<ntiles, 256, 0, s>>>(d_pts, offs, list, (float2*)d_U, (int)ny, (int)nx, p,
                                           lambda, plane_z, p_limit, square);
    }
    return check_launch("wrp accumulate");
}

extern "C" int olx_wrp_method_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                                  int64_t ny, int64_t nx, double p, double lambda, double zmin,
                                  double p_limit, float s, int mode, uint32_t flags,
                                  double* plane_z_out) {
    const int square = (flags & OLX_WRP_SQUARE) ? 1 : 0;
    const double plane_z = olx_wrp_plane_z(ny, nx, p, lambda, zmin, p_limit, s, square);
    if (plane_z_out) *plane_z_out = plane_z;
    OLX_TRY(olx_wrp_c64(ctx, d_pts, n, d_U, ny, nx, p, lambda, plane_z, p_limit, flags));
    // WRPStep2, include/golcgh.h:494-517 (AUTO falls back to ASM, not band-limited Fresnel)
    int m = mode;
    if (m == OLX_MODE_AUTO) {
        if (olx_asm_prop_check(ny, nx, p, p, lambda, plane_z) ||
            !olx_fresnel_prop_check(ny, nx, p, p, lambda, plane_z))
            m = OLX_MODE_ASM;
        else
            m = OLX_MODE_FRESNEL;
    }
    if (m == OLX_MODE_ASM) return olx_asm_c64(ctx, d_U, ny, nx, p, p, lambda, plane_z, 0);
    if (m == OLX_MODE_FRESNEL) return olx_fresnel_c64(ctx, d_U, ny, nx, p, p, lambda, plane_z, 0);
    OLX_FAIL(OLX_EINVAL, "wrp_method: unknown mode %d", mode);
}
