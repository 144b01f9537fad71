// olx_utils.hip -- array helpers of the reference's olutils.h / golutils.h on HBM:
// window copies (zeropadding, del_zero, cut and their g* twins), the quadrant swap
// (fftshift / gfftshift), elementwise complex products and sums (mul_complex,
// mult, add) and real scaling (mul_scalar, standardization).
//
// One window-copy kernel serves padding, cropping and cutting: dst is dny x dnx,
// dst[h][w] = src[h - oy][w - ox] where that lies inside the sny x snx source and
// 0 elsewhere, for elements of 1, 2, 4, 8 or 16 bytes.  Each thread moves one
// element; rows of the destination are contiguous, so both sides coalesce.
#include <algorithm>

#include "olx_common.hpp"

namespace olx {

template <class T>
__global__ void k_window_copy(const T* __restrict__ src, int64_t sny, int64_t snx,
                              T* __restrict__ dst, int64_t dny, int64_t dnx, int64_t oy,
                              int64_t ox) {
    const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t h = blockIdx.y;
    if (w >= dnx || h >= dny) return;
    const int64_t sh = h - oy, sw = w - ox;
    T v{};
    if (sh >= 0 && sh < sny && sw >= 0 && sw < snx) v = src[sh * snx + sw];
    dst[h * dnx + w] = v;
}

// fftshift, include/olutils.h:95-117 / gfftshift, include/golutils.h:80-110: for
// h < ny/2, w < nx/2 swap (h, w) with (h + ny/2, w + nx/2) and (h, w + nx/2) with
// (h + ny/2, w) (the half-sizes floor, so odd sides keep their last row / column)
template <class T>
__global__ void k_fftshift(T* __restrict__ u, int64_t ny, int64_t nx) {
    const int64_t w = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t h = blockIdx.y;
    const int64_t hny = ny / 2, hnx = nx / 2;
    if (w >= hnx || h >= hny) return;
    int64_t i1 = h * nx + w, i2 = (h + hny) * nx + (w + hnx);
    T t = u[i1];
    u[i1] = u[i2];
    u[i2] = t;
    i1 = h * nx + (w + hnx);
    i2 = (h + hny) * nx + w;
    t = u[i1];
    u[i1] = u[i2];
    u[i2] = t;
}

__global__ void k_mul_c64(const float2* a, const float2* b, float2* c, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float2 x = a[i], y = b[i];
        c[i] = make_float2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
    }
}

__global__ void k_add_c64(const float2* a, const float2* b, float2* c, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float2 x = a[i], y = b[i];
        c[i] = make_float2(x.x + y.x, x.y + y.y);
    }
}

__global__ void k_scale_c64(float2* u, int64_t n, float s) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float2 x = u[i];
        u[i] = make_float2(x.x * s, x.y * s);
    }
}

__global__ void k_mul_c128(const double2* a, const double2* b, double2* c, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double2 x = a[i], y = b[i];
        c[i] = make_double2(x.x * y.x - x.y * y.y, x.x * y.y + x.y * y.x);
    }
}

__global__ void k_scale_c128(double2* u, int64_t n, double s) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double2 x = u[i];
        u[i] = make_double2(x.x * s, x.y * s);
    }
}

static unsigned grid_of(int64_t n) { return (unsigned)std::min<int64_t>((n + 255) / 256, 8192); }

template <class T>
static void launch_window(olx_ctx* ctx, const void* src, int64_t sny, int64_t snx, void* dst,
                          int64_t dny, int64_t dnx, int64_t oy, int64_t ox) {
    dim3 grid((unsigned)((dnx + 255) / 256), (unsigned)dny);
    k_window_copy<T><<<grid, 256, 0, ctx->stream>>>((const T*)src, sny, snx, (T*)dst, dny, dnx,
                                                     oy, ox);
}

template <class T>
static void launch_shift(olx_ctx* ctx, void* u, int64_t ny, int64_t nx) {
    dim3 grid((unsigned)((nx / 2 + 255) / 256), (unsigned)(ny / 2));
    k_fftshift<T><<<grid, 256, 0, ctx->stream>>>((T*)u, ny, nx);
}

}  // namespace olx

using namespace olx;

extern "C" int olx_window_copy(olx_ctx* ctx, const void* d_src, int64_t sny, int64_t snx,
                               void* d_dst, int64_t dny, int64_t dnx, int64_t oy, int64_t ox,
                               size_t elem_bytes) {
    if (!ctx || !d_src || !d_dst) OLX_FAIL(OLX_EINVAL, "window copy: null argument");
    if (sny < 0 || snx < 0 || dny < 0 || dnx < 0 || dny > 65535)
        OLX_FAIL(OLX_EINVAL, "window copy: bad sizes (%lld x %lld -> %lld x %lld)",
                 (long long)sny, (long long)snx, (long long)dny, (long long)dnx);
    if (d_src == d_dst) OLX_FAIL(OLX_EINVAL, "window copy: source and destination must differ");
    if (dny == 0 || dnx == 0) return OLX_OK;
    switch (elem_bytes) {
        case 1: launch_window<uint8_t>(ctx, d_src, sny, snx, d_dst, dny, dnx, oy, ox); break;
        case 2: launch_window<uint16_t>(ctx, d_src, sny, snx, d_dst, dny, dnx, oy, ox); break;
        case 4: launch_window<uint32_t>(ctx, d_src, sny, snx, d_dst, dny, dnx, oy, ox); break;
        case 8: launch_window<uint2>(ctx, d_src, sny, snx, d_dst, dny, dnx, oy, ox); break;
        case 16: launch_window<uint4>(ctx, d_src, sny, snx, d_dst, dny, dnx, oy, ox); break;
        default: OLX_FAIL(OLX_EINVAL, "window copy: element size %zu not 1, 2, 4, 8 or 16", elem_bytes);
    }
    return check_launch("window copy");
}

extern "C" int olx_fftshift(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, size_t elem_bytes) {
    if (!ctx || !d_u || ny < 0 || nx < 0 || ny > 131070) OLX_FAIL(OLX_EINVAL, "fftshift: bad argument");
    if (ny < 2 || nx < 2) return OLX_OK;
    switch (elem_bytes) {
        case 4: launch_shift<uint32_t>(ctx, d_u, ny, nx); break;
        case 8: launch_shift<uint2>(ctx, d_u, ny, nx); break;
        case 16: launch_shift<uint4>(ctx, d_u, ny, nx); break;
        default: OLX_FAIL(OLX_EINVAL, "fftshift: element size %zu not 4, 8 or 16", elem_bytes);
    }
    return check_launch("fftshift");
}

extern "C" int olx_mul_c64(olx_ctx* ctx, const void* d_a, const void* d_b, void* d_c, int64_t n) {
    if (!ctx || !d_a || !d_b || !d_c || n < 0) OLX_FAIL(OLX_EINVAL, "mul: bad argument");
    if (n == 0) return OLX_OK;
    k_mul_c64<<<grid_of(n), 256, 0, ctx->stream>>>((const float2*)d_a, (const float2*)d_b,
                                                   (float2*)d_c, n);
    return check_launch("mul");
}

extern "C" int olx_add_c64(olx_ctx* ctx, const void* d_a, const void* d_b, void* d_c, int64_t n) {
    if (!ctx || !d_a || !d_b || !d_c || n < 0) OLX_FAIL(OLX_EINVAL, "add: bad argument");
    if (n == 0) return OLX_OK;
    k_add_c64<<<grid_of(n), 256, 0, ctx->stream>>>((const float2*)d_a, (const float2*)d_b,
                                                   (float2*)d_c, n);
    return check_launch("add");
}

extern "C" int olx_scale_c64(olx_ctx* ctx, void* d_u, int64_t n, float s) {
    if (!ctx || !d_u || n < 0) OLX_FAIL(OLX_EINVAL, "scale: bad argument");
    if (n == 0) return OLX_OK;
    k_scale_c64<<<grid_of(n), 256, 0, ctx->stream>>>((float2*)d_u, n, s);
    return check_launch("scale");
}

extern "C" int olx_mul_c128(olx_ctx* ctx, const void* d_a, const void* d_b, void* d_c, int64_t n) {
    if (!ctx || !d_a || !d_b || !d_c || n < 0) OLX_FAIL(OLX_EINVAL, "mul: bad argument");
    if (n == 0) return OLX_OK;
    k_mul_c128<<<grid_of(n), 256, 0, ctx->stream>>>((const double2*)d_a, (const double2*)d_b,
                                                    (double2*)d_c, n);
    return check_launch("mul");
}

extern "C" int olx_scale_c128(olx_ctx* ctx, void* d_u, int64_t n, double s) {
    if (!ctx || !d_u || n < 0) OLX_FAIL(OLX_EINVAL, "scale: bad argument");
    if (n == 0) return OLX_OK;
    k_scale_c128<<<grid_of(n), 256, 0, ctx->stream>>>((double2*)d_u, n, s);
    return check_launch("scale");
}
