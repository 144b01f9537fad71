// olx_img.hip -- 8-bit hologram encodings (phase8bit, quant8bit, complex2img,
// complex2real) on the device and 8-bit grey BMP files on the host,
// include/olimg.h.  Arithmetic follows the reference's: FP64 with one rounding per
// operation (no fused multiply-add), round half away from zero, and the x86
// conversion of the rounded double to uint8 (through int32, modulo 256).
#include <cmath>
#include <cstdio>
#include <cstring>

#include "olx_common.hpp"
#include "img_common.hpp"

namespace olx {

int device_bounds(olx_ctx* ctx, const olx_point* d, int64_t n, double* mn, double* mx);

// four values per thread (16 bytes of float in, 4 bytes out); the last thread
// takes the tail one value at a time
template <typename T>
__global__ void k_phase8(const T* __restrict__ ph, uint8_t* __restrict__ img, int64_t n,
                         int offset) {
    const int64_t j = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) * 4;
    if (j >= n) return;
    if (sizeof(T) == 4 && j + 4 <= n && ((uintptr_t)(ph + j) & 15) == 0 && ((uintptr_t)(img + j) & 3) == 0) {
        const float4 x = *(const float4*)(ph + j);
        uchar4 o;
        o.x = phase8_one<T>((T)x.x, offset);
        o.y = phase8_one<T>((T)x.y, offset);
        o.z = phase8_one<T>((T)x.z, offset);
        o.w = phase8_one<T>((T)x.w, offset);
        *(uchar4*)(img + j) = o;
        return;
    }
    for (int64_t i = j; i < j + 4 && i < n; i++) img[i] = phase8_one<T>(ph[i], offset);
}

// ------------------------------------------------------------------ extremes
// min / max of v(j) as doubles: per-block partials part[2b], part[2b+1] from a
// grid-stride loop, then one work-group folds them.  fmin / fmax skip NaN, as the
// reference's > / < comparisons do for every element after the first.
constexpr int QT = 256, QB = 2048;

// LDS fold of one (mn, mx) pair per thread; every thread returns the block's pair
__device__ inline void block_fold(double& mn, double& mx) {
    __shared__ double smin[QT], smax[QT];
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int w = QT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
        }
        __syncthreads();
    }
    mn = smin[0];
    mx = smax[0];
}

template <typename F>
__device__ inline void block_minmax(F v, int64_t n, double* __restrict__ part) {
    double mn = INFINITY, mx = -INFINITY;
    for (int64_t j = blockIdx.x * (int64_t)QT + threadIdx.x; j < n; j += (int64_t)gridDim.x * QT) {
        const double x = v(j);
        mn = fmin(mn, x);
        mx = fmax(mx, x);
    }
    block_fold(mn, mx);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = mn;
        part[2 * blockIdx.x + 1] = mx;
    }
}

// the partials of nb blocks folded by one work-group of QT threads
__device__ inline void fold_partials(const double* __restrict__ part, int nb, double& mn, double& mx) {
    mn = INFINITY;
    mx = -INFINITY;
    for (int b = threadIdx.x; b < nb; b += QT) {
        mn = fmin(mn, part[2 * b]);
        mx = fmax(mx, part[2 * b + 1]);
    }
    block_fold(mn, mx);
}

template <typename T>
__global__ void __launch_bounds__(QT) k_real_minmax(const T* __restrict__ r, int64_t n,
                                                    double* __restrict__ part) {
    block_minmax([=](int64_t j) { return (double)r[j]; }, n, part);
}

// folds the partials into part[0..1] (written after every thread has read them).
// The reference seeds min and max with element 0, so a NaN there sticks (and every
// output is then 0); a NaN elsewhere is skipped.
template <typename T>
__global__ void __launch_bounds__(QT) k_minmax_final(double* __restrict__ part, int nb,
                                                     const T* __restrict__ r) {
    double mn, mx;
    fold_partials(part, nb, mn, mx);
    if (threadIdx.x != 0) return;
    if (r[0] != r[0]) mn = mx = (double)r[0];
    part[0] = mn;
    part[1] = mx;
}

// quant8bit: round((v - min) / (max - min) * 255)
template <typename T>
__global__ void k_quant8(const T* __restrict__ v, uint8_t* __restrict__ img, int64_t n,
                         const double* __restrict__ mm) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double mn = mm[0], mx = mm[1];
    const double t = __dmul_rn(__ddiv_rn(__dsub_rn((double)v[j], mn), __dsub_rn(mx, mn)), 255.0);
    img[j] = to_u8(round(t));
}

// ------------------------------------------------------------------ complex2img / real
template <typename T> struct C2;
template <> struct C2<float> { using V = float2; };
template <> struct C2<double> { using V = double2; };

__device__ inline float sq_sum(float2 u) { return __fadd_rn(__fmul_rn(u.x, u.x), __fmul_rn(u.y, u.y)); }
__device__ inline double sq_sum(double2 u) { return __dadd_rn(__dmul_rn(u.x, u.x), __dmul_rn(u.y, u.y)); }

// the per-element value of complex2img / complex2real.  The reference's headers
// include <math.h> before these templates, so unqualified sqrt / atan2 on float
// arguments are the float overloads: for complex<float> the modulus is the float
// square root of the float sum of float products, the phase the float angle; for
// complex<double> both are FP64.  The float angle here is the FP64 angle rounded to
// float (the correctly rounded atan2f, up to an FP64 ulp at a float tie).
// the correctly rounded float root: the FP64 root rounded to float (for a square
// root the double rounding is exact; the hardware FP32 root is not correctly rounded)
__device__ inline float sqrt_f32_rn(float s) { return (float)sqrt((double)s); }
__device__ inline double c2v_amp(float2 u) { return (double)sqrt_f32_rn(sq_sum(u)); }
__device__ inline double c2v_amp(double2 u) { return sqrt(sq_sum(u)); }
__device__ inline double c2v_phase(float2 u) { return (double)(float)atan2((double)u.y, (double)u.x); }
__device__ inline double c2v_phase(double2 u) { return atan2(u.y, u.x); }

template <typename V>
__device__ inline double c2v(V u, int mode) {
    switch (mode) {
        case OLX_IMG_REAL:
        case OLX_IMG_REAL_LEGACY:
            return (double)u.x;
        case OLX_IMG_IMAG:
            return (double)u.y;
        case OLX_IMG_AMP:
            return c2v_amp(u);
        default:
            return c2v_phase(u);
    }
}

// the value whose extremes give complex2img's bounds: for AMP the squared modulus
// (the correctly rounded square root is monotone, so the root of the extreme is the
// extreme of the roots, and one root per bound replaces one per element)
template <typename V>
__device__ inline double c2v_order(V u, int mode) {
    return mode == OLX_IMG_AMP ? (double)sq_sum(u) : c2v(u, mode);
}

template <typename T>
__global__ void __launch_bounds__(QT) k_c2_minmax(const typename C2<T>::V* __restrict__ u, int64_t n,
                                                  int mode, double* __restrict__ part) {
    block_minmax([=](int64_t j) { return c2v_order(u[j], mode); }, n, part);
}

// complex2img's bounds into mm[0..1] (non-PHASE modes): the extremes (square roots
// taken for AMP); the legacy real mode's maximum starts at 0.  A NaN value at
// element 0 seeds the reference's minimum and sticks there.
template <typename T>
__global__ void __launch_bounds__(QT) k_img_bounds(double* __restrict__ mm, int nb, int mode,
                                                   const typename C2<T>::V* __restrict__ u) {
    double mn, mx;
    fold_partials(mm, nb, mn, mx);
    if (threadIdx.x != 0) return;
    if (mode == OLX_IMG_AMP) {
        if (sizeof(T) == 4) {  // the float root of the (float) squared modulus
            mn = (double)sqrt_f32_rn((float)mn);
            mx = (double)sqrt_f32_rn((float)mx);
        } else {
            mn = sqrt(mn);
            mx = sqrt(mx);
        }
    }
    if (mode == OLX_IMG_REAL_LEGACY) mx = fmax(mx, 0.0);
    const double v0 = c2v(u[0], mode);
    if (v0 != v0) mn = v0;
    mm[0] = mn;
    mm[1] = mx;
}

// img = round(255 (v - min) / (max - min)); PHASE takes its fixed bounds [-pi, pi]
// in registers and, for float fields, an FP32 atan2 fast path (its error, under
// 1e-6 rad, moves q = 255 (v + pi) / 2pi by under 1e-4)
template <typename T, bool PHASE>
__global__ void k_c2img(const typename C2<T>::V* __restrict__ u, uint8_t* __restrict__ img,
                        int64_t n, int mode, const double* __restrict__ mm) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    const typename C2<T>::V z = u[j];
    if (PHASE && sizeof(T) == 4) {
        const float qf = (atan2f((float)z.y, (float)z.x) + 3.14159265f) * 40.5845105f;
        int k;
        if (fast_round(qf, 1e-3f, &k)) {
            img[j] = (uint8_t)(uint32_t)k;
            return;
        }
    }
    const double mn = PHASE ? -M_PI : mm[0], mx = PHASE ? M_PI : mm[1];
    const double v = PHASE ? c2v(z, OLX_IMG_PHASE) : c2v(z, mode);
    img[j] = to_u8(round(__ddiv_rn(__dmul_rn(255.0, __dsub_rn(v, mn)), __dsub_rn(mx, mn))));
}

template <typename T>
__global__ void k_c2real(const typename C2<T>::V* __restrict__ u, T* __restrict__ r, int64_t n,
                         int mode) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    r[j] = (T)c2v(u[j], mode);
}

template <typename T>
static int quant8(olx_ctx* ctx, const T* d_real, uint8_t* d_img, int64_t n) {
    const int nb = (int)std::min<int64_t>(QB, (n + QT - 1) / QT);
    OLX_TRY(ctx->obj_tmp.reserve(sizeof(double) * 2 * (size_t)nb));
    double* part = (double*)ctx->obj_tmp.p;
    hipStream_t s = ctx->stream;
    k_real_minmax<T><<<nb, QT, 0, s>>>(d_real, n, part);
    k_minmax_final<T><<<1, QT, 0, s>>>(part, nb, d_real);
    k_quant8<T><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(d_real, d_img, n, part);
    return check_launch("quant8bit");
}

template <typename T>
static int c2img(olx_ctx* ctx, const void* d_u, uint8_t* d_img, int64_t n, int mode) {
    using V = typename C2<T>::V;
    const V* u = (const V*)d_u;
    hipStream_t s = ctx->stream;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (mode == OLX_IMG_PHASE) {
        k_c2img<T, true><<<g, 256, 0, s>>>(u, d_img, n, mode, nullptr);
        return check_launch("complex2img");
    }
    const int nb = (int)std::min<int64_t>(QB, (n + QT - 1) / QT);
    OLX_TRY(ctx->obj_tmp.reserve(sizeof(double) * 2 * (size_t)nb));
    double* part = (double*)ctx->obj_tmp.p;
    k_c2_minmax<T><<<nb, QT, 0, s>>>(u, n, mode, part);
    k_img_bounds<T><<<1, QT, 0, s>>>(part, nb, mode, u);
    k_c2img<T, false><<<g, 256, 0, s>>>(u, d_img, n, mode, part);
    return check_launch("complex2img");
}

// ------------------------------------------------------------------ BMP files
#pragma pack(push, 1)
struct BmpFile {
    uint16_t type;
    uint32_t size;
    uint16_t r1, r2;
    uint32_t off;
};
struct BmpInfo {
    uint32_t size;
    int32_t w, h;
    uint16_t planes, bits;
    uint32_t comp, isize;
    int32_t xppm, yppm;
    uint32_t used, important;
};
#pragma pack(pop)
static_assert(sizeof(BmpFile) == 14 && sizeof(BmpInfo) == 40, "BMP header layout");

}  // namespace olx

using namespace olx;

extern "C" int olx_phase8bit(olx_ctx* ctx, const void* d_phase, uint8_t* d_img, int64_t n,
                             uint32_t flags) {
    if (!ctx || (n > 0 && (!d_phase || !d_img))) OLX_FAIL(OLX_EINVAL, "phase8bit: null argument");
    if (flags & ~(OLX_PHASE8_OFFSET | OLX_IN_F64)) OLX_FAIL(OLX_EINVAL, "phase8bit: unknown flags");
    if (n <= 0) return OLX_OK;
    const unsigned g = (unsigned)((n + 1023) / 1024);  // 4 values per thread
    const int off = (flags & OLX_PHASE8_OFFSET) ? 1 : 0;
    if (flags & OLX_IN_F64)
        k_phase8<double><<<g, 256, 0, ctx->stream>>>((const double*)d_phase, d_img, n, off);
    else
        k_phase8<float><<<g, 256, 0, ctx->stream>>>((const float*)d_phase, d_img, n, off);
    return check_launch("phase8bit");
}

extern "C" int olx_quant8bit(olx_ctx* ctx, const void* d_real, uint8_t* d_img, int64_t n,
                             uint32_t flags) {
    if (!ctx || (n > 0 && (!d_real || !d_img))) OLX_FAIL(OLX_EINVAL, "quant8bit: null argument");
    if (flags & ~OLX_IN_F64) OLX_FAIL(OLX_EINVAL, "quant8bit: unknown flags");
    if (n <= 0) return OLX_OK;
    return (flags & OLX_IN_F64) ? quant8<double>(ctx, (const double*)d_real, d_img, n)
                                : quant8<float>(ctx, (const float*)d_real, d_img, n);
}

extern "C" int olx_complex2img(olx_ctx* ctx, const void* d_u, uint8_t* d_img, int64_t n, int mode) {
    if (!ctx || (n > 0 && (!d_u || !d_img))) OLX_FAIL(OLX_EINVAL, "complex2img: null argument");
    const bool f64 = (mode & OLX_IN_F64) != 0;
    const int m = mode & ~OLX_IN_F64;
    if (m < OLX_IMG_PHASE || m > OLX_IMG_REAL_LEGACY)
        OLX_FAIL(OLX_EINVAL, "complex2img: unknown mode %d", mode);
    if (n <= 0) return OLX_OK;
    return f64 ? c2img<double>(ctx, d_u, d_img, n, m) : c2img<float>(ctx, d_u, d_img, n, m);
}

extern "C" int olx_complex2real(olx_ctx* ctx, const void* d_u, void* d_real, int64_t n, int mode) {
    if (!ctx || (n > 0 && (!d_u || !d_real))) OLX_FAIL(OLX_EINVAL, "complex2real: null argument");
    const bool f64 = (mode & OLX_IN_F64) != 0;
    const int m = mode & ~OLX_IN_F64;
    if (m < OLX_IMG_PHASE || m > OLX_IMG_AMP)
        OLX_FAIL(OLX_EINVAL, "complex2real: unknown mode %d", mode);
    if (n <= 0) return OLX_OK;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (f64)
        k_c2real<double><<<g, 256, 0, ctx->stream>>>((const double2*)d_u, (double*)d_real, n, m);
    else
        k_c2real<float><<<g, 256, 0, ctx->stream>>>((const float2*)d_u, (float*)d_real, n, m);
    return check_launch("complex2real");
}

extern "C" int olx_bmp_write(const char* path, const uint8_t* img, int ny, int nx) {
    if (!path || !img || ny <= 0 || nx <= 0) OLX_FAIL(OLX_EINVAL, "bmp_write: bad argument");
    BmpFile f = {19778, 14 + 40 + 1024 + (uint32_t)nx * (uint32_t)ny, 0, 0, 14 + 40 + 1024};
    BmpInfo i = {40, nx, ny, 1, 8, 0, 0, 0, 0, 0, 0};
    uint8_t pal[1024];
    for (int k = 0; k < 256; k++) {
        pal[4 * k] = pal[4 * k + 1] = pal[4 * k + 2] = (uint8_t)k;
        pal[4 * k + 3] = 0;
    }
    FILE* fp = fopen(path, "wb");
    if (!fp) OLX_FAIL(OLX_EIO, "bmp_write: cannot open %s", path);
    const uint8_t zeros[4] = {0, 0, 0, 0};
    const int pad = nx % 4;  // the reference's row padding
    bool ok = fwrite(&f, sizeof f, 1, fp) == 1 && fwrite(&i, sizeof i, 1, fp) == 1 &&
              fwrite(pal, sizeof pal, 1, fp) == 1;
    for (int r = 0; ok && r < ny; r++)
        ok = fwrite(img + (size_t)r * nx, 1, nx, fp) == (size_t)nx &&
             (pad == 0 || fwrite(zeros, 1, pad, fp) == (size_t)pad);
    if (fclose(fp) != 0) ok = false;
    if (!ok) OLX_FAIL(OLX_EIO, "bmp_write: write to %s failed", path);
    return OLX_OK;
}

extern "C" int olx_bmp_read(const char* path, uint8_t* img, int ny, int nx) {
    if (!path || !img || ny <= 0 || nx <= 0) OLX_FAIL(OLX_EINVAL, "bmp_read: bad argument");
    FILE* fp = fopen(path, "rb");
    if (!fp) OLX_FAIL(OLX_EIO, "bmp_read: cannot open %s", path);
    BmpFile f;
    BmpInfo i;
    uint8_t pal[1024];
    bool ok = fread(&f, sizeof f, 1, fp) == 1 && fread(&i, sizeof i, 1, fp) == 1 &&
              fread(pal, sizeof pal, 1, fp) == 1;
    const int pad = nx % 4;
    for (int r = 0; ok && r < ny; r++)
        ok = fread(img + (size_t)r * nx, 1, nx, fp) == (size_t)nx && fseek(fp, pad, SEEK_CUR) == 0;
    fclose(fp);
    if (!ok) OLX_FAIL(OLX_EIO, "bmp_read: %s is shorter than %d x %d", path, ny, nx);
    return OLX_OK;
}
