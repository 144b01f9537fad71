// olx_obj.hip -- point-cloud ingest (.3d / .3df files) and the device forms of
// the reference's cloud helpers: bounding box, objset (fit to a range) and
// objsort (descending z), include/olobject.h.
#include <cstdio>
#include <cstring>

#include "olx_common.hpp"

namespace olx {

// ------------------------------------------------------------------ reading
static int read_cloud(const char* path, int format, olx_point* pts, int64_t cap, int64_t* n,
                      double scale, double ox, double oy, double oz) {
    if (format == OLX_OBJ_AUTO) {
        const char* ext = strrchr(path, '.');
        if (ext && strcmp(ext, ".3d") == 0) format = OLX_OBJ_3D;
        else if (ext && strcmp(ext, ".3df") == 0) format = OLX_OBJ_3DF;
        else OLX_FAIL(OLX_EINVAL, "obj_read: %s: extension is neither .3d nor .3df", path);
    }
    if (format != OLX_OBJ_3D && format != OLX_OBJ_3DF)
        OLX_FAIL(OLX_EINVAL, "obj_read: unknown format %d", format);
    FILE* fp = fopen(path, "rb");
    if (!fp) OLX_FAIL(OLX_EIO, "obj_read: cannot open %s", path);
    int32_t N = 0;
    if (fread(&N, 4, 1, fp) != 1 || N < 0) {
        fclose(fp);
        OLX_FAIL(OLX_EIO, "obj_read: %s: missing or negative point count", path);
    }
    *n = N;
    if (pts) {
        const int64_t m = cap < N ? cap : N;
        for (int64_t j = 0; j < m; j++) {
            double v[3];
            if (format == OLX_OBJ_3D) {
                int32_t q[3];
                if (fread(q, 4, 3, fp) != 3) {
                    fclose(fp);
                    OLX_FAIL(OLX_EIO, "obj_read: %s: short read at point %lld", path, (long long)j);
                }
                for (int c = 0; c < 3; c++) v[c] = (double)q[c];
            } else {
                float q[3];
                if (fread(q, 4, 3, fp) != 3) {
                    fclose(fp);
                    OLX_FAIL(OLX_EIO, "obj_read: %s: short read at point %lld", path, (long long)j);
                }
                for (int c = 0; c < 3; c++) v[c] = (double)q[c];
            }
            pts[j].x = v[0] * scale + ox;
            pts[j].y = v[1] * scale + oy;
            pts[j].z = v[2] * scale + oz;
            pts[j].amp = 1.0f;
            pts[j].phase = 0.0f;
        }
    }
    fclose(fp);
    return OLX_OK;
}

// ------------------------------------------------------------------ bounds
constexpr int RED_T = 256, RED_B = 1024;

// per-block minima and maxima of x, y, z: part[b * 6 + {0,1,2}] = min, {3,4,5} = max
__global__ void __launch_bounds__(RED_T) k_obj_bounds(const olx_point* __restrict__ p, int64_t n,
                                                      double* __restrict__ part) {
    __shared__ double s[6][RED_T];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t j = blockIdx.x * (int64_t)RED_T + threadIdx.x; j < n; j += (int64_t)gridDim.x * RED_T) {
        const double v[3] = {p[j].x, p[j].y, p[j].z};
#pragma unroll
        for (int c = 0; c < 3; c++) {
            mn[c] = fmin(mn[c], v[c]);
            mx[c] = fmax(mx[c], v[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        s[c][threadIdx.x] = mn[c];
        s[3 + c][threadIdx.x] = mx[c];
    }
    __syncthreads();
    for (int w = RED_T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                s[c][threadIdx.x] = fmin(s[c][threadIdx.x], s[c][threadIdx.x + w]);
                s[3 + c][threadIdx.x] = fmax(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// folds the nb per-block partials with one work-group of RED_T threads
__global__ void __launch_bounds__(RED_T) k_obj_bounds_final(const double* __restrict__ part, int nb,
                                                            double* __restrict__ out) {
    __shared__ double s[6][RED_T];
    double r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nb; b += RED_T) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            r[c] = fmin(r[c], part[b * 6 + c]);
            r[3 + c] = fmax(r[3 + c], part[b * 6 + 3 + c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 6; c++) s[c][threadIdx.x] = r[c];
    __syncthreads();
    for (int w = RED_T / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                s[c][threadIdx.x] = fmin(s[c][threadIdx.x], s[c][threadIdx.x + w]);
                s[3 + c][threadIdx.x] = fmax(s[3 + c][threadIdx.x], s[3 + c][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

int device_bounds(olx_ctx* ctx, const olx_point* d, int64_t n, double* mn, double* mx) {
    const int nb = (int)std::min<int64_t>(RED_B, (n + RED_T - 1) / RED_T);
    OLX_TRY(ctx->obj_tmp.reserve(sizeof(double) * (6 * (size_t)nb + 6)));
    double* part = (double*)ctx->obj_tmp.p;
    double* out = part + 6 * nb;
    k_obj_bounds<<<nb, RED_T, 0, ctx->stream>>>(d, n, part);
    k_obj_bounds_final<<<1, RED_T, 0, ctx->stream>>>(part, nb, out);
    OLX_TRY(check_launch("obj_bounds"));
    double h[6];
    OLX_HIP(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    OLX_HIP(hipStreamSynchronize(ctx->stream));
    for (int c = 0; c < 3; c++) {
        mn[c] = h[c];
        mx[c] = h[3 + c];
    }
    return OLX_OK;
}

// ------------------------------------------------------------------ objset
// x' = x scale + o, rounded after the product and after the sum (no fused
// multiply-add), as the reference's arithmetic on Object<double> (T = double) or
// Object<float> (T = float, coordinates rounded to float first)
template <typename T>
__global__ void k_obj_affine(olx_point* __restrict__ p, int64_t n, T scale, T ox, T oy, T oz) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    olx_point q = p[j];
    if constexpr (sizeof(T) == 8) {
        q.x = __dadd_rn(__dmul_rn(q.x, scale), ox);
        q.y = __dadd_rn(__dmul_rn(q.y, scale), oy);
        q.z = __dadd_rn(__dmul_rn(q.z, scale), oz);
    } else {
        q.x = (double)__fadd_rn(__fmul_rn((float)q.x, scale), ox);
        q.y = (double)__fadd_rn(__fmul_rn((float)q.y, scale), oy);
        q.z = (double)__fadd_rn(__fmul_rn((float)q.z, scale), oz);
    }
    p[j] = q;
}

// The scalar part of objset (include/olobject.h:197-221 single range, :233-257 two
// ranges) in the object's element type T.  Host arithmetic, one rounding per
// operation; mn / mx are the exact extremes of the (T-valued) coordinates.
template <typename T>
static void objset_scalars(const double* mn, const double* mx, float x_range, float y_range,
                           bool two, int zmode, T* o, T* scale_out) {
    const T xmin = (T)mn[0], xmax = (T)mx[0], ymin = (T)mn[1], ymax = (T)mx[1];
    const T zmin = (T)mn[2], zmax = (T)mx[2];
    const T rx = xmax - xmin, ry = ymax - ymin;
    T scale;
    if (!two) {
        scale = rx > ry ? rx : ry;
        scale = (T)x_range / scale;
    } else {
        const T sx = (T)x_range / rx, sy = (T)y_range / ry;
        scale = sx < sy ? sx : sy;
    }
    T t = (xmax + xmin) / 2;
    o[0] -= t * scale;
    t = (ymax + ymin) / 2;
    o[1] -= t * scale;
    if (zmode == OLX_OBJSET_ZMIN) {
        o[2] -= zmin * scale;
    } else {
        t = (zmax + zmin) / 2;
        o[2] -= t * scale;
    }
    *scale_out = scale;
}

// ------------------------------------------------------------------ objsort
// Bitonic sort of (key, index) pairs, key = z mapped to an unsigned integer whose
// ascending order is descending z; ties are ordered by index, so the result is the
// stable order.  Padding entries carry the largest key and sort last.
struct SortKey {
    uint64_t k;
    uint32_t i;
    uint32_t pad;
};

__device__ inline uint64_t desc_key(double z) {
    // -0 and +0 compare equal under the reference's z > z comparator: one key
    uint64_t u = __double_as_longlong(z == 0.0 ? 0.0 : z);
    u = (u >> 63) ? ~u : (u | 0x8000000000000000ull);  // ascending-z order
    return ~u;                                           // descending z
}

__device__ inline bool key_less(const SortKey& a, const SortKey& b) {
    return a.k < b.k || (a.k == b.k && a.i < b.i);
}

__global__ void k_sort_init(const olx_point* __restrict__ p, int64_t n, int64_t n2,
                            SortKey* __restrict__ key) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n2) return;
    SortKey s;
    s.k = j < n ? desc_key(p[j].z) : ~0ull;
    s.i = (uint32_t)j;
    s.pad = 0;
    key[j] = s;
}

// Bitonic sort of the (key, index) pairs: a total order, so the result is the
// stable order.
// * SORT_CHUNK = 2048 pairs are sorted in LDS by one 256-thread work-group (a 1 M-point
//   cloud gives 512 of them, several per CU).  Steps j >= 8 go through LDS; steps
//   j = 4, 2, 1 run in registers on 8 consecutive pairs per thread, with LDS rows of
//   8 pairs padded by one so those reads are free of bank conflicts.
// * Distances j >= SORT_CHUNK run as global passes of three steps (j, j/2, j/4) on
//   8-pair groups, or of one or two steps for the remainder.
constexpr int SORT_CHUNK = 2048, SORT_NT = 256;
constexpr int SORT_LDS = SORT_CHUNK + SORT_CHUNK / 8;

template <bool PAD>
__device__ __forceinline__ int sort_lp(int i) { return PAD ? i + (i >> 3) : i; }

__device__ __forceinline__ void sort_cas(SortKey& a, SortKey& b, bool up) {
    if (key_less(b, a) == up) {
        const SortKey t = a;
        a = b;
        b = t;
    }
}

// one compare-exchange step (k: bitonic sequence length, j: partner distance)
__global__ void k_sort_step(SortKey* __restrict__ key, int64_t n2, int64_t k, int64_t j) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n2) return;
    const int64_t l = i ^ j;
    if (l <= i) return;
    const bool up = (i & k) == 0;
    SortKey a = key[i], b = key[l];
    if (key_less(b, a) == up) {
        key[i] = b;
        key[l] = a;
    }
}

// S steps j, j/2, .. in one pass: thread t owns the 2^S pairs whose indices differ
// from b (t with zero bits inserted at j/2^(S-1) .. j) only in those bits; all
// share bit k (k >= 2j), so one direction
template <int S>
__global__ void k_sort_steps(SortKey* __restrict__ key, int64_t n2, int64_t k, int64_t j) {
    constexpr int G = 1 << S;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n2 / G) return;
    const int64_t lo = j >> (S - 1);  // smallest distance of the pass
    const int64_t b = (t & (lo - 1)) | ((t & ~(lo - 1)) << S);
    const bool up = (b & k) == 0;
    SortKey r[G];
#pragma unroll
    for (int e = 0; e < G; e++) r[e] = key[b + e * lo];
#pragma unroll
    for (int d = G / 2; d >= 1; d >>= 1)
#pragma unroll
        for (int e = 0; e < G; e++)
            if ((e & d) == 0) sort_cas(r[e], r[e + d], up);
#pragma unroll
    for (int e = 0; e < G; e++) key[b + e * lo] = r[e];
}

// steps J, J/2, .., 1 on the thread's 8 consecutive pairs (index base + 8 t + e)
template <int J>
__device__ __forceinline__ void sort_reg_steps(SortKey (&r)[8], int64_t first, int64_t k) {
#pragma unroll
    for (int d = J; d >= 1; d >>= 1)
#pragma unroll
        for (int e = 0; e < 8; e++)
            if ((e & d) == 0) sort_cas(r[e], r[e + d], ((first + e) & k) == 0);
}

// steps jmax, jmax/2, .., 1 of sequence length k on one SORT_CHUNK block in LDS
// (REG: the steps j <= 4 in registers, on the padded layout)
template <bool REG>
__device__ __forceinline__ void sort_local_steps(SortKey* s, int64_t base, int64_t k, int jmax) {
    int j = jmax;
    for (; j >= (REG ? 8 : 1); j >>= 1) {
#pragma unroll
        for (int q = 0; q < SORT_CHUNK / 2 / SORT_NT; q++) {
            const int p = threadIdx.x + q * SORT_NT;               // pair index
            const int i = ((p & ~(j - 1)) << 1) | (p & (j - 1));   // bit j clear
            const int ia = sort_lp<REG>(i), ib = sort_lp<REG>(i + j);
            const SortKey a = s[ia], b = s[ib];
            if (key_less(b, a) == (((base + i) & k) == 0)) {
                s[ia] = b;
                s[ib] = a;
            }
        }
        __syncthreads();
    }
    if constexpr (REG) {
        SortKey r[8];
        const int o = 8 * threadIdx.x;
#pragma unroll
        for (int e = 0; e < 8; e++) r[e] = s[sort_lp<true>(o + e)];
        if (j == 4) sort_reg_steps<4>(r, base + o, k);
        else if (j == 2) sort_reg_steps<2>(r, base + o, k);
        else sort_reg_steps<1>(r, base + o, k);
#pragma unroll
        for (int e = 0; e < 8; e++) s[sort_lp<true>(o + e)] = r[e];
        __syncthreads();
    }
}

template <bool PAD>
__device__ __forceinline__ void sort_chunk_load(SortKey* s, const SortKey* key, int64_t base) {
#pragma unroll
    for (int q = 0; q < SORT_CHUNK / SORT_NT; q++) {
        const int i = threadIdx.x + q * SORT_NT;
        s[sort_lp<PAD>(i)] = key[base + i];
    }
    __syncthreads();
}

template <bool PAD>
__device__ __forceinline__ void sort_chunk_store(const SortKey* s, SortKey* key, int64_t base) {
#pragma unroll
    for (int q = 0; q < SORT_CHUNK / SORT_NT; q++) {
        const int i = threadIdx.x + q * SORT_NT;
        key[base + i] = s[sort_lp<PAD>(i)];
    }
}

// every sequence length k <= SORT_CHUNK, on each block
__global__ void __launch_bounds__(SORT_NT) k_sort_chunks(SortKey* __restrict__ key) {
    __shared__ SortKey s[SORT_LDS];
    const int64_t base = blockIdx.x * (int64_t)SORT_CHUNK;
    sort_chunk_load<true>(s, key, base);
    for (int k = 2; k <= SORT_CHUNK; k <<= 1) sort_local_steps<true>(s, base, k, k >> 1);
    sort_chunk_store<true>(s, key, base);
}

// the steps j < SORT_CHUNK of sequence length k > SORT_CHUNK
__global__ void __launch_bounds__(SORT_NT) k_sort_local(SortKey* __restrict__ key, int64_t k) {
    __shared__ SortKey s[SORT_LDS];
    const int64_t base = blockIdx.x * (int64_t)SORT_CHUNK;
    sort_chunk_load<true>(s, key, base);
    sort_local_steps<true>(s, base, k, SORT_CHUNK / 2);
    sort_chunk_store<true>(s, key, base);
}

__global__ void k_sort_gather(const olx_point* __restrict__ src, const SortKey* __restrict__ key,
                              int64_t n, olx_point* __restrict__ dst) {
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    dst[j] = src[key[j].i];
}

}  // namespace olx

using namespace olx;

extern "C" int olx_obj_read(const char* path, int format, olx_point* pts, int64_t cap, int64_t* n,
                            double scale, double ox, double oy, double oz) {
    if (!path || !n) OLX_FAIL(OLX_EINVAL, "obj_read: null argument");
    if (pts && cap < 0) OLX_FAIL(OLX_EINVAL, "obj_read: negative capacity");
    return read_cloud(path, format, pts, cap, n, scale, ox, oy, oz);
}

extern "C" int olx_obj_bounds(olx_ctx* ctx, const olx_point* d_pts, int64_t n, double* mn,
                              double* mx) {
    if (!ctx || !d_pts || !mn || !mx) OLX_FAIL(OLX_EINVAL, "obj_bounds: null argument");
    if (n <= 0) OLX_FAIL(OLX_EINVAL, "obj_bounds: empty cloud");
    return device_bounds(ctx, d_pts, n, mn, mx);
}

extern "C" int olx_obj_set(olx_ctx* ctx, olx_point* d_pts, int64_t n, double x_range,
                           double y_range, double ox, double oy, double oz, int mode) {
    if (!ctx || !d_pts) OLX_FAIL(OLX_EINVAL, "obj_set: null argument");
    if (n <= 0) OLX_FAIL(OLX_EINVAL, "obj_set: empty cloud");
    const bool f32 = (mode & OLX_OBJSET_F32) != 0;
    const int zmode = mode & ~OLX_OBJSET_F32;
    if (zmode != OLX_OBJSET_ZMIN && zmode != OLX_OBJSET_ZMIDDLE)
        OLX_FAIL(OLX_EINVAL, "obj_set: unknown mode %d", mode);
    double mn[3], mx[3];
    OLX_TRY(device_bounds(ctx, d_pts, n, mn, mx));
    const bool two = y_range > 0;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (f32) {
        // Object<float>: the extremes of the float coordinates, float offset
        for (int c = 0; c < 3; c++) {
            mn[c] = (double)(float)mn[c];
            mx[c] = (double)(float)mx[c];
        }
        float o[3] = {(float)ox, (float)oy, (float)oz}, scale;
        objset_scalars<float>(mn, mx, (float)x_range, (float)y_range, two, zmode, o, &scale);
        k_obj_affine<float><<<g, 256, 0, ctx->stream>>>(d_pts, n, scale, o[0], o[1], o[2]);
    } else {
        double o[3] = {ox, oy, oz}, scale;
        objset_scalars<double>(mn, mx, (float)x_range, (float)y_range, two, zmode, o, &scale);
        k_obj_affine<double><<<g, 256, 0, ctx->stream>>>(d_pts, n, scale, o[0], o[1], o[2]);
    }
    return check_launch("obj_set");
}

extern "C" int olx_obj_sort_z(olx_ctx* ctx, olx_point* d_pts, int64_t n) {
    if (!ctx || (n > 0 && !d_pts)) OLX_FAIL(OLX_EINVAL, "obj_sort_z: null argument");
    if (n < 0 || n > (int64_t)UINT32_MAX) OLX_FAIL(OLX_EUNSUPPORTED, "obj_sort_z: bad count");
    if (n < 2) return OLX_OK;
    int64_t n2 = SORT_CHUNK;
    while (n2 < n) n2 <<= 1;
    OLX_TRY(ctx->obj_keys.reserve(sizeof(SortKey) * (size_t)n2 + sizeof(olx_point) * (size_t)n));
    SortKey* key = (SortKey*)ctx->obj_keys.p;
    olx_point* copy = (olx_point*)(key + n2);
    hipStream_t s = ctx->stream;
    OLX_HIP(hipMemcpyAsync(copy, d_pts, sizeof(olx_point) * (size_t)n, hipMemcpyDeviceToDevice, s));
    const unsigned g = (unsigned)((n2 + 255) / 256);
    k_sort_init<<<g, 256, 0, s>>>(copy, n, n2, key);
    const unsigned nchunk = (unsigned)(n2 / SORT_CHUNK), g4 = (unsigned)((n2 / 4 + 255) / 256),
                   g8 = (unsigned)((n2 / 8 + 255) / 256);
    k_sort_chunks<<<nchunk, SORT_NT, 0, s>>>(key);
    for (int64_t k = 2 * SORT_CHUNK; k <= n2; k <<= 1) {
        int64_t j = k / 2;
        for (; j >= 4 * SORT_CHUNK; j >>= 3) k_sort_steps<3><<<g8, 256, 0, s>>>(key, n2, k, j);
        if (j == 2 * SORT_CHUNK) k_sort_steps<2><<<g4, 256, 0, s>>>(key, n2, k, j);
        else if (j == SORT_CHUNK) k_sort_step<<<g, 256, 0, s>>>(key, n2, k, j);
        k_sort_local<<<nchunk, SORT_NT, 0, s>>>(key, k);
    }
    k_sort_gather<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(copy, key, n, d_pts);
    return check_launch("obj_sort_z");
}
