// olx_wrp.hip -- wavefront-recording-plane accumulation and the WRP method.
//
// WRPStep1 (include/golcgh.h:303-367 circle, :403-468 square) loops every point
// for every pixel.  Here points are binned instead:
//   1. per point: window radius W_j = |z_j - z_wrp| / sqrt(4 (p_lim/lambda)^2 - 1)
//      (square: D_j = sqrt(2) W_j) and the 16 x 16 pixel tiles its window touches;
//   2. counting sort of (tile, point) pairs: count, exclusive scan, fill;
//   3. one work-group per tile: load the tile's point list, sort it by point index
//      (fixed summation order), stage each point's constants (position, window
//      limit, 1/(2 dz lambda), 1/|dz|) in LDS 128 at a time, and accumulate
//      exp(i pi rho^2/(lambda dz))/|dz| for the pixels inside each window (two
//      pixels of one column per thread).  A list longer than the LDS sort holds
//      (dense clouds) is not used: that tile scans the whole cloud in index order
//      instead, keeping the points whose window touches it (the same test that
//      built the lists), so every tile sums its points in ascending index order and
//      the result does not depend on the order the atomic fill left.
// This file is built without FMA contraction: dx*dx + dy*dy rounds each product,
// as the reference's host arithmetic does, so window membership is decided alike.
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

// Exclusive scan of the per-tile counts in three kernels:
//   k_scan_blocks: each 1024-thread block scans SCAN_B = 4096 counts (4 per thread,
//                  loaded contiguously) into block-local offsets and a block total;
//   k_scan_top:    one block scans the block totals (64-bit);
//   k_scan_add:    adds each block's base; offs[ntiles] = total, or -1 beyond 2^31-1.
constexpr int SCAN_T = 1024, SCAN_V = 4, SCAN_B = SCAN_T * SCAN_V;

// exclusive block-wide scan of one 64-bit value per thread (1024 threads)
__device__ __forceinline__ long long block_excl_scan(long long v, long long* part, long long* total) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    long long x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const long long y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) part[wv] = x;
    __syncthreads();
    if (tid < 64) {
        long long t = tid < SCAN_T / 64 ? part[tid] : 0;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const long long y = __shfl_up(t, off, 64);
            if (lane >= off) t += y;
        }
        if (tid < SCAN_T / 64) part[SCAN_T / 64 + tid] = t;  // inclusive wave bases
    }
    __syncthreads();
    const long long wbase = wv ? part[SCAN_T / 64 + wv - 1] : 0;
    *total = part[SCAN_T / 64 + SCAN_T / 64 - 1];
    return wbase + x - v;
}

__global__ void __launch_bounds__(SCAN_T)
k_scan_blocks(const int* __restrict__ counts, int* __restrict__ offs, long long* __restrict__ bsum,
              int ntiles, int* __restrict__ ndense, int* __restrict__ dlist, int small) {
    __shared__ long long part[2 * SCAN_T / 64];
    const int i0 = blockIdx.x * SCAN_B + threadIdx.x * SCAN_V;
    int c[SCAN_V];
#pragma unroll
    for (int k = 0; k < SCAN_V; k++) {
        c[k] = i0 + k < ntiles ? counts[i0 + k] : 0;
        // tiles whose list passes the sparse launch's sort space, for the dense launch
        if (c[k] > small) dlist[atomicAdd(ndense, 1)] = i0 + k;
    }
    long long v = 0;
#pragma unroll
    for (int k = 0; k < SCAN_V; k++) v += c[k];
    long long tot;
    long long run = block_excl_scan(v, part, &tot);
#pragma unroll
    for (int k = 0; k < SCAN_V; k++) {
        if (i0 + k < ntiles) offs[i0 + k] = (int)run;  // block-local; the base is added later
        run += c[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_T)
k_scan_top(long long* __restrict__ bsum, int nblk, int* __restrict__ total_out) {
    __shared__ long long part[2 * SCAN_T / 64];
    __shared__ long long carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nblk; b0 += SCAN_T) {
        const int i = b0 + threadIdx.x;
        const long long v = i < nblk ? bsum[i] : 0;
        long long tot;
        const long long e = block_excl_scan(v, part, &tot);
        const long long c = carry;
        if (i < nblk) bsum[i] = c + e;  // exclusive block base, in place
        __syncthreads();
        if (threadIdx.x == 0) carry = c + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *total_out = carry > INT_MAX ? -1 : (int)carry;
}

__global__ void __launch_bounds__(256)
k_scan_add(int* __restrict__ offs, const long long* __restrict__ bbase, int ntiles) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < ntiles) offs[i] = (int)(offs[i] + bbase[i / SCAN_B]);
}

constexpr int WRP_SORT_CAP = 4096;
// sort space of the launch for sparse tiles (see k_wrp_accum): at 200,000 points into
// 4096^2 (about 27 points per tile) 0.41 ms against 0.53 ms with 4096 for every tile
// (2048: 0.45 ms; 512: 0.41 ms)
constexpr int WRP_SMALL_CAP = 1024;

struct WrpPt {     // per-point constants of one tile's list, staged in LDS
    double x, y;
    double lim;    // circle: W^2 (tmp < W^2); square: W (2 |dx| <= W)
    double rc;     // 1 / (2 dz lambda): phase in cycles = tmp * rc
    float inv;     // 1 / |dz|
    float pad;
};

__device__ __forceinline__ WrpPt wrp_stage(const olx_point& q, double lambda, double plane_z,
                                           double plim, int square) {
    WrpPt s;
    const double Wd = wrp_radius(q.z, plane_z, plim, lambda, square);
    const double dz = q.z - plane_z;
    s.x = q.x;
    s.y = q.y;
    s.lim = square ? Wd : Wd * Wd;
    s.rc = 1.0 / (2.0 * dz * lambda);
    s.inv = (float)(1.0 / fabs(dz));
    s.pad = 0.f;
    return s;
}

// One work-group of WRP_NT = 128 threads per 16 x 16 tile: thread t owns column t % 16
// and rows t / 16 and t / 16 + 8, which share dx and each point read from LDS.
constexpr int WRP_NT = 128;
static_assert(WRP_NT == 128, "the dense-tile compaction counts two waves");

template <bool SQUARE>
__device__ __forceinline__ void wrp_pixel(const WrpPt& P, double dx, double dy, float& re,
                                          float& im) {
    const double tmp = dx * dx + dy * dy;
    const bool inside = SQUARE ? (fabs(dx) * 2 <= P.lim && fabs(dy) * 2 <= P.lim) : (tmp < P.lim);
    if (inside) {
        double cyc = tmp * P.rc;
        cyc -= rint(cyc);
        const float t = (float)cyc;
        re += __builtin_amdgcn_cosf(t) * P.inv;
        im += __builtin_amdgcn_sinf(t) * P.inv;
    }
}

// does the window of point q touch tile (tx, ty)?  The test of k_wrp_count / k_wrp_fill.
__device__ __forceinline__ bool wrp_touches(const olx_point& q, int tx, int ty, int ny, int nx,
                                            double p, double lambda, double plane_z, double plim,
                                            int square) {
    const double R = window_extent(wrp_radius(q.z, plane_z, plim, lambda, square), square);
    int x0, x1, y0, y1;
    tile_span(q.x, R, p, nx, &x0, &x1);
    tile_span(q.y, R, p, ny, &y0, &y1);
    return tx >= x0 && tx <= x1 && ty >= y0 && ty <= y1;
}

// Tiles whose list holds LO < count <= CAP entries are summed from their sorted list;
// with DENSE, tiles beyond CAP scan the whole cloud; other tiles are left to the other
// launch.  Two launches share the tiles: <1024, -1, false> takes the lists of up to 1024
// points with 4 KiB of sort space, so that more work-groups fit per CU (LDS-bound with
// 16 KiB), and <4096, 1024, true> the rest.  Every tile sums in point-index order either
// way, so the split does not change the result.
template <bool SQUARE, int CAP, int LO, bool DENSE>
__device__ __forceinline__ void wrp_tile(int tile, int* ids, WrpPt* sp, int* wcnt,
                                         const olx_point* __restrict__ pts, int n,
                                         const int* __restrict__ offs, const int* __restrict__ list,
                                         float2* __restrict__ U, int ny, int nx, double p,
                                         double lambda, double plane_z, double plim) {
    const int tiles_x = (nx + WT - 1) / WT;
    const int tx = tile % tiles_x, ty = tile / tiles_x;
    const int w = tx * WT + threadIdx.x % WT, h0 = ty * WT + threadIdx.x / WT, h1 = h0 + WT / 2;
    const int beg = offs[tile], cnt = offs[tile + 1] - beg;
    if (cnt <= LO || (!DENSE && cnt > CAP)) return;  // the other launch's tile (uniform)
    const double xw = p * w, yh0 = p * h0, yh1 = p * h1;
    float re0 = 0.f, im0 = 0.f, re1 = 0.f, im1 = 0.f;
    const bool sorted = cnt <= CAP;
    if (sorted) {
        int np2 = 1;
        while (np2 < cnt) np2 <<= 1;
        for (int i = threadIdx.x; i < np2; i += WRP_NT) ids[i] = i < cnt ? list[beg + i] : INT_MAX;
        __syncthreads();
        // bitonic sort, ascending point index: a fixed summation order
        for (int k = 2; k <= np2; k <<= 1)
            for (int jj = k >> 1; jj > 0; jj >>= 1) {
                for (int i = threadIdx.x; i < np2; i += WRP_NT) {
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
    if (!sorted) {
        // dense tile: the whole cloud in index order, 128 points a step, compacted
        // (order kept) to those whose window touches this tile
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        for (int base = 0; base < n; base += WRP_NT) {
            const int j = base + threadIdx.x;
            olx_point q;
            bool take = false;
            if (j < n) {
                q = pts[j];
                take = wrp_touches(q, tx, ty, ny, nx, p, lambda, plane_z, plim, SQUARE);
            }
            const unsigned long long mk = __ballot(take);
            if (lane == 0) wcnt[wv] = __popcll(mk);
            __syncthreads();
            const int before = wv ? wcnt[0] : 0;  // WRP_NT = 128: two waves
            const int m = wcnt[0] + wcnt[1];
            if (take) sp[before + __popcll(mk & ((1ull << lane) - 1ull))] =
                wrp_stage(q, lambda, plane_z, plim, SQUARE);
            __syncthreads();
            for (int k = 0; k < m; k++) {
                const WrpPt P = sp[k];
                const double dx = xw - P.x;
                wrp_pixel<SQUARE>(P, dx, yh0 - P.y, re0, im0);
                wrp_pixel<SQUARE>(P, dx, yh1 - P.y, re1, im1);
            }
            __syncthreads();
        }
    }
    for (int base = 0; sorted && base < cnt; base += WRP_NT) {
        const int m = min(WRP_NT, cnt - base);
        if ((int)threadIdx.x < m) {
            const int j = ids[base + threadIdx.x];
            sp[threadIdx.x] = wrp_stage(pts[j], lambda, plane_z, plim, SQUARE);
        }
        __syncthreads();
        for (int q = 0; q < m; q++) {
            const WrpPt P = sp[q];
            const double dx = xw - P.x;
            wrp_pixel<SQUARE>(P, dx, yh0 - P.y, re0, im0);
            wrp_pixel<SQUARE>(P, dx, yh1 - P.y, re1, im1);
        }
        __syncthreads();
    }
    if (w < nx && h0 < ny) U[(size_t)h0 * nx + w] = make_float2(re0, im0);
    if (w < nx && h1 < ny) U[(size_t)h1 * nx + w] = make_float2(re1, im1);
}

// one work-group per tile, or (LIST) a grid-stride loop over the nt tiles listed in
// dlist (the dense launch, which runs only when there are such tiles)
template <bool SQUARE, int CAP, int LO, bool DENSE, bool LIST>
__global__ void __launch_bounds__(WRP_NT)
k_wrp_accum(const olx_point* __restrict__ pts, int n, const int* __restrict__ offs,
            const int* __restrict__ list, float2* __restrict__ U, int ny, int nx, double p,
            double lambda, double plane_z, double plim, int nt, const int* __restrict__ dlist) {
    __shared__ int ids[CAP];
    __shared__ WrpPt sp[WRP_NT];
    __shared__ int wcnt[WRP_NT / 64];
    if constexpr (LIST) {
        for (int e = blockIdx.x; e < nt; e += gridDim.x) {
            wrp_tile<SQUARE, CAP, LO, DENSE>(dlist[e], ids, sp, wcnt, pts, n, offs, list, U, ny, nx,
                                             p, lambda, plane_z, plim);
            __syncthreads();  // the next tile reuses the LDS arrays
        }
    } else {
        wrp_tile<SQUARE, CAP, LO, DENSE>(blockIdx.x, ids, sp, wcnt, pts, n, offs, list, U, ny, nx, p,
                                         lambda, plane_z, plim);
    }
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
    const int nblk = (ntiles + SCAN_B - 1) / SCAN_B;
    const size_t ints = ((size_t)4 * ntiles + 2 + 1) & ~(size_t)1;  // keep the 64-bit part aligned
    OLX_TRY(ctx->wrp_counts.reserve(sizeof(int) * ints + sizeof(long long) * (size_t)nblk));
    int* counts = (int*)ctx->wrp_counts.p;
    int* offs = counts + ntiles;       // ntiles + 1 (the total last)
    int* ndense = offs + ntiles + 1;   // 1, beside the total: one copy reads both
    int* cursor = ndense + 1;          // ntiles
    int* dlist = cursor + ntiles;      // ntiles
    long long* bsum = (long long*)(counts + ints);  // nblk
    hipStream_t s = ctx->stream;
    OLX_HIP(hipMemsetAsync(counts, 0, sizeof(int) * (size_t)ntiles, s));
    OLX_HIP(hipMemsetAsync(ndense, 0, sizeof(int) * ((size_t)ntiles + 1), s));
    const int nb = (int)((n + 255) / 256);
    {
        TimedLaunch tl(ctx, "wrp_bin");
        if (n > 0)
            k_wrp_count<<<nb, 256, 0, s>>>(d_pts, (int)n, counts, (int)ny, (int)nx, p, lambda,
                                           plane_z, p_limit, square);
        k_scan_blocks<<<nblk, SCAN_T, 0, s>>>(counts, offs, bsum, ntiles, ndense, dlist, WRP_SMALL_CAP);
        k_scan_top<<<1, SCAN_T, 0, s>>>(bsum, nblk, offs + ntiles);
        k_scan_add<<<(ntiles + 255) / 256, 256, 0, s>>>(offs, bsum, ntiles);
    }
    OLX_TRY(check_launch("wrp count/scan"));
    int tot2[2] = {0, 0};  // (tile, point) pairs, dense tiles
    OLX_HIP(hipMemcpyAsync(tot2, offs + ntiles, sizeof tot2, hipMemcpyDeviceToHost, s));
    OLX_HIP(hipStreamSynchronize(s));
    const int total = tot2[0], nd = tot2[1];
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
        auto go = [&](auto SQ) {
            constexpr bool Q = decltype(SQ)::value;
            k_wrp_accum<Q, WRP_SMALL_CAP, -1, false, false><<<ntiles, WRP_NT, 0, s>>>(
                d_pts, (int)n, offs, list, (float2*)d_U, (int)ny, (int)nx, p, lambda, plane_z, p_limit,
                ntiles, nullptr);
            if (nd > 0)
                k_wrp_accum<Q, WRP_SORT_CAP, WRP_SMALL_CAP, true, true><<<std::min(nd, 4096), WRP_NT, 0, s>>>(
                    d_pts, (int)n, offs, list, (float2*)d_U, (int)ny, (int)nx, p, lambda, plane_z,
                    p_limit, nd, dlist);
        };
        if (square) go(std::true_type());
        else go(std::false_type());
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
