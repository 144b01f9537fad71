// olx_layered.hip -- layered Fresnel CGH as complex GEMMs on MFMA (gfx950).
//
// Points are snapped to the centres z_l of nlayers uniform layers.  For a point
// j in layer l (c_l = 1 / (lambda z_l)) the Fresnel term factorises:
//   exp(i pi ((x_w - x_j)^2 + (y_h - y_j)^2) c_l + i phi_j) A_j
//     = [cy_l(h) e^{-2 pi i h beta_j} a_j] [cx_l(w) e^{-2 pi i w alpha_j}]
// with the layer chirps cx_l(w) = e^{i pi (px w)^2 c_l}, cy_l(h) likewise, the
// cycles per pixel alpha_j = px x_j c_l, beta_j = py y_j c_l, and the point
// factor a_j = A_j e^{i (phi_j + pi (x_j^2 + y_j^2) c_l)}.  Hence
//   U = Y X^T,  Y[h, j] = cy_l(h) e^{-2 pi i h beta_j} a_j,  X[w, j] = cx_l(w) e^{-2 pi i w alpha_j},
// one complex GEMM with K = number of points, each layer's chirp folded into its
// points' factors (so no per-layer epilogue).  The complex product is two real
// MFMA GEMMs over K' = 2K interleaved (re, im) columns:
//   Re U = A Br^T,  Im U = A Bi^T,  A = [Re Y, Im Y],  Br = [Re X, -Im X],  Bi = [Im X, Re X].
// Phases are reduced exactly before quantisation: chirps in FP64 tables, and
// w alpha_j mod 1 as w alpha_hi (exact in FP32, alpha_hi has few fraction bits)
// plus w alpha_lo.  Factors are stored FP16, BF16, or split BF16 (hi + lo, three
// products Yh Xh + Yh Xl + Yl Xh); accumulation is FP32 throughout.
#include <climits>

#include "olx_common.hpp"

namespace olx {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef float f16v_t __attribute__((ext_vector_type(16)));

struct LayPt {       // 32 bytes
    int layer;
    float ahi, alo;  // alpha_j = ahi + alo (cycles per pixel along w), mod 1
    float bhi, blo;  // beta_j along h
    float2 a;        // point factor a_j
    float pad;
};

constexpr int GT = 128;    // output tile side (pixels)
constexpr int KS = 64;     // K step (halves) per LDS stage = 32 points
constexpr int KP = KS + 8; // padded LDS row (halves)
constexpr int KC_PTS = 32768;  // points per factor chunk

__host__ __device__ inline int layer_of(double z, double zmin, double zmax, int nl) {
    double f = floor((z - zmin) / (zmax - zmin) * nl);
    if (!(f >= 0.0)) f = 0.0;
    if (f > nl - 1) f = nl - 1;
    return (int)f;
}

// split v mod 1 into hi (nb fraction bits) + lo
__device__ inline void split_cycles(double v, int nb, float* hi, float* lo) {
    v -= floor(v);
    const double s = ldexp(1.0, nb);
    const double h = floor(v * s) / s;
    *hi = (float)h;
    *lo = (float)(v - h);
}

__global__ void k_lay_prep(const olx_point* __restrict__ pts, int n, LayPt* __restrict__ out,
                           double px, double py, double lambda, double zmin, double zmax, int nl,
                           int nb) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const olx_point p = pts[j];
    const int l = layer_of(p.z, zmin, zmax, nl);
    const double zl = zmin + ((double)l + 0.5) * (zmax - zmin) / nl;
    const double c = 1.0 / (lambda * zl);
    LayPt q;
    q.layer = l;
    split_cycles(px * p.x * c, nb, &q.ahi, &q.alo);
    split_cycles(py * p.y * c, nb, &q.bhi, &q.blo);
    double cyc = 0.5 * (p.x * p.x + p.y * p.y) * c + (double)p.phase * (0.5 / M_PI);
    cyc -= rint(cyc);
    double s, co;
    sincospi(2.0 * cyc, &s, &co);
    q.a = make_float2((float)(p.amp * co), (float)(p.amp * s));
    q.pad = 0.f;
    out[j] = q;
}

// per-layer chirp tables: ch[l * n + i] = exp(i pi (d i)^2 c_l), and the chirp's
// half-step in cycles per pixel^2, cd2[l] = d^2 c_l / 2
__global__ void k_lay_chirp(float2* __restrict__ ch, float* __restrict__ cd2, int n, int nl,
                            double d, double lambda, double zmin, double zmax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x, l = blockIdx.y;
    if (i >= n) return;
    const double zl = zmin + ((double)l + 0.5) * (zmax - zmin) / nl;
    if (i == 0) cd2[l] = (float)(0.5 * d * d / (lambda * zl));
    const double x = d * i;
    double cyc = 0.5 * x * x / (lambda * zl);
    cyc -= rint(cyc);
    double s, c;
    sincospi(2.0 * cyc, &s, &c);
    ch[(size_t)l * n + i] = make_float2((float)c, (float)s);
}

template <typename T>
__device__ inline uint32_t pack2(float a, float b, float* ra, float* rb);
template <>
__device__ inline uint32_t pack2<_Float16>(float a, float b, float* ra, float* rb) {
    const _Float16 x = (_Float16)a, y = (_Float16)b;
    *ra = (float)x;
    *rb = (float)y;
    return (uint32_t)__builtin_bit_cast(uint16_t, x) | ((uint32_t)__builtin_bit_cast(uint16_t, y) << 16);
}
template <>
__device__ inline uint32_t pack2<__bf16>(float a, float b, float* ra, float* rb) {
    const __bf16 x = (__bf16)a, y = (__bf16)b;
    *ra = (float)x;
    *rb = (float)y;
    return (uint32_t)__builtin_bit_cast(uint16_t, x) | ((uint32_t)__builtin_bit_cast(uint16_t, y) << 16);
}

constexpr int GEN_RB = 16;  // rows per thread (one exact anchor, then a recurrence)

// factor rows: for rows r (h for Y, w for X) and chunk points j in [0, kc):
//   Y: A[r][2j] = Re, A[r][2j+1] = Im of cy_l(r) e^{-2 pi i r beta} a
//   X: B[r][2j] = Re, B[r][2j+1] = -Im of cx_l(r) e^{-2 pi i r alpha}
// A thread owns one point and GEN_RB consecutive rows: the phase e^{-2 pi i r alpha}
// is evaluated exactly (reduced) at the first row and advanced by the exact step
// e^{-2 pi i alpha} (rounding drift ~1e-6 over the run, far below FP16/BF16 ulps).
// MODE 0: one factor matrix; MODE 1: hi in out, lo = value - hi in out2 (split bf16)
template <typename T, bool ISY, int MODE>
__global__ void k_lay_gen(const LayPt* __restrict__ pts, int j0, int n, int kc,
                          const float2* __restrict__ chirp, const float* __restrict__ chd2,
                          int rows, uint32_t* __restrict__ out, uint32_t* __restrict__ out2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;  // chunk-local point
    const int rb = blockIdx.y * GEN_RB;
    if (j >= kc) return;
    const bool live = j0 + j < n;
    LayPt p;
    if (live) p = pts[j0 + j];
    else { p.layer = 0; p.ahi = p.alo = p.bhi = p.blo = 0.f; p.a = make_float2(0.f, 0.f); }
    const float hi = ISY ? p.bhi : p.ahi, lo = ISY ? p.blo : p.alo;
    // anchor e^{-2 pi i rb alpha} and step e^{-2 pi i alpha}
    const float fr = (float)rb;
    float t = fr * hi;  // exact: hi has few fraction bits
    t = (t - floorf(t)) + fr * lo;
    t = -(t - rintf(t));
    float2 e = make_float2(__builtin_amdgcn_cosf(t), __builtin_amdgcn_sinf(t));
    float ts = hi + lo;
    ts = -(ts - rintf(ts));
    const float2 stp = make_float2(__builtin_amdgcn_cosf(ts), __builtin_amdgcn_sinf(ts));
    // layer chirp e^{i pi c (d r)^2}: the FP64-built table gives the anchor at rb;
    // the next rows follow from ch(r+1) = ch(r) g(r), g(r) = e^{i pi c d^2 (2r+1)},
    // g(r+1) = g(r) e^{2 pi i c d^2}; their phases are at most a few cycles, exact
    // enough in FP32 (c d^2 = cd2 cycles per pixel^2 over 2)
    float2 ch = chirp[(size_t)p.layer * rows + rb];
    const float cd2 = chd2[p.layer];
    float tg = cd2 * (2.0f * fr + 1.0f);
    tg -= rintf(tg);
    float2 gr = make_float2(__builtin_amdgcn_cosf(tg), __builtin_amdgcn_sinf(tg));
    float tgg = 2.0f * cd2;
    tgg -= rintf(tgg);
    const float2 gg = make_float2(__builtin_amdgcn_cosf(tgg), __builtin_amdgcn_sinf(tgg));
    for (int q = 0; q < GEN_RB; q++) {
        const int r = rb + q;
        if (r >= rows) break;
        float2 v = make_float2(ch.x * e.x - ch.y * e.y, ch.x * e.y + ch.y * e.x);
        if (ISY) v = make_float2(v.x * p.a.x - v.y * p.a.y, v.x * p.a.y + v.y * p.a.x);
        else v.y = -v.y;
        if (!live) v = make_float2(0.f, 0.f);
        const size_t o = (size_t)r * kc + j;  // in 32-bit words (2 halves each)
        float hx, hy;
        out[o] = pack2<T>(v.x, v.y, &hx, &hy);
        if (MODE == 1) {
            float dx_, dy_;
            out2[o] = pack2<T>(v.x - hx, v.y - hy, &dx_, &dy_);
        }
        e = make_float2(e.x * stp.x - e.y * stp.y, e.x * stp.y + e.y * stp.x);
        ch = make_float2(ch.x * gr.x - ch.y * gr.y, ch.x * gr.y + ch.y * gr.x);
        gr = make_float2(gr.x * gg.x - gr.y * gg.y, gr.x * gg.y + gr.y * gg.x);
    }
}

// Bi operand from Br: per dword (lo = Re X, hi = -Im X) -> (lo = Im X, hi = Re X)
__device__ inline uint32_t bi_of_br(uint32_t w) {
    const uint32_t sw = (w >> 16) | (w << 16);  // (lo = -Im X, hi = Re X)
    return sw ^ 0x00008000u;                    // negate the low half
}

template <typename T>
__device__ inline f16v_t mfma(const uint4& a, const uint4& b, f16v_t c);
template <>
__device__ inline f16v_t mfma<_Float16>(const uint4& a, const uint4& b, f16v_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(half8_t, a),
                                                  __builtin_bit_cast(half8_t, b), c, 0, 0, 0);
}
template <>
__device__ inline f16v_t mfma<__bf16>(const uint4& a, const uint4& b, f16v_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8_t, a),
                                                   __builtin_bit_cast(bf8_t, b), c, 0, 0, 0);
}

// U[h][w] (+)= sum_k A[h][k] Br[w][k]  (re),  A[h][k] Bi[w][k]  (im), k < ldk.
// Work-group: 256 threads, 128 x 128 output tile; wave (wr, wc) owns 64 x 64.
// The factor matrices hold whole tiles (rows past ny / nx padded, any content: they
// only reach outputs past the hologram's edge, which are not stored).
template <typename T>
__global__ void __launch_bounds__(256, 2)
k_lay_gemm(const T* __restrict__ A, const T* __restrict__ Br, int ldk, float2* __restrict__ U,
           int ny, int nx, int accumulate) {
    __shared__ __align__(16) T sA[2][GT * KP];
    __shared__ __align__(16) T sB[2][GT * KP];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wr = wave >> 1, wc = wave & 1;
    const int h0 = blockIdx.y * GT, w0 = blockIdx.x * GT;
    // global -> LDS: 128 rows x 64 halves = 1024 x 16 B per matrix; 4 per thread
    const T* gA = A + (size_t)h0 * ldk;
    const T* gB = Br + (size_t)w0 * ldk;
    uint4 ra0, ra1, ra2, ra3, rb0, rb1, rb2, rb3;
    // element (row, seg) of the q-th 16-byte piece of this thread
#define OLX_GLOAD(k0)                                                                         \
    do {                                                                                      \
        const int r0_ = tid >> 3, sg_ = (tid & 7) * 8;                                        \
        ra0 = *(const uint4*)(gA + (size_t)(r0_)*ldk + (k0) + sg_);                           \
        ra1 = *(const uint4*)(gA + (size_t)(r0_ + 32) * ldk + (k0) + sg_);                    \
        ra2 = *(const uint4*)(gA + (size_t)(r0_ + 64) * ldk + (k0) + sg_);                    \
        ra3 = *(const uint4*)(gA + (size_t)(r0_ + 96) * ldk + (k0) + sg_);                    \
        rb0 = *(const uint4*)(gB + (size_t)(r0_)*ldk + (k0) + sg_);                           \
        rb1 = *(const uint4*)(gB + (size_t)(r0_ + 32) * ldk + (k0) + sg_);                    \
        rb2 = *(const uint4*)(gB + (size_t)(r0_ + 64) * ldk + (k0) + sg_);                    \
        rb3 = *(const uint4*)(gB + (size_t)(r0_ + 96) * ldk + (k0) + sg_);                    \
    } while (0)
#define OLX_SSTORE(buf)                                                                       \
    do {                                                                                      \
        const int r0_ = tid >> 3, sg_ = (tid & 7) * 8;                                        \
        *(uint4*)(&sA[buf][(r0_)*KP + sg_]) = ra0;                                            \
        *(uint4*)(&sA[buf][(r0_ + 32) * KP + sg_]) = ra1;                                     \
        *(uint4*)(&sA[buf][(r0_ + 64) * KP + sg_]) = ra2;                                     \
        *(uint4*)(&sA[buf][(r0_ + 96) * KP + sg_]) = ra3;                                     \
        *(uint4*)(&sB[buf][(r0_)*KP + sg_]) = rb0;                                            \
        *(uint4*)(&sB[buf][(r0_ + 32) * KP + sg_]) = rb1;                                     \
        *(uint4*)(&sB[buf][(r0_ + 64) * KP + sg_]) = rb2;                                     \
        *(uint4*)(&sB[buf][(r0_ + 96) * KP + sg_]) = rb3;                                     \
    } while (0)
    f16v_t accr[2][2], acci[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) accr[i][j][v] = acci[i][j][v] = 0.f;

    const int nsteps = ldk / KS;
    OLX_GLOAD(0);
    OLX_SSTORE(0);
    __syncthreads();
    for (int st = 0; st < nsteps; st++) {
        const int buf = st & 1;
        if (st + 1 < nsteps) OLX_GLOAD((st + 1) * KS);
#pragma unroll
        for (int kk = 0; kk < KS / 16; kk++) {
            const int koff = kk * 16 + 8 * (lane >> 5);
            uint4 av[2], bv[2], biv[2];
#pragma unroll
            for (int i = 0; i < 2; i++)
                av[i] = *(const uint4*)(&sA[buf][(wr * 64 + i * 32 + (lane & 31)) * KP + koff]);
#pragma unroll
            for (int j = 0; j < 2; j++) {
                bv[j] = *(const uint4*)(&sB[buf][(wc * 64 + j * 32 + (lane & 31)) * KP + koff]);
                biv[j] = make_uint4(bi_of_br(bv[j].x), bi_of_br(bv[j].y), bi_of_br(bv[j].z),
                                    bi_of_br(bv[j].w));
            }
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) {
                    accr[i][j] = mfma<T>(av[i], bv[j], accr[i][j]);
                    acci[i][j] = mfma<T>(av[i], biv[j], acci[i][j]);
                }
        }
        if (st + 1 < nsteps) OLX_SSTORE(buf ^ 1);
        __syncthreads();
    }
#undef OLX_GLOAD
#undef OLX_SSTORE
    // epilogue: lane holds column w = l % 32 and rows 8 (v/4) + 4 (l/32) + v % 4
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int h = h0 + wr * 64 + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int w = w0 + wc * 64 + j * 32 + (lane & 31);
                if (h >= ny || w >= nx) continue;
                float2* p = &U[(size_t)h * nx + w];
                float2 o = make_float2(accr[i][j][v], acci[i][j][v]);
                if (accumulate) {
                    const float2 u = *p;
                    o.x += u.x;
                    o.y += u.y;
                }
                *p = o;
            }
}

// The same product with a 256 x 128 work-group tile for hologram sides that are
// multiples of 256 (rows) and 128 (columns): 8 waves in a 4 x 2 grid, each owning
// 64 x 64 outputs (8 MFMAs per 16-wide K step), two waves per SIMD.
// Stages go from global memory straight into LDS (global_load_lds_dwordx4: lane l of
// a wave writes 16 bytes at the wave's LDS base + 16 l), so no registers hold them.
// LDS rows are 64 halves (128 B) without padding; the 16-byte chunk c of row r sits
// in slot c ^ ((r >> 1) & 7).  A ds_read_b128 serves 16 lanes (16 rows, one chunk)
// from the 64 banks at a time: rows of one parity take 8 distinct slots in one half
// of a 256-byte bank line and rows of the other parity the other half, so no two
// lanes of a group share a bank.  Three stages of 256 + 128 rows (144 KiB): loads
// run two stages ahead of the MFMAs.  Each K step's fragments are read while the
// previous step's MFMAs run, and its Bi operand is formed between them.  The K order
// of the sums is that of k_lay_gemm, so both kernels give identical results.
// (A 4-wave form with 128 x 64 wave tiles, one wave per SIMD, measured 2% slower.)
constexpr int GT2M = 256, GT2N = 128, GT2_WAVES = 8;
constexpr int ROWB = KS * 2;  // bytes per LDS row

typedef __attribute__((address_space(3))) void* lds_ptr_t;

// Work-group barrier without the release fence of __syncthreads(), which would wait
// for every outstanding global load (the direct loads of the next stage included).
// Callers wait for the loads that must have landed with s_waitcnt first; the asm
// statements keep the compiler from moving memory accesses across the barrier.
__device__ __forceinline__ void stage_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// fragment of K step kk from the swizzled stage (sa, sb): lane holds row (lane % 32),
// chunk 2 kk + lane / 32, found at byte soff = (lane % 32) * ROWB + 16 (chunk ^ ((lane >> 1) & 7))
// (row bases are multiples of 32, so (row >> 1) & 7 = ((lane % 32) >> 1) & 7)
template <int MI>
__device__ __forceinline__ void lay_lfrag(const char* sa, const char* sb, int soff, int wr, int wc,
                                          uint4 (&fa)[MI], uint4 (&fb)[2]) {
#pragma unroll
    for (int i = 0; i < MI; i++) fa[i] = *(const uint4*)(sa + (wr * 32 * MI + i * 32) * ROWB + soff);
#pragma unroll
    for (int j = 0; j < 2; j++) fb[j] = *(const uint4*)(sb + (wc * 64 + j * 32) * ROWB + soff);
}

__device__ __forceinline__ void lay_bi(const uint4 (&fb)[2], uint4 (&bi)[2]) {
#pragma unroll
    for (int j = 0; j < 2; j++)
        bi[j] = make_uint4(bi_of_br(fb[j].x), bi_of_br(fb[j].y), bi_of_br(fb[j].z), bi_of_br(fb[j].w));
}

// The 4 MI MFMAs of one K step from (fa, fb, bi).  With NEXT, the Bi operand of the next step
// (whose fragments were read at the start of this one) is formed during the second
// half, each MFMA followed by an equal share of its 16 VALU operations, so the VALU
// work issues while the matrix unit is busy instead of in front of the next step.
template <typename T, bool NEXT, int MI>
__device__ __forceinline__ void lay_step(const uint4 (&fa)[MI], const uint4 (&fb)[2],
                                         const uint4 (&bi)[2], const uint4 (&fbn)[2],
                                         uint4 (&bin)[2], f16v_t (&accr)[MI][2],
                                         f16v_t (&acci)[MI][2]) {
#pragma unroll
    for (int i = 0; i < MI; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            accr[i][j] = mfma<T>(fa[i], fb[j], accr[i][j]);
            acci[i][j] = mfma<T>(fa[i], bi[j], acci[i][j]);
        }
    if constexpr (NEXT) {
        lay_bi(fbn, bin);
        // half of the MFMAs alone (the next fragments are still arriving), then
        // (MFMA, VALU ...) pairs carrying the 16 VALU operations
        constexpr int H = 2 * MI;
        __builtin_amdgcn_sched_group_barrier(0x008, H, 0);
#pragma unroll
        for (int q = 0; q < H; q++) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
            __builtin_amdgcn_sched_group_barrier(0x002, 16 / H, 0);
        }
    }
}

// direct loads of one stage (K columns k0 .. k0 + 63) into (da, db): the instruction for
// 8-row group g writes bytes g*1024 .. +1023; lane l fills row r = 8g + l/8, slot l%8
// with global chunk (l%8) ^ ((r >> 1) & 7) = (l%8) ^ (4 (g & 1) + l/16) -- g = wave + 4q
// keeps the parity of the wave index (loff is that lane's element offset)
template <typename T, int WAVES>
__device__ __forceinline__ void lay_dload(const T* gA, const T* gB, size_t ldk, size_t k0,
                                          size_t loff, int wave, char* da, char* db) {
#pragma unroll
    for (int q = 0; q < GT2M / 8 / WAVES; q++) {
        const int g = wave + WAVES * q;
        __builtin_amdgcn_global_load_lds((const void*)(gA + (size_t)(8 * g) * ldk + k0 + loff),
                                         (lds_ptr_t)(da + g * 8 * ROWB), 16, 0, 0);
    }
#pragma unroll
    for (int q = 0; q < GT2N / 8 / WAVES; q++) {
        const int g = wave + WAVES * q;
        __builtin_amdgcn_global_load_lds((const void*)(gB + (size_t)(8 * g) * ldk + k0 + loff),
                                         (lds_ptr_t)(db + g * 8 * ROWB), 16, 0, 0);
    }
}

// s_waitcnt immediates (expcnt and, except in the last, lgkmcnt left open): all of this
// wave's loads landed, or all but the newest stage's (loads complete in issue order);
// the last also waits for all of this wave's LDS reads
constexpr int WAIT_VM0 = 0x0F70;
// direct loads per wave and stage: 12 with 4 waves, 6 with 8
template <int WAVES> constexpr int wait_vm_stage() { return 0x0F70 | ((GT2M + GT2N) / 8 / WAVES); }
template <int WAVES> constexpr int wait_vm_stage_lgkm0() { return 0x0070 | ((GT2M + GT2N) / 8 / WAVES); }
static_assert((GT2M + GT2N) / 8 / 4 == 12 && (GT2M + GT2N) / 8 / 8 == 6, "vmcnt fits 4 bits");

// One stage: read (sa, sb), load stage st + 2 into (da, db) -- the buffers of stage
// st - 1, whose reads all completed before that stage's barrier -- and finish with
// stage st + 1 landed in (na, nb) and its first fragments and Bi operand in (fa0, fb0,
// bi0).  The barrier sits before the last step's MFMAs.  The six LDS pointers are restrict:
// after inlining, the reads carry alias scopes that tell them apart from the direct
// loads in flight, so the compiler waits for those loads only where they are read.
// Nothing is conditional (branches would merge the compiler's wait bookkeeping into
// full waits): past the end, the last stage is loaded again into a buffer that is not
// read any more, and the fragments read are not used.
template <typename T, int WAVES, int MI>
__device__ __forceinline__ void lay_stage(
    const T* gA, const T* gB, size_t ldk, size_t loff, int wave, int wr, int wc, const int (&soff)[4],
    int st, int nsteps, const char* __restrict__ sa, const char* __restrict__ sb,
    const char* __restrict__ na, const char* __restrict__ nb, char* __restrict__ da,
    char* __restrict__ db, uint4 (&fa0)[MI], uint4 (&fb0)[2], uint4 (&fa1)[MI], uint4 (&fb1)[2],
    uint4 (&bi0)[2], uint4 (&bi1)[2], f16v_t (&accr)[MI][2], f16v_t (&acci)[MI][2]) {
    lay_dload<T, WAVES>(gA, gB, ldk, (size_t)min(st + 2, nsteps - 1) * KS, loff, wave, da, db);
    // sched_barrier(0) fences: each step's fragment reads for the next step are issued
    // before its MFMAs (8 per wave, 16 per SIMD), which cover their latency
    lay_lfrag(sa, sb, soff[1], wr, wc, fa1, fb1);
    __builtin_amdgcn_sched_barrier(0);
    lay_step<T, true, MI>(fa0, fb0, bi0, fb1, bi1, accr, acci);
    __builtin_amdgcn_sched_barrier(0);
    lay_lfrag(sa, sb, soff[2], wr, wc, fa0, fb0);
    __builtin_amdgcn_sched_barrier(0);
    lay_step<T, true, MI>(fa1, fb1, bi1, fb0, bi0, accr, acci);
    __builtin_amdgcn_sched_barrier(0);
    lay_lfrag(sa, sb, soff[3], wr, wc, fa1, fb1);
    __builtin_amdgcn_sched_barrier(0);
    lay_step<T, true, MI>(fa0, fb0, bi0, fb1, bi1, accr, acci);
    __builtin_amdgcn_sched_barrier(0);
    // The last step's operands are in registers: once this wave's reads of the stage
    // are complete and stage st + 1 has landed, the barrier can come before the last
    // step's MFMAs, which then cover the next stage's first reads (after the barrier
    // the other waves may start loading into this stage's buffers).
    __builtin_amdgcn_s_waitcnt(wait_vm_stage_lgkm0<WAVES>());
    stage_barrier();
    lay_lfrag(na, nb, soff[0], wr, wc, fa0, fb0);
    __builtin_amdgcn_sched_barrier(0);
    lay_step<T, true, MI>(fa1, fb1, bi1, fb0, bi0, accr, acci);
    __builtin_amdgcn_sched_barrier(0);
}

template <typename T, int WAVES>
__global__ void __launch_bounds__(64 * WAVES, 1)
k_lay_gemm2(const T* __restrict__ A, const T* __restrict__ Br, int ldk, float2* __restrict__ U,
            int ny, int nx, int accumulate, int swz) {
    __shared__ __align__(16) char sA0[GT2M * ROWB], sA1[GT2M * ROWB], sA2[GT2M * ROWB];
    __shared__ __align__(16) char sB0[GT2N * ROWB], sB1[GT2N * ROWB], sB2[GT2N * ROWB];
    char *a0 = sA0, *a1 = sA1, *a2 = sA2, *b0 = sB0, *b1 = sB1, *b2 = sB2;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int MI = 16 / WAVES;  // A fragments (32 rows each) per wave
    const int wr = wave >> 1, wc = wave & 1;
    int bx = blockIdx.x, by = blockIdx.y;
    if (swz > 0) {
        // work-groups are dealt to the 8 XCDs in turn: give each XCD a contiguous run of
        // tiles, taken swz tile rows at a time, so that the tiles running together on one
        // XCD share few A rows and B columns in its L2
        const int gx = gridDim.x, G = gx * gridDim.y;
        const int id = blockIdx.y * gx + blockIdx.x;
        const int t = (id % 8) * (G / 8) + id / 8;
        const int grp = t / (swz * gx), in = t % (swz * gx);
        by = grp * swz + in % swz;
        bx = in / swz;
    }
    const int h0 = by * GT2M, w0 = bx * GT2N;
    const T* gA = A + (size_t)h0 * ldk;
    const T* gB = Br + (size_t)w0 * ldk;
    const size_t loff = (size_t)(lane >> 3) * ldk +
                        (size_t)(((lane & 7) ^ (4 * (wave & 1) + (lane >> 4))) * 8);
    static_assert(WAVES % 2 == 0, "8-row groups g = wave + WAVES q keep the wave's parity");
    int soff[KS / 16];
#pragma unroll
    for (int kk = 0; kk < KS / 16; kk++)
        soff[kk] = (lane & 31) * ROWB + (((2 * kk + (lane >> 5)) ^ ((lane >> 1) & 7)) * 16);
    uint4 fa0[MI], fb0[2], fa1[MI], fb1[2], bi0[2], bi1[2];
    f16v_t accr[MI][2], acci[MI][2];
#pragma unroll
    for (int i = 0; i < MI; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) accr[i][j][v] = acci[i][j][v] = 0.f;

    static_assert(KS == 64, "four 16-wide K steps of 8 16-byte chunks per stage");
    const int nsteps = ldk / KS;
    lay_dload<T, WAVES>(gA, gB, (size_t)ldk, 0, loff, wave, a0, b0);
    lay_dload<T, WAVES>(gA, gB, (size_t)ldk, (size_t)min(1, nsteps - 1) * KS, loff, wave, a1, b1);
    __builtin_amdgcn_s_waitcnt(wait_vm_stage<WAVES>());
    stage_barrier();
    lay_lfrag(a0, b0, soff[0], wr, wc, fa0, fb0);
    lay_bi(fb0, bi0);
    int st = 0;
    for (; st + 3 <= nsteps; st += 3) {
        lay_stage<T, WAVES, MI>(gA, gB, ldk, loff, wave, wr, wc, soff, st, nsteps, a0, b0, a1, b1, a2, b2,
                     fa0, fb0, fa1, fb1, bi0, bi1, accr, acci);
        lay_stage<T, WAVES, MI>(gA, gB, ldk, loff, wave, wr, wc, soff, st + 1, nsteps, a1, b1, a2, b2, a0, b0,
                     fa0, fb0, fa1, fb1, bi0, bi1, accr, acci);
        lay_stage<T, WAVES, MI>(gA, gB, ldk, loff, wave, wr, wc, soff, st + 2, nsteps, a2, b2, a0, b0, a1, b1,
                     fa0, fb0, fa1, fb1, bi0, bi1, accr, acci);
    }
    if (st < nsteps)
        lay_stage<T, WAVES, MI>(gA, gB, ldk, loff, wave, wr, wc, soff, st, nsteps, a0, b0, a1, b1, a2, b2,
                     fa0, fb0, fa1, fb1, bi0, bi1, accr, acci);
    if (st + 1 < nsteps)
        lay_stage<T, WAVES, MI>(gA, gB, ldk, loff, wave, wr, wc, soff, st + 1, nsteps, a1, b1, a2, b2, a0, b0,
                     fa0, fb0, fa1, fb1, bi0, bi1, accr, acci);
    // no direct load may still be writing LDS when the work-group ends
    __builtin_amdgcn_s_waitcnt(WAIT_VM0);
    // epilogue: lane holds column w = l % 32 and rows 8 (v/4) + 4 (l/32) + v % 4
#pragma unroll
    for (int i = 0; i < MI; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const int h = h0 + wr * 32 * MI + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const int w = w0 + wc * 64 + j * 32 + (lane & 31);
                if (h >= ny || w >= nx) continue;
                float2* p = &U[(size_t)h * nx + w];
                float2 o = make_float2(accr[i][j][v], acci[i][j][v]);
                if (accumulate) {
                    const float2 u = *p;
                    o.x += u.x;
                    o.y += u.y;
                }
                *p = o;
            }
}

static int nbits_for(int64_t P) {
    int lg = 0;
    while ((1ll << lg) < P) lg++;
    return 24 - lg;
}

}  // namespace olx

using namespace olx;

extern "C" double olx_layer_depth(double z, double zmin, double zmax, int nlayers) {
    const int l = layer_of(z, zmin, zmax, nlayers);
    return zmin + ((double)l + 0.5) * (zmax - zmin) / nlayers;
}

extern "C" int olx_cgh_layered_c64(olx_ctx* ctx, const olx_point* d_pts, int64_t n, void* d_U,
                                   int64_t ny, int64_t nx, double py, double px, double lambda,
                                   int nlayers, double zmin, double zmax, int precision,
                                   uint32_t flags) {
    if (!ctx || !d_U) OLX_FAIL(OLX_EINVAL, "layered: null argument");
    if (n < 0 || (n > 0 && !d_pts)) OLX_FAIL(OLX_EINVAL, "layered: bad point array");
    if (n > INT32_MAX / 2) OLX_FAIL(OLX_EUNSUPPORTED, "layered: too many points");
    if (ny <= 0 || nx <= 0 || ny > 8192 || nx > 8192)
        OLX_FAIL(OLX_EUNSUPPORTED, "layered: hologram sides must be in [1, 8192]");
    if (!(py > 0) || !(px > 0) || !(lambda > 0) || nlayers < 1 || !(zmax > zmin) || !(zmin > 0))
        OLX_FAIL(OLX_EINVAL, "layered: need pitch, lambda > 0, nlayers >= 1, 0 < zmin < zmax");
    if (precision < OLX_LAYERED_F16 || precision > OLX_LAYERED_BF16X3)
        OLX_FAIL(OLX_EINVAL, "layered: unknown precision %d", precision);
    hipStream_t s = ctx->stream;
    const bool acc0 = (flags & OLX_CGH_ACCUMULATE) != 0;
    if (n == 0) {
        if (!acc0) OLX_HIP(hipMemsetAsync(d_U, 0, sizeof(float2) * ny * nx, s));
        return OLX_OK;
    }
    // point records and chirp tables
    const size_t pts_bytes = sizeof(LayPt) * (size_t)n;
    const size_t cx_off = (pts_bytes + 255) & ~(size_t)255;
    const size_t cy_off = cx_off + sizeof(float2) * (size_t)nlayers * nx;
    const size_t dx_off = cy_off + sizeof(float2) * (size_t)nlayers * ny;
    const size_t dy_off = dx_off + sizeof(float) * (size_t)nlayers;
    const size_t meta_bytes = dy_off + sizeof(float) * (size_t)nlayers;
    OLX_TRY(ctx->lay_pts.reserve(meta_bytes));
    char* meta = (char*)ctx->lay_pts.p;
    LayPt* lp = (LayPt*)meta;
    float2* cx = (float2*)(meta + cx_off);
    float2* cy = (float2*)(meta + cy_off);
    float* cdx = (float*)(meta + dx_off);
    float* cdy = (float*)(meta + dy_off);
    const int nb = nbits_for(std::max(nx, ny));
    {
        TimedLaunch tl(ctx, "layered_prep");
        k_lay_prep<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(d_pts, (int)n, lp, px, py, lambda,
                                                               zmin, zmax, nlayers, nb);
        k_lay_chirp<<<dim3((unsigned)((nx + 255) / 256), nlayers), 256, 0, s>>>(
            cx, cdx, (int)nx, nlayers, px, lambda, zmin, zmax);
        k_lay_chirp<<<dim3((unsigned)((ny + 255) / 256), nlayers), 256, 0, s>>>(
            cy, cdy, (int)ny, nlayers, py, lambda, zmin, zmax);
    }
    OLX_TRY(check_launch("layered prep"));
    // factor buffers for one chunk of points (KC multiple of KS/2 points)
    const int kc = (int)std::min<int64_t>(KC_PTS, ((n + KS / 2 - 1) / (KS / 2)) * (KS / 2));
    const int ldk = 2 * kc;
    const int nmat = precision == OLX_LAYERED_BF16X3 ? 4 : 2;
    // factor rows padded to whole work-group tiles: 256 x 128 tiles unless that pads
    // the rows by more than an eighth beyond the 128 x 128 tiles (or those are asked for)
    auto up = [](int64_t v, int64_t m) { return (v + m - 1) / m * m; };
    const bool big = !(flags & OLX_LAYERED_TILE128) && 8 * up(ny, GT2M) <= 9 * up(ny, GT);
    const int64_t nyp = big ? up(ny, GT2M) : up(ny, GT), nxp = up(nx, GT);
    const size_t ybytes = sizeof(uint16_t) * (size_t)nyp * ldk;
    const size_t xbytes = sizeof(uint16_t) * (size_t)nxp * ldk;
    OLX_TRY(ctx->lay_counts.reserve((nmat / 2) * (ybytes + xbytes)));
    char* fb = (char*)ctx->lay_counts.p;
    void* Yh = fb;
    void* Xh = fb + ybytes;
    void* Yl = fb + ybytes + xbytes;
    void* Xl = fb + 2 * ybytes + xbytes;
    const dim3 ggrid((unsigned)(nxp / GT), (unsigned)(nyp / GT));
    const dim3 ggrid2((unsigned)(nxp / GT2N), (unsigned)(nyp / GT2M));

    // XCD tile swizzle for the 256 x 128 tiles: each XCD takes a contiguous run of
    // tiles, 4 tile rows at a time (OLX_LAYSWZ overrides; 0 = launch order).  It needs
    // a tile count divisible by 8 and whole groups of rows.  At 4096^2 (1 M points,
    // 64 layers, f16): 109.3-109.6 ms against 111.1-111.2 ms in launch order (2, 8 rows:
    // 110.5-111.4 ms); the outputs are bitwise the same.
    static const char* swz_env = getenv("OLX_LAYSWZ");
    int swz = swz_env ? atoi(swz_env) : 4;
    if (swz > 0 && ((ggrid2.x * ggrid2.y) % 8 != 0 || (ggrid2.y % swz) != 0)) swz = 0;
    auto gemm = [&](const void* Ya, const void* Xb, int acc) {
        if (precision == OLX_LAYERED_F16) {
            if (big)
                k_lay_gemm2<_Float16, GT2_WAVES><<<ggrid2, 64 * GT2_WAVES, 0, s>>>((const _Float16*)Ya, (const _Float16*)Xb, ldk, (float2*)d_U, (int)ny, (int)nx, acc, swz);
            else
                k_lay_gemm<_Float16><<<ggrid, 256, 0, s>>>((const _Float16*)Ya, (const _Float16*)Xb, ldk, (float2*)d_U, (int)ny, (int)nx, acc);
        } else {
            if (big)
                k_lay_gemm2<__bf16, GT2_WAVES><<<ggrid2, 64 * GT2_WAVES, 0, s>>>((const __bf16*)Ya, (const __bf16*)Xb, ldk, (float2*)d_U, (int)ny, (int)nx, acc, swz);
            else
                k_lay_gemm<__bf16><<<ggrid, 256, 0, s>>>((const __bf16*)Ya, (const __bf16*)Xb, ldk, (float2*)d_U, (int)ny, (int)nx, acc);
        }
    };
    bool first = true;
    for (int64_t j0 = 0; j0 < n; j0 += kc) {
        const dim3 gy((unsigned)((kc + 255) / 256), (unsigned)((ny + GEN_RB - 1) / GEN_RB)),
            gx((unsigned)((kc + 255) / 256), (unsigned)((nx + GEN_RB - 1) / GEN_RB));
        {
            TimedLaunch tl(ctx, "layered_factors");
            if (precision == OLX_LAYERED_F16) {
                k_lay_gen<_Float16, true, 0><<<gy, 256, 0, s>>>(lp, (int)j0, (int)n, kc, cy, cdy, (int)ny, (uint32_t*)Yh, nullptr);
                k_lay_gen<_Float16, false, 0><<<gx, 256, 0, s>>>(lp, (int)j0, (int)n, kc, cx, cdx, (int)nx, (uint32_t*)Xh, nullptr);
            } else if (precision == OLX_LAYERED_BF16) {
                k_lay_gen<__bf16, true, 0><<<gy, 256, 0, s>>>(lp, (int)j0, (int)n, kc, cy, cdy, (int)ny, (uint32_t*)Yh, nullptr);
                k_lay_gen<__bf16, false, 0><<<gx, 256, 0, s>>>(lp, (int)j0, (int)n, kc, cx, cdx, (int)nx, (uint32_t*)Xh, nullptr);
            } else {
                k_lay_gen<__bf16, true, 1><<<gy, 256, 0, s>>>(lp, (int)j0, (int)n, kc, cy, cdy, (int)ny, (uint32_t*)Yh, (uint32_t*)Yl);
                k_lay_gen<__bf16, false, 1><<<gx, 256, 0, s>>>(lp, (int)j0, (int)n, kc, cx, cdx, (int)nx, (uint32_t*)Xh, (uint32_t*)Xl);
            }
        }
        OLX_TRY(check_launch("layered factors"));
        {
            TimedLaunch tl(ctx, "layered_gemm");
            const int acc = (first && !acc0) ? 0 : 1;
            gemm(Yh, Xh, acc);
            if (precision == OLX_LAYERED_BF16X3) {
                gemm(Yh, Xl, 1);
                gemm(Yl, Xh, 1);
            }
        }
        OLX_TRY(check_launch("layered gemm"));
        first = false;
    }
    return OLX_OK;
}
