// olx_prop.hip -- fused 3-pass angular-spectrum / Fresnel / ARSS propagation.
//
// One propagation of an ny x nx field through the 2x-padded (My x Mx) transform
// domain is three kernels over one ny x Mx intermediate W in HBM:
//   A  rows:    load the nx input values of a row (pre-chirp for ARSS), pad in
//               registers (the zero half is never loaded), forward FFT of
//               length Mx, store the Mx-row of W.
//   B  columns: load the ny stored values of a column, forward FFT of length My,
//               multiply by the transfer function (ASM: computed on the fly with an
//               FP64 phase; Fresnel/ARSS: product of two cached 1-D spectra),
//               inverse FFT, store the centre ny values back into W.
//   C  rows:    load an Mx-row of W, inverse FFT, keep the centre nx values,
//               scale by 1/(My Mx) (post-chirp for ARSS), store.
// The padded rows that are all zero are never materialised, so HBM traffic is
// 8 + 16 + 16 + 16 + 16 + 8 = 80 ny nx bytes per propagation.
#include <climits>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "fft_core.hpp"
#include "olx_common.hpp"
#include "prop_common.hpp"

namespace olx {


// index of (row h, padded column m) in the column-blocked intermediate
__device__ __forceinline__ size_t widx(const PropArgs& a, int h, int m) {
    const int g = m >> a.cbl_log, c = m & (a.cbl - 1);
    return ((size_t)g * a.rows + h) * a.cbl + c;
}

template <int L>
constexpr int row_rpw() { return Plan<L>::T >= 256 ? 1 : 256 / Plan<L>::T; }
template <int L>
constexpr int row_nt() { return row_rpw<L>() * Plan<L>::T; }
template <int L>
constexpr int col_cb() { return L >= 16384 ? 1 : (L >= 1024 ? 16384 / L : 16); }
template <int L>
constexpr int col_nt() { return col_cb<L>() * Plan<L>::T; }

// ------------------------------------------------------------------ pass A

template <int L>
__global__ void __launch_bounds__(row_nt<L>())
k_rows_pad_fwd(const float2* __restrict__ src, float2* __restrict__ W, PropArgs a,
               const float2* __restrict__ tw) {
    using P = Plan<L>;
    constexpr int RPW = row_rpw<L>();
    constexpr int R0 = P::radix(0), RL = P::radix(P::NST - 1);
    __shared__ float2 lds[RPW * P::LP];
    const int tid = threadIdx.x, rr = tid / P::T, t = tid - rr * P::T;
    const int row = xcd_remap(blockIdx.x, gridDim.x) * RPW + rr;
    const bool valid = row < a.rows;
    const int nx = a.nx;
    float2 ay = make_float2(1.f, 0.f);
    if (a.ay && valid) ay = a.ay[row];
    float2 v[P::E];
#pragma unroll
    for (int b = 0; b < P::E / R0; b++)
#pragma unroll
        for (int r = 0; r < R0; r++) {
            float2 x = make_float2(0.f, 0.f);
            if (first_nonzero<P>(b, r) && valid) {
                const int col = first_index<P>(t, b, r) - L / 4;
                x = src[(size_t)row * a.ss + col];
                if (a.ax) x = cmul(x, cmul(a.ax[col], ay));
                if (a.pre) x = cmul(x, a.pre[(size_t)(row % a.ny) * nx + col]);
            }
            v[b * R0 + r] = x;
        }
    run_stages<P, false, TWN, 1>(v, t, lds, rr * P::LP, tw);
    if (valid) {
#pragma unroll
        for (int b = 0; b < P::E / RL; b++)
#pragma unroll
            for (int r = 0; r < RL; r++) W[widx(a, row, last_index<P>(t, b, r))] = v[b * RL + r];
    }
}


// ------------------------------------------------------------------ pass B
template <int L, int CB>
constexpr int col_waves_per_eu() {
    // two work-groups per CU when one is at most 512 threads and its LDS fits twice
    return (CB * Plan<L>::T <= 512 && 2 * CB * Plan<L>::LP * 8 <= 160 * 1024) ? 2 * CB * Plan<L>::T / 256 : 1;
}

template <int L, int CB, int MODE>
__global__ void __launch_bounds__(CB * Plan<L>::T, (col_waves_per_eu<L, CB>()))
k_cols_conv(float2* __restrict__ W, PropArgs a, const float2* __restrict__ tw) {
    using P = Plan<L>;
    constexpr int R0 = P::radix(0), RL = P::radix(P::NST - 1);
    __shared__ float2 lds[CB * P::LP];
    const int tid = threadIdx.x, c = tid % CB, t = tid / CB;
    const int Mx = 2 * a.nx;
    const int grp = xcd_remap(blockIdx.x, gridDim.x);
    const int col = grp * CB + c;
    const bool valid = col < Mx;
    // blocked layout: this group's contiguous block; row-major: stride Mx
    const bool rowmajor = a.cbl == Mx;
    const size_t fo = (size_t)blockIdx.y * a.ny;  // this field's first row
    float2* Wb = rowmajor ? W + fo * Mx + col : W + ((size_t)grp * a.rows + fo) * CB + c;
    const size_t wst = rowmajor ? (size_t)Mx : (size_t)CB;
    float2 v[P::E];
#pragma unroll
    for (int b = 0; b < P::E / R0; b++)
#pragma unroll
        for (int r = 0; r < R0; r++) {
            float2 x = make_float2(0.f, 0.f);
            if (first_nonzero<P>(b, r) && valid)
                x = Wb[(size_t)(first_index<P>(t, b, r) - L / 4) * wst];
            v[b * R0 + r] = x;
        }
    run_stages<P, false, TWN, CB>(v, t, lds, c, tw);
#pragma unroll
    for (int b = 0; b < P::E / RL; b++)
#pragma unroll
        for (int r = 0; r < RL; r++) lds[c + lpad(last_index<P>(t, b, r)) * CB] = v[b * RL + r];
    __syncthreads();
    // transfer function, then inverse transform
    const ColH ch = col_h<MODE>(a, valid ? col : 0, Mx);
    // multiply the spectrum in LDS; a short loop keeps few FP64 temporaries live
#pragma unroll 2
    for (int q = 0; q < P::E; q++) {
        const int i = t + P::T * q;
        float2* p = &lds[c + lpad(i) * CB];
        *p = cmul(*p, trans_h<MODE>(a, ch, i, L));
    }
    __syncthreads();
    stage_load_lds<P, 0, CB>(v, t, lds, c);
    __syncthreads();
    run_stages<P, true, TWN, CB>(v, t, lds, c, tw);
#pragma unroll
    for (int b = 0; b < P::E / RL; b++)
#pragma unroll
        for (int r = 0; r < RL; r++) {
            if (!last_kept<P>(b, r)) continue;
            const int i = last_index<P>(t, b, r);
            if (valid && i >= L / 4 && i < 3 * L / 4)
                Wb[(size_t)(i - L / 4) * wst] = v[b * RL + r];
        }
}

// ------------------------------------------------------------------ pass C
template <int L>
__global__ void __launch_bounds__(row_nt<L>())
k_rows_crop_inv(const float2* __restrict__ W, float2* __restrict__ dst, PropArgs a,
                const float2* __restrict__ tw) {
    using P = Plan<L>;
    constexpr int RPW = row_rpw<L>();
    constexpr int R0 = P::radix(0), RL = P::radix(P::NST - 1);
    __shared__ float2 lds[RPW * P::LP];
    const int tid = threadIdx.x, rr = tid / P::T, t = tid - rr * P::T;
    const int row = xcd_remap(blockIdx.x, gridDim.x) * RPW + rr;
    const bool valid = row < a.rows;
    const int nx = a.nx;
    float2 v[P::E];
#pragma unroll
    for (int b = 0; b < P::E / R0; b++)
#pragma unroll
        for (int r = 0; r < R0; r++)
            v[b * R0 + r] = valid ? W[widx(a, row, first_index<P>(t, b, r))] : make_float2(0.f, 0.f);
    run_stages<P, true, TWN, 1>(v, t, lds, rr * P::LP, tw);
    if (!valid) return;
    const bool post = a.cx != nullptr && abs(row - a.ny / 2) < a.chl;
    const float2 cy = post ? a.cy[row] : make_float2(1.f, 0.f);
#pragma unroll
    for (int b = 0; b < P::E / RL; b++)
#pragma unroll
        for (int r = 0; r < RL; r++) {
            if (!last_kept<P>(b, r)) continue;
            const int i = last_index<P>(t, b, r);
            if (i < L / 4 || i >= 3 * L / 4) continue;
            const int col = i - L / 4;
            float2 x = cscale(v[b * RL + r], a.scale);
            if (post && abs(col - nx / 2) < a.cwl) x = cmul(x, cmul(a.cx[col], cy));
            if (a.post) x = cmul(x, a.post[(size_t)(row % a.ny) * nx + col]);
            float2* o = &dst[(size_t)row * a.ds + col];
            if (a.accum) x = make_float2(o->x + x.x, o->y + x.y);
            *o = x;
            if (a.img8) store_img8(a, (size_t)row * a.ds + col, x);
        }
}

// ------------------------------------------------------------------ tiled passes
// For padded lengths >= 2048 the three passes use a tiled intermediate.  A
// work-group carries 2 KR transforms in KR T threads (T = L/16): thread (c, t),
// c = tid % KR, holds transform c in value set a and transform KR + c in set b,
// and LDS holds the KR transforms of one set at a time.
//   W layout: column tiles of CBL columns (a power of two >= the column pass's
//   2 KR), each tile a contiguous ny x CBL block (widx with cbl = CBL).
// Pass B reads and writes whole tile columns; passes A and C handle 2 KR
// consecutive rows, so each tile receives a 2 KR x CBL block per work-group.  The
// HBM byte count is the 80 ny nx of the row-major form; what changes is that the
// column pass no longer reads one 8-byte value per 128-byte line.

// Global accesses of the tiled passes are raw buffer operations: a per-thread
// 32-bit byte offset in one VGPR, the per-value part (uniform) in an SGPR.  The
// descriptor carries the buffer's size, so an access past the end reads zero and a
// store past the end is dropped instead of touching other memory.
struct Rsrc {
    __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ Rsrc make_rsrc(const void* p, size_t bytes) {
    // raw buffer, stride 0, gfx9 data format word
    return Rsrc{__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000)};
}
__device__ __forceinline__ float2 bload(Rsrc b, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(b.r, voff, soff, 0));
}
__device__ __forceinline__ void bstore(Rsrc b, uint32_t voff, uint32_t soff, float2 v) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(b.r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, v), b.r, voff, soff, 0);
}
__device__ __forceinline__ float bloadf(Rsrc b, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, voff, soff, 0));
}
__device__ __forceinline__ void bstoref(Rsrc b, uint32_t voff, uint32_t soff, float v) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b32(b.r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(V, v), b.r, voff, soff, 0);
}
// two adjacent complex values in one 16-byte access
__device__ __forceinline__ void bload2(Rsrc b, uint32_t voff, uint32_t soff, float2& x, float2& y) {
    const float4 q = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(b.r, voff, soff, 0));
    x = make_float2(q.x, q.y);
    y = make_float2(q.z, q.w);
}
__device__ __forceinline__ void bstore2(Rsrc b, uint32_t voff, uint32_t soff, float2 x, float2 y) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b128(b.r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, make_float4(x.x, x.y, y.x, y.y)),
                                           b.r, voff, soff, 0);
}

constexpr int PAIR_MIN = 2048;  // shortest padded length on the tiled path
// largest KR: 1024 threads and KR padded transforms of LDS (136 KiB)
template <int L>
constexpr int kr_max() { return 16384 / L; }
// the work-group shapes instantiated: KR = kr_max / 2^d, d < 4, at least 1
template <int L, int D>
constexpr int kr_div() { return kr_max<L>() >> D >= 1 ? kr_max<L>() >> D : 1; }

// calls f(std::integral_constant<int, KR>) with the largest instantiated KR <= kr
template <int L, class F>
static void with_kr(int kr, F&& f) {
    if (kr >= kr_div<L, 0>()) f(std::integral_constant<int, kr_div<L, 0>()>());
    else if (kr >= kr_div<L, 1>()) f(std::integral_constant<int, kr_div<L, 1>()>());
    else if (kr >= kr_div<L, 2>()) f(std::integral_constant<int, kr_div<L, 2>()>());
    else f(std::integral_constant<int, kr_div<L, 3>()>());
}
// waves per SIMD the register budget must allow: two work-groups per CU when a
// work-group is at most 512 threads and its LDS fits twice
template <int L, int KR>
constexpr int kr_waves() {
    return (KR * Plan<L>::T <= 512 && 2 * KR * Plan<L>::LP * 8 <= 160 * 1024) ? 2 * KR * Plan<L>::T / 256 : 1;
}


// LG: the loss-gradient prologue (PropArgs::lg_*) on the source values
template <int L, int KR, int ORD, bool LG = false>
__global__ void __launch_bounds__(KR * Plan<L>::T, (kr_waves<L, KR>()))
k_rows_pad_fwd2(const float2* __restrict__ src, float2* __restrict__ W, PropArgs a,
                const float2* __restrict__ tw) {
    using P = Plan<L, ORD>;
    constexpr int R0 = P::radix(0), RL = P::radix(P::NST - 1);
    static_assert(P::T % 16 == 0, "pair plan");
    __shared__ float2 lds[KR * P::LP];
    const int tid = threadIdx.x, c = tid % KR, t = tid / KR;
    const int h0 = xcd_remap(blockIdx.x, gridDim.x) * 2 * KR;
    const int ha = h0 + c, hb = h0 + KR + c;
    const int nx = a.nx;
    const Rsrc sb = make_rsrc(src, ((size_t)(a.rows - 1) * a.ss + nx) * sizeof(float2));
    // the factor field: ny x nx, dense, shared by every field of a batch
    const Rsrc pb = make_rsrc(a.pre, a.pre ? (size_t)a.ny * nx * sizeof(float2) : 0);
    const Rsrc wb = make_rsrc(W, (size_t)a.rows * 2 * nx * sizeof(float2));
    const uint32_t rowb = (uint32_t)(KR * a.ss * sizeof(float2));  // set b rows past set a's
    const uint32_t so = (uint32_t)(((size_t)ha * a.ss + t) * sizeof(float2));
    const uint32_t fo = (uint32_t)(((size_t)(ha % a.ny) * nx + t) * sizeof(float2));
    const uint32_t frowb = (uint32_t)(KR * nx * sizeof(float2));
    float2 aya = make_float2(1.f, 0.f), ayb = aya;
    if (a.ay) {
        aya = a.ay[ha];
        ayb = a.ay[hb];
    }
    Rsrc tb;
    double mi = 0.0, mt = 0.0, c1 = 0.0, c2 = 0.0;
    if constexpr (LG) {
        tb = make_rsrc(a.lg_t, ((size_t)(a.rows - 1) * a.ss + nx) * sizeof(float));
        mi = a.lg_c[0];
        mt = a.lg_c[1];
        c1 = a.lg_c[2];
        c2 = a.lg_c[3];
    }
    float2 va[P::E], vb[P::E];
#pragma unroll
    for (int b = 0; b < P::E / R0; b++)
#pragma unroll
        for (int r = 0; r < R0; r++) {
            float2 xa = make_float2(0.f, 0.f), xb = xa;
            if (first_nonzero<P>(b, r)) {
                // column t + D of rows ha / hb
                const int D = first_index<P>(0, b, r) - L / 4;
                xa = bload(sb, so, D * sizeof(float2));
                xb = bload(sb, so, D * sizeof(float2) + rowb);
                if constexpr (LG) {  // float offsets are half the complex ones
                    xa = loss_grad(xa, bloadf(tb, so / 2, D * sizeof(float)), mi, mt, c1, c2);
                    xb = loss_grad(xb, bloadf(tb, so / 2, D * sizeof(float) + rowb / 2), mi, mt, c1, c2);
                }
                if (a.ax) {
                    const float2 ax = a.ax[t + D];
                    xa = cmul(xa, cmul(ax, aya));
                    xb = cmul(xb, cmul(ax, ayb));
                }
                if (a.pre) {
                    xa = cmul(xa, bload(pb, fo, D * sizeof(float2)));
                    xb = cmul(xb, bload(pb, fo, D * sizeof(float2) + frowb));
                }
            }
            va[b * R0 + r] = xa;
            vb[b * R0 + r] = xb;
        }
    run_stages_pipe<P, false, TWN, KR>(va, vb, t, lds, c, tw);
    // element t + D of row h sits at widx(h, t) + D ny (D a multiple of the tile width)
    const uint32_t wo = (uint32_t)(widx(a, ha, t) * sizeof(float2));
    const uint32_t setb = (uint32_t)(KR * a.cbl * sizeof(float2));
#pragma unroll
    for (int b = 0; b < P::E / RL; b++)
#pragma unroll
        for (int r = 0; r < RL; r++) {
            const uint32_t so2 = (uint32_t)(last_index<P>(0, b, r) * a.rows * sizeof(float2));
            bstore(wb, wo, so2, va[b * RL + r]);
            bstore(wb, wo, so2 + setb, vb[b * RL + r]);
        }
}

template <int L, int KR, int MODE, int ORDF, int ORDI>
__global__ void __launch_bounds__(KR * Plan<L>::T, (kr_waves<L, KR>()))
k_cols_conv2(float2* __restrict__ W, PropArgs a, const float2* __restrict__ tw) {
    using PF = Plan<L, ORDF>;  // forward transform
    using PI = Plan<L, ORDI>;  // inverse transform
    constexpr int R0 = PF::radix(0), RLF = PF::radix(PF::NST - 1);
    constexpr int R0I = PI::radix(0), RLI = PI::radix(PI::NST - 1);
    constexpr int E = PF::E, T = PF::T;
    static_assert(T % 16 == 0, "pair plan");
    __shared__ __align__(16) float2 lds[KR * PF::LP];
    const int tid = threadIdx.x, c = tid % KR, t = tid / KR;
    const int cbl = a.cbl;
    // sets a and b are two neighbouring columns: one 16-byte access per row serves both
    const int cola = xcd_remap(blockIdx.x, gridDim.x) * 2 * KR + 2 * c, colb = cola + 1;
    const Rsrc wr = make_rsrc(W, (size_t)a.rows * 2 * a.nx * sizeof(float2));
    // tile base plus this field's first row (blockIdx.y); column within the tile
    const uint32_t gb = (uint32_t)(((size_t)(cola >> a.cbl_log) * a.rows + (size_t)blockIdx.y * a.ny) *
                                   cbl * sizeof(float2));
    const uint32_t wo = (uint32_t)((t * cbl + (cola & (cbl - 1))) * sizeof(float2));
    const uint32_t rowst = (uint32_t)(cbl * sizeof(float2));
    float2 va[E], vb[E];
#pragma unroll
    for (int b = 0; b < E / R0; b++)
#pragma unroll
        for (int r = 0; r < R0; r++) {
            float2 xa = make_float2(0.f, 0.f), xb = xa;
            if (first_nonzero<PF>(b, r))
                bload2(wr, wo, gb + (first_index<PF>(0, b, r) - L / 4) * rowst, xa, xb);
            va[b * R0 + r] = xa;
            vb[b * R0 + r] = xb;
        }
    run_stages_pipe<PF, false, TWN, KR>(va, vb, t, lds, c, tw);
    // transfer function on the spectrum held in registers
    if constexpr (MODE == K_TABLE) {
        const float2 sxa = a.sx[cola], sxb = a.sx[colb];
        const Rsrc syb = make_rsrc(a.sy, L * sizeof(float2));
        const uint32_t to = (uint32_t)(t * sizeof(float2));
#pragma unroll
        for (int b = 0; b < E / RLF; b++)
#pragma unroll
            for (int r = 0; r < RLF; r++) {
                const float2 sy = bload(syb, to, last_index<PF>(0, b, r) * sizeof(float2));
                float2 ha = cmul(sxa, sy), hb = cmul(sxb, sy);
                if (a.adjoint) {
                    ha = cconj(ha);
                    hb = cconj(hb);
                }
                va[b * RLF + r] = cmul(va[b * RLF + r], ha);
                vb[b * RLF + r] = cmul(vb[b * RLF + r], hb);
            }
    } else {
        constexpr bool SH = MODE == K_ASHIFT;
        constexpr bool F32 = MODE == K_ASM32;
        const int Mx = 2 * a.nx;
        const int fxa = cola < Mx / 2 ? cola : cola - Mx, fxb = colb < Mx / 2 ? colb : colb - Mx;
        const double ua = (double)fxa * a.du, ub = (double)fxb * a.du;
        const float ufa = (float)fxa * a.duf, ufb = (float)fxb * a.duf;
        const double Aa = a.inv_l2 - ua * ua, Ab = a.inv_l2 - ub * ub;
        const bool cuta = a.bandlimit && !(fabs(SH ? ua - a.u0 : ua) * 2 < a.uw);
        const bool cutb = a.bandlimit && !(fabs(SH ? ub - a.u0 : ub) * 2 < a.uw);
        const double pca = SH ? a.ox * ua : 0.0, pcb = SH ? a.ox * ub : 0.0;
        // K_ASM: H depends on v^2 and |v| only, so row i and its mirror L - i share
        // one value.  The thread's rows below L/2 (slots q < E/2 of i = t + T q) are
        // evaluated and published in LDS (free here: the forward transform ended with
        // a barrier after its last LDS reads), and the rows past L/2 read their
        // mirror's value.  Row L/2 is its own mirror: every thread evaluates it and
        // stores it at entry t + (t == 0 ? L/2 : 0), computed without a select (a
        // branch here makes the register allocator spill); for t > 0 that entry is
        // then overwritten, in program order, with row t's value.  Entry (row, c)
        // holds sets a and b.  The shifted form's window and phase are not
        // symmetric: there every row is evaluated.
        float4* tab = reinterpret_cast<float4*>(lds);
        static_assert((L / 2 + 1) * 2 <= PF::LP, "mirror table fits in LDS");
        if constexpr (!SH) {
            const float2 hma = F32 ? asm_h_f32(a, ufa, cuta, L / 2, L) : asm_h(a, Aa, cuta, L / 2, L);
            __builtin_amdgcn_sched_barrier(0);
            const float2 hmb = F32 ? asm_h_f32(a, ufb, cutb, L / 2, L) : asm_h(a, Ab, cutb, L / 2, L);
            const int e0 = t + (((t - 1) >> 31) & (L / 2));
            tab[e0 * KR + c] = make_float4(hma.x, hma.y, hmb.x, hmb.y);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int b = 0; b < E / RLF; b++)
#pragma unroll
            for (int r = 0; r < RLF; r++) {
                const int i = last_index<PF>(t, b, r);
                if (!SH && b + r * (E / RLF) >= E / 2) continue;
                const float2 ha = F32 ? asm_h_f32(a, ufa, cuta, i, L) : asm_h<SH>(a, Aa, cuta, i, L, pca);
                const float2 hb = F32 ? asm_h_f32(a, ufb, cutb, i, L) : asm_h<SH>(a, Ab, cutb, i, L, pcb);
                va[b * RLF + r] = cmul(va[b * RLF + r], ha);
                vb[b * RLF + r] = cmul(vb[b * RLF + r], hb);
                if (!SH) tab[i * KR + c] = make_float4(ha.x, ha.y, hb.x, hb.y);
                // keep the FP64 temporaries of one element pair live at a time
                __builtin_amdgcn_sched_barrier(0);
            }
        if constexpr (!SH) {
            __syncthreads();
#pragma unroll
            for (int b = 0; b < E / RLF; b++)
#pragma unroll
                for (int r = 0; r < RLF; r++) {
                    if (b + r * (E / RLF) < E / 2) continue;
                    const float4 h = tab[(L - last_index<PF>(t, b, r)) * KR + c];
                    va[b * RLF + r] = cmul(va[b * RLF + r], make_float2(h.x, h.y));
                    vb[b * RLF + r] = cmul(vb[b * RLF + r], make_float2(h.z, h.w));
                }
            // the inverse transform's first exchange overwrites the table
            __syncthreads();
        }
    }
    float2 wa[E], wb[E];
#pragma unroll
    for (int b = 0; b < E / R0I; b++)
#pragma unroll
        for (int r = 0; r < R0I; r++) {
            wa[b * R0I + r] = va[spec_slot<PF, PI>(b, r)];
            wb[b * R0I + r] = vb[spec_slot<PF, PI>(b, r)];
        }
    // (the forward transform ended with a barrier after its last LDS reads)
    run_stages_pipe<PI, true, TWN, KR>(wa, wb, t, lds, c, tw);
#pragma unroll
    for (int b = 0; b < E / RLI; b++)
#pragma unroll
        for (int r = 0; r < RLI; r++) {
            if (!last_kept<PI>(b, r)) continue;
            bstore2(wr, wo, gb + (last_index<PI>(0, b, r) - L / 4) * rowst, wa[b * RLI + r],
                    wb[b * RLI + r]);
        }
}

// ACC: add into the destination (split propagation sums the blocks' contributions)
__device__ __forceinline__ void bstore_acc(Rsrc b, uint32_t voff, uint32_t soff, float2 v, bool acc) {
    if (acc) {
        const float2 o = bload(b, voff, soff);
        v = make_float2(o.x + v.x, o.y + v.y);
    }
    bstore(b, voff, soff, v);
}

// PG: the phase gradient (PropArgs::pg_*) in place of the output field
// LS: the loss sums (PropArgs::ls_*) beside it
template <int L, int KR, int ORD, bool ACC = false, bool PG = false, bool LS = false>
__global__ void __launch_bounds__(KR * Plan<L>::T, (kr_waves<L, KR>()))
k_rows_crop_inv2(const float2* __restrict__ W, float2* __restrict__ dst, PropArgs a,
                 const float2* __restrict__ tw) {
    using P = Plan<L, ORD>;
    constexpr int R0 = P::radix(0), RL = P::radix(P::NST - 1);
    static_assert(P::T % 16 == 0, "pair plan");
    __shared__ float2 lds[KR * P::LP];
    const int tid = threadIdx.x, c = tid % KR, t = tid / KR;
    const int h0 = xcd_remap(blockIdx.x, gridDim.x) * 2 * KR;
    const int ha = h0 + c, hb = h0 + KR + c;
    const int nx = a.nx;
    // element t + D of row h sits at widx(h, t) + D ny (D a multiple of the tile width)
    const Rsrc wb = make_rsrc(W, (size_t)a.rows * 2 * nx * sizeof(float2));
    const Rsrc db = make_rsrc(dst, ((size_t)(a.rows - 1) * a.ds + nx) * sizeof(float2));
    const Rsrc qb = make_rsrc(a.post, a.post ? (size_t)a.ny * nx * sizeof(float2) : 0);
    const uint32_t wo = (uint32_t)(widx(a, ha, t) * sizeof(float2));
    const uint32_t setb = (uint32_t)(KR * a.cbl * sizeof(float2));
    float2 va[P::E], vb[P::E];
#pragma unroll
    for (int b = 0; b < P::E / R0; b++)
#pragma unroll
        for (int r = 0; r < R0; r++) {
            const uint32_t so = (uint32_t)(first_index<P>(0, b, r) * a.rows * sizeof(float2));
            va[b * R0 + r] = bload(wb, wo, so);
            vb[b * R0 + r] = bload(wb, wo, so + setb);
        }
    run_stages_pair<P, true, TWN, KR>(va, vb, t, lds, c, tw);
    const uint32_t so = (uint32_t)(((size_t)ha * a.ds + t) * sizeof(float2));
    const uint32_t rowb = (uint32_t)(KR * a.ds * sizeof(float2));
    const uint32_t fo = (uint32_t)(((size_t)(ha % a.ny) * nx + t) * sizeof(float2));
    const uint32_t frowb = (uint32_t)(KR * nx * sizeof(float2));
    if constexpr (PG) {
        // float offsets are half the complex ones
        const Rsrc ub = make_rsrc(a.pg_u, ((size_t)(a.rows - 1) * a.ds + nx) * sizeof(float2));
        const Rsrc gb = make_rsrc(a.pg_grad, ((size_t)(a.rows - 1) * a.ds + nx) * sizeof(float));
#pragma unroll
        for (int b = 0; b < P::E / RL; b++)
#pragma unroll
            for (int r = 0; r < RL; r++) {
                if (!last_kept<P>(b, r)) continue;
                const uint32_t D = (last_index<P>(0, b, r) - L / 4) * sizeof(float2);
                const float2 xa = cscale(va[b * RL + r], a.scale), xb = cscale(vb[b * RL + r], a.scale);
                const float2 ua = bload(ub, so, D), ubv = bload(ub, so, D + rowb);
                float ga = ua.x * xa.y - ua.y * xa.x, gbv = ubv.x * xb.y - ubv.y * xb.x;
                if (a.pg_acc) {
                    ga += bloadf(gb, so / 2, D / 2);
                    gbv += bloadf(gb, so / 2, (D + rowb) / 2);
                }
                bstoref(gb, so / 2, D / 2, ga);
                bstoref(gb, so / 2, (D + rowb) / 2, gbv);
            }
        return;
    }
    if (a.cx == nullptr) {
        double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        Rsrc tb;
        if constexpr (LS) tb = make_rsrc(a.ls_t, ((size_t)(a.rows - 1) * a.ds + nx) * sizeof(float));
#pragma unroll
        for (int b = 0; b < P::E / RL; b++)
#pragma unroll
            for (int r = 0; r < RL; r++) {
                if (!last_kept<P>(b, r)) continue;
                // column t + D
                const uint32_t D = (last_index<P>(0, b, r) - L / 4) * sizeof(float2);
                float2 xa = cscale(va[b * RL + r], a.scale), xb = cscale(vb[b * RL + r], a.scale);
                if (a.post) {
                    xa = cmul(xa, bload(qb, fo, D));
                    xb = cmul(xb, bload(qb, fo, D + frowb));
                }
                bstore_acc(db, so, D, xa, ACC);
                bstore_acc(db, so, D + rowb, xb, ACC);
                if (a.img8) {  // not with ACC (the entry points exclude it)
                    const size_t e = (so + D) / sizeof(float2);
                    store_img8(a, e, xa);
                    store_img8(a, e + (size_t)KR * a.ds, xb);
                }
                if constexpr (LS) {  // float offsets are half the complex ones
                    const double Ia = (double)xa.x * xa.x + (double)xa.y * xa.y;
                    const double Ib = (double)xb.x * xb.x + (double)xb.y * xb.y;
                    const double Ta = (double)bloadf(tb, so / 2, D / 2);
                    const double Tb = (double)bloadf(tb, so / 2, (D + rowb) / 2);
                    acc[0] += Ia + Ib;
                    acc[1] += Ta + Tb;
                    acc[2] += Ia * Ia + Ib * Ib;
                    acc[3] += Ia * Ta + Ib * Tb;
                    acc[4] += Ta * Ta + Tb * Tb;
                }
            }
        if constexpr (LS) {
            // work-group sums in a fixed order through LDS (free again: the inverse
            // transform's last LDS reads are behind it, then this barrier)
            constexpr int NT = KR * P::T;
            static_assert(5 * NT * sizeof(double) <= KR * P::LP * sizeof(float2), "sums fit in LDS");
            double* sh = reinterpret_cast<double*>(lds);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 5; k++) sh[k * NT + tid] = acc[k];
            __syncthreads();
            for (int off = NT / 2; off > 0; off >>= 1) {
                if (tid < off) {
#pragma unroll
                    for (int k = 0; k < 5; k++) sh[k * NT + tid] += sh[k * NT + tid + off];
                }
                __syncthreads();
            }
            if (tid == 0)
                for (int k = 0; k < 5; k++) a.ls_part[k * gridDim.x + blockIdx.x] = sh[k * NT];
        }
        return;
    }
    const bool posta = abs(ha - a.ny / 2) < a.chl;
    const bool postb = abs(hb - a.ny / 2) < a.chl;
    const float2 cya = posta ? a.cy[ha] : make_float2(1.f, 0.f);
    const float2 cyb = postb ? a.cy[hb] : make_float2(1.f, 0.f);
#pragma unroll
    for (int b = 0; b < P::E / RL; b++)
#pragma unroll
        for (int r = 0; r < RL; r++) {
            if (!last_kept<P>(b, r)) continue;
            const int D = last_index<P>(0, b, r) - L / 4;  // column t + D
            const int col = t + D;
            float2 xa = cscale(va[b * RL + r], a.scale), xb = cscale(vb[b * RL + r], a.scale);
            if (abs(col - nx / 2) < a.cwl) {
                const float2 cx = a.cx[col];
                if (posta) xa = cmul(xa, cmul(cx, cya));
                if (postb) xb = cmul(xb, cmul(cx, cyb));
            }
            if (a.post) {
                xa = cmul(xa, bload(qb, fo, D * sizeof(float2)));
                xb = cmul(xb, bload(qb, fo, D * sizeof(float2) + frowb));
            }
            bstore_acc(db, so, D * sizeof(float2), xa, ACC);
            bstore_acc(db, so, D * sizeof(float2) + rowb, xb, ACC);
            if (a.img8) {
                const size_t e = so / sizeof(float2) + D;
                store_img8(a, e, xa);
                store_img8(a, e + (size_t)KR * a.ds, xb);
            }
        }
}

// work-group shapes and radix orders of the tiled path; OLX_TILE="ka,kb,kc,cbl"
// and OLX_ORDER="oa,of,oi,oc" override them for tuning (each KR is taken down to
// kr_max / 2^d, d < 4; each order 0 = remainder radix first, 1 = last)
struct TileCfg {
    int kra, krb, krc, cbl;
    int oa, of, oi, oc;
};

template <int L, int KR>
static void launch_rows_fwd(const TileCfg& g, const PropArgs& a, const float2* src, float2* W,
                            const float2* tw, hipStream_t s) {
    const int grid = a.rows / (2 * KR), nt = KR * Plan<L>::T;
    if (a.lg_t) {
        if (g.oa) k_rows_pad_fwd2<L, KR, REM_LAST, true><<<grid, nt, 0, s>>>(src, W, a, tw);
        else k_rows_pad_fwd2<L, KR, REM_FIRST, true><<<grid, nt, 0, s>>>(src, W, a, tw);
    } else {
        if (g.oa) k_rows_pad_fwd2<L, KR, REM_LAST><<<grid, nt, 0, s>>>(src, W, a, tw);
        else k_rows_pad_fwd2<L, KR, REM_FIRST><<<grid, nt, 0, s>>>(src, W, a, tw);
    }
}
template <int L, int KR, int MODE>
static void launch_cols_mode(const TileCfg& g, const PropArgs& a, float2* W, const float2* tw,
                             hipStream_t s) {
    const dim3 grid(2 * a.nx / (2 * KR), a.nb);
    const int nt = KR * Plan<L>::T;
    if (g.of && g.oi) k_cols_conv2<L, KR, MODE, REM_LAST, REM_LAST><<<grid, nt, 0, s>>>(W, a, tw);
    else if (g.of) k_cols_conv2<L, KR, MODE, REM_LAST, REM_FIRST><<<grid, nt, 0, s>>>(W, a, tw);
    else if (g.oi) k_cols_conv2<L, KR, MODE, REM_FIRST, REM_LAST><<<grid, nt, 0, s>>>(W, a, tw);
    else k_cols_conv2<L, KR, MODE, REM_FIRST, REM_FIRST><<<grid, nt, 0, s>>>(W, a, tw);
}
template <int L, int KR>
static void launch_cols_conv(const TileCfg& g, const PropArgs& a, float2* W, const float2* tw,
                             hipStream_t s) {
    if (a.mode == K_ASM) launch_cols_mode<L, KR, K_ASM>(g, a, W, tw, s);
    else if (a.mode == K_ASHIFT) launch_cols_mode<L, KR, K_ASHIFT>(g, a, W, tw, s);
    else if (a.mode == K_ASM32) {
        // the default radix orders only (a compatibility mode, not tuned)
        const dim3 grid(2 * a.nx / (2 * KR), a.nb);
        const int nt = KR * Plan<L>::T;
        k_cols_conv2<L, KR, K_ASM32, REM_LAST, REM_LAST><<<grid, nt, 0, s>>>(W, a, tw);
    }
    else launch_cols_mode<L, KR, K_TABLE>(g, a, W, tw, s);
}
template <int L, int KR>
static void launch_rows_inv(const TileCfg& g, const PropArgs& a, const float2* W, float2* dst,
                            const float2* tw, hipStream_t s) {
    const int grid = a.rows / (2 * KR), nt = KR * Plan<L>::T;
    if (a.ls_t) {
        if (g.oc) k_rows_crop_inv2<L, KR, REM_LAST, false, false, true><<<grid, nt, 0, s>>>(W, dst, a, tw);
        else k_rows_crop_inv2<L, KR, REM_FIRST, false, false, true><<<grid, nt, 0, s>>>(W, dst, a, tw);
    } else if (a.pg_grad) {
        if (g.oc) k_rows_crop_inv2<L, KR, REM_LAST, false, true><<<grid, nt, 0, s>>>(W, dst, a, tw);
        else k_rows_crop_inv2<L, KR, REM_FIRST, false, true><<<grid, nt, 0, s>>>(W, dst, a, tw);
    } else if (a.accum) {
        if (g.oc) k_rows_crop_inv2<L, KR, REM_LAST, true><<<grid, nt, 0, s>>>(W, dst, a, tw);
        else k_rows_crop_inv2<L, KR, REM_FIRST, true><<<grid, nt, 0, s>>>(W, dst, a, tw);
    } else {
        if (g.oc) k_rows_crop_inv2<L, KR, REM_LAST><<<grid, nt, 0, s>>>(W, dst, a, tw);
        else k_rows_crop_inv2<L, KR, REM_FIRST><<<grid, nt, 0, s>>>(W, dst, a, tw);
    }
}

// ------------------------------------------------------------------ plain 2-D FFT
template <int L, bool INV>
__global__ void __launch_bounds__(row_nt<L>())
k_rows_fft(float2* __restrict__ u, int rows, const float2* __restrict__ tw) {
    using P = Plan<L>;
    constexpr int RPW = row_rpw<L>();
    constexpr int R0 = P::radix(0), RL = P::radix(P::NST - 1);
    __shared__ float2 lds[RPW * P::LP];
    const int tid = threadIdx.x, rr = tid / P::T, t = tid - rr * P::T;
    const int row = blockIdx.x * RPW + rr;
    const bool valid = row < rows;
    float2* p = u + (size_t)row * L;
    float2 v[P::E];
#pragma unroll
    for (int b = 0; b < P::E / R0; b++)
#pragma unroll
        for (int r = 0; r < R0; r++)
            v[b * R0 + r] = valid ? p[first_index<P>(t, b, r)] : make_float2(0.f, 0.f);
    run_stages<P, INV, TWN, 1>(v, t, lds, rr * P::LP, tw);
    if (!valid) return;
#pragma unroll
    for (int b = 0; b < P::E / RL; b++)
#pragma unroll
        for (int r = 0; r < RL; r++) p[last_index<P>(t, b, r)] = v[b * RL + r];
}

template <int L, bool INV>
__global__ void __launch_bounds__(col_nt<L>())
k_cols_fft(float2* __restrict__ u, int cols, const float2* __restrict__ tw) {
    using P = Plan<L>;
    constexpr int CB = col_cb<L>();
    constexpr int R0 = P::radix(0), RL = P::radix(P::NST - 1);
    __shared__ float2 lds[CB * P::LP];
    const int tid = threadIdx.x, c = tid % CB, t = tid / CB;
    // neighbouring column groups share 128-byte lines: keep them on one XCD's L2
    const int col = xcd_remap(blockIdx.x, gridDim.x) * CB + c;
    const bool valid = col < cols;
    float2 v[P::E];
#pragma unroll
    for (int b = 0; b < P::E / R0; b++)
#pragma unroll
        for (int r = 0; r < R0; r++)
            v[b * R0 + r] = valid ? u[(size_t)first_index<P>(t, b, r) * cols + col]
                                  : make_float2(0.f, 0.f);
    run_stages<P, INV, TWN, CB>(v, t, lds, c, tw);
    if (!valid) return;
#pragma unroll
    for (int b = 0; b < P::E / RL; b++)
#pragma unroll
        for (int r = 0; r < RL; r++) u[(size_t)last_index<P>(t, b, r) * cols + col] = v[b * RL + r];
}

// ------------------------------------------------------------------ tables
// centred chirp in FP64: h[m] = exp(i K (qa x^2 + qb x + qc)), x = (m - n/2) d,
// zero where |m - n/2| >= lim
// (rot: stored at (m + n/2) mod n, the fftshifted order an FFT of the table takes)
__global__ void k_chirp_f64(double2* __restrict__ out, int n, double d, double K, double qa,
                            double qb, double qc, long long lim, int rot = 0) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int at = rot ? (m + n / 2) % n : m;
    const long long off = (long long)m - n / 2;
    if (llabs(off) >= lim) {
        out[at] = make_double2(0.0, 0.0);
        return;
    }
    const double x = (double)off * d;
    const double ph = K * (qa * x * x + qb * x + qc);
    double s, c;
    sincos(ph, &s, &c);
    out[at] = make_double2(c, s);
}

__global__ void k_spec_to_f32(float2* __restrict__ out, const double2* __restrict__ in, int n,
                              double scale, int adjoint) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    out[m] = make_float2((float)(in[m].x * scale), (float)((adjoint ? -in[m].y : in[m].y) * scale));
}

__global__ void k_to_f32(float2* __restrict__ out, const double2* __restrict__ in, int n,
                         int adjoint) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    out[m] = make_float2((float)in[m].x, (float)(adjoint ? -in[m].y : in[m].y));
}

// natural-order spectrum of the fftshift of a centred table, FP64 direct DFT:
// S[k] = scale * sum_n hc[(n + M/2) mod M] exp(-2 pi i n k / M)
__global__ void k_dft_f64(float2* __restrict__ S, const double2* __restrict__ hc, int M,
                          double scale, int adjoint) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    double re = 0.0, im = 0.0;
    for (int n = 0; n < M; n++) {
        const double2 h = hc[(n + M / 2) % M];
        if (h.x == 0.0 && h.y == 0.0) continue;
        const long long nk = ((long long)n * k) % M;
        double s, c;
        sincospi(-2.0 * (double)nk / (double)M, &s, &c);
        re += h.x * c - h.y * s;
        im += h.x * s + h.y * c;
    }
    S[k] = make_float2((float)(re * scale), (float)((adjoint ? -im : im) * scale));
}

// centred H / h tables (the reference's AsmTransferF / FresnelResponse layouts)
__global__ void k_asm_transfer(float2* __restrict__ H, int ny, int nx, double dv, double du,
                               double inv_l2, double z, int bl, double uw, double vw) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y;
    if (m >= nx) return;
    const double v = (double)(n - ny / 2) * dv, u = (double)(m - nx / 2) * du;
    const double arg = inv_l2 - u * u - v * v;
    float2 h = make_float2(0.f, 0.f);
    if (arg >= 0.0 && (!bl || (fabs(u) * 2 < uw && fabs(v) * 2 < vw))) {
        double cyc = sqrt(arg) * z;
        cyc -= rint(cyc);
        double s, c;
        sincospi(2.0 * cyc, &s, &c);
        h = make_float2((float)c, (float)s);
    }
    H[(size_t)n * nx + m] = h;
}

// shiftedAsmTransferF (include/olprop.h:250-305) / gshiftedAsmTransferF (include/golprop.h:
// 336-419): H = exp(2 pi i (w z + ox u + oy v)), centred; bl: the window
// |u - u0| * 2 < uw, |v - v0| * 2 < vw of the CPU form; evanescent entries are 0
__global__ void k_shifted_transfer(float2* __restrict__ H, int ny, int nx, double dv, double du,
                                   double inv_l2, double z, double oy, double ox, int bl,
                                   double u0, double uw, double v0, double vw) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y;
    if (m >= nx) return;
    const double v = (double)(n - ny / 2) * dv, u = (double)(m - nx / 2) * du;
    const double arg = inv_l2 - u * u - v * v;
    float2 h = make_float2(0.f, 0.f);
    if (arg >= 0.0 && (!bl || (fabs(u - u0) * 2 < uw && fabs(v - v0) * 2 < vw))) {
        double cyc = sqrt(arg) * z + ox * u + oy * v;
        cyc -= rint(cyc);
        double s, c;
        sincospi(2.0 * cyc, &s, &c);
        h = make_float2((float)c, (float)s);
    }
    H[(size_t)n * nx + m] = h;
}

__global__ void k_fresnel_response(float2* __restrict__ h, int ny, int nx, double dy, double dx,
                                   double tmp, long long wlim, long long hlim) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x, n = blockIdx.y;
    if (m >= nx) return;
    float2 o = make_float2(0.f, 0.f);
    if (llabs((long long)m - nx / 2) < wlim && llabs((long long)n - ny / 2) < hlim) {
        const double y = (double)(n - ny / 2) * dy, x = (double)(m - nx / 2) * dx;
        double cyc = 0.5 * (x * x + y * y) * tmp;  // pi (x^2 + y^2) tmp radians
        cyc -= rint(cyc);
        double s, c;
        sincospi(2.0 * cyc, &s, &c);
        o = make_float2((float)c, (float)s);
    }
    h[(size_t)n * nx + m] = o;
}

static int64_t win_limit(double v) {
    const double f = std::floor(v);
    if (!(f == f) || f >= (double)INT_MAX) return INT_MAX;
    if (f <= (double)INT_MIN) return INT_MIN;
    return (int64_t)f;
}

static std::string key_of(const char* kind, int n, double d, double K, double qa, double qb,
                          double qc, int64_t lim, int adj) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s|%d|%.17g|%.17g|%.17g|%.17g|%.17g|%lld|%d", kind, n, d, K, qa,
             qb, qc, (long long)lim, adj);
    return buf;
}

// float2 chirp table of length n (pre/post chirps)
static const float2* chirp_table(olx_ctx* ctx, int n, double d, double K, double qa, double qb,
                                 double qc, int64_t lim, int* rc) {
    bool fresh = false;
    float2* t = table(ctx, key_of("chirp", n, d, K, qa, qb, qc, lim, 0), n, &fresh, rc);
    if (!t || !fresh) return t;
    if ((*rc = ctx->scratch2.reserve(sizeof(double2) * n)) != OLX_OK) return nullptr;
    double2* tmp = (double2*)ctx->scratch2.p;
    k_chirp_f64<<<(n + 255) / 256, 256, 0, ctx->stream>>>(tmp, n, d, K, qa, qb, qc, lim);
    k_to_f32<<<(n + 255) / 256, 256, 0, ctx->stream>>>(t, tmp, n, 0);
    if ((*rc = check_launch("chirp table")) != OLX_OK) return nullptr;
    return t;
}

// spectrum (length M, natural order, scaled by 1/M) of the centred chirp
// K (qa x^2 + qb x), zero outside the window lim
const float2* spectrum_table(olx_ctx* ctx, int M, double d, double K, double qa, int64_t lim,
                             int adj, int* rc, double qb) {
    bool fresh = false;
    float2* t = table(ctx, key_of("spec", M, d, K, qa, qb, 0, lim, adj), M, &fresh, rc);
    if (!t || !fresh) return t;
    if ((*rc = ctx->scratch2.reserve(sizeof(double2) * M)) != OLX_OK) return nullptr;
    double2* tmp = (double2*)ctx->scratch2.p;
    // the chirp's FFT in FP64 (a 8192-point table: tens of microseconds against about
    // 2 ms for the direct DFT, which takes the lengths the FFT does not: above 16384)
    if (M % 2 == 0) {
        k_chirp_f64<<<(M + 255) / 256, 256, 0, ctx->stream>>>(tmp, M, d, K, qa, qb, 0.0, lim, 1);
        if (fft64_one(ctx, tmp, M) == OLX_OK) {
            k_spec_to_f32<<<(M + 255) / 256, 256, 0, ctx->stream>>>(t, tmp, M, 1.0 / M, adj);
            if ((*rc = check_launch("spectrum table")) != OLX_OK) return nullptr;
            return t;
        }
    }
    k_chirp_f64<<<(M + 255) / 256, 256, 0, ctx->stream>>>(tmp, M, d, K, qa, qb, 0.0, lim);
    k_dft_f64<<<(M + 127) / 128, 128, 0, ctx->stream>>>(t, tmp, M, 1.0 / M, adj);
    if ((*rc = check_launch("spectrum table")) != OLX_OK) return nullptr;
    return t;
}

// ------------------------------------------------------------------ launchers
#define OLX_LEN_SWITCH(LOG2, ...)             \
    switch (LOG2) {                           \
        case 3: { constexpr int L = 8; __VA_ARGS__; } break;     \
        case 4: { constexpr int L = 16; __VA_ARGS__; } break;    \
        case 5: { constexpr int L = 32; __VA_ARGS__; } break;    \
        case 6: { constexpr int L = 64; __VA_ARGS__; } break;    \
        case 7: { constexpr int L = 128; __VA_ARGS__; } break;   \
        case 8: { constexpr int L = 256; __VA_ARGS__; } break;   \
        case 9: { constexpr int L = 512; __VA_ARGS__; } break;   \
        case 10: { constexpr int L = 1024; __VA_ARGS__; } break; \
        case 11: { constexpr int L = 2048; __VA_ARGS__; } break; \
        case 12: { constexpr int L = 4096; __VA_ARGS__; } break; \
        case 13: { constexpr int L = 8192; __VA_ARGS__; } break; \
        case 14: { constexpr int L = 16384; __VA_ARGS__; } break;\
        default: OLX_FAIL(OLX_EUNSUPPORTED, "transform length 2^%d unsupported", (int)(LOG2)); \
    }

static int ilog2(int64_t n) { return 63 - __builtin_clzll((unsigned long long)n); }

// columns per work-group of pass B: CB * L/16 threads and CB padded columns of
// LDS must fit (1024 threads, 160 KiB); OLX_COLCB overrides the default for tuning
template <int L>
constexpr int col_cb_max() { return L >= 16384 ? 1 : (L >= 1024 ? 16384 / L : 16); }

template <int L>
static int col_batch() {
    static int env = [] {
        const char* e = getenv("OLX_COLCB");
        return e ? atoi(e) : 0;
    }();
    int cb = env > 0 ? env : col_cb<L>();
    if (cb > col_cb_max<L>()) cb = col_cb_max<L>();
    int p2 = 1;
    while (p2 * 2 <= cb) p2 *= 2;
    return p2;
}

template <int L, int CB>
static void launch_cols_cb(const PropArgs& a, float2* W, const float2* tw, hipStream_t s) {
    const int Mx = 2 * a.nx;
    const dim3 grid((Mx + CB - 1) / CB, a.nb);
    if (a.mode == K_ASM)
        k_cols_conv<L, CB, K_ASM><<<grid, CB * Plan<L>::T, 0, s>>>(W, a, tw);
    else if (a.mode == K_ASHIFT)
        k_cols_conv<L, CB, K_ASHIFT><<<grid, CB * Plan<L>::T, 0, s>>>(W, a, tw);
    else if (a.mode == K_ASM32)
        k_cols_conv<L, CB, K_ASM32><<<grid, CB * Plan<L>::T, 0, s>>>(W, a, tw);
    else
        k_cols_conv<L, CB, K_TABLE><<<grid, CB * Plan<L>::T, 0, s>>>(W, a, tw);
}

template <int L>
static int launch_cols(int cb, const PropArgs& a, float2* W, const float2* tw, hipStream_t s) {
    constexpr int CMAX = col_cb_max<L>();
    if (cb >= 16 && CMAX >= 16) launch_cols_cb<L, (CMAX >= 16 ? 16 : 1)>(a, W, tw, s);
    else if (cb >= 8 && CMAX >= 8) launch_cols_cb<L, (CMAX >= 8 ? 8 : 1)>(a, W, tw, s);
    else if (cb >= 4 && CMAX >= 4) launch_cols_cb<L, (CMAX >= 4 ? 4 : 1)>(a, W, tw, s);
    else if (cb >= 2 && CMAX >= 2) launch_cols_cb<L, (CMAX >= 2 ? 2 : 1)>(a, W, tw, s);
    else launch_cols_cb<L, 1>(a, W, tw, s);
    return OLX_OK;
}

static int kr_max_of(int L) { return 16384 / L; }

// KR for a pass of `lines` transforms: half the largest (measured at 4096^2, padded
// 8192: 512-thread work-groups, so two fit on a CU and one's memory phases overlap
// the other's arithmetic), halved while the pass would have fewer than `groups`
// work-groups.  Measured on smaller fields: the row passes, which stream, run
// fastest with at least 1024 groups; the column pass, which computes, with 512 and
// larger groups (1024^2: 13.7 + 25.9 + 10.5 us against 21.1 + 27.4 + 15.6 us with
// the 4096^2 shapes; 2048^2: 22.7 + 55.7 + 24.3 us against 23.5 + 55.9 + 26.4 us).
static int kr_default(int L, int64_t lines, int64_t groups) {
    int kr = std::max(1, kr_max_of(L) / 2);
    while (kr > 1 && kr_max_of(L) / kr < 8 && lines / (2 * kr) < groups) kr /= 2;
    return kr;
}

// Defaults: KR per pass as above; tiles 8 columns wide, so a row pass's two rows
// fill whole 128-byte lines; the row passes with the remainder radix first; the
// column pass's inverse transform with it last (no twiddles in its first radix-16
// stage), and its forward transform with it last too, except with one transform per
// LDS round (KR = 1, padded columns of 8192 and more), where the remainder first
// (one nonzero input per radix-2 butterfly under the zero padding) measured faster
// with the pipelined exchanges: 4096^2 ASM 0.4094 against 0.4143 ms per propagation
// (at 1024^2 and 2048^2 the two orders are within 0.5%).
static TileCfg tile_cfg(int Mx, int My) {
    TileCfg c;
    c.kra = c.krc = kr_default(Mx, My / 2, 1024);
    c.krb = kr_default(My, Mx, 512);
    c.cbl = std::max(8, 2 * c.krb);
    c.oa = c.oc = REM_FIRST;
    c.of = c.krb == 1 ? REM_FIRST : REM_LAST;
    c.oi = REM_LAST;
    static const char* oenv = getenv("OLX_ORDER");
    if (oenv) {
        int o[4];
        if (sscanf(oenv, "%d,%d,%d,%d", &o[0], &o[1], &o[2], &o[3]) == 4) {
            c.oa = o[0] ? REM_LAST : REM_FIRST;
            c.of = o[1] ? REM_LAST : REM_FIRST;
            c.oi = o[2] ? REM_LAST : REM_FIRST;
            c.oc = o[3] ? REM_LAST : REM_FIRST;
        }
    }
    static const char* env = getenv("OLX_TILE");
    if (env) {
        int v[4];
        if (sscanf(env, "%d,%d,%d,%d", &v[0], &v[1], &v[2], &v[3]) == 4) {
            auto pick = [](int want, int mx) {
                int k = mx;
                while (k > 1 && k > want && mx / k < 8) k /= 2;
                return k;
            };
            c.kra = pick(v[0], kr_max_of(Mx));
            c.krb = pick(v[1], kr_max_of(My));
            c.krc = pick(v[2], kr_max_of(Mx));
            int cb = 2 * c.krb;
            while (cb < v[3] && cb < 16) cb *= 2;
            c.cbl = cb;
        }
    }
    return c;
}

// the loss-gradient prologue in place over a packed ny x nx field (untiled path)
__global__ void k_loss_grad(float2* __restrict__ x, const float* __restrict__ t, int64_t n,
                            const double* __restrict__ c) {
    const double mi = c[0], mt = c[1], c1 = c[2], c2 = c[3];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = loss_grad(x[i], t[i], mi, mt, c1, c2);
}

// pg_grad (+)= Im(conj(u) x) over a packed ny x nx field x (the untiled path's
// phase-gradient epilogue)
__global__ void k_phigrad(const float2* __restrict__ u, const float2* __restrict__ x, int64_t n,
                          float* __restrict__ grad, int accumulate) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const float2 p = u[i], q = x[i];
        const float g = p.x * q.y - p.y * q.x;
        grad[i] = accumulate ? grad[i] + g : g;
    }
}

// Wext: a caller-provided intermediate of at least ny x (2nx rounded up to 8) complex64 values in
// place of the context's workspace (which is then neither grown nor touched)
static int run_three_pass(olx_ctx* ctx, const float2* src, float2* dst, PropArgs& a,
                          float2* Wext = nullptr) {
    const int64_t Mx = 2 * (int64_t)a.nx, My = 2 * (int64_t)a.ny;
    float2* W = Wext;
    if (!W) {
        // Mx rounded up to whole 8-column tiles (the mixed-radix passes' layout)
        OLX_TRY(ctx->work.reserve(sizeof(float2) * (size_t)a.rows * (size_t)((Mx + 7) & ~7)));
        W = (float2*)ctx->work.p;
    }
    a.scale = (float)(a.extra_scale / ((double)Mx * (double)My));
    // W layout: row-major (the default), or blocked by the column batch when
    // OLX_WLAYOUT=blocked (contiguous column pass, scattered row passes)
    static const bool blocked = [] {
        const char* e = getenv("OLX_WLAYOUT");
        return e && strcmp(e, "blocked") == 0;
    }();
    static const bool pair_off = [] {
        const char* e = getenv("OLX_PAIR");
        return e && strcmp(e, "0") == 0;
    }();
    hipStream_t s = ctx->stream;
    if (!is_pow2(a.nx) || !is_pow2(a.ny) || a.nx < 4 || a.ny < 4) {
        // mixed-radix passes; the loss-gradient prologue and the phase-gradient
        // epilogue run as kernels of their own, as on the untiled path
        const int64_t n = (int64_t)a.ny * a.nx;
        const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
        if (a.lg_t) {
            k_loss_grad<<<blocks, 256, 0, s>>>(const_cast<float2*>(src), a.lg_t, n, a.lg_c);
            OLX_TRY(check_launch("loss gradient"));
        }
        OLX_TRY(mr_three_pass(ctx, src, dst, a, W));
        if (a.pg_grad) {
            k_phigrad<<<blocks, 256, 0, s>>>(a.pg_u, dst, n, a.pg_grad, a.pg_acc);
            return check_launch("phase gradient");
        }
        return OLX_OK;
    }
    if (!pair_off && Mx >= PAIR_MIN && My >= PAIR_MIN) {
        TileCfg cfg = tile_cfg((int)Mx, (int)My);
        a.cbl = cfg.cbl;
        a.cbl_log = ilog2(a.cbl);
        {
            TimedLaunch tl(ctx, "prop_rows_pad_fwd");
            OLX_LEN_SWITCH(ilog2(Mx), {
                if constexpr (L >= PAIR_MIN) {
                    with_kr<L>(cfg.kra, [&](auto K) { launch_rows_fwd<L, decltype(K)::value>(cfg, a, src, W, ctx->twiddle, s); });
                }
            });
        }
        OLX_TRY(check_launch("prop pass A"));
        {
            TimedLaunch tl(ctx, "prop_cols_conv");
            OLX_LEN_SWITCH(ilog2(My), {
                if constexpr (L >= PAIR_MIN) {
                    with_kr<L>(cfg.krb, [&](auto K) { launch_cols_conv<L, decltype(K)::value>(cfg, a, W, ctx->twiddle, s); });
                }
            });
        }
        OLX_TRY(check_launch("prop pass B"));
        {
            TimedLaunch tl(ctx, "prop_rows_crop_inv");
            OLX_LEN_SWITCH(ilog2(Mx), {
                if constexpr (L >= PAIR_MIN) {
                    with_kr<L>(cfg.krc, [&](auto K) {
                        launch_rows_inv<L, decltype(K)::value>(cfg, a, W, dst, ctx->twiddle, s);
                        if (a.ls_t) a.ls_groups = a.ny / (2 * decltype(K)::value);
                    });
                }
            });
        }
        return check_launch("prop pass C");
    }
    if (a.lg_t) {  // the prologue in place (the caller's source is writable in this mode)
        const int64_t n = (int64_t)a.ny * a.nx;
        const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
        k_loss_grad<<<blocks, 256, 0, s>>>(const_cast<float2*>(src), a.lg_t, n, a.lg_c);
        OLX_TRY(check_launch("loss gradient"));
    }
    int cb = 1;
    OLX_LEN_SWITCH(ilog2(My), { cb = col_batch<L>(); });
    if (cb > Mx) cb = (int)Mx;
    a.cbl = blocked ? cb : (int)Mx;
    a.cbl_log = ilog2(a.cbl);
    {
        TimedLaunch tl(ctx, "prop_rows_pad_fwd");
        OLX_LEN_SWITCH(ilog2(Mx), {
            const int grid = (a.rows + row_rpw<L>() - 1) / row_rpw<L>();
            k_rows_pad_fwd<L><<<grid, row_nt<L>(), 0, s>>>(src, W, a, ctx->twiddle);
        });
    }
    OLX_TRY(check_launch("prop pass A"));
    {
        TimedLaunch tl(ctx, "prop_cols_conv");
        OLX_LEN_SWITCH(ilog2(My), {
            OLX_TRY(launch_cols<L>(cb, a, W, ctx->twiddle, s));
        });
    }
    OLX_TRY(check_launch("prop pass B"));
    {
        TimedLaunch tl(ctx, "prop_rows_crop_inv");
        OLX_LEN_SWITCH(ilog2(Mx), {
            const int grid = (a.rows + row_rpw<L>() - 1) / row_rpw<L>();
            k_rows_crop_inv<L><<<grid, row_nt<L>(), 0, s>>>(W, dst, a, ctx->twiddle);
        });
    }
    OLX_TRY(check_launch("prop pass C"));
    if (a.pg_grad) {
        const int64_t n = (int64_t)a.ny * a.nx;
        const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
        k_phigrad<<<blocks, 256, 0, s>>>(a.pg_u, dst, n, a.pg_grad, a.pg_acc);
        return check_launch("phase gradient");
    }
    return OLX_OK;
}

// Power-of-two sides in [4, 8192] take the fused passes; other sides in [2, 8192]
// whose padded lengths 2n factor into 2, 3, 5, 7, 11 and 13 the mixed-radix ones.
static int check_field(olx_ctx* ctx, const void* p, int64_t ny, int64_t nx) {
    if (!ctx) OLX_FAIL(OLX_EINVAL, "null context");
    if (!p) OLX_FAIL(OLX_EINVAL, "null field pointer");
    if (is_pow2(ny) && is_pow2(nx) && ny >= 4 && nx >= 4 && ny <= 8192 && nx <= 8192)
        return OLX_OK;
    if (ny < 2 || nx < 2 || ny > 8192 || nx > 8192 || !mr_supported(2 * ny) ||
        !mr_supported(2 * nx))
        OLX_FAIL(OLX_EUNSUPPORTED,
                 "field %lld x %lld: sides must be in [2, 8192] with prime factors 2, 3, 5, 7, "
                 "11 and 13 only (powers of two from 4)", (long long)ny, (long long)nx);
    return OLX_OK;
}

static PropArgs base_args(int64_t ny, int64_t nx) {
    PropArgs a;
    memset(&a, 0, sizeof a);
    a.ny = (int)ny;
    a.nx = (int)nx;
    a.nb = 1;
    a.rows = (int)ny;
    a.ss = a.ds = (int)nx;
    a.extra_scale = 1.0;
    return a;
}

}  // namespace olx

using namespace olx;

static int asm_args(PropArgs& a, int64_t ny, int64_t nx, double dy, double dx, double lambda,
                    double z, uint32_t flags) {
    if (!(lambda > 0) || !(dx > 0) || !(dy > 0) || !std::isfinite(z))
        OLX_FAIL(OLX_EINVAL, "asm: lambda, dx, dy must be > 0 and z finite");
    a = base_args(ny, nx);
    a.mode = K_ASM;
    a.adjoint = (flags & OLX_ADJOINT) ? 1 : 0;
    const int64_t nx2 = 2 * nx, ny2 = 2 * ny;
    a.du = 1 / (dx * nx2);
    a.dv = 1 / (dy * ny2);
    a.inv_l2 = 1 / (lambda * lambda);
    a.zs = a.adjoint ? -z : z;
    a.bandlimit = (flags & OLX_BANDLIMIT) ? 1 : 0;
    if (flags & OLX_EXTRA_INV_M2) a.extra_scale = 1.0 / ((double)nx2 * (double)ny2);
    if (flags & OLX_PREC_F32) {
        // include/olprop.h:227-233 with PREC_T = float: du = 1 / (dx * width2) and
        // 1 / lambda^2 in float (AsmTransferF :176)
        a.mode = K_ASM32;
        const float dxf = (float)dx, dyf = (float)dy, lf = (float)lambda;
        a.duf = 1 / (dxf * (float)nx2);
        a.dvf = 1 / (dyf * (float)ny2);
        a.tmpf = 1 / (lf * lf);
        a.zf = (float)z;
    }
    if (a.bandlimit) {
        // shiftedAsmTransferF at ox = oy = 0 (include/olprop.h:258-270): uw = ulp + ulm
        const double ulm = 1.0 / std::sqrt(std::pow(0.0 - 0.5f / a.du, -2) * z * z + 1.0f) / lambda;
        const double ulp = 1.0 / std::sqrt(std::pow(0.0 + 0.5f / a.du, -2) * z * z + 1.0f) / lambda;
        const double vlm = 1.0 / std::sqrt(std::pow(0.0 - 0.5f / a.dv, -2) * z * z + 1.0f) / lambda;
        const double vlp = 1.0 / std::sqrt(std::pow(0.0 + 0.5f / a.dv, -2) * z * z + 1.0f) / lambda;
        a.uw = ulp + ulm;
        a.vw = vlp + vlm;
    }
    return OLX_OK;
}

int olx::asm_prop(olx_ctx* ctx, const void* d_src, void* d_dst, int64_t ny, int64_t nx, double dy,
                  double dx, double lambda, double z, uint32_t flags) {
    OLX_TRY(check_field(ctx, d_src, ny, nx));
    if (!d_dst) OLX_FAIL(OLX_EINVAL, "asm: null destination");
    PropArgs a;
    OLX_TRY(asm_args(a, ny, nx, dy, dx, lambda, z, flags));
    return run_three_pass(ctx, (const float2*)d_src, (float2*)d_dst, a);
}

int olx::asm_prop_sums(olx_ctx* ctx, const void* d_src, void* d_dst, int64_t ny, int64_t nx,
                       double dy, double dx, double lambda, double z, uint32_t flags,
                       const float* d_t, double* d_part, int* groups) {
    OLX_TRY(check_field(ctx, d_src, ny, nx));
    if (!d_dst || !d_t || !d_part || !groups) OLX_FAIL(OLX_EINVAL, "asm sums: null argument");
    PropArgs a;
    OLX_TRY(asm_args(a, ny, nx, dy, dx, lambda, z, flags));
    a.ls_t = d_t;
    a.ls_part = d_part;
    OLX_TRY(run_three_pass(ctx, (const float2*)d_src, (float2*)d_dst, a));
    *groups = a.ls_groups;
    return OLX_OK;
}

int olx::asm_phase_grad(olx_ctx* ctx, void* d_g, const void* d_u, float* d_grad, int accumulate,
                        int64_t ny, int64_t nx, double dy, double dx, double lambda, double z,
                        uint32_t flags, const float* d_t, const double* d_c) {
    OLX_TRY(check_field(ctx, d_g, ny, nx));
    if (!d_u || !d_grad) OLX_FAIL(OLX_EINVAL, "asm phase gradient: null argument");
    PropArgs a;
    OLX_TRY(asm_args(a, ny, nx, dy, dx, lambda, z, flags));
    a.pg_u = (const float2*)d_u;
    a.pg_grad = d_grad;
    a.pg_acc = accumulate ? 1 : 0;
    if (d_t && d_c) {
        a.lg_t = d_t;
        a.lg_c = d_c;
    }
    return run_three_pass(ctx, (const float2*)d_g, (float2*)d_g, a);
}

extern "C" int olx_asm_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy,
                           double dx, double lambda, double z, uint32_t flags) {
    return asm_prop(ctx, d_u, d_u, ny, nx, dy, dx, lambda, z, flags);
}

// ---------------------------------------------------------------- no-allocation ASM
// The intermediate W of olx_asm_c64_noalloc lives in this file-scope device array
// (one copy per GPU, held until the process ends).  Host code never passes the
// array's name to a kernel: its device address is fetched once per device by
// noalloc_init (called from olx_ctx_create with the device current) and kept in
// g_noalloc_ptr.
constexpr size_t NOALLOC_BYTES = (size_t)64 << 20;  // ny * 2nx * 8 bytes: up to 2048 x 2048
static __device__ __align__(256) float2 scratch[NOALLOC_BYTES / sizeof(float2)];
constexpr int NOALLOC_MAX_DEVICES = 64;
static void* g_noalloc_ptr[NOALLOC_MAX_DEVICES];

int olx::noalloc_init(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES)
        OLX_FAIL(OLX_EUNSUPPORTED, "no-allocation workspace: device %d beyond %d", device,
                 NOALLOC_MAX_DEVICES);
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (g_noalloc_ptr[device]) return OLX_OK;
    void* p = nullptr;
    OLX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) OLX_FAIL(OLX_EDEVICE, "no-allocation workspace: null symbol address on device %d", device);
    g_noalloc_ptr[device] = p;
    return OLX_OK;
}

extern "C" int olx_asm_c64_noalloc(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy,
                                   double dx, double lambda, double z, uint32_t flags) {
    OLX_TRY(check_field(ctx, d_u, ny, nx));
    const size_t need = sizeof(float2) * (size_t)ny * (size_t)((2 * nx + 7) & ~7);
    if (need > NOALLOC_BYTES)
        OLX_FAIL(OLX_EINVAL, "asm noalloc: %lld x %lld needs %zu workspace bytes, above the "
                 "static %zu (at most 2048 x 2048 elements)", (long long)ny, (long long)nx, need,
                 NOALLOC_BYTES);
    void* w = ctx->device >= 0 && ctx->device < NOALLOC_MAX_DEVICES ? g_noalloc_ptr[ctx->device]
                                                                   : nullptr;
    if (!w) OLX_FAIL(OLX_EDEVICE, "asm noalloc: workspace address of device %d not set up",
                     ctx->device);
    PropArgs a;
    OLX_TRY(asm_args(a, ny, nx, dy, dx, lambda, z, flags));
    return run_three_pass(ctx, (const float2*)d_u, (float2*)d_u, a, (float2*)w);
}

// Fresnel arguments: the separable spectrum tables, built or fetched from the cache
static int fresnel_args(olx_ctx* ctx, PropArgs& a, int64_t ny, int64_t nx, double dy, double dx,
                        double lambda, double z, uint32_t flags) {
    if (!(lambda > 0) || !(dx > 0) || !(dy > 0) || !(z != 0) || !std::isfinite(z))
        OLX_FAIL(OLX_EINVAL, "fresnel: lambda, dx, dy must be > 0 and z finite, nonzero");
    if ((flags & OLX_FRESNEL_BL_CPU) && (flags & OLX_FRESNEL_BL_GPU))
        OLX_FAIL(OLX_EINVAL, "fresnel: choose one band-limit variant");
    a = base_args(ny, nx);
    a.mode = K_TABLE;
    const int adj = (flags & OLX_ADJOINT) ? 1 : 0;
    int64_t wlim = INT_MAX, hlim = INT_MAX;
    if (flags & OLX_FRESNEL_BL_CPU) {  // include/olprop.h:35-38
        wlim = win_limit(std::fabs(lambda * z * 0.5 / dx) / dx);
        hlim = win_limit(std::fabs(lambda * z * 0.5 / dy) / dy);
    } else if (flags & OLX_FRESNEL_BL_GPU) {  // include/golprop.h:229-231
        wlim = hlim = win_limit(std::fabs(lambda * z * 0.5 / dx) / dx);
    }
    const double K = M_PI * (1 / (lambda * z));
    int rc = OLX_OK;
    a.sx = spectrum_table(ctx, (int)(2 * nx), dx, K, 1.0, wlim, adj, &rc);
    if (!a.sx) return rc;
    a.sy = spectrum_table(ctx, (int)(2 * ny), dy, K, 1.0, hlim, adj, &rc);
    if (!a.sy) return rc;
    return OLX_OK;
}

extern "C" int olx_fresnel_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy,
                               double dx, double lambda, double z, uint32_t flags) {
    OLX_TRY(check_field(ctx, d_u, ny, nx));
    PropArgs a;
    OLX_TRY(fresnel_args(ctx, a, ny, nx, dy, dx, lambda, z, flags));
    return run_three_pass(ctx, (const float2*)d_u, (float2*)d_u, a);
}

// nb fields stacked in d_u, each replaced by post * Prop(pre * u_k): the row passes
// run over one tall field of nb ny rows, the column pass over the nb fields as
// blockIdx.y.  The 32-bit buffer offsets of the tiled passes bound the
// intermediate (nb ny 2nx complex64) to 2 GiB.
extern "C" int olx_prop_batch_c64(olx_ctx* ctx, int mode, void* d_u, int64_t nb, int64_t ny,
                                  int64_t nx, double dy, double dx, double lambda, double z,
                                  uint32_t flags, const void* d_pre, const void* d_post,
                                  uint8_t* d_img, uint32_t img_flags) {
    OLX_TRY(check_field(ctx, d_u, ny, nx));
    if (nb < 1 || nb > 65535)
        OLX_FAIL(OLX_EINVAL, "prop batch: nb = %lld not in [1, 65535]", (long long)nb);
    const size_t wbytes = sizeof(float2) * (size_t)nb * (size_t)ny * 2 * (size_t)nx;
    if (wbytes > ((size_t)2 << 30))
        OLX_FAIL(OLX_EINVAL, "prop batch: %lld fields of %lld x %lld need a %zu-byte "
                 "intermediate, above 2 GiB", (long long)nb, (long long)ny, (long long)nx, wbytes);
    PropArgs a;
    if (mode == OLX_MODE_ASM) {
        OLX_TRY(asm_args(a, ny, nx, dy, dx, lambda, z, flags));
    } else if (mode == OLX_MODE_FRESNEL) {
        OLX_TRY(fresnel_args(ctx, a, ny, nx, dy, dx, lambda, z, flags));
    } else {
        OLX_FAIL(OLX_EINVAL, "prop batch: mode %d is neither OLX_MODE_ASM nor OLX_MODE_FRESNEL",
                 mode);
    }
    a.nb = (int)nb;
    a.rows = (int)(nb * ny);
    a.pre = (const float2*)d_pre;
    a.post = (const float2*)d_post;
    if (img_flags & ~OLX_PHASE8_OFFSET) OLX_FAIL(OLX_EINVAL, "prop batch: unknown image flags");
    a.img8 = d_img;
    a.img8_off = (img_flags & OLX_PHASE8_OFFSET) ? 1 : 0;
    return run_three_pass(ctx, (const float2*)d_u, (float2*)d_u, a);
}

// ARSS arguments (tables built or fetched from the cache); the caller sets strides
static int arss_args(olx_ctx* ctx, PropArgs& a, int64_t ny, int64_t nx, double dy, double dx,
                     double lambda, double z, double s, double oy, double ox, uint32_t flags) {
    if (!(lambda > 0) || !(dx > 0) || !(dy > 0) || !(z != 0) || !std::isfinite(z) ||
        !(s != 0) || !std::isfinite(s))
        OLX_FAIL(OLX_EINVAL, "arss: lambda, dx, dy > 0, z and s finite and nonzero");
    if (flags & OLX_ADJOINT)
        OLX_FAIL(OLX_EUNSUPPORTED, "arss: adjoint not provided (chirp order reverses)");
    a = base_args(ny, nx);
    a.mode = K_TABLE;
    const double K = M_PI / (lambda * z);
    int rc = OLX_OK;
    // CPU shiftedFresnelProp (include/olprop.h:338-408): no windows anywhere
    const bool cpu = (flags & OLX_ARSS_CPU) != 0;
    // stage u window and pre-chirp (include/golprop.h:849-854, 831-840)
    const int64_t uwl = cpu ? INT_MAX : win_limit(lambda * std::fabs(z) * 0.5 / (s * std::fabs(s - 1) * dx) / dx);
    const int64_t uhl = cpu ? INT_MAX : win_limit(lambda * std::fabs(z) * 0.5 / (s * std::fabs(s - 1) * dy) / dy);
    a.ax = chirp_table(ctx, (int)nx, dx, K, s * s - s, -2 * s * ox, 0.0, uwl, &rc);
    if (!a.ax) return rc;
    a.ay = chirp_table(ctx, (int)ny, dy, K, s * s - s, -2 * s * oy, 0.0, uhl, &rc);
    if (!a.ay) return rc;
    // stage h: kernel chirp, x-limit on both axes (include/golprop.h:890-896)
    const int64_t hl = cpu ? INT_MAX : win_limit(std::fabs(lambda * z * 0.5 / (dx * s)) / dx);
    a.sx = spectrum_table(ctx, (int)(2 * nx), dx, K, s, hl, 0, &rc);
    if (!a.sx) return rc;
    a.sy = spectrum_table(ctx, (int)(2 * ny), dy, K, s, hl, 0, &rc);
    if (!a.sy) return rc;
    // stage Cz window and post-chirp (include/golprop.h:928-933, 911-920)
    const int64_t cwl = cpu ? INT_MAX : win_limit(lambda * std::fabs(z) * 0.5 / (std::fabs(1 - s) * dx) / dx);
    const int64_t chl = cpu ? INT_MAX : win_limit(lambda * std::fabs(z) * 0.5 / (std::fabs(1 - s) * dy) / dy);
    a.cx = chirp_table(ctx, (int)nx, dx, K, 1 - s, 2 * ox, ox * ox, INT_MAX, &rc);
    if (!a.cx) return rc;
    a.cy = chirp_table(ctx, (int)ny, dy, K, 1 - s, 2 * oy, oy * oy, INT_MAX, &rc);
    if (!a.cy) return rc;
    a.cwl = (int)std::max<int64_t>(std::min<int64_t>(cwl, INT_MAX), INT_MIN);
    a.chl = (int)std::max<int64_t>(std::min<int64_t>(chl, INT_MAX), INT_MIN);
    // the CPU variant scales u by 1/M^2 twice and the output once more (olprop.h:367,
    // 384, 389) against the GPU variant's 1/M^2 on u and on h: one extra 1/M^2
    a.extra_scale = cpu ? 1.0 / (4.0 * (double)nx * (double)ny) : 1.0;
    return OLX_OK;
}

// shiftedAsmTransferF's window on one axis (include/olprop.h:255-285): centre c and
// width w of the band |f - c| * 2 < w for a destination offset o, spectral pitch d
static void shift_window(double o, double d, double z, double lambda, double* c, double* w) {
    const double sx = 0.5f / d;
    const double lm = 1 / std::sqrt(std::pow(o - 0.5f / d, -2) * z * z + 1.0f) / lambda;
    const double lp = 1 / std::sqrt(std::pow(o + 0.5f / d, -2) * z * z + 1.0f) / lambda;
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

// shifted ASM arguments: H = exp(2 pi i (w z + ox u + oy v)), window per flags
static int shifted_asm_args(PropArgs& a, int64_t ny, int64_t nx, double dy, double dx,
                            double lambda, double z, double oy, double ox, uint32_t flags) {
    if (!(lambda > 0) || !(dx > 0) || !(dy > 0) || !std::isfinite(z) || !std::isfinite(oy) ||
        !std::isfinite(ox))
        OLX_FAIL(OLX_EINVAL, "shifted asm: lambda, dx, dy must be > 0 and z, oy, ox finite");
    a = base_args(ny, nx);
    a.mode = K_ASHIFT;
    a.adjoint = (flags & OLX_ADJOINT) ? 1 : 0;
    const int64_t nx2 = 2 * nx, ny2 = 2 * ny;
    a.du = 1 / (dx * nx2);
    a.dv = 1 / (dy * ny2);
    a.inv_l2 = 1 / (lambda * lambda);
    // the adjoint conjugates H: every phase term changes sign
    a.zs = a.adjoint ? -z : z;
    a.ox = a.adjoint ? -ox : ox;
    a.oy = a.adjoint ? -oy : oy;
    a.bandlimit = (flags & (OLX_BANDLIMIT | OLX_SHIFTED_CPU)) ? 1 : 0;
    if (a.bandlimit) {
        shift_window(ox, a.du, z, lambda, &a.u0, &a.uw);
        shift_window(oy, a.dv, z, lambda, &a.v0, &a.vw);
    }
    // the CPU shiftedAsmProp scales by 1/M^2 before the inverse FFT and again after
    // the crop (include/olprop.h:325, 329): one 1/M^2 more than the exact result
    a.extra_scale = (flags & OLX_SHIFTED_CPU) ? 1.0 / ((double)nx2 * (double)ny2) : 1.0;
    return OLX_OK;
}

extern "C" int olx_arss_c64(olx_ctx* ctx, const void* d_src, void* d_dst, int64_t ny,
                            int64_t nx, double dy, double dx, double lambda, double z, double s,
                            double oy, double ox, uint32_t flags) {
    OLX_TRY(check_field(ctx, d_src, ny, nx));
    if (!d_dst) OLX_FAIL(OLX_EINVAL, "arss: null destination");
    PropArgs a;
    OLX_TRY(arss_args(ctx, a, ny, nx, dy, dx, lambda, z, s, oy, ox, flags));
    return run_three_pass(ctx, (const float2*)d_src, (float2*)d_dst, a);
}

extern "C" int olx_shifted_asm_c64(olx_ctx* ctx, const void* d_src, void* d_dst, int64_t ny,
                                   int64_t nx, double dy, double dx, double lambda, double z,
                                   double oy, double ox, uint32_t flags) {
    OLX_TRY(check_field(ctx, d_src, ny, nx));
    if (!d_dst) OLX_FAIL(OLX_EINVAL, "shifted asm: null destination");
    PropArgs a;
    OLX_TRY(shifted_asm_args(a, ny, nx, dy, dx, lambda, z, oy, ox, flags));
    return run_three_pass(ctx, (const float2*)d_src, (float2*)d_dst, a);
}

extern "C" int olx_split_prop_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy,
                                  double dx, double lambda, double z, int mode, int by, int bx,
                                  uint32_t flags) {
    if (!ctx || !d_u) OLX_FAIL(OLX_EINVAL, "split prop: null argument");
    if (by <= 0 || bx <= 0 || ny % by || nx % bx)
        OLX_FAIL(OLX_EINVAL, "split prop: %d x %d blocks do not tile %lld x %lld", by, bx,
                 (long long)ny, (long long)nx);
    if (mode != OLX_MODE_ASM && mode != OLX_MODE_FRESNEL)
        OLX_FAIL(OLX_EINVAL, "split prop: mode must be ASM or FRESNEL");
    if ((flags & OLX_SHIFTED_CPU) && mode != OLX_MODE_ASM)
        OLX_FAIL(OLX_EINVAL, "split prop: the CPU form exists for ASM only");
    const int64_t bh = ny / by, bw = nx / bx;
    OLX_TRY(check_field(ctx, d_u, bh, bw));
    OLX_TRY(ctx->split_acc.reserve(sizeof(float2) * (size_t)(ny * nx)));
    float2* u = (float2*)d_u;
    float2* acc = (float2*)ctx->split_acc.p;
    OLX_HIP(hipMemsetAsync(acc, 0, sizeof(float2) * (size_t)(ny * nx), ctx->stream));
    // include/olprop.h:614-651 (splitAsmProp) and include/golprop.h:773-813
    // (gsplitProp): every source block to every destination block, shifted by
    // (src - dst) block offsets computed in float as there, summed
    for (int gn = 0; gn < by; gn++)
        for (int gm = 0; gm < bx; gm++)
            for (int dn = 0; dn < by; dn++)
                for (int dm = 0; dm < bx; dm++) {
                    const float oy = (float)((gn - dn) * (int)bh) * (float)dy;
                    const float ox = (float)((gm - dm) * (int)bw) * (float)dx;
                    PropArgs a;
                    if (mode == OLX_MODE_ASM)
                        OLX_TRY(shifted_asm_args(a, bh, bw, dy, dx, lambda, z, oy, ox,
                                                 flags & OLX_SHIFTED_CPU));
                    else
                        OLX_TRY(arss_args(ctx, a, bh, bw, dy, dx, lambda, z, 1.0, oy, ox, 0));
                    a.ss = a.ds = (int)nx;
                    a.accum = 1;
                    OLX_TRY(run_three_pass(ctx, u + gn * bh * nx + gm * bw,
                                           acc + dn * bh * nx + dm * bw, a));
                }
    OLX_HIP(hipMemcpyAsync(u, acc, sizeof(float2) * (size_t)(ny * nx), hipMemcpyDeviceToDevice,
                           ctx->stream));
    return OLX_OK;
}


extern "C" int olx_asm_transfer_c64(olx_ctx* ctx, void* d_H, int64_t ny, int64_t nx, double dv,
                                    double du, double lambda, double z, uint32_t flags) {
    if (!ctx || !d_H || ny <= 0 || nx <= 0 || ny > 65535 || nx > (1 << 24))
        OLX_FAIL(OLX_EINVAL, "asm_transfer: bad argument");
    double uw = 0, vw = 0;
    if (flags & OLX_BANDLIMIT) {
        const double ulm = 1.0 / std::sqrt(std::pow(0.0 - 0.5f / du, -2) * z * z + 1.0f) / lambda;
        const double ulp = 1.0 / std::sqrt(std::pow(0.0 + 0.5f / du, -2) * z * z + 1.0f) / lambda;
        const double vlm = 1.0 / std::sqrt(std::pow(0.0 - 0.5f / dv, -2) * z * z + 1.0f) / lambda;
        const double vlp = 1.0 / std::sqrt(std::pow(0.0 + 0.5f / dv, -2) * z * z + 1.0f) / lambda;
        uw = ulp + ulm;
        vw = vlp + vlm;
    }
    k_asm_transfer<<<dim3((unsigned)((nx + 255) / 256), (unsigned)ny), 256, 0, ctx->stream>>>(
        (float2*)d_H, (int)ny, (int)nx, dv, du, 1 / (lambda * lambda), z,
        (flags & OLX_BANDLIMIT) ? 1 : 0, uw, vw);
    return check_launch("asm_transfer");
}

extern "C" int olx_shifted_asm_transfer_c64(olx_ctx* ctx, void* d_H, int64_t ny, int64_t nx,
                                            double dv, double du, double lambda, double z,
                                            double oy, double ox, uint32_t flags) {
    if (!ctx || !d_H || ny <= 0 || nx <= 0 || ny > 65535 || nx > (1 << 24))
        OLX_FAIL(OLX_EINVAL, "shifted_asm_transfer: bad argument");
    if (!(lambda > 0) || !(du > 0) || !(dv > 0) || !std::isfinite(z) || !std::isfinite(oy) ||
        !std::isfinite(ox))
        OLX_FAIL(OLX_EINVAL, "shifted_asm_transfer: lambda, du, dv must be > 0 and z, oy, ox finite");
    double u0 = 0, uw = 0, v0 = 0, vw = 0;
    const int bl = (flags & OLX_BANDLIMIT) ? 1 : 0;
    if (bl) {
        shift_window(ox, du, z, lambda, &u0, &uw);
        shift_window(oy, dv, z, lambda, &v0, &vw);
    }
    k_shifted_transfer<<<dim3((unsigned)((nx + 255) / 256), (unsigned)ny), 256, 0, ctx->stream>>>(
        (float2*)d_H, (int)ny, (int)nx, dv, du, 1 / (lambda * lambda), z, oy, ox, bl, u0, uw, v0, vw);
    return check_launch("shifted_asm_transfer");
}

extern "C" int olx_fresnel_response_c64(olx_ctx* ctx, void* d_h, int64_t ny, int64_t nx,
                                        double dy, double dx, double lambda, double z,
                                        int64_t wlim, int64_t hlim) {
    if (!ctx || !d_h || ny <= 0 || nx <= 0 || ny > 65535 || nx > (1 << 24) || z == 0)
        OLX_FAIL(OLX_EINVAL, "fresnel_response: bad argument");
    k_fresnel_response<<<dim3((unsigned)((nx + 255) / 256), (unsigned)ny), 256, 0, ctx->stream>>>(
        (float2*)d_h, (int)ny, (int)nx, dy, dx, 1 / (lambda * z), (long long)wlim, (long long)hlim);
    return check_launch("fresnel_response");
}

extern "C" int olx_asm_prop_check(int64_t ny, int64_t nx, double dy, double dx, double lambda,
                                  double z) {
    // AsmPropCheck, include/olprop.h:209-225
    const int64_t nx2 = 2 * nx, ny2 = 2 * ny;
    const double du = 1 / (dx * nx2), dv = 1 / (dy * ny2);
    const double umax = du * nx2 / 2, vmax = dv * ny2 / 2;
    const double tmpsqrt = std::sqrt(1 / (lambda * lambda) - umax * umax - vmax * vmax);
    const double zmax_u = tmpsqrt / (2 * umax * du), zmax_v = tmpsqrt / (2 * vmax * dv);
    if (zmax_u < std::fabs(z) || zmax_v < std::fabs(z)) return 0;
    return 1;
}

extern "C" int olx_fresnel_prop_check(int64_t ny, int64_t nx, double dy, double dx,
                                      double lambda, double z) {
    // FresnelPropCheck, include/olprop.h:109-123
    const int64_t nx2 = 2 * nx, ny2 = 2 * ny;
    const double xmax = dx * nx2 / 2, ymax = dy * ny2 / 2;
    const double zmin_x = 2 * xmax * dx / lambda, zmin_y = 2 * ymax * dy / lambda;
    if (zmin_x > std::fabs(z) || zmin_y > std::fabs(z)) return 0;
    return 1;
}

extern "C" int olx_prop_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy,
                            double dx, double lambda, double z, int mode, int* chosen) {
    // Prop, include/olprop.h:656-682 (silent: the reference prints the chosen mode)
    int c;
    if (mode == OLX_MODE_FRESNEL) c = OLX_MODE_FRESNEL;
    else if (mode == OLX_MODE_ASM) c = OLX_MODE_ASM;
    else if (mode == OLX_MODE_AUTO) {
        if (olx_asm_prop_check(ny, nx, dy, dx, lambda, z)) c = OLX_MODE_ASM;
        else if (!olx_fresnel_prop_check(ny, nx, dy, dx, lambda, z)) c = OLX_MODE_FRESNEL_BL;
        else c = OLX_MODE_FRESNEL;
    } else OLX_FAIL(OLX_EINVAL, "prop: unknown mode %d", mode);
    if (chosen) *chosen = c;
    if (c == OLX_MODE_ASM) return olx_asm_c64(ctx, d_u, ny, nx, dy, dx, lambda, z, 0);
    if (c == OLX_MODE_FRESNEL) return olx_fresnel_c64(ctx, d_u, ny, nx, dy, dx, lambda, z, 0);
    return olx_fresnel_c64(ctx, d_u, ny, nx, dy, dx, lambda, z, OLX_FRESNEL_BL_CPU);
}

extern "C" int olx_fft2d_c64(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, int sign) {
    if (!ctx || !d_u) OLX_FAIL(OLX_EINVAL, "fft2d: null argument");
    float2* u = (float2*)d_u;
    if (!is_pow2(ny) || !is_pow2(nx) || ny < 8 || nx < 8 || ny > 16384 || nx > 16384) {
        if (!mr_supported(ny) || !mr_supported(nx))
            OLX_FAIL(OLX_EUNSUPPORTED, "fft2d: sides must be in [2, 16384] with prime factors 2, "
                     "3, 5, 7, 11 and 13 only");
        return mr_fft2d(ctx, u, ny, nx, sign > 0);
    }
    hipStream_t s = ctx->stream;
    const bool inv = sign > 0;
    {
        TimedLaunch tl(ctx, "fft_rows");
        OLX_LEN_SWITCH(ilog2(nx), {
            const int grid = (int)((ny + row_rpw<L>() - 1) / row_rpw<L>());
            if (inv) k_rows_fft<L, true><<<grid, row_nt<L>(), 0, s>>>(u, (int)ny, ctx->twiddle);
            else k_rows_fft<L, false><<<grid, row_nt<L>(), 0, s>>>(u, (int)ny, ctx->twiddle);
        });
    }
    OLX_TRY(check_launch("fft rows"));
    {
        TimedLaunch tl(ctx, "fft_cols");
        OLX_LEN_SWITCH(ilog2(ny), {
            const int grid = (int)((nx + col_cb<L>() - 1) / col_cb<L>());
            if (inv) k_cols_fft<L, true><<<grid, col_nt<L>(), 0, s>>>(u, (int)nx, ctx->twiddle);
            else k_cols_fft<L, false><<<grid, col_nt<L>(), 0, s>>>(u, (int)nx, ctx->twiddle);
        });
    }
    return check_launch("fft cols");
}
