// fft_core.hpp -- LDS-staged Stockham FFT building blocks for gfx950.
//
// A length-L transform (L = 2^p, 8 <= L <= 16384) is carried by T = L/E threads,
// each holding E = min(16, L) complex values in registers.  Stages use radix 16
// and one remainder radix 2, 4 or 8, placed first or last (the plan's ORDER);
// between stages the data goes through LDS (one float2 of padding every 16
// elements against bank conflicts).  The first stage loads straight from HBM and
// the last stage stores straight to HBM, so the caller can fuse zero-padding
// (statically skipped loads), pointwise factors and cropping (statically skipped
// stores) into them.  Every helper takes the plan type P.
//
// Stockham step for stage s (radix R, NS = product of earlier radices):
//   butterfly j in [0, L/R):  v[r] = x[j + r L/R],  k = j mod NS,
//   v[r] *= W_{NS R}^{k r},   v = DFT_R(v),   y[(j/NS) NS R + k + r NS] = v[r]
// which leaves y in natural order after the last stage.
#pragma once
#include <hip/hip_runtime.h>

namespace olx {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 mul_mi(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// cos / sin of 2 pi m / 16 (forward twiddle is cos - i sin)
__device__ constexpr float kCos16[16] = {
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
    0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
    0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};
__device__ constexpr float kSin16[16] = {
    0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f,
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
    0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f};

// multiply by W_16^m (forward e^{-2 pi i m/16}, inverse conjugate); m compile-time
template <int M, bool INV>
__device__ __forceinline__ float2 tw16(float2 a) {
    constexpr int m = M & 15;
    if constexpr (m == 0) return a;
    else if constexpr (m == 4) return mul_mi<INV>(a);
    else if constexpr (m == 8) return make_float2(-a.x, -a.y);
    else if constexpr (m == 12) return mul_mi<!INV>(a);
    else {
        constexpr float c = kCos16[m];
        constexpr float s = INV ? kSin16[m] : -kSin16[m];
        return make_float2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

template <bool INV>
__device__ __forceinline__ void dft2(float2& a0, float2& a1) {
    float2 t = a0;
    a0 = cadd(t, a1);
    a1 = csub(t, a1);
}

template <bool INV>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
    float2 t2 = cadd(a1, a3), t3 = mul_mi<INV>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

// in-place DFT of R values, natural-order output
template <int R, bool INV>
__device__ __forceinline__ void dft(float2* v) {
    if constexpr (R == 2) {
        dft2<INV>(v[0], v[1]);
    } else if constexpr (R == 4) {
        dft4<INV>(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        // R = 4 x 2: n = 2 n1 + n2, k = k1 + 4 k2
        float2 b0 = v[0], b1 = v[2], b2 = v[4], b3 = v[6];  // n2 = 0
        float2 c0 = v[1], c1 = v[3], c2 = v[5], c3 = v[7];  // n2 = 1
        dft4<INV>(b0, b1, b2, b3);
        dft4<INV>(c0, c1, c2, c3);
        c1 = tw16<2, INV>(c1);
        c2 = tw16<4, INV>(c2);
        c3 = tw16<6, INV>(c3);
        v[0] = cadd(b0, c0); v[4] = csub(b0, c0);
        v[1] = cadd(b1, c1); v[5] = csub(b1, c1);
        v[2] = cadd(b2, c2); v[6] = csub(b2, c2);
        v[3] = cadd(b3, c3); v[7] = csub(b3, c3);
    } else if constexpr (R == 16) {
        // R = 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2
        float2 b[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; n2++) {
            b[n2][0] = v[n2];
            b[n2][1] = v[4 + n2];
            b[n2][2] = v[8 + n2];
            b[n2][3] = v[12 + n2];
            dft4<INV>(b[n2][0], b[n2][1], b[n2][2], b[n2][3]);
        }
        b[1][1] = tw16<1, INV>(b[1][1]);
        b[1][2] = tw16<2, INV>(b[1][2]);
        b[1][3] = tw16<3, INV>(b[1][3]);
        b[2][1] = tw16<2, INV>(b[2][1]);
        b[2][2] = tw16<4, INV>(b[2][2]);
        b[2][3] = tw16<6, INV>(b[2][3]);
        b[3][1] = tw16<3, INV>(b[3][1]);
        b[3][2] = tw16<6, INV>(b[3][2]);
        b[3][3] = tw16<9, INV>(b[3][3]);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            float2 a0 = b[0][k1], a1 = b[1][k1], a2 = b[2][k1], a3 = b[3][k1];
            dft4<INV>(a0, a1, a2, a3);
            v[k1] = a0;
            v[k1 + 4] = a1;
            v[k1 + 8] = a2;
            v[k1 + 12] = a3;
        }
    }
}

// ---------------------------------------------------------------- plan
// ORDER = REM_FIRST: remainder radix first.  With centred 2x zero padding its
// butterflies see one nonzero input each, and the final radix-16 stage of a
// cropped transform has whole outputs that are never stored (dead code).
// ORDER = REM_LAST: remainder radix last.  The first radix-16 stage then needs no
// twiddles (NS = 1) and sees half-zero input under padding, and a cropped last
// radix-2 stage keeps one output per butterfly.
enum { REM_FIRST = 0, REM_LAST = 1 };

template <int L_, int ORDER = REM_FIRST>
struct Plan {
    static constexpr int L = L_;
    static_assert(L >= 8 && L <= 16384 && (L & (L - 1)) == 0, "unsupported length");
    static constexpr int LOG = __builtin_ctz(L);
    static constexpr int E = L < 16 ? L : 16;  // values per thread
    static constexpr int T = L / E;            // threads per transform
    static constexpr int N16 = L < 16 ? 0 : LOG / 4;
    static constexpr int REM = L < 16 ? L : (1 << (LOG % 4));
    static constexpr int NST = N16 + (REM > 1 ? 1 : 0);
    static constexpr int LP = L + L / 16;  // padded LDS length (the largest padding, lpadn<4>)
    __host__ __device__ static constexpr int radix(int s) {
        if (REM == 1) return 16;
        if (L < 16) return L;
        return ORDER == REM_FIRST ? (s == 0 ? REM : 16) : (s == NST - 1 ? REM : 16);
    }
    __host__ __device__ static constexpr int ns(int s) {
        int p = 1;
        for (int i = 0; i < s; i++) p *= radix(i);
        return p;
    }
};

__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }  // lpadn<4>

// w1, w2, w4, w8 of a radix-16 stage (W_{16 NS}^{k}, ^{2k}, ^{4k}, ^{8k}) from the stage
// table after the 16384-entry twiddle table: entry 4 (NS - 1) + 4 k + i, two contiguous
// 16-byte loads per lane.  The table's entries are bitwise the twiddle table's at
// index 2^i k 16384 / (16 NS); read from the twiddle table, lanes of one wave touched
// up to 64 cache lines per load.
template <int NS>
__device__ __forceinline__ void stage16_twiddles(int k, const float2* __restrict__ tw, float2* w) {
    const float4* st = reinterpret_cast<const float4*>(tw + 16384 + 4 * (NS - 1)) + 2 * k;
    const float4 a = st[0], b = st[1];
    w[0] = make_float2(a.x, a.y);
    w[1] = make_float2(a.z, a.w);
    w[2] = make_float2(b.x, b.y);
    w[3] = make_float2(b.z, b.w);
}


// twiddles W_{NS R}^{k r}, r = 1..R-1, from the length-TWN table (forward sign)
template <int R, bool INV>
__device__ __forceinline__ void apply_twiddles(float2* v, int base, const float2* __restrict__ tw) {
    if constexpr (R == 2) {
        float2 w1 = tw[base];
        if (INV) w1 = cconj(w1);
        v[1] = cmul(v[1], w1);
    } else if constexpr (R == 4) {
        float2 w1 = tw[base], w2 = tw[2 * base];
        if (INV) { w1 = cconj(w1); w2 = cconj(w2); }
        v[1] = cmul(v[1], w1);
        v[2] = cmul(v[2], w2);
        v[3] = cmul(v[3], cmul(w1, w2));
    } else {
        // w1, w2, w4, w8 from the table; the others as products, applied as soon as
        // they exist so that few twiddles are live at once
        float2 w1 = tw[base], w2 = tw[2 * base], w4 = tw[4 * base];
        if (INV) { w1 = cconj(w1); w2 = cconj(w2); w4 = cconj(w4); }
        v[1] = cmul(v[1], w1);
        v[2] = cmul(v[2], w2);
        v[4] = cmul(v[4], w4);
        const float2 w3 = cmul(w1, w2);
        v[3] = cmul(v[3], w3);
        const float2 w5 = cmul(w1, w4);
        v[5] = cmul(v[5], w5);
        const float2 w6 = cmul(w2, w4);
        v[6] = cmul(v[6], w6);
        const float2 w7 = cmul(w3, w4);
        v[7] = cmul(v[7], w7);
        if constexpr (R == 16) {
            float2 w8 = tw[8 * base];
            if (INV) w8 = cconj(w8);
            v[8] = cmul(v[8], w8);
            v[9] = cmul(v[9], cmul(w1, w8));
            v[10] = cmul(v[10], cmul(w2, w8));
            v[11] = cmul(v[11], cmul(w3, w8));
            v[12] = cmul(v[12], cmul(w4, w8));
            v[13] = cmul(v[13], cmul(w5, w8));
            v[14] = cmul(v[14], cmul(w6, w8));
            v[15] = cmul(v[15], cmul(w7, w8));
        }
    }
}

// the same twiddles applied to two value sets of one thread (same k, same W)
template <int R, bool INV>
__device__ __forceinline__ void apply_twiddles2(float2* a, float2* b, int base,
                                                const float2* __restrict__ tw) {
    if constexpr (R == 2) {
        float2 w1 = tw[base];
        if (INV) w1 = cconj(w1);
        a[1] = cmul(a[1], w1);
        b[1] = cmul(b[1], w1);
    } else if constexpr (R == 4) {
        float2 w1 = tw[base], w2 = tw[2 * base];
        if (INV) { w1 = cconj(w1); w2 = cconj(w2); }
        const float2 w3 = cmul(w1, w2);
        a[1] = cmul(a[1], w1); b[1] = cmul(b[1], w1);
        a[2] = cmul(a[2], w2); b[2] = cmul(b[2], w2);
        a[3] = cmul(a[3], w3); b[3] = cmul(b[3], w3);
    } else {
        float2 w1 = tw[base], w2 = tw[2 * base], w4 = tw[4 * base];
        if (INV) { w1 = cconj(w1); w2 = cconj(w2); w4 = cconj(w4); }
        a[1] = cmul(a[1], w1); b[1] = cmul(b[1], w1);
        a[2] = cmul(a[2], w2); b[2] = cmul(b[2], w2);
        a[4] = cmul(a[4], w4); b[4] = cmul(b[4], w4);
        const float2 w3 = cmul(w1, w2);
        a[3] = cmul(a[3], w3); b[3] = cmul(b[3], w3);
        const float2 w5 = cmul(w1, w4);
        a[5] = cmul(a[5], w5); b[5] = cmul(b[5], w5);
        const float2 w6 = cmul(w2, w4);
        a[6] = cmul(a[6], w6); b[6] = cmul(b[6], w6);
        const float2 w7 = cmul(w3, w4);
        a[7] = cmul(a[7], w7); b[7] = cmul(b[7], w7);
        if constexpr (R == 16) {
            float2 w8 = tw[8 * base];
            if (INV) w8 = cconj(w8);
            a[8] = cmul(a[8], w8); b[8] = cmul(b[8], w8);
            float2 w = cmul(w1, w8);
            a[9] = cmul(a[9], w); b[9] = cmul(b[9], w);
            w = cmul(w2, w8);
            a[10] = cmul(a[10], w); b[10] = cmul(b[10], w);
            w = cmul(w3, w8);
            a[11] = cmul(a[11], w); b[11] = cmul(b[11], w);
            w = cmul(w4, w8);
            a[12] = cmul(a[12], w); b[12] = cmul(b[12], w);
            w = cmul(w5, w8);
            a[13] = cmul(a[13], w); b[13] = cmul(b[13], w);
            w = cmul(w6, w8);
            a[14] = cmul(a[14], w); b[14] = cmul(b[14], w);
            w = cmul(w7, w8);
            a[15] = cmul(a[15], w); b[15] = cmul(b[15], w);
        }
    }
}

// the products of apply_twiddles(2) for R = 16 from w1, w2, w4, w8 already loaded
template <bool INV, bool TWO>
__device__ __forceinline__ void apply16_from(float2* a, float2* b, const float2* l) {
    float2 w1 = l[0], w2 = l[1], w4 = l[2], w8 = l[3];
    if (INV) { w1 = cconj(w1); w2 = cconj(w2); w4 = cconj(w4); w8 = cconj(w8); }
    a[1] = cmul(a[1], w1); if constexpr (TWO) b[1] = cmul(b[1], w1);
    a[2] = cmul(a[2], w2); if constexpr (TWO) b[2] = cmul(b[2], w2);
    a[4] = cmul(a[4], w4); if constexpr (TWO) b[4] = cmul(b[4], w4);
    const float2 w3 = cmul(w1, w2);
    a[3] = cmul(a[3], w3); if constexpr (TWO) b[3] = cmul(b[3], w3);
    const float2 w5 = cmul(w1, w4);
    a[5] = cmul(a[5], w5); if constexpr (TWO) b[5] = cmul(b[5], w5);
    const float2 w6 = cmul(w2, w4);
    a[6] = cmul(a[6], w6); if constexpr (TWO) b[6] = cmul(b[6], w6);
    const float2 w7 = cmul(w3, w4);
    a[7] = cmul(a[7], w7); if constexpr (TWO) b[7] = cmul(b[7], w7);
    a[8] = cmul(a[8], w8); if constexpr (TWO) b[8] = cmul(b[8], w8);
    float2 w = cmul(w1, w8);
    a[9] = cmul(a[9], w); if constexpr (TWO) b[9] = cmul(b[9], w);
    w = cmul(w2, w8);
    a[10] = cmul(a[10], w); if constexpr (TWO) b[10] = cmul(b[10], w);
    w = cmul(w3, w8);
    a[11] = cmul(a[11], w); if constexpr (TWO) b[11] = cmul(b[11], w);
    w = cmul(w4, w8);
    a[12] = cmul(a[12], w); if constexpr (TWO) b[12] = cmul(b[12], w);
    w = cmul(w5, w8);
    a[13] = cmul(a[13], w); if constexpr (TWO) b[13] = cmul(b[13], w);
    w = cmul(w6, w8);
    a[14] = cmul(a[14], w); if constexpr (TWO) b[14] = cmul(b[14], w);
    w = cmul(w7, w8);
    a[15] = cmul(a[15], w); if constexpr (TWO) b[15] = cmul(b[15], w);
}

// twiddle + DFT for every butterfly of stage S held by thread t
template <class P, int S, bool INV, int TWLEN>
__device__ __forceinline__ void stage_compute(float2* v, int t, const float2* __restrict__ tw) {
    constexpr int R = P::radix(S), NS = P::ns(S), NB = P::E / R;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        if constexpr (NS > 1) {
            int j = t + P::T * b;
            int k = j & (NS - 1);
            if constexpr (R == 16 && TWLEN == 16384 && NS <= 1024) {
                float2 l[4];
                stage16_twiddles<NS>(k, tw, l);
                apply16_from<INV, false>(v + b * R, nullptr, l);
            } else {
                apply_twiddles<R, INV>(v + b * R, k * (TWLEN / (NS * R)), tw);
            }
        }
        dft<R, INV>(v + b * R);
    }
}

// stage S on two value sets of the same thread, twiddles computed once
template <class P, int S, bool INV, int TWLEN>
__device__ __forceinline__ void stage_compute2(float2* va, float2* vb, int t,
                                               const float2* __restrict__ tw) {
    constexpr int R = P::radix(S), NS = P::ns(S), NB = P::E / R;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        if constexpr (NS > 1) {
            int j = t + P::T * b;
            int k = j & (NS - 1);
            if constexpr (R == 16 && TWLEN == 16384 && NS <= 1024) {
                float2 l[4];
                stage16_twiddles<NS>(k, tw, l);
                apply16_from<INV, true>(va + b * R, vb + b * R, l);
            } else {
                apply_twiddles2<R, INV>(va + b * R, vb + b * R, k * (TWLEN / (NS * R)), tw);
            }
        }
        dft<R, INV>(va + b * R);
        dft<R, INV>(vb + b * R);
    }
}

// LDS positions: element i of a transform sits at lpadn<SH>(i) = i + (i >> SH), one
// float2 of padding every 2^SH elements (lpad: SH = 4).  lpadn(x + y) = lpadn(x) +
// lpadn(y) whenever x % 2^SH + y % 2^SH < 2^SH, which holds for every split below
// when T (threads per transform) is a multiple of 2^SH: each access is then one
// per-thread base plus a compile-time offset, so the compiler needs a single address
// register per stage instead of one per value.
//
// The period against bank conflicts: gfx950's LDS serves a b64 access 32 lanes per
// cycle from 64 banks, so the 32 lanes of a half-wave must hit 32 distinct 8-byte
// bank pairs.  With one transform per LDS round (STRIDE 1), the contiguous reads of
// every stage are conflict-free only with SH = 5 (with SH = 4 lanes 16-31 are shifted
// by one and wrap onto lanes 0-1), and only the second stage's writes then conflict,
// 2-way; modelled over an 8192-point transform, 1792 cycles per half-wave group
// against 2560 (SH = 4) and 1536 ideal.  With transforms interleaved (STRIDE >= 2),
// SH = 4 is the better choice.
template <int STRIDE>
constexpr int pad_shift() {
    return STRIDE == 1 ? 5 : 4;
}
template <int SH>
__device__ __forceinline__ constexpr int lpadn(int i) {
    return i + (i >> SH);
}

// stage S output (b, r) goes to position (t - k) R + k + r NS + T R b, k = t mod NS
// (T is a multiple of NS, so k does not depend on b)
template <class P, int S, int SH = 4>
__device__ __forceinline__ int store_base(int t) {
    constexpr int R = P::radix(S), NS = P::ns(S);
    const int k = t & (NS - 1);
    return lpadn<SH>((t - k) * R) + lpadn<SH>(k);
}
template <class P, int S, int SH = 4>
__device__ __forceinline__ constexpr int store_off(int b, int r) {
    constexpr int R = P::radix(S), NS = P::ns(S);
    return lpadn<SH>(r * NS) + lpadn<SH>(P::T * R * b);
}

// store stage S outputs to LDS (element i of this transform at off + lpadn<SH>(i)*STRIDE)
template <class P, int S, int STRIDE, int SH = 4>
__device__ __forceinline__ void stage_store_lds(const float2* v, int t, float2* lds, int off) {
    constexpr int R = P::radix(S), NS = P::ns(S), NB = P::E / R;
    if constexpr (P::T % (1 << SH) == 0) {
        float2* q = lds + off + store_base<P, S, SH>(t) * STRIDE;
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int r = 0; r < R; r++) q[store_off<P, S, SH>(b, r) * STRIDE] = v[b * R + r];
    } else {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            int j = t + P::T * b;
            int k = j & (NS - 1);
            int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) lds[off + lpadn<SH>(base + r * NS) * STRIDE] = v[b * R + r];
        }
    }
}

// load stage S inputs from LDS: (b, r) reads element t + T b + r L/R
template <class P, int S, int STRIDE, int SH = 4>
__device__ __forceinline__ void stage_load_lds(float2* v, int t, const float2* lds, int off) {
    constexpr int R = P::radix(S), NB = P::E / R;
    if constexpr (P::T % (1 << SH) == 0) {
        const float2* q = lds + off + lpadn<SH>(t) * STRIDE;
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int r = 0; r < R; r++) v[b * R + r] = q[lpadn<SH>(P::T * b + r * (P::L / R)) * STRIDE];
    } else {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            int j = t + P::T * b;
#pragma unroll
            for (int r = 0; r < R; r++) v[b * R + r] = lds[off + lpadn<SH>(j + r * (P::L / R)) * STRIDE];
        }
    }
}

// Run stages 0..NST-1 on values already loaded for stage 0; intermediate
// exchanges go through LDS.  On return v holds the LAST stage's outputs:
// v[b R + r] is element last_index(t, b, r) of the natural-order result.
template <class P, bool INV, int TWLEN, int STRIDE, int S = 0>
__device__ __forceinline__ void run_stages(float2* v, int t, float2* lds, int off,
                                           const float2* __restrict__ tw) {
    stage_compute<P, S, INV, TWLEN>(v, t, tw);
    if constexpr (S + 1 < P::NST) {
        stage_store_lds<P, S, STRIDE>(v, t, lds, off);
        __syncthreads();
        stage_load_lds<P, S + 1, STRIDE>(v, t, lds, off);
        __syncthreads();
        run_stages<P, INV, TWLEN, STRIDE, S + 1>(v, t, lds, off, tw);
    }
}

// Two transforms per thread (value sets a and b), each laid out exactly as in
// run_stages.  Both sets are computed together (their twiddles are the same), and
// LDS holds one set at a time, so every exchange is done in two rounds.  On
// return both sets hold their last-stage outputs.
template <class P, bool INV, int TWLEN, int STRIDE, int S = 0>
__device__ __forceinline__ void run_stages_pair(float2* va, float2* vb, int t, float2* lds,
                                                int off, const float2* __restrict__ tw) {
    // an opaque copy of t per call: twiddle offsets are recomputed (one shift each)
    // rather than kept live, or spilled, between identical uses
    int tt = t;
    asm volatile("" : "+v"(tt));
    if constexpr (S == 0) stage_compute2<P, 0, INV, TWLEN>(va, vb, tt, tw);
    if constexpr (S + 1 < P::NST) {
        stage_store_lds<P, S, STRIDE, pad_shift<STRIDE>()>(va, t, lds, off);
        __syncthreads();
        stage_load_lds<P, S + 1, STRIDE, pad_shift<STRIDE>()>(va, t, lds, off);
        __syncthreads();
        stage_store_lds<P, S, STRIDE, pad_shift<STRIDE>()>(vb, t, lds, off);
        __syncthreads();
        // set b's reads overlap the next stage's twiddles and set a's arithmetic
        stage_load_lds<P, S + 1, STRIDE, pad_shift<STRIDE>()>(vb, t, lds, off);
        stage_compute2<P, S + 1, INV, TWLEN>(va, vb, tt, tw);
        __syncthreads();
        run_stages_pair<P, INV, TWLEN, STRIDE, S + 1>(va, vb, t, lds, off, tw);
    }
}

// ---------------------------------------------------------------- pipelined pair
// Stage S twiddles of one thread, split in two steps so that the table loads can be
// issued an exchange before their use: TwBase holds the loaded w1, w2, w4, w8 of each
// butterfly, TwFull the R twiddles formed from them (the products of apply_twiddles2,
// in the same order, so the results are bitwise those of run_stages_pair).
template <class P, int S, bool INV, int TWLEN>
struct TwBase {
    static constexpr int R = P::radix(S), NS = P::ns(S), NB = P::E / R;
    static constexpr int NL = NS > 1 ? (R == 2 ? 1 : R == 4 ? 2 : R == 8 ? 3 : 4) : 0;
    float2 l[NB][NL > 0 ? NL : 1];
    __device__ __forceinline__ void load(int t, const float2* __restrict__ tw) {
        if constexpr (NL > 0) {
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int k = (t + P::T * b) & (NS - 1);
                if constexpr (R == 16 && TWLEN == 16384 && NS <= 1024) {
                    stage16_twiddles<NS>(k, tw, l[b]);
                } else {
                    const int base = k * (TWLEN / (NS * R));
#pragma unroll
                    for (int i = 0; i < NL; i++) l[b][i] = tw[(1 << i) * base];
                }
            }
        }
    }
};

template <class P, int S, bool INV, int TWLEN>
struct TwFull {
    static constexpr int R = P::radix(S), NS = P::ns(S), NB = P::E / R;
    float2 w[NB][R];  // w[b][0] = 1 is not stored or applied
    __device__ __forceinline__ void expand(const TwBase<P, S, INV, TWLEN>& tb) {
        if constexpr (NS > 1) {
#pragma unroll
            for (int b = 0; b < NB; b++) {
                float2 w1 = tb.l[b][0];
                if (INV) w1 = cconj(w1);
                w[b][1] = w1;
                if constexpr (R >= 4) {
                    float2 w2 = tb.l[b][1];
                    if (INV) w2 = cconj(w2);
                    w[b][2] = w2;
                    w[b][3] = cmul(w1, w2);
                }
                if constexpr (R >= 8) {
                    float2 w4 = tb.l[b][2];
                    if (INV) w4 = cconj(w4);
                    w[b][4] = w4;
                    w[b][5] = cmul(w1, w4);
                    w[b][6] = cmul(w[b][2], w4);
                    w[b][7] = cmul(w[b][3], w4);
                }
                if constexpr (R == 16) {
                    float2 w8 = tb.l[b][3];
                    if (INV) w8 = cconj(w8);
                    w[b][8] = w8;
#pragma unroll
                    for (int r = 1; r < 8; r++) w[b][8 + r] = cmul(w[b][r], w8);
                }
            }
        }
    }
    // twiddles then DFT on every butterfly of one value set
    __device__ __forceinline__ void compute(float2* v) const {
#pragma unroll
        for (int b = 0; b < NB; b++) {
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) v[b * R + r] = cmul(v[b * R + r], w[b][r]);
            }
            dft<R, INV>(v + b * R);
        }
    }
};

// run_stages_pair with the exchanges pipelined.  LDS holds one value set at a
// time, so each exchange is two rounds (write, barrier, read, barrier); here the
// set not in LDS computes during the other set's writes:
//   write a(S) | compute b(S) ; barrier ; read a(S+1) ; barrier ;
//   write b(S) | compute a(S+1) ; barrier ; read b(S+1) ; barrier ; ...
// and stage S+1's twiddle loads are issued at the start of the exchange before it.
// The twiddles of a stage are formed once and kept for the second set.
// sched_barrier pins each computation on its side of the work-group barrier (the
// scheduler otherwise sinks the arithmetic below it, serialising LDS and VALU).
// Computing half of a set on each side of the read barrier instead measured slower:
// the read rounds leave the VALU to the CU's other work-group.
// On entry both sets hold stage-0 inputs; on return both hold last-stage outputs,
// and the last LDS reads are followed by a barrier.
template <class P, bool INV, int TWLEN, int STRIDE, int S>
__device__ __forceinline__ void pipe_rest(float2* va, float2* vb, int t, int tt, float2* lds,
                                          int off, const float2* __restrict__ tw,
                                          const TwFull<P, S, INV, TWLEN>& ws) {
    if constexpr (S + 1 < P::NST) {
        TwBase<P, S + 1, INV, TWLEN> tb;
        tb.load(tt, tw);
        stage_store_lds<P, S, STRIDE, pad_shift<STRIDE>()>(va, t, lds, off);
        ws.compute(vb);
        __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        stage_load_lds<P, S + 1, STRIDE, pad_shift<STRIDE>()>(va, t, lds, off);
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        stage_store_lds<P, S, STRIDE, pad_shift<STRIDE>()>(vb, t, lds, off);
        TwFull<P, S + 1, INV, TWLEN> wn;
        wn.expand(tb);
        wn.compute(va);
        __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        stage_load_lds<P, S + 1, STRIDE, pad_shift<STRIDE>()>(vb, t, lds, off);
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);
        pipe_rest<P, INV, TWLEN, STRIDE, S + 1>(va, vb, t, tt, lds, off, tw, wn);
    } else {
        ws.compute(vb);
    }
}

template <class P, bool INV, int TWLEN, int STRIDE>
__device__ __forceinline__ void run_stages_pipe(float2* va, float2* vb, int t, float2* lds,
                                                int off, const float2* __restrict__ tw) {
    int tt = t;
    asm volatile("" : "+v"(tt));
    TwBase<P, 0, INV, TWLEN> tb;
    tb.load(tt, tw);
    TwFull<P, 0, INV, TWLEN> w0;
    w0.expand(tb);
    w0.compute(va);
    pipe_rest<P, INV, TWLEN, STRIDE, 0>(va, vb, t, tt, lds, off, tw, w0);
}

// element index held in v[b R + r] after run_stages (last stage: NS R = L, so the
// output index is j + r NS with j = t + T b < NS)
template <class P>
__device__ __forceinline__ constexpr int last_index(int t, int b, int r) {
    constexpr int R = P::radix(P::NST - 1);
    return t + P::T * b + r * (P::L / R);
}

// element index read by v[b R0 + r] in stage 0
template <class P>
__device__ __forceinline__ constexpr int first_index(int t, int b, int r) {
    constexpr int R = P::radix(0);
    return t + P::T * b + r * (P::L / R);
}

// Both indices above are t + T q with q = b + r E/R: a spectrum left by a forward
// transform of plan PF feeds an inverse transform of plan PI without an exchange,
// stage-0 slot b R0 + r of PI being PF's last-stage slot spec_slot(b, r).
template <class PF, class PI>
__device__ __forceinline__ constexpr int spec_slot(int b, int r) {
    constexpr int R0 = PI::radix(0), RL = PF::radix(PF::NST - 1);
    const int q = b + r * (PI::E / R0);
    const int nbl = PF::E / RL;  // butterflies per thread in PF's last stage
    return (q % nbl) * RL + q / nbl;
}

// For a transform whose input is nonzero only on [L/4, 3L/4) (centred 2x zero
// padding): whether stage-0 element (b, r) can be nonzero.  Thread t of
// butterfly block b reads j + r L/R0 with j = t + T b, t < T, and L/4 is a
// multiple of T, so the answer does not depend on t.
template <class P>
__device__ __forceinline__ constexpr bool first_nonzero(int b, int r) {
    const int lo = first_index<P>(0, b, r);
    return lo >= P::L / 4 && lo + P::T <= 3 * P::L / 4;
}

// Whether last-stage element (b, r) falls in the kept centre [L/4, 3L/4) of a
// cropped transform (same argument).
template <class P>
__device__ __forceinline__ constexpr bool last_kept(int b, int r) {
    const int lo = last_index<P>(0, b, r);
    return lo >= P::L / 4 && lo + P::T <= 3 * P::L / 4;
}

}  // namespace olx
