// fft_core.hpp -- LDS-staged Stockham FFT building blocks for gfx950.
//
// A length-L transform (L = 2^p, 8 <= L <= 16384) is carried by T = L/E threads,
// each holding E = min(16, L) complex values in registers.  Stages use radix 16,
// the remainder radix 2, 4 or 8 first; between stages the data goes through LDS
// (one float2 of padding every 16 elements against bank conflicts).  The first
// stage loads straight from HBM and the last stage stores straight to HBM, so the
// caller can fuse zero-padding (statically skipped loads), pointwise factors and
// cropping (statically skipped stores) into them.
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
template <int L>
struct Plan {
    static_assert(L >= 8 && L <= 16384 && (L & (L - 1)) == 0, "unsupported length");
    static constexpr int LOG = __builtin_ctz(L);
    static constexpr int E = L < 16 ? L : 16;  // values per thread
    static constexpr int T = L / E;            // threads per transform
    static constexpr int N16 = L < 16 ? 0 : LOG / 4;
    static constexpr int REM = L < 16 ? L : (1 << (LOG % 4));
    static constexpr int NST = N16 + (REM > 1 ? 1 : 0);
    static constexpr int LP = L + L / 16;  // padded LDS length
    // remainder radix first: with centred 2x padding its butterflies see one
    // nonzero input each, and the final radix-16 stage of a cropped inverse
    // transform has whole outputs that are never stored (dead code)
    __host__ __device__ static constexpr int radix(int s) {
        return REM > 1 ? (s == 0 ? REM : 16) : 16;
    }
    __host__ __device__ static constexpr int ns(int s) {
        int p = 1;
        for (int i = 0; i < s; i++) p *= radix(i);
        return p;
    }
};

__device__ __forceinline__ int lpad(int i) { return i + (i >> 4); }

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

// twiddle + DFT for every butterfly of stage S held by thread t
template <int L, int S, bool INV, int TWLEN>
__device__ __forceinline__ void stage_compute(float2* v, int t, const float2* __restrict__ tw) {
    using P = Plan<L>;
    constexpr int R = P::radix(S), NS = P::ns(S), NB = P::E / R;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        if constexpr (NS > 1) {
            int j = t + P::T * b;
            int k = j & (NS - 1);
            apply_twiddles<R, INV>(v + b * R, k * (TWLEN / (NS * R)), tw);
        }
        dft<R, INV>(v + b * R);
    }
}

// store stage S outputs to LDS (element i of this transform at off + lpad(i)*STRIDE)
template <int L, int S, int STRIDE>
__device__ __forceinline__ void stage_store_lds(const float2* v, int t, float2* lds, int off) {
    using P = Plan<L>;
    constexpr int R = P::radix(S), NS = P::ns(S), NB = P::E / R;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        int j = t + P::T * b;
        int k = j & (NS - 1);
        int base = (j - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; r++) lds[off + lpad(base + r * NS) * STRIDE] = v[b * R + r];
    }
}

// load stage S inputs from LDS
template <int L, int S, int STRIDE>
__device__ __forceinline__ void stage_load_lds(float2* v, int t, const float2* lds, int off) {
    using P = Plan<L>;
    constexpr int R = P::radix(S), NB = P::E / R;
#pragma unroll
    for (int b = 0; b < NB; b++) {
        int j = t + P::T * b;
#pragma unroll
        for (int r = 0; r < R; r++) v[b * R + r] = lds[off + lpad(j + r * (L / R)) * STRIDE];
    }
}

// Run stages 0..NST-1 on values already loaded for stage 0; intermediate
// exchanges go through LDS.  On return v holds the LAST stage's outputs:
// v[b R + r] is element (j + r L/R) of the natural-order result, j = t + T b,
// R = last radix (the caller stores them, see last_index).
template <int L, bool INV, int TWLEN, int STRIDE, int S = 0>
__device__ __forceinline__ void run_stages(float2* v, int t, float2* lds, int off,
                                           const float2* __restrict__ tw) {
    using P = Plan<L>;
    stage_compute<L, S, INV, TWLEN>(v, t, tw);
    if constexpr (S + 1 < P::NST) {
        stage_store_lds<L, S, STRIDE>(v, t, lds, off);
        __syncthreads();
        stage_load_lds<L, S + 1, STRIDE>(v, t, lds, off);
        __syncthreads();
        run_stages<L, INV, TWLEN, STRIDE, S + 1>(v, t, lds, off, tw);
    }
}

// element index held in v[b R + r] after run_stages (last stage: NS R = L, so the
// output index is j + r NS with j < NS)
template <int L>
__device__ __forceinline__ int last_index(int t, int b, int r) {
    using P = Plan<L>;
    constexpr int R = P::radix(P::NST - 1);
    constexpr int NS = L / R;
    return t + P::T * b + r * NS;
}

// element index read by v[b R0 + r] in stage 0
template <int L>
__device__ __forceinline__ int first_index(int t, int b, int r) {
    using P = Plan<L>;
    constexpr int R = P::radix(0);
    return t + P::T * b + r * (L / R);
}

// For a transform whose input is nonzero only on [L/4, 3L/4) (centred 2x zero
// padding): whether stage-0 element (b, r) can be nonzero.  Thread t of
// butterfly block b reads j + r L/R0 with j = t + T b, t < T, and L/4 is a
// multiple of T, so the answer does not depend on t.
template <int L>
__device__ __forceinline__ constexpr bool first_nonzero(int b, int r) {
    constexpr int R0 = Plan<L>::radix(0), T = Plan<L>::T;
    const int lo = T * b + r * (L / R0);
    return lo >= L / 4 && lo + T <= 3 * L / 4;
}

// Whether last-stage element (b, r) falls in the kept centre [L/4, 3L/4) of a
// cropped transform (same argument: the output index is j + r L/RL).
template <int L>
__device__ __forceinline__ constexpr bool last_kept(int b, int r) {
    constexpr int RL = Plan<L>::radix(Plan<L>::NST - 1), T = Plan<L>::T;
    const int lo = T * b + r * (L / RL);
    return lo >= L / 4 && lo + T <= 3 * L / 4;
}

}  // namespace olx
