// olx_prop64.hip -- complex128 propagation and 2-D FFT in FP64 throughout: the
// reference's fftw3 (double) path, ol::AsmProp / FresnelProp / FresnelPropBandLimit
// on std::complex<double> (include/olprop.h:125-170, 227-248; FFTW plans
// include/olutils.h:47-64) and ol::fft / ifft on complex<double> (:19-66).
//
// The same three passes as the complex64 mixed-radix path over a row-major
// intermediate W of ny x Mx complex128 values: A pads and transforms rows, B
// transforms columns (ny of My values nonzero), multiplies by the transfer function
// and transforms back, keeping the centre rows, C transforms rows back, crops and
// scales by 1/(My Mx).  Each transform is a runtime-planned Stockham FFT in LDS
// (radix 16 with a remainder 8 / 4 / 2, then 3, 5, 7, 11, 13), twiddles from an FP64
// table of W_L.  A transform is held in LDS whole, so padded lengths are at most
// 8192 (sides up to 4096; the plain FFT takes sides up to 8192).
//
// Transfer functions: ASM with H = exp(2 pi i z sqrt(1/lambda^2 - u^2 - v^2)) in
// FP64 (phase reduced mod 1 cycle, then sincospi), the band-limit window of
// shiftedAsmTransferF at zero offset (include/olprop.h:250-305), evanescent waves
// cut (0; the reference's sqrt of a negative gives NaN); Fresnel as the product of
// the FP64 spectra of the centred 1-D chirps (the FFT of the reference's
// FresnelResponse is separable), windows as in the complex64 path.
#include <climits>
#include <cmath>
#include <cstring>

#include "olx_common.hpp"

namespace olx {

namespace {

__device__ __forceinline__ double2 dadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 dsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 dmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 dconj(double2 a) { return make_double2(a.x, -a.y); }
template <bool INV>
__device__ __forceinline__ double2 dmul_mi(double2 a) {
    return INV ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

constexpr double kPi64 = 3.141592653589793238462643383279502884;
constexpr double c_sin(double x) {
    while (x > kPi64) x -= 2 * kPi64;
    while (x < -kPi64) x += 2 * kPi64;
    double term = x, sum = x;
    for (int i = 1; i < 40; i++) {
        term *= -x * x / ((2.0 * i) * (2.0 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double c_cos(double x) { return c_sin(x + kPi64 / 2); }

// cos / sin of 2 pi m / R, m in [0, R), evaluated by the compiler; the odd-radix DFT
// indexes it with its unrolled loop counters, so every entry folds to an immediate
// (the series called there directly is not a constant expression and ran at run time)
template <int R>
struct OddRoots64 {
    double c[R], s[R];
    constexpr OddRoots64() : c(), s() {
        for (int m = 0; m < R; m++) {
            c[m] = c_cos(2.0 * kPi64 * m / R);
            s[m] = c_sin(2.0 * kPi64 * m / R);
        }
    }
};
template <int R>
__device__ constexpr OddRoots64<R> kOddRoots64{};

// multiply by W_R^m (forward e^{-2 pi i m / R}); m, R compile-time
template <int M, int R, bool INV>
__device__ __forceinline__ double2 dtw(double2 a) {
    constexpr int m = ((M % R) + R) % R;
    if constexpr (m == 0) return a;
    else if constexpr (4 * m == R) return dmul_mi<INV>(a);
    else if constexpr (2 * m == R) return make_double2(-a.x, -a.y);
    else if constexpr (4 * m == 3 * R) return dmul_mi<!INV>(a);
    else {
        constexpr double c = c_cos(2.0 * kPi64 * m / R);
        constexpr double s = INV ? c_sin(2.0 * kPi64 * m / R) : -c_sin(2.0 * kPi64 * m / R);
        return make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

template <bool INV>
__device__ __forceinline__ void ddft4(double2& a0, double2& a1, double2& a2, double2& a3) {
    const double2 t0 = dadd(a0, a2), t1 = dsub(a0, a2);
    const double2 t2 = dadd(a1, a3), t3 = dmul_mi<INV>(dsub(a1, a3));
    a0 = dadd(t0, t2);
    a2 = dsub(t0, t2);
    a1 = dadd(t1, t3);
    a3 = dsub(t1, t3);
}

// R-point DFT in registers, natural order in and out
template <int R, bool INV>
__device__ __forceinline__ void ddft(double2* v) {
    if constexpr (R == 2) {
        const double2 t = v[0];
        v[0] = dadd(t, v[1]);
        v[1] = dsub(t, v[1]);
    } else if constexpr (R == 4) {
        ddft4<INV>(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8 || R == 16) {
        // R = 4 x Q: n = Q n1 + n2, k = k1 + 4 k2 (Q = 2 or 4)
        constexpr int Q = R / 4;
        double2 b[Q][4];
#pragma unroll
        for (int n2 = 0; n2 < Q; n2++) {
#pragma unroll
            for (int n1 = 0; n1 < 4; n1++) b[n2][n1] = v[Q * n1 + n2];
            ddft4<INV>(b[n2][0], b[n2][1], b[n2][2], b[n2][3]);
        }
#pragma unroll
        for (int n2 = 1; n2 < Q; n2++) {
            b[n2][1] = n2 == 1 ? dtw<1, R, INV>(b[n2][1]) : (n2 == 2 ? dtw<2, R, INV>(b[n2][1]) : dtw<3, R, INV>(b[n2][1]));
            b[n2][2] = n2 == 1 ? dtw<2, R, INV>(b[n2][2]) : (n2 == 2 ? dtw<4, R, INV>(b[n2][2]) : dtw<6, R, INV>(b[n2][2]));
            b[n2][3] = n2 == 1 ? dtw<3, R, INV>(b[n2][3]) : (n2 == 2 ? dtw<6, R, INV>(b[n2][3]) : dtw<9, R, INV>(b[n2][3]));
        }
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            double2 c[Q];
#pragma unroll
            for (int n2 = 0; n2 < Q; n2++) c[n2] = b[n2][k1];
            if constexpr (Q == 2) {
                v[k1] = dadd(c[0], c[1]);
                v[k1 + 4] = dsub(c[0], c[1]);
            } else {
                ddft4<INV>(c[0], c[1], c[2], c[3]);
                v[k1] = c[0];
                v[k1 + 4] = c[1];
                v[k1 + 8] = c[2];
                v[k1 + 12] = c[3];
            }
        }
    } else {
        // odd R: symmetric pairs a_n = x_n + x_{R-n}, b_n = x_n - x_{R-n}
        constexpr int H = (R - 1) / 2;
        double2 a[H], b[H];
        double2 y0 = v[0];
#pragma unroll
        for (int n = 1; n <= H; n++) {
            a[n - 1] = dadd(v[n], v[R - n]);
            b[n - 1] = dsub(v[n], v[R - n]);
            y0 = dadd(y0, a[n - 1]);
        }
        double2 out[R];
#pragma unroll
        for (int k = 1; k <= H; k++) {
            double2 A = v[0], B = make_double2(0.0, 0.0);
#pragma unroll
            for (int n = 1; n <= H; n++) {
                const double c = kOddRoots64<R>.c[(n * k) % R];
                const double s = kOddRoots64<R>.s[(n * k) % R];
                A = make_double2(fma(a[n - 1].x, c, A.x), fma(a[n - 1].y, c, A.y));
                B = make_double2(fma(b[n - 1].x, s, B.x), fma(b[n - 1].y, s, B.y));
            }
            if (INV) {
                out[k] = make_double2(A.x - B.y, A.y + B.x);
                out[R - k] = make_double2(A.x + B.y, A.y - B.x);
            } else {
                out[k] = make_double2(A.x + B.y, A.y - B.x);
                out[R - k] = make_double2(A.x - B.y, A.y + B.x);
            }
        }
        v[0] = y0;
#pragma unroll
        for (int k = 1; k < R; k++) v[k] = out[k];
    }
}

constexpr int ST_MAX = 20;
constexpr int LMAX64 = 8192;  // longest transform held in LDS
constexpr int NT64 = 512;

struct Stage64 {
    int R, ns, lr, tws;
    int off;  // this stage's twiddles in the stage table: k (R - 1) + r - 1 from here
};
struct Plan64 {
    int L, nst;
    const double2* tw;    // W_L^m, m in [0, L)
    const double2* stab;  // per stage (ns > 1): W_L^{k r tws}, k-major, a butterfly's R - 1
                          // twiddles contiguous
    Stage64 st[ST_MAX];
};

__device__ __forceinline__ int lp64(int i) { return i + (i >> 4); }
__host__ __device__ inline int ldsl64(int L) { return L + L / 16 + 1; }

template <int R, bool INV>
__device__ __forceinline__ void stage64(double2* lds, int LP, int nb, const Stage64& S,
                                        const double2* __restrict__ tw, int tid) {
    constexpr int Q = (LMAX64 / R + NT64 - 1) / NT64;
    const int tot = nb * S.lr;
    double2 v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int g = tid + NT64 * q;
        if (g < tot) {
            const int tr = g / S.lr, j = g - tr * S.lr;
            const double2* p = lds + tr * LP;
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = p[lp64(j + r * S.lr)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int g = tid + NT64 * q;
        if (g < tot) {
            const int tr = g / S.lr, j = g - tr * S.lr;
            const int k = j % S.ns;
            if (S.ns > 1) {
                const double2* wk = tw + S.off + k * (R - 1);
#pragma unroll
                for (int r = 1; r < R; r++) {
                    double2 w = wk[r - 1];
                    if (INV) w = dconj(w);
                    v[q][r] = dmul(v[q][r], w);
                }
            }
            ddft<R, INV>(v[q]);
            double2* p = lds + tr * LP;
            const int o = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) p[lp64(o + r * S.ns)] = v[q][r];
        }
    }
    __syncthreads();
}

// inlined into each kernel: as a called function it kept a 156-byte stack frame in
// scratch memory
template <bool INV>
__device__ __forceinline__ void run64(double2* lds, int LP, int nb, const Plan64* __restrict__ P, int tid) {
    const double2* tw = P->stab;
    for (int s = 0; s < P->nst; s++) {
        const Stage64 S = P->st[s];
        switch (S.R) {
            case 2: stage64<2, INV>(lds, LP, nb, S, tw, tid); break;
            case 3: stage64<3, INV>(lds, LP, nb, S, tw, tid); break;
            case 4: stage64<4, INV>(lds, LP, nb, S, tw, tid); break;
            case 5: stage64<5, INV>(lds, LP, nb, S, tw, tid); break;
            case 7: stage64<7, INV>(lds, LP, nb, S, tw, tid); break;
            case 8: stage64<8, INV>(lds, LP, nb, S, tw, tid); break;
            case 9: stage64<9, INV>(lds, LP, nb, S, tw, tid); break;
            case 15: stage64<15, INV>(lds, LP, nb, S, tw, tid); break;
            case 11: stage64<11, INV>(lds, LP, nb, S, tw, tid); break;
            case 13: stage64<13, INV>(lds, LP, nb, S, tw, tid); break;
            default: stage64<16, INV>(lds, LP, nb, S, tw, tid); break;
        }
    }
}

constexpr int LDS64 = LMAX64 + LMAX64 / 16 + 32;

struct Args64 {
    int ny, nx;
    int table;           // 1: H = sx[col] sy[row] (Fresnel), 0: ASM
    int adjoint, bandlimit;
    double du, dv, inv_l2, zs, uw, vw;
    const double2* sx;
    const double2* sy;
    double scale;        // 1 / (My Mx), times any extra factor
};

// pass A: rows padded to Mx (centred at nx / 2), forward
__global__ void __launch_bounds__(NT64)
k64_rows_fwd(const double2* __restrict__ src, double2* __restrict__ W, Args64 a,
             const Plan64* __restrict__ P, int nb) {
    __shared__ double2 lds[LDS64];
    const int tid = threadIdx.x, L = P->L, LP = ldsl64(L), off = a.nx / 2;
    const int row0 = blockIdx.x * nb;
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        for (int i = tid; i < L; i += NT64) {
            const int col = i - off;
            lds[tr * LP + lp64(i)] = (row < a.ny && col >= 0 && col < a.nx)
                                         ? src[(size_t)row * a.nx + col] : make_double2(0.0, 0.0);
        }
    }
    __syncthreads();
    run64<false>(lds, LP, nb, P, tid);
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        if (row >= a.ny) break;
        for (int i = tid; i < L; i += NT64) W[(size_t)row * L + i] = lds[tr * LP + lp64(i)];
    }
}

// transfer function at spectrum row i of column col (natural order)
__device__ __forceinline__ double2 h64(const Args64& a, int col, int i, int Mx, int My) {
    if (a.table) {
        const double2 h = dmul(a.sx[col], a.sy[i]);
        return a.adjoint ? dconj(h) : h;
    }
    const int fx = col < Mx / 2 ? col : col - Mx, fy = i < My / 2 ? i : i - My;
    const double u = (double)fx * a.du, v = (double)fy * a.dv;
    const double w2 = a.inv_l2 - u * u - v * v;
    if (!(w2 >= 0.0)) return make_double2(0.0, 0.0);
    if (a.bandlimit && !(fabs(u) * 2 < a.uw && fabs(v) * 2 < a.vw)) return make_double2(0.0, 0.0);
    double cyc = sqrt(w2) * a.zs;
    cyc -= floor(cyc);
    double s, c;
    sincospi(2.0 * cyc, &s, &c);
    return make_double2(c, s);
}

// pass B: columns (ny of My nonzero, centred at ny / 2), forward, H, inverse, centre
__global__ void __launch_bounds__(NT64)
k64_cols(double2* __restrict__ W, Args64 a, const Plan64* __restrict__ P, int nb) {
    __shared__ double2 lds[LDS64];
    const int tid = threadIdx.x, My = P->L, LP = ldsl64(My), ny = a.ny, off = ny / 2;
    const int Mx = 2 * a.nx;
    const int col0 = blockIdx.x * nb;
    // adjacent threads read adjacent columns of one row
    for (int e = tid; e < nb * My; e += NT64) {
        const int i = e / nb, tr = e - i * nb, col = col0 + tr, h = i - off;
        lds[tr * LP + lp64(i)] = (col < Mx && h >= 0 && h < ny) ? W[(size_t)h * Mx + col]
                                                                 : make_double2(0.0, 0.0);
    }
    __syncthreads();
    run64<false>(lds, LP, nb, P, tid);
    for (int e = tid; e < nb * My; e += NT64) {
        const int i = e / nb, tr = e - i * nb, col = col0 + tr;
        if (col < Mx) lds[tr * LP + lp64(i)] = dmul(lds[tr * LP + lp64(i)], h64(a, col, i, Mx, My));
    }
    __syncthreads();
    run64<true>(lds, LP, nb, P, tid);
    for (int e = tid; e < nb * ny; e += NT64) {
        const int h = e / nb, tr = e - h * nb, col = col0 + tr;
        if (col < Mx) W[(size_t)h * Mx + col] = lds[tr * LP + lp64(h + off)];
    }
}

// pass C: rows inverse, crop, scale
__global__ void __launch_bounds__(NT64)
k64_rows_inv(const double2* __restrict__ W, double2* __restrict__ dst, Args64 a,
             const Plan64* __restrict__ P, int nb) {
    __shared__ double2 lds[LDS64];
    const int tid = threadIdx.x, L = P->L, LP = ldsl64(L), off = a.nx / 2;
    const int row0 = blockIdx.x * nb;
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        for (int i = tid; i < L; i += NT64)
            lds[tr * LP + lp64(i)] = row < a.ny ? W[(size_t)row * L + i] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    run64<true>(lds, LP, nb, P, tid);
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        if (row >= a.ny) break;
        for (int c = tid; c < a.nx; c += NT64) {
            const double2 x = lds[tr * LP + lp64(c + off)];
            dst[(size_t)row * a.nx + c] = make_double2(x.x * a.scale, x.y * a.scale);
        }
    }
}

// plain 2-D FFT: rows in place, then columns in place
template <bool INV>
__global__ void __launch_bounds__(NT64)
k64_fft_rows(double2* __restrict__ u, int rows, const Plan64* __restrict__ P, int nb) {
    __shared__ double2 lds[LDS64];
    const int tid = threadIdx.x, L = P->L, LP = ldsl64(L);
    const int row0 = blockIdx.x * nb;
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        for (int i = tid; i < L; i += NT64)
            lds[tr * LP + lp64(i)] = row < rows ? u[(size_t)row * L + i] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    run64<INV>(lds, LP, nb, P, tid);
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        if (row >= rows) break;
        for (int i = tid; i < L; i += NT64) u[(size_t)row * L + i] = lds[tr * LP + lp64(i)];
    }
}

template <bool INV>
__global__ void __launch_bounds__(NT64)
k64_fft_cols(double2* __restrict__ u, int cols, const Plan64* __restrict__ P, int nb) {
    __shared__ double2 lds[LDS64];
    const int tid = threadIdx.x, L = P->L, LP = ldsl64(L);
    const int col0 = blockIdx.x * nb;
    for (int e = tid; e < nb * L; e += NT64) {
        const int i = e / nb, tr = e - i * nb, col = col0 + tr;
        lds[tr * LP + lp64(i)] = col < cols ? u[(size_t)i * cols + col] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    run64<INV>(lds, LP, nb, P, tid);
    for (int e = tid; e < nb * L; e += NT64) {
        const int i = e / nb, tr = e - i * nb, col = col0 + tr;
        if (col < cols) u[(size_t)i * cols + col] = lds[tr * LP + lp64(i)];
    }
}

// the stage tables: entry S.off + k (R - 1) + r - 1 = W_L^{k r tws}, by the same FP64
// expression as k64_twiddle at that index
__global__ void k64_stage_tab(double2* __restrict__ stab, int L, Plan64 plan, int total) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    int s = -1;
    for (int q = 0; q < plan.nst; q++) {
        const Stage64& Q = plan.st[q];
        if (Q.ns > 1 && e >= Q.off && e < Q.off + Q.ns * (Q.R - 1)) {
            s = q;
            break;
        }
    }
    if (s < 0) return;
    const Stage64& S = plan.st[s];
    const int loc = e - S.off, k = loc / (S.R - 1), r = loc % (S.R - 1) + 1;
    const int m = k * r * S.tws;
    double sn, cs;
    sincospi(-2.0 * (double)m / (double)L, &sn, &cs);
    stab[e] = make_double2(cs, sn);
}

__global__ void k64_twiddle(double2* __restrict__ tw, int L, Plan64 plan, Plan64* __restrict__ dst) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m == 0) *dst = plan;
    if (m >= L) return;
    double s, c;
    sincospi(-2.0 * (double)m / (double)L, &s, &c);
    tw[m] = make_double2(c, s);
}

// spectrum (natural order, scaled by 1/M) of the centred chirp exp(i K x^2), x = (m -
// M/2) d, zero for |m - M/2| >= lim: a direct FP64 DFT (cached per parameter set)
__global__ void k64_chirp_spectrum(double2* __restrict__ S, int M, double d, double K,
                                   long long lim, int adjoint) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    double re = 0.0, im = 0.0;
    for (int n = 0; n < M; n++) {
        // h[(n + M/2) mod M] of the fftshifted response: offset n - M/2 from the centre,
        // i.e. x = ((n + M/2) mod M - M/2) d
        const int m = (n + M / 2) % M;
        const long long o = (long long)m - M / 2;
        if (llabs(o) >= lim) continue;
        const double x = (double)o * d;
        double hs, hc;
        sincos(K * x * x, &hs, &hc);
        const long long nk = ((long long)n * k) % M;
        double s, c;
        sincospi(-2.0 * (double)nk / (double)M, &s, &c);
        re += hc * c - hs * s;
        im += hc * s + hs * c;
    }
    const double sc = 1.0 / M;
    S[k] = make_double2(re * sc, (adjoint ? -im : im) * sc);
}

bool factor64(int L, int* radix, int* nst) {
    int n = L, k = 0;
    while (n % 16 == 0 && k < ST_MAX) { radix[k++] = 16; n /= 16; }
    for (int r : {8, 4, 2})
        if (n % r == 0 && k < ST_MAX) { radix[k++] = r; n /= r; }
    while (n % 9 == 0 && k < ST_MAX) { radix[k++] = 9; n /= 9; }
    if (n % 15 == 0 && k < ST_MAX) { radix[k++] = 15; n /= 15; }
    for (int p : {3, 5, 7, 11, 13})
        while (n % p == 0 && k < ST_MAX) { radix[k++] = p; n /= p; }
    *nst = k;
    return n == 1 && L >= 2;
}

int plan64(olx_ctx* ctx, int L, const Plan64** out) {
    int radix[ST_MAX], nst = 0;
    if (L > LMAX64 || !factor64(L, radix, &nst))
        OLX_FAIL(OLX_EUNSUPPORTED, "complex128 transform length %d: at most %d with prime "
                 "factors 2, 3, 5, 7, 11, 13", L, LMAX64);
    char key[64];
    snprintf(key, sizeof key, "plan64|%d", L);
    bool fresh = false;
    int rc = OLX_OK;
    // L double2 twiddles (2 L float2), the stage tables, then the plan
    const size_t plan_f2 = (sizeof(Plan64) + sizeof(float2) - 1) / sizeof(float2);
    int ntab = 0;
    {
        int ns = 1;
        for (int s = 0; s < nst; s++) {
            if (ns > 1) ntab += ns * (radix[s] - 1);
            ns *= radix[s];
        }
    }
    float2* mem = table(ctx, key, 2 * (size_t)L + 2 * (size_t)ntab + plan_f2, &fresh, &rc);
    if (!mem) return rc;
    double2* tw = reinterpret_cast<double2*>(mem);
    double2* stab = tw + L;
    Plan64* dplan = reinterpret_cast<Plan64*>(mem + 2 * (size_t)L + 2 * (size_t)ntab);
    if (fresh) {
        Plan64 P;
        memset(&P, 0, sizeof P);
        P.L = L;
        P.nst = nst;
        P.tw = tw;
        P.stab = stab;
        int ns = 1, off = 0;
        for (int s = 0; s < nst; s++) {
            Stage64& S = P.st[s];
            S.R = radix[s];
            S.ns = ns;
            S.lr = L / S.R;
            S.tws = L / (ns * S.R);
            S.off = off;
            if (ns > 1) off += ns * (S.R - 1);
            ns *= S.R;
        }
        k64_twiddle<<<(L + 255) / 256, 256, 0, ctx->stream>>>(tw, L, P, dplan);
        if (ntab > 0)
            k64_stage_tab<<<(ntab + 255) / 256, 256, 0, ctx->stream>>>(stab, L, P, ntab);
        OLX_TRY(check_launch("complex128 plan"));
    }
    *out = dplan;
    return OLX_OK;
}

int nb64(int L) {
    const int nb = LMAX64 / L;
    return nb < 1 ? 1 : (nb > 8 ? 8 : nb);
}

// the fftshifted chirp of k64_chirp_spectrum, g[n] = h[(n + M/2) mod M], for an FFT
__global__ void k64_chirp_rot(double2* __restrict__ g, int M, double d, double K, long long lim) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= M) return;
    const long long o = (long long)((n + M / 2) % M) - M / 2;
    if (llabs(o) >= lim) {
        g[n] = make_double2(0.0, 0.0);
        return;
    }
    const double x = (double)o * d;
    double hs, hc;
    sincos(K * x * x, &hs, &hc);
    g[n] = make_double2(hc, hs);
}

__global__ void k64_scale_conj(double2* __restrict__ S, int M, double sc, int adjoint) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= M) return;
    const double2 v = S[k];
    S[k] = make_double2(v.x * sc, (adjoint ? -v.y : v.y) * sc);
}

const double2* spectrum64(olx_ctx* ctx, int M, double d, double K, int64_t lim, int adj, int* rc) {
    char key[256];
    snprintf(key, sizeof key, "spec64|%d|%.17g|%.17g|%lld|%d", M, d, K, (long long)lim, adj);
    bool fresh = false;
    float2* t = table(ctx, key, 2 * (size_t)M, &fresh, rc);
    if (!t || !fresh) return reinterpret_cast<const double2*>(t);
    double2* S = reinterpret_cast<double2*>(t);
    // the chirp's FFT in FP64 (a 8192-point table: tens of microseconds against about
    // 1.5 ms for the direct DFT); the direct DFT for lengths the FFT does not take
    k64_chirp_rot<<<(M + 255) / 256, 256, 0, ctx->stream>>>(S, M, d, K, (long long)lim);
    if (fft64_one(ctx, S, M) == OLX_OK) {
        k64_scale_conj<<<(M + 255) / 256, 256, 0, ctx->stream>>>(S, M, 1.0 / M, adj);
    } else {
        k64_chirp_spectrum<<<(M + 127) / 128, 128, 0, ctx->stream>>>(S, M, d, K, (long long)lim, adj);
    }
    if ((*rc = check_launch("complex128 spectrum table")) != OLX_OK) return nullptr;
    return reinterpret_cast<const double2*>(t);
}

int64_t win64(double v) {
    const double f = std::floor(v);
    if (!(f == f) || f >= (double)INT_MAX) return INT_MAX;
    if (f <= (double)INT_MIN) return INT_MIN;
    return (int64_t)f;
}

int check64(olx_ctx* ctx, const void* p, int64_t ny, int64_t nx, int64_t maxside, const char* what) {
    if (!ctx) OLX_FAIL(OLX_EINVAL, "null context");
    if (!p) OLX_FAIL(OLX_EINVAL, "%s: null field pointer", what);
    int r[ST_MAX], k;
    const int64_t my = maxside == LMAX64 ? ny : 2 * ny, mx = maxside == LMAX64 ? nx : 2 * nx;
    if (ny < 2 || nx < 2 || ny > maxside || nx > maxside || !factor64((int)my, r, &k) ||
        !factor64((int)mx, r, &k))
        OLX_FAIL(OLX_EUNSUPPORTED, "%s: complex128 field %lld x %lld: sides must be in [2, %lld] "
                 "with prime factors 2, 3, 5, 7, 11, 13 only", what, (long long)ny, (long long)nx,
                 (long long)maxside);
    return OLX_OK;
}

int run64_prop(olx_ctx* ctx, double2* u, Args64& a) {
    const int Mx = 2 * a.nx, My = 2 * a.ny;
    const Plan64 *px, *py;
    OLX_TRY(plan64(ctx, Mx, &px));
    OLX_TRY(plan64(ctx, My, &py));
    OLX_TRY(ctx->work.reserve(sizeof(double2) * (size_t)a.ny * Mx));
    double2* W = (double2*)ctx->work.p;
    hipStream_t s = ctx->stream;
    {
        TimedLaunch tl(ctx, "prop64_rows_pad_fwd");
        const int nb = nb64(Mx);
        k64_rows_fwd<<<(a.ny + nb - 1) / nb, NT64, 0, s>>>(u, W, a, px, nb);
    }
    OLX_TRY(check_launch("complex128 pass A"));
    {
        TimedLaunch tl(ctx, "prop64_cols_conv");
        const int nb = nb64(My);
        k64_cols<<<(Mx + nb - 1) / nb, NT64, 0, s>>>(W, a, py, nb);
    }
    OLX_TRY(check_launch("complex128 pass B"));
    {
        TimedLaunch tl(ctx, "prop64_rows_crop_inv");
        const int nb = nb64(Mx);
        k64_rows_inv<<<(a.ny + nb - 1) / nb, NT64, 0, s>>>(W, u, a, px, nb);
    }
    return check_launch("complex128 pass C");
}

}  // namespace

// radix-2 split of a length-2H sequence: eo[0, H) its even samples, eo[H, 2H) its odd ones
__global__ void k64_split2(const double2* __restrict__ in, double2* __restrict__ eo, int H) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= H) return;
    eo[i] = in[2 * i];
    eo[H + i] = in[2 * i + 1];
}
// out[k] = E[k] + W^k O[k], out[k + H] = E[k] - W^k O[k], W = exp(-2 pi i / 2H)
__global__ void k64_join2(double2* __restrict__ out, const double2* __restrict__ eo, int H) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= H) return;
    double s, c;
    sincospi(-(double)k / (double)H, &s, &c);
    const double2 e = eo[k], o = eo[H + k];
    const double2 wo = make_double2(o.x * c - o.y * s, o.x * s + o.y * c);
    out[k] = make_double2(e.x + wo.x, e.y + wo.y);
    out[k + H] = make_double2(e.x - wo.x, e.y - wo.y);
}

int fft64_one(olx_ctx* ctx, double2* u, int L) {
    int radix[ST_MAX], nst = 0;
    hipStream_t s = ctx->stream;
    if (L >= 2 && L <= LMAX64 && factor64(L, radix, &nst)) {
        const Plan64* P = nullptr;
        OLX_TRY(plan64(ctx, L, &P));
        k64_fft_rows<false><<<1, NT64, 0, s>>>(u, 1, P, 1);
        return check_launch("FP64 FFT");
    }
    const int H = L / 2;
    if (L % 2 != 0 || H > LMAX64 || L > 2 * LMAX64 || !factor64(H, radix, &nst)) return OLX_EUNSUPPORTED;
    OLX_TRY(ctx->fft64_tmp.reserve(sizeof(double2) * (size_t)L));
    double2* eo = (double2*)ctx->fft64_tmp.p;
    const Plan64* P = nullptr;
    OLX_TRY(plan64(ctx, H, &P));
    k64_split2<<<(H + 255) / 256, 256, 0, s>>>(u, eo, H);
    k64_fft_rows<false><<<2, NT64, 0, s>>>(eo, 2, P, 1);
    k64_join2<<<(H + 255) / 256, 256, 0, s>>>(u, eo, H);
    return check_launch("FP64 FFT (split)");
}

}  // namespace olx

using namespace olx;

extern "C" int olx_asm_c128(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy, double dx,
                            double lambda, double z, uint32_t flags) {
    OLX_TRY(check64(ctx, d_u, ny, nx, LMAX64 / 2, "asm"));
    if (!(lambda > 0) || !(dx > 0) || !(dy > 0) || !std::isfinite(z))
        OLX_FAIL(OLX_EINVAL, "asm: lambda, dx, dy must be > 0 and z finite");
    if (flags & ~(OLX_BANDLIMIT | OLX_ADJOINT | OLX_EXTRA_INV_M2))
        OLX_FAIL(OLX_EINVAL, "asm complex128: flags must be BANDLIMIT, ADJOINT, EXTRA_INV_M2");
    Args64 a;
    memset(&a, 0, sizeof a);
    a.ny = (int)ny;
    a.nx = (int)nx;
    const double nx2 = 2.0 * nx, ny2 = 2.0 * ny;
    a.du = 1 / (dx * nx2);
    a.dv = 1 / (dy * ny2);
    a.inv_l2 = 1 / (lambda * lambda);
    a.adjoint = (flags & OLX_ADJOINT) ? 1 : 0;
    a.zs = a.adjoint ? -z : z;
    a.bandlimit = (flags & OLX_BANDLIMIT) ? 1 : 0;
    if (a.bandlimit) {
        // shiftedAsmTransferF at zero offset (include/olprop.h:258-270): width ulp + ulm
        const double ulm = 1.0 / std::sqrt(std::pow(0.0 - 0.5f / a.du, -2) * z * z + 1.0f) / lambda;
        const double ulp = 1.0 / std::sqrt(std::pow(0.0 + 0.5f / a.du, -2) * z * z + 1.0f) / lambda;
        const double vlm = 1.0 / std::sqrt(std::pow(0.0 - 0.5f / a.dv, -2) * z * z + 1.0f) / lambda;
        const double vlp = 1.0 / std::sqrt(std::pow(0.0 + 0.5f / a.dv, -2) * z * z + 1.0f) / lambda;
        a.uw = ulp + ulm;
        a.vw = vlp + vlm;
    }
    a.scale = 1.0 / (nx2 * ny2);
    if (flags & OLX_EXTRA_INV_M2) a.scale /= nx2 * ny2;
    return run64_prop(ctx, (double2*)d_u, a);
}

extern "C" int olx_fresnel_c128(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy,
                                double dx, double lambda, double z, uint32_t flags) {
    OLX_TRY(check64(ctx, d_u, ny, nx, LMAX64 / 2, "fresnel"));
    if (!(lambda > 0) || !(dx > 0) || !(dy > 0) || !(z != 0) || !std::isfinite(z))
        OLX_FAIL(OLX_EINVAL, "fresnel: lambda, dx, dy must be > 0 and z finite, nonzero");
    if ((flags & OLX_FRESNEL_BL_CPU) && (flags & OLX_FRESNEL_BL_GPU))
        OLX_FAIL(OLX_EINVAL, "fresnel: choose one band-limit variant");
    if (flags & ~(OLX_FRESNEL_BL_CPU | OLX_FRESNEL_BL_GPU | OLX_ADJOINT))
        OLX_FAIL(OLX_EINVAL, "fresnel complex128: flags must be FRESNEL_BL_CPU / _GPU, ADJOINT");
    Args64 a;
    memset(&a, 0, sizeof a);
    a.ny = (int)ny;
    a.nx = (int)nx;
    a.table = 1;
    a.adjoint = (flags & OLX_ADJOINT) ? 1 : 0;
    int64_t wlim = INT_MAX, hlim = INT_MAX;
    if (flags & OLX_FRESNEL_BL_CPU) {  // include/olprop.h:35-38
        wlim = win64(std::fabs(lambda * z * 0.5 / dx) / dx);
        hlim = win64(std::fabs(lambda * z * 0.5 / dy) / dy);
    } else if (flags & OLX_FRESNEL_BL_GPU) {  // include/golprop.h:229-231
        wlim = hlim = win64(std::fabs(lambda * z * 0.5 / dx) / dx);
    }
    const double K = M_PI * (1 / (lambda * z));
    int rc = OLX_OK;
    // the adjoint is applied as conj(H) per element (a.adjoint), the tables stay forward
    a.sx = spectrum64(ctx, (int)(2 * nx), dx, K, wlim, 0, &rc);
    if (!a.sx) return rc;
    a.sy = spectrum64(ctx, (int)(2 * ny), dy, K, hlim, 0, &rc);
    if (!a.sy) return rc;
    a.scale = 1.0 / (4.0 * (double)nx * (double)ny);
    return run64_prop(ctx, (double2*)d_u, a);
}

extern "C" int olx_fft2d_c128(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, int sign) {
    OLX_TRY(check64(ctx, d_u, ny, nx, LMAX64, "fft"));
    if (sign != -1 && sign != 1) OLX_FAIL(OLX_EINVAL, "fft: sign must be -1 or +1");
    const Plan64 *px, *py;
    OLX_TRY(plan64(ctx, (int)nx, &px));
    OLX_TRY(plan64(ctx, (int)ny, &py));
    double2* u = (double2*)d_u;
    hipStream_t s = ctx->stream;
    const int nbx = nb64((int)nx), nby = nb64((int)ny);
    if (sign > 0) {
        k64_fft_rows<true><<<(unsigned)((ny + nbx - 1) / nbx), NT64, 0, s>>>(u, (int)ny, px, nbx);
        k64_fft_cols<true><<<(unsigned)((nx + nby - 1) / nby), NT64, 0, s>>>(u, (int)nx, py, nby);
    } else {
        k64_fft_rows<false><<<(unsigned)((ny + nbx - 1) / nbx), NT64, 0, s>>>(u, (int)ny, px, nbx);
        k64_fft_cols<false><<<(unsigned)((nx + nby - 1) / nby), NT64, 0, s>>>(u, (int)nx, py, nby);
    }
    return check_launch("complex128 fft");
}

extern "C" int olx_prop_c128(olx_ctx* ctx, void* d_u, int64_t ny, int64_t nx, double dy,
                             double dx, double lambda, double z, int mode, int* chosen) {
    // Prop, include/olprop.h:656-682, choosing as olx_prop_c64 does
    int c;
    if (mode == OLX_MODE_FRESNEL) c = OLX_MODE_FRESNEL;
    else if (mode == OLX_MODE_ASM) c = OLX_MODE_ASM;
    else if (mode == OLX_MODE_AUTO) {
        if (olx_asm_prop_check(ny, nx, dy, dx, lambda, z)) c = OLX_MODE_ASM;
        else if (!olx_fresnel_prop_check(ny, nx, dy, dx, lambda, z)) c = OLX_MODE_FRESNEL_BL;
        else c = OLX_MODE_FRESNEL;
    } else OLX_FAIL(OLX_EINVAL, "prop: unknown mode %d", mode);
    if (chosen) *chosen = c;
    if (c == OLX_MODE_ASM) return olx_asm_c128(ctx, d_u, ny, nx, dy, dx, lambda, z, 0);
    if (c == OLX_MODE_FRESNEL) return olx_fresnel_c128(ctx, d_u, ny, nx, dy, dx, lambda, z, 0);
    return olx_fresnel_c128(ctx, d_u, ny, nx, dy, dx, lambda, z, OLX_FRESNEL_BL_CPU);
}
