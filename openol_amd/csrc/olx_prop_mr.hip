// olx_prop_mr.hip -- propagation and 2-D FFT for field sides that are not powers of
// two (the reference pads any ny x nx field to 2ny x 2nx, include/olutils.h:133-178,
// and transforms it with FFTW / cuFFT, which are mixed radix).
//
// Transforms of length L whose prime factors are 2, 3, 5, 7, 11 and 13 run as a
// runtime-planned Stockham FFT in LDS: stage s of radix R with NS = product of
// the earlier radices takes butterfly j in [0, L/R) from x[j + r L/R], multiplies
// by W_{NS R}^{k r} (k = j mod NS, from an FP64-built table of W_L), applies the
// R-point DFT and writes y[(j - k) R + k + r NS], leaving natural order after the
// last stage.  Radix 16 (remainder 8, 4, 2) takes the factors of two; the odd
// factors go to radix 9 while 9 divides, then one radix 15 where 15 divides, then
// 3, 5, 7, 11, 13, all by the symmetric-pair DFT (valid for any odd R).  Each stage
// is an LDS round trip, so fewer, larger odd radices cost less than their extra
// arithmetic: 4320 = 16 2 9 15 takes four stages where 16 2 3 3 3 5 took six.  Each stage loads every butterfly a thread owns into
// registers, waits, then writes, so one LDS image per transform suffices.
//
// The three propagation passes mirror the power-of-two ones over an intermediate W
// of ny x Mx values in column tiles of 8 (tile g = m / 8 holds columns 8 g .. 8 g + 7
// of every row contiguously, W[(g rows + h) 8 + m % 8]): A pads, transforms and
// stores rows; B loads columns (ny of My values nonzero), transforms, multiplies by
// the transfer function, transforms back and stores the centre rows; C transforms
// rows back, crops, scales and applies the post-chirp.  With tiles, the column pass's
// one value per row shares its 64-byte line with the seven neighbouring columns,
// which neighbouring work-groups read at the same time (a row-major W puts every
// value in a line of its own).  Zero padding is centred at offset
// (M - n) / 2 = n / 2, as the reference's zeropadding / del_zero.
#include <climits>
#include <cstring>

#include "olx_common.hpp"
#include "prop_common.hpp"

namespace olx {

constexpr int MR_MAXST = 20;

struct MrStage {
    int R;        // radix
    int ns;       // product of the earlier radices
    int lr;       // L / R (butterflies per transform)
    int tws;      // twiddle stride L / (ns R)
    uint32_t mns; // fast-division magic of ns (unused when ns == 1)
    uint32_t mlr; // fast-division magic of lr
    int off;      // this stage's twiddles in the stage table: k (R - 1) + r - 1 from here
};

// a plan lives in device memory right after its length's twiddle table, so kernels
// index its stages through a pointer (uniform scalar loads)
struct MrPlan {
    int L, nst;
    const float2* tw;    // W_L^m = exp(-2 pi i m / L), m in [0, L)
    const float2* stab;  // per stage (ns > 1): W_L^{k r tws} for k < ns, r = 1 .. R-1, k-major:
                         // a butterfly's R - 1 twiddles are contiguous (from tw, lanes touched
                         // scattered lines)
    MrStage st[MR_MAXST];
};

// n / d for n, d < 2^16 (n d < 2^32): magic m = floor(2^32 / d) + 1
__host__ __device__ constexpr uint32_t fdiv_magic(int d) {
    return d <= 1 ? 0u : (uint32_t)((0x100000000ull / (uint64_t)d) + 1);
}
__device__ __forceinline__ int fdiv(int n, int d, uint32_t m) {
    return d == 1 ? n : (int)__umulhi((uint32_t)n, m);
}

// compile-time cos / sin of 2 pi n / R (Taylor series after reduction to
// [-pi, pi], double precision)
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double ct_sin(double x) {
    while (x > kPi) x -= 2 * kPi;
    while (x < -kPi) x += 2 * kPi;
    double term = x, sum = x;
    for (int i = 1; i < 40; i++) {
        term *= -x * x / ((2.0 * i) * (2.0 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double ct_cos(double x) { return ct_sin(x + kPi / 2); }

// cos / sin of 2 pi m / R for m in [0, R), evaluated by the compiler.  The DFT below
// indexes it with its unrolled loop counters, so every entry folds to an immediate
// (calling ct_cos / ct_sin there directly is not a constant expression, and the
// series then ran at run time: 132 FP64 divisions in the column kernel).
template <int R>
struct OddRoots {
    float c[R], s[R];
    constexpr OddRoots() : c(), s() {
        for (int m = 0; m < R; m++) {
            c[m] = (float)ct_cos(2.0 * kPi * m / R);
            s[m] = (float)ct_sin(2.0 * kPi * m / R);
        }
    }
};
template <int R>
__device__ constexpr OddRoots<R> kOddRoots{};

// R-point DFT in registers, natural-order output; forward e^{-2 pi i nk/R}
template <int R, bool INV>
__device__ __forceinline__ void dft_any(float2* v) {
    if constexpr (R == 2 || R == 4 || R == 8 || R == 16) {
        dft<R, INV>(v);
    } else {
        // odd R: a_n = x_n + x_{R-n}, b_n = x_n - x_{R-n} (n = 1 .. H), then
        // y_k = A_k -/+ i B_k, y_{R-k} = A_k +/- i B_k with A_k = x_0 + sum a_n cos,
        // B_k = sum b_n sin (angles 2 pi n k / R; upper signs forward)
        constexpr int H = (R - 1) / 2;
        float2 a[H], b[H];
        float2 y0 = v[0];
#pragma unroll
        for (int n = 1; n <= H; n++) {
            a[n - 1] = cadd(v[n], v[R - n]);
            b[n - 1] = csub(v[n], v[R - n]);
            y0 = cadd(y0, a[n - 1]);
        }
        float2 out[R];
#pragma unroll
        for (int k = 1; k <= H; k++) {
            float2 A = v[0], B = make_float2(0.f, 0.f);
#pragma unroll
            for (int n = 1; n <= H; n++) {
                const float c = kOddRoots<R>.c[(n * k) % R];
                const float s = kOddRoots<R>.s[(n * k) % R];
                A = make_float2(fmaf(a[n - 1].x, c, A.x), fmaf(a[n - 1].y, c, A.y));
                B = make_float2(fmaf(b[n - 1].x, s, B.x), fmaf(b[n - 1].y, s, B.y));
            }
            if (INV) {
                out[k] = make_float2(A.x - B.y, A.y + B.x);
                out[R - k] = make_float2(A.x + B.y, A.y - B.x);
            } else {
                out[k] = make_float2(A.x + B.y, A.y - B.x);
                out[R - k] = make_float2(A.x - B.y, A.y + B.x);
            }
        }
        v[0] = y0;
#pragma unroll
        for (int k = 1; k < R; k++) v[k] = out[k];
    }
}

// one Stockham stage over nb transforms (LDS stride LP apart) of nb * lr butterflies
template <int R, bool INV, int NT, int LCAP>
__device__ __forceinline__ void mr_stage(float2* lds, int LP, int nb, const MrStage& S,
                                         const float2* __restrict__ stab, int tid) {
    constexpr int Q = (LCAP / R + NT - 1) / NT;
    const int tot = nb * S.lr;
    float2 v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int g = tid + NT * q;
        if (g < tot) {
            const int tr = fdiv(g, S.lr, S.mlr), j = g - tr * S.lr;
            const float2* p = lds + tr * LP;
#pragma unroll
            for (int r = 0; r < R; r++) v[q][r] = p[lpad(j + r * S.lr)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int g = tid + NT * q;
        if (g < tot) {
            const int tr = fdiv(g, S.lr, S.mlr), j = g - tr * S.lr;
            const int k = j - fdiv(j, S.ns, S.mns) * S.ns;
            if (S.ns > 1) {
                const float2* wk = stab + S.off + k * (R - 1);
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = wk[r - 1];
                    if (INV) w = cconj(w);
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            dft_any<R, INV>(v[q]);
            float2* p = lds + tr * LP;
            const int o = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) p[lpad(o + r * S.ns)] = v[q][r];
        }
    }
    __syncthreads();
}

// the whole transform of nb LDS-resident sequences (natural order in and out);
// inlined: as a called function it kept its own frame and spilled (580 B of scratch)
// P7: the plan may hold radix 7, 11 or 13.  Their DFTs set the register budget of the
// whole switch (1024-thread kernels spilled 207 VGPRs with them), so lengths without
// those factors run a copy of the switch that leaves them out.
template <bool INV, int NT, int LCAP, bool P7>
__device__ __forceinline__ void mr_run(float2* lds, int LP, int nb, const MrPlan* __restrict__ P, int tid) {
    const int nst = P->nst;
    const float2* tw = P->stab;
    for (int s = 0; s < nst; s++) {
        const MrStage S = P->st[s];
        switch (S.R) {
            case 2: mr_stage<2, INV, NT, LCAP>(lds, LP, nb, S, tw, tid); break;
            case 3: mr_stage<3, INV, NT, LCAP>(lds, LP, nb, S, tw, tid); break;
            case 4: mr_stage<4, INV, NT, LCAP>(lds, LP, nb, S, tw, tid); break;
            case 5: mr_stage<5, INV, NT, LCAP>(lds, LP, nb, S, tw, tid); break;
            case 7:
                if constexpr (P7) mr_stage<7, INV, NT, LCAP>(lds, LP, nb, S, tw, tid);
                break;
            case 8: mr_stage<8, INV, NT, LCAP>(lds, LP, nb, S, tw, tid); break;
            case 9: mr_stage<9, INV, NT, LCAP>(lds, LP, nb, S, tw, tid); break;
            case 15: mr_stage<15, INV, NT, LCAP>(lds, LP, nb, S, tw, tid); break;
            case 11:
                if constexpr (P7) mr_stage<11, INV, NT, LCAP>(lds, LP, nb, S, tw, tid);
                break;
            case 13:
                if constexpr (P7) mr_stage<13, INV, NT, LCAP>(lds, LP, nb, S, tw, tid);
                break;
            default: mr_stage<16, INV, NT, LCAP>(lds, LP, nb, S, tw, tid); break;
        }
    }
}

// LDS capacity of a work-group: LCAP values plus the padding of every 16th
template <int LCAP>
constexpr int mr_lds() { return LCAP + LCAP / 16 + 32; }
__host__ __device__ inline int mr_lp(int L) { return L + L / 16 + 1; }

// element (row, m) of the tiled intermediate over `rows` rows
__device__ __forceinline__ size_t mrw(int rows, int row, int m) {
    return ((size_t)(m >> 3) * rows + row) * 8 + (m & 7);
}

// ------------------------------------------------------------------ pass A
template <int NT, int LCAP, bool P7>
__global__ void __launch_bounds__(NT)
k_mr_rows_fwd(const float2* __restrict__ src, float2* __restrict__ W, PropArgs a, const MrPlan* __restrict__ P, int nb) {
    __shared__ float2 lds[mr_lds<LCAP>()];
    const int tid = threadIdx.x, L = P->L, LP = mr_lp(L), nx = a.nx, off = nx / 2;
    const int row0 = xcd_remap(blockIdx.x, gridDim.x) * nb;
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        const float2 ay = (a.ay && row < a.rows) ? a.ay[row] : make_float2(1.f, 0.f);
        for (int i = tid; i < L; i += NT) {
            float2 x = make_float2(0.f, 0.f);
            const int col = i - off;
            if (row < a.rows && col >= 0 && col < nx) {
                x = src[(size_t)row * a.ss + col];
                if (a.ax) x = cmul(x, cmul(a.ax[col], ay));
                if (a.pre) x = cmul(x, a.pre[(size_t)(row % a.ny) * nx + col]);
            }
            lds[tr * LP + lpad(i)] = x;
        }
    }
    __syncthreads();
    mr_run<false, NT, LCAP, P7>(lds, LP, nb, P, tid);
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        if (row >= a.rows) break;
        for (int i = tid; i < L; i += NT) W[mrw(a.rows, row, i)] = lds[tr * LP + lpad(i)];
    }
}

// ------------------------------------------------------------------ pass B
template <int NT, int LCAP, int MODE, bool P7>
__global__ void __launch_bounds__(NT)
k_mr_cols_conv(float2* __restrict__ W, PropArgs a, const MrPlan* __restrict__ P, int nb) {
    __shared__ float2 lds[mr_lds<LCAP>()];
    const int tid = threadIdx.x, My = P->L, LP = mr_lp(My), ny = a.ny, off = ny / 2;
    const int Mx = 2 * a.nx;
    const int col0 = xcd_remap(blockIdx.x, gridDim.x) * nb;
    const int fo = blockIdx.y * ny;  // this field's first row
    for (int tr = 0; tr < nb; tr++) {
        const int col = col0 + tr;
        for (int i = tid; i < My; i += NT) {
            const int h = i - off;
            float2 x = make_float2(0.f, 0.f);
            if (col < Mx && h >= 0 && h < ny) x = W[mrw(a.rows, fo + h, col)];
            lds[tr * LP + lpad(i)] = x;
        }
    }
    __syncthreads();
    mr_run<false, NT, LCAP, P7>(lds, LP, nb, P, tid);
    for (int tr = 0; tr < nb; tr++) {
        const int col = col0 + tr;
        if (col >= Mx) break;
        const ColH ch = col_h<MODE>(a, col, Mx);
        for (int i = tid; i < My; i += NT) {
            float2* p = &lds[tr * LP + lpad(i)];
            *p = cmul(*p, trans_h<MODE>(a, ch, i, My));
        }
    }
    __syncthreads();
    mr_run<true, NT, LCAP, P7>(lds, LP, nb, P, tid);
    for (int tr = 0; tr < nb; tr++) {
        const int col = col0 + tr;
        if (col >= Mx) break;
        for (int h = tid; h < ny; h += NT) W[mrw(a.rows, fo + h, col)] = lds[tr * LP + lpad(h + off)];
    }
}

// ------------------------------------------------------------------ pass C
template <int NT, int LCAP, bool P7>
__global__ void __launch_bounds__(NT)
k_mr_rows_inv(const float2* __restrict__ W, float2* __restrict__ dst, PropArgs a, const MrPlan* __restrict__ P, int nb) {
    __shared__ float2 lds[mr_lds<LCAP>()];
    const int tid = threadIdx.x, L = P->L, LP = mr_lp(L), nx = a.nx, off = nx / 2;
    const int row0 = xcd_remap(blockIdx.x, gridDim.x) * nb;
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        for (int i = tid; i < L; i += NT)
            lds[tr * LP + lpad(i)] = row < a.rows ? W[mrw(a.rows, row, i)] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    mr_run<true, NT, LCAP, P7>(lds, LP, nb, P, tid);
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        if (row >= a.rows) break;
        const bool post = a.cx != nullptr && abs(row - a.ny / 2) < a.chl;
        const float2 cy = post ? a.cy[row] : make_float2(1.f, 0.f);
        for (int col = tid; col < nx; col += NT) {
            float2 x = cscale(lds[tr * LP + lpad(col + off)], a.scale);
            if (post && abs(col - nx / 2) < a.cwl) x = cmul(x, cmul(a.cx[col], cy));
            if (a.post) x = cmul(x, a.post[(size_t)(row % a.ny) * nx + col]);
            float2* o = &dst[(size_t)row * a.ds + col];
            if (a.accum) x = make_float2(o->x + x.x, o->y + x.y);
            *o = x;
            if (a.img8) store_img8(a, (size_t)row * a.ds + col, x);
        }
    }
}

// ------------------------------------------------------------------ plain 2-D FFT
template <int NT, int LCAP, bool INV, bool P7>
__global__ void __launch_bounds__(NT)
k_mr_fft_rows(float2* __restrict__ u, int rows, const MrPlan* __restrict__ P, int nb) {
    __shared__ float2 lds[mr_lds<LCAP>()];
    const int tid = threadIdx.x, L = P->L, LP = mr_lp(L);
    const int row0 = blockIdx.x * nb;
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        for (int i = tid; i < L; i += NT)
            lds[tr * LP + lpad(i)] = row < rows ? u[(size_t)row * L + i] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    mr_run<INV, NT, LCAP, P7>(lds, LP, nb, P, tid);
    for (int tr = 0; tr < nb; tr++) {
        const int row = row0 + tr;
        if (row >= rows) break;
        for (int i = tid; i < L; i += NT) u[(size_t)row * L + i] = lds[tr * LP + lpad(i)];
    }
}

template <int NT, int LCAP, bool INV, bool P7>
__global__ void __launch_bounds__(NT)
k_mr_fft_cols(float2* __restrict__ u, int cols, const MrPlan* __restrict__ P, int nb) {
    __shared__ float2 lds[mr_lds<LCAP>()];
    const int tid = threadIdx.x, L = P->L, LP = mr_lp(L);
    const int col0 = xcd_remap(blockIdx.x, gridDim.x) * nb;
    // adjacent threads read adjacent columns of one row
    for (int e = tid; e < nb * L; e += NT) {
        const int i = e / nb, tr = e - i * nb, col = col0 + tr;
        lds[tr * LP + lpad(i)] = col < cols ? u[(size_t)i * cols + col] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    mr_run<INV, NT, LCAP, P7>(lds, LP, nb, P, tid);
    for (int e = tid; e < nb * L; e += NT) {
        const int i = e / nb, tr = e - i * nb, col = col0 + tr;
        if (col < cols) u[(size_t)i * cols + col] = lds[tr * LP + lpad(i)];
    }
}

// W_L^m = exp(-2 pi i m / L) in FP64, rounded once; thread 0 also stores the plan
__global__ void k_mr_twiddle(float2* __restrict__ tw, int L, MrPlan plan, MrPlan* __restrict__ dst) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m == 0) *dst = plan;
    if (m >= L) return;
    double s, c;
    sincospi(-2.0 * (double)m / (double)L, &s, &c);
    tw[m] = make_float2((float)c, (float)s);
}

// the stage tables: entry S.off + k (R - 1) + r - 1 = W_L^{k r tws}, by the same FP64
// expression as k_mr_twiddle at that index (bitwise the table's value)
__global__ void k_mr_stage_tab(float2* __restrict__ stab, int L, MrPlan plan, int total) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    int s = -1;
    for (int q = 0; q < plan.nst; q++) {
        const MrStage& Q = plan.st[q];
        if (Q.ns > 1 && e >= Q.off && e < Q.off + Q.ns * (Q.R - 1)) {
            s = q;
            break;
        }
    }
    if (s < 0) return;
    const MrStage& S = plan.st[s];
    const int loc = e - S.off, k = loc / (S.R - 1), r = loc % (S.R - 1) + 1;
    const int m = k * r * S.tws;
    double sn, cs;
    sincospi(-2.0 * (double)m / (double)L, &sn, &cs);
    stab[e] = make_float2((float)cs, (float)sn);
}

// ------------------------------------------------------------------ host side
// Stage radices of L: sixteens then the remaining 8 / 4 / 2, then nines, one 15,
// then 3, 5, 7, 11, 13.
// Returns false when L has another prime factor or needs too many stages.
static bool mr_factor(int L, int* radix, int* nst) {
    int n = L, k = 0;
    while (n % 16 == 0 && k < MR_MAXST) { radix[k++] = 16; n /= 16; }
    for (int r : {8, 4, 2})
        if (n % r == 0 && k < MR_MAXST) { radix[k++] = r; n /= r; }
    while (n % 9 == 0 && k < MR_MAXST) { radix[k++] = 9; n /= 9; }
    if (n % 15 == 0 && k < MR_MAXST) { radix[k++] = 15; n /= 15; }
    for (int p : {3, 5, 7, 11, 13})
        while (n % p == 0 && k < MR_MAXST) { radix[k++] = p; n /= p; }
    *nst = k;
    return n == 1 && L >= 2;
}

bool mr_supported(int64_t L) {
    if (L < 2 || L > 16384) return false;
    int r[MR_MAXST], k;
    return mr_factor((int)L, r, &k);
}

// the device-resident plan of length L (built with its twiddles on first use)
static int mr_plan(olx_ctx* ctx, int L, const MrPlan** out) {
    int radix[MR_MAXST], nst = 0;
    if (L > 16384 || !mr_factor(L, radix, &nst))
        OLX_FAIL(OLX_EUNSUPPORTED, "transform length %d: prime factors must be 2, 3, 5, 7, 11, 13 "
                 "and the length at most 16384", L);
    char key[64];
    snprintf(key, sizeof key, "mrplan|%d", L);
    bool fresh = false;
    int rc = OLX_OK;
    constexpr size_t plan_f2 = (sizeof(MrPlan) + sizeof(float2) - 1) / sizeof(float2);
    // stage-table length: sum over the stages with ns > 1 of ns (R - 1)
    int ntab = 0;
    {
        int ns = 1;
        for (int s = 0; s < nst; s++) {
            if (ns > 1) ntab += ns * (radix[s] - 1);
            ns *= radix[s];
        }
    }
    const size_t tabf2 = ((size_t)ntab + 1) & ~(size_t)1;  // keep the plan 16-byte aligned
    float2* tw = table(ctx, key, (size_t)L + tabf2 + plan_f2, &fresh, &rc);
    if (!tw) return rc;
    float2* stab = tw + L;
    MrPlan* dplan = reinterpret_cast<MrPlan*>(stab + tabf2);
    if (fresh) {
        MrPlan P;
        memset(&P, 0, sizeof P);
        P.L = L;
        P.nst = nst;
        P.tw = tw;
        P.stab = stab;
        int ns = 1, off = 0;
        for (int s = 0; s < nst; s++) {
            MrStage& S = P.st[s];
            S.R = radix[s];
            S.ns = ns;
            S.lr = L / S.R;
            S.tws = L / (ns * S.R);
            S.mns = fdiv_magic(ns);
            S.mlr = fdiv_magic(S.lr);
            S.off = off;
            if (ns > 1) off += ns * (S.R - 1);
            ns *= S.R;
        }
        k_mr_twiddle<<<(L + 255) / 256, 256, 0, ctx->stream>>>(tw, L, P, dplan);
        if (ntab > 0)
            k_mr_stage_tab<<<(ntab + 255) / 256, 256, 0, ctx->stream>>>(stab, L, P, ntab);
        OLX_TRY(check_launch("mixed-radix plan"));
    }
    *out = dplan;
    return OLX_OK;
}

// transforms per work-group: as many as fill LCAP values, at most 8
template <int LCAP>
static int mr_nb(int L) {
    int nb = LCAP / L;
    return nb < 1 ? 1 : (nb > 8 ? 8 : nb);
}

// calls f(integral NT, integral LCAP) for the work-group shape of length L.  A stage
// holds each thread's ceil(LCAP / R / NT) butterflies in registers between its LDS
// read and write, so long transforms take 1024 threads: with 256 the radix-3 stage of
// an 8192-value image held 33 values per thread and the kernel spilled (289 VGPRs,
// one wave per SIMD).
template <class F>
static void mr_shape2(int L, F&& f) {
    if (L <= 2048) f(std::integral_constant<int, 256>(), std::integral_constant<int, 8192>());
    else if (L <= 8192) {
        // 512 threads; lengths up to 4352 hold two transforms per work-group (capacity
        // 8704), so the radix-16 stages of a 4320-point column keep more threads busy.
        // OLX_MRNT = 256, 512 or 1024 (capacity 8192) overrides for tuning.
        static const char* e = getenv("OLX_MRNT");
        const int nt = e ? atoi(e) : 0;
        if (nt == 256) f(std::integral_constant<int, 256>(), std::integral_constant<int, 8192>());
        else if (nt == 1024) f(std::integral_constant<int, 1024>(), std::integral_constant<int, 8192>());
        else if (nt == 0 && L <= 4352) f(std::integral_constant<int, 512>(), std::integral_constant<int, 8704>());
        else f(std::integral_constant<int, 512>(), std::integral_constant<int, 8192>());
    }
    else f(std::integral_constant<int, 1024>(), std::integral_constant<int, 16384>());
}
// also passes P7 (whether L has a factor 7, 11 or 13) as a compile-time constant
template <class F>
static void mr_shape(int L, F&& f) {
    const bool p7 = L % 7 == 0 || L % 11 == 0 || L % 13 == 0;
    mr_shape2(L, [&](auto NT, auto LC) {
        if (p7) f(NT, LC, std::true_type());
        else f(NT, LC, std::false_type());
    });
}

int mr_three_pass(olx_ctx* ctx, const float2* src, float2* dst, PropArgs& a, float2* W) {
    const int Mx = 2 * a.nx, My = 2 * a.ny;
    const MrPlan *px, *py;
    OLX_TRY(mr_plan(ctx, Mx, &px));
    OLX_TRY(mr_plan(ctx, My, &py));
    hipStream_t s = ctx->stream;
    {
        TimedLaunch tl(ctx, "prop_rows_pad_fwd");
        mr_shape(Mx, [&](auto NT, auto LC, auto P7) {
            const int nb = mr_nb<LC.value>(Mx);
            k_mr_rows_fwd<NT.value, LC.value, P7.value><<<(a.rows + nb - 1) / nb, NT.value, 0, s>>>(src, W, a, px, nb);
        });
    }
    OLX_TRY(check_launch("mixed-radix pass A"));
    {
        TimedLaunch tl(ctx, "prop_cols_conv");
        mr_shape(My, [&](auto NT, auto LC, auto P7) {
            const int nb = mr_nb<LC.value>(My);
            const dim3 grid((Mx + nb - 1) / nb, a.nb);
            if (a.mode == K_ASM)
                k_mr_cols_conv<NT.value, LC.value, K_ASM, P7.value><<<grid, NT.value, 0, s>>>(W, a, py, nb);
            else if (a.mode == K_ASHIFT)
                k_mr_cols_conv<NT.value, LC.value, K_ASHIFT, P7.value><<<grid, NT.value, 0, s>>>(W, a, py, nb);
            else if (a.mode == K_ASM32)
                k_mr_cols_conv<NT.value, LC.value, K_ASM32, P7.value><<<grid, NT.value, 0, s>>>(W, a, py, nb);
            else
                k_mr_cols_conv<NT.value, LC.value, K_TABLE, P7.value><<<grid, NT.value, 0, s>>>(W, a, py, nb);
        });
    }
    OLX_TRY(check_launch("mixed-radix pass B"));
    {
        TimedLaunch tl(ctx, "prop_rows_crop_inv");
        mr_shape(Mx, [&](auto NT, auto LC, auto P7) {
            const int nb = mr_nb<LC.value>(Mx);
            k_mr_rows_inv<NT.value, LC.value, P7.value><<<(a.rows + nb - 1) / nb, NT.value, 0, s>>>(W, dst, a, px, nb);
        });
    }
    return check_launch("mixed-radix pass C");
}

int mr_fft2d(olx_ctx* ctx, float2* u, int64_t ny, int64_t nx, bool inv) {
    const MrPlan *px, *py;
    OLX_TRY(mr_plan(ctx, (int)nx, &px));
    OLX_TRY(mr_plan(ctx, (int)ny, &py));
    hipStream_t s = ctx->stream;
    {
        TimedLaunch tl(ctx, "fft_rows");
        mr_shape((int)nx, [&](auto NT, auto LC, auto P7) {
            const int nb = mr_nb<LC.value>((int)nx);
            const int grid = (int)((ny + nb - 1) / nb);
            if (inv) k_mr_fft_rows<NT.value, LC.value, true, P7.value><<<grid, NT.value, 0, s>>>(u, (int)ny, px, nb);
            else k_mr_fft_rows<NT.value, LC.value, false, P7.value><<<grid, NT.value, 0, s>>>(u, (int)ny, px, nb);
        });
    }
    OLX_TRY(check_launch("mixed-radix fft rows"));
    {
        TimedLaunch tl(ctx, "fft_cols");
        mr_shape((int)ny, [&](auto NT, auto LC, auto P7) {
            const int nb = mr_nb<LC.value>((int)ny);
            const int grid = (int)((nx + nb - 1) / nb);
            if (inv) k_mr_fft_cols<NT.value, LC.value, true, P7.value><<<grid, NT.value, 0, s>>>(u, (int)nx, py, nb);
            else k_mr_fft_cols<NT.value, LC.value, false, P7.value><<<grid, NT.value, 0, s>>>(u, (int)nx, py, nb);
        });
    }
    return check_launch("mixed-radix fft cols");
}

}  // namespace olx
