// prop_common.hpp -- propagation arguments and the transfer function, shared by
// the power-of-two passes (olx_prop.hip) and the mixed-radix passes (olx_prop_mr.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_core.hpp"
#include "img_common.hpp"

namespace olx {

// transfer functions: ASM (FP64 phase), table product (Fresnel / ARSS), shifted ASM,
// ASM with the reference's PREC_T = float arithmetic (OLX_PREC_F32)
enum { K_ASM = 0, K_TABLE = 1, K_ASHIFT = 2, K_ASM32 = 3 };

struct PropArgs {
    int ny, nx;  // unpadded field
    int nb;      // fields in the batch, stacked row-wise: the row passes see a field of
    int rows;    // rows = nb ny rows, the column pass takes field blockIdx.y
    int mode;    // K_ASM / K_TABLE / K_ASHIFT
    int adjoint;
    int bandlimit;
    double du, dv, inv_l2, zs, uw, vw;  // ASM transfer function (zs: z, -z for the
                                        // adjoint) and band limit
    double u0, v0, ox, oy;  // K_ASHIFT: window centres and the shift's linear phase
    float duf, dvf, tmpf, zf;  // K_ASM32: du, dv, 1/lambda^2 and z rounded as PREC_T = float
    int ss, ds;             // row strides (elements) of the source and destination
    const float2* pre;      // ny x nx factor fields (dense, shared by every field of a
    const float2* post;     // batch) multiplied into the source as pass A loads it and
                            // into the result as pass C stores it
    int accum;              // pass C adds into the destination instead of storing
    uint8_t* img8;          // pass C also stores phase8bit(atan2(x)) of each output value
    int img8_off;           // (include/olimg.h:205-213; img8_off: the +pi form)
    const float2* sx;                  // K_TABLE: natural-order spectra, length Mx / My
    const float2* sy;
    const float2* ax;  // pre-chirp tables (length nx / ny), zero outside the window
    const float2* ay;
    const float2* cx;  // post-chirp tables, applied inside |w-nx/2|<cwl, |h-ny/2|<chl
    const float2* cy;
    int cwl, chl;
    float scale;
    double extra_scale;  // multiplies the 1/(My Mx) output scale
    int cbl;  // column-block width of W (power of two): W[(m / cbl) ny + h][m % cbl]
    int cbl_log;
    // phase gradient in place of the output field: pg_grad (+)= Im(conj(pg_u) x) for
    // each output pixel x (pg_acc: add), by pass C on the tiled path, else by k_phigrad
    const float2* pg_u;
    float* pg_grad;
    int pg_acc;
    // loss-gradient prologue: each source value x becomes g x with
    // g = c1 (|x|^2 m_i - t m_t) - c2, lg_c = {m_i, m_t, c1, c2} (FP64, on the device)
    // and t from lg_t; by pass A on the tiled path, else by k_loss_grad
    const float* lg_t;
    const double* lg_c;
    // loss sums beside the output field: pass C (tiled path) writes each work-group's
    // sums of I, T, I^2, I T, T^2 (I = |x|^2, T from ls_t) to ls_part[k * G + group]
    // and sets ls_groups = G; the untiled path leaves ls_groups 0
    const float* ls_t;
    double* ls_part;
    int ls_groups;
};

// Work-groups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Neighbouring column groups share 128-byte lines of W, so the block index is
// remapped to give every XCD a contiguous run of column groups.
__device__ __forceinline__ int xcd_remap(int b, int nblk) {
    if (nblk % 8) return b;
    return (b % 8) * (nblk / 8) + b / 8;
}

// transfer function at spectrum row i (natural order) of a column whose
// 1/lambda^2 - u^2 is A (FP64).  w = sqrt(A - v^2) comes from an FP32 rsqrt seed
// and one FP64 Newton step (relative error ~1e-14); the phase w z is reduced to
// [0, 1) cycle in FP64 and the sine / cosine are taken by the hardware unit in
// revolutions (absolute error ~1e-7).  zs is z, negated for the adjoint (which
// conjugates H).  Branch-free: every lane evaluates and a 0/1 factor zeroes the
// cut and evanescent ones (a select of these values is turned into a branch by
// the compiler); the argument of the root is |w2| + 1e-30, so every lane's
// value is finite and the product is exact zero where masked.
// The shifted form (SHIFT, shiftedAsmTransferF include/olprop.h:250-305) adds the
// linear phase ox u + oy v of a destination window offset by (ox, oy) (pc = ox u
// of the column, sign-flipped with z for the adjoint) and centres the band-limit
// window of each axis on u0 / v0.
template <bool SHIFT = false>
__device__ __forceinline__ float2 asm_h(const PropArgs& a, double A, bool ucut, int i, int My,
                                        double pc = 0.0) {
    const int fy = i < My / 2 ? i : i - My;
    const double v = (double)fy * a.dv;
    const double w2 = A - v * v;
    const double vc = SHIFT ? v - a.v0 : v;
    const bool keep = (!ucut) & ((!a.bandlimit) | (fabs(vc) * 2 < a.vw)) & (w2 >= 0.0);
    const double w2c = fabs(w2) + 1e-30;
    const double y0 = (double)__builtin_amdgcn_rsqf((float)w2c);
    const double s0 = w2c * y0;
    const double w = fma(fma(-s0, s0, w2c), 0.5 * y0, s0);
    const float t = SHIFT ? (float)__builtin_amdgcn_fract(fma(w, a.zs, fma(a.oy, v, pc)))
                          : (float)__builtin_amdgcn_fract(w * a.zs);
    const float m = keep ? 1.f : 0.f;
    return make_float2(__builtin_amdgcn_cosf(t) * m, __builtin_amdgcn_sinf(t) * m);
}

// AsmTransferF with PREC_T = float (include/olprop.h:172-189): v = (n - ny/2) dv,
// w = sqrt(1/lambda^2 - u^2 - v^2) in float, each operation rounded (no contraction),
// the phase 2 pi w z formed in double and stored to float, then cos / sin of that
// float phase.  uf is the column's float u.  The band-limit window (on the FP64
// frequencies, as asm_h) and the evanescent cut zero the value; the adjoint takes
// the conjugate.
__device__ __forceinline__ float2 asm_h_f32(const PropArgs& a, float uf, bool ucut, int i, int My) {
#pragma clang fp contract(off)
    const int fy = i < My / 2 ? i : i - My;
    const float v = (float)fy * a.dvf;
    const float arg = a.tmpf - uf * uf - v * v;
    const double vd = (double)fy * a.dv;
    const bool keep = (!ucut) & ((!a.bandlimit) | (fabs(vd) * 2 < a.vw)) & (arg >= 0.f);
    const float w = sqrtf(fmaxf(arg, 0.f));
    const float ph = (float)(2.0 * M_PI * (double)w * (double)a.zf);
    // sin / cos of that float phase (up to about 1e6 rad): reduced to revolutions mod 1
    // in FP64 (1e-10 rad), then the hardware sine and cosine (about 1e-7 absolute, far
    // below the 0.06 rad spacing of float phases of that size, which is what this mode
    // reproduces).  sincosf's large-argument reduction needed so many registers that the
    // column pass spilled, and an FP64 sincos more still.
    double cyc = (double)ph * 0.15915494309189535;
    cyc -= rint(cyc);
    const float t = (float)cyc;
    const float s = __builtin_amdgcn_sinf(t), c = __builtin_amdgcn_cosf(t);
    const float m = keep ? 1.f : 0.f;
    return make_float2(c * m, (a.adjoint ? -s : s) * m);
}

// one column's transfer-function state and the value at spectrum row i, for every
// mode (the untiled and mixed-radix column passes)
struct ColH {
    double A = 0.0;   // 1/lambda^2 - u^2 (FP64 modes)
    double pc = 0.0;  // K_ASHIFT: ox u
    float uf = 0.f;   // K_ASM32: the float u
    bool ucut = false;
    float2 sx = {0.f, 0.f};  // K_TABLE: the column's spectrum value
};
template <int MODE>
__device__ __forceinline__ ColH col_h(const PropArgs& a, int col, int Mx) {
    ColH c;
    if constexpr (MODE == K_TABLE) {
        c.sx = a.sx[col];
    } else {
        const int fx = col < Mx / 2 ? col : col - Mx;
        const double u = (double)fx * a.du;
        c.A = a.inv_l2 - u * u;
        c.ucut = a.bandlimit && !(fabs(MODE == K_ASHIFT ? u - a.u0 : u) * 2 < a.uw);
        c.pc = a.ox * u;
        c.uf = (float)fx * a.duf;
    }
    return c;
}
template <int MODE>
__device__ __forceinline__ float2 trans_h(const PropArgs& a, const ColH& c, int i, int My) {
    if constexpr (MODE == K_TABLE) {
        const float2 h = cmul(c.sx, a.sy[i]);
        return a.adjoint ? cconj(h) : h;
    } else if constexpr (MODE == K_ASHIFT) {
        return asm_h<true>(a, c.A, c.ucut, i, My, c.pc);
    } else if constexpr (MODE == K_ASM32) {
        return asm_h_f32(a, c.uf, c.ucut, i, My);
    } else {
        return asm_h(a, c.A, c.ucut, i, My);
    }
}

__device__ __forceinline__ void store_img8(const PropArgs& a, size_t idx, float2 x) {
    a.img8[idx] = phase8_one<float>(atan2f(x.y, x.x), a.img8_off);
}

__device__ __forceinline__ float2 loss_grad(float2 x, float t, double mi, double mt, double c1,
                                            double c2) {
    const double I = (double)x.x * x.x + (double)x.y * x.y;
    const double g = c1 * (I * mi - (double)t * mt) - c2;
    return make_float2((float)(g * x.x), (float)(g * x.y));
}

}  // namespace olx
