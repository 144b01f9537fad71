// prop_common.hpp -- propagation arguments and the transfer function, shared by
// the power-of-two passes (olx_prop.hip) and the mixed-radix passes (olx_prop_mr.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fft_core.hpp"

namespace olx {

enum { K_ASM = 0, K_TABLE = 1, K_ASHIFT = 2 };

struct PropArgs {
    int ny, nx;  // unpadded field
    int mode;    // K_ASM / K_TABLE / K_ASHIFT
    int adjoint;
    int bandlimit;
    double du, dv, inv_l2, zs, uw, vw;  // ASM transfer function (zs: z, -z for the
                                        // adjoint) and band limit
    double u0, v0, ox, oy;  // K_ASHIFT: window centres and the shift's linear phase
    int ss, ds;             // row strides (elements) of the source and destination
    int accum;              // pass C adds into the destination instead of storing
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

__device__ __forceinline__ float2 loss_grad(float2 x, float t, double mi, double mt, double c1,
                                            double c2) {
    const double I = (double)x.x * x.x + (double)x.y * x.y;
    const double g = c1 * (I * mi - (double)t * mt) - c2;
    return make_float2((float)(g * x.x), (float)(g * x.y));
}

}  // namespace olx
