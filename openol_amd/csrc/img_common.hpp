// img_common.hpp -- the 8-bit conversions shared by the encodings (olx_img.hip)
// and the fused phase-byte epilogues of the CGH and propagation kernels.  One
// rounding per FP64 operation (explicit _rn intrinsics, so the file's contraction
// setting does not matter), C round() and the reference's x86-64 uint8 conversion.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace olx {

// the reference's x86-64 conversion of a rounded double to uint8: cvttsd2si into
// int32 (INT32_MIN, low byte 0, for NaN and for values outside the int32 range),
// then the low byte
__device__ inline uint8_t to_u8(double v) {
    if (!(v >= -2147483648.0 && v < 2147483648.0)) return 0;
    return (uint8_t)(uint32_t)(int32_t)v;
}

// Fast path for an encoding q = f(x) whose C round() is wanted: q_fast is an FP32
// estimate within err of the FP64 value the reference computes.  When q_fast lies
// farther than err from every half-integer, rounding it gives the same integer as
// rounding the FP64 value; otherwise (and for NaN / huge values) the caller takes
// the FP64 path.
__device__ inline bool fast_round(float q_fast, float err, int* k) {
    if (!(fabsf(q_fast) < 65536.f)) return false;
    const float f = q_fast - floorf(q_fast);
    if (!(fabsf(f - 0.5f) > err)) return false;
    *k = (int)rintf(q_fast);
    return true;
}

// one phase value: the FP32 fast path, else the reference's FP64 expression
template <typename T>
__device__ inline uint8_t phase8_one(T x, int offset) {
    int k;
    if (sizeof(T) == 4) {
        // |q| < 65536: FP32 error of the three operations below plus the rounding of
        // pi and 255 / 2pi stays under 2^-12 + 4e-7 |q|
        const float pf = (float)x;
        const float qf = (offset ? pf + 3.14159265f : pf - 3.14159265f) * 40.5845105f;
        if (fast_round(qf, 2.5e-4f + 4e-7f * fabsf(qf), &k)) return (uint8_t)(uint32_t)k;
    }
    const double p = (double)x;
    const double t = offset ? __dadd_rn(p, M_PI) : __dsub_rn(p, M_PI);
    return to_u8(round(__dmul_rn(__ddiv_rn(t, 2 * M_PI), 255.0)));
}

}  // namespace olx
