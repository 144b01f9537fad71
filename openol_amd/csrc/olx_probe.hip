// olx_probe.hip -- on-device ceilings: FP32 / FP64 FMA, hardware sine and FP16 /
// BF16 MFMA issue rates, each from a kernel that does nothing else, timed with
// HIP events on the context's stream.  These are the rates the kernels of this
// library are compared against next to the published peaks.
#include <hip/hip_runtime.h>

#include "olx_common.hpp"

namespace olx {

typedef _Float16 ph8_t __attribute__((ext_vector_type(8)));
typedef __bf16 pb8_t __attribute__((ext_vector_type(8)));
typedef float pacc_t __attribute__((ext_vector_type(16)));

constexpr int PROBE_BLOCKS = 4096, PROBE_THREADS = 256, PROBE_ITERS = 2048;

__global__ void __launch_bounds__(PROBE_THREADS) k_probe_fma32(float* out, float x, float y) {
    float a[16];
#pragma unroll
    for (int i = 0; i < 16; i++) a[i] = (float)(threadIdx.x + i);
    for (int it = 0; it < PROBE_ITERS; it++)
#pragma unroll
        for (int i = 0; i < 16; i++) a[i] = fmaf(a[i], x, y);
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) s += a[i];
    out[blockIdx.x * PROBE_THREADS + threadIdx.x] = s;
}

__global__ void __launch_bounds__(PROBE_THREADS) k_probe_fma64(double* out, double x, double y) {
    double a[8];
#pragma unroll
    for (int i = 0; i < 8; i++) a[i] = (double)(threadIdx.x + i);
    for (int it = 0; it < PROBE_ITERS; it++)
#pragma unroll
        for (int i = 0; i < 8; i++) a[i] = fma(a[i], x, y);
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 8; i++) s += a[i];
    out[blockIdx.x * PROBE_THREADS + threadIdx.x] = s;
}

__global__ void __launch_bounds__(PROBE_THREADS) k_probe_sin(float* out) {
    float a[8];
#pragma unroll
    for (int i = 0; i < 8; i++) a[i] = 0.001f * (float)(threadIdx.x + i);
    for (int it = 0; it < PROBE_ITERS / 4; it++)
#pragma unroll
        for (int i = 0; i < 8; i++) a[i] = __builtin_amdgcn_sinf(a[i]);
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 8; i++) s += a[i];
    out[blockIdx.x * PROBE_THREADS + threadIdx.x] = s;
}

template <bool BF>
__global__ void __launch_bounds__(PROBE_THREADS) k_probe_mfma(float* out) {
    pacc_t c[4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int e = 0; e < 16; e++) c[i][e] = 0.f;
    const uint4 av = make_uint4(0x3c003c00u ^ threadIdx.x, 0x3c003c00u, 0x3c003c00u, 0x3c003c00u);
    const uint4 bv = make_uint4(0x3c003c00u, 0x3c003c00u ^ threadIdx.x, 0x3c003c00u, 0x3c003c00u);
    for (int it = 0; it < PROBE_ITERS / 16; it++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if constexpr (BF)
                c[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(pb8_t, av),
                                                               __builtin_bit_cast(pb8_t, bv), c[i], 0, 0, 0);
            else
                c[i] = __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(ph8_t, av),
                                                              __builtin_bit_cast(ph8_t, bv), c[i], 0, 0, 0);
        }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int e = 0; e < 16; e++) s += c[i][e];
    out[blockIdx.x * PROBE_THREADS + threadIdx.x] = s;
}

}  // namespace olx

using namespace olx;

extern "C" int olx_peak_probe(olx_ctx* ctx, int kind, double* rate) {
    if (!ctx || !rate) OLX_FAIL(OLX_EINVAL, "peak_probe: null argument");
    if (kind < OLX_PROBE_FMA_F32 || kind > OLX_PROBE_MFMA_BF16)
        OLX_FAIL(OLX_EINVAL, "peak_probe: unknown kind %d", kind);
    const size_t nthr = (size_t)PROBE_BLOCKS * PROBE_THREADS;
    OLX_TRY(ctx->scratch2.reserve(nthr * sizeof(double)));
    void* out = ctx->scratch2.p;
    hipStream_t s = ctx->stream;
    double ops = 0.0;  // operations per launch
    auto launch = [&]() {
        switch (kind) {
            case OLX_PROBE_FMA_F32:
                k_probe_fma32<<<PROBE_BLOCKS, PROBE_THREADS, 0, s>>>((float*)out, 0.999f, 1e-3f);
                ops = 2.0 * 16 * PROBE_ITERS * (double)nthr;
                break;
            case OLX_PROBE_FMA_F64:
                k_probe_fma64<<<PROBE_BLOCKS, PROBE_THREADS, 0, s>>>((double*)out, 0.999, 1e-3);
                ops = 2.0 * 8 * PROBE_ITERS * (double)nthr;
                break;
            case OLX_PROBE_SIN_F32:
                k_probe_sin<<<PROBE_BLOCKS, PROBE_THREADS, 0, s>>>((float*)out);
                ops = 8.0 * (PROBE_ITERS / 4) * (double)nthr;
                break;
            case OLX_PROBE_MFMA_F16:
                k_probe_mfma<false><<<PROBE_BLOCKS, PROBE_THREADS, 0, s>>>((float*)out);
                ops = 4.0 * (PROBE_ITERS / 16) * (double)(nthr / 64) * 32 * 32 * 16 * 2;
                break;
            default:
                k_probe_mfma<true><<<PROBE_BLOCKS, PROBE_THREADS, 0, s>>>((float*)out);
                ops = 4.0 * (PROBE_ITERS / 16) * (double)(nthr / 64) * 32 * 32 * 16 * 2;
                break;
        }
    };
    launch();  // warm-up
    OLX_TRY(check_launch("peak_probe"));
    hipEvent_t e0, e1;
    OLX_HIP(hipEventCreate(&e0));
    OLX_HIP(hipEventCreate(&e1));
    OLX_HIP(hipEventRecord(e0, s));
    launch();
    OLX_HIP(hipEventRecord(e1, s));
    OLX_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    OLX_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *rate = ops / (ms * 1e-3);
    return check_launch("peak_probe");
}
