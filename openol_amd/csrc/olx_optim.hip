// olx_optim.hip -- multi-plane phase-only hologram loss and gradient in one pass.
//
// For u = exp(i phi) and planes p with f_p = AsmProp_{z_p}(u), I_p = |f_p|^2,
//   L = sum_p mean((I_p / mean I_p - T_p / mean T_p)^2)          (N pixels per plane)
//   dL/dI_k = (2/N) (r_k / mI - (sum_j r_j I_j) / (N mI^2)),      r = I / mI - T / mT
//   dL/df_k = 2 f_k dL/dI_k  (gradient w.r.t. Re f plus i times w.r.t. Im f)
//   dL/dphi = sum_p Im(conj(u) A_p^H dL/df_p)                    (A^H: the adjoint ASM)
// Each plane is propagated, reduced, turned into its gradient and propagated back
// before the next one, so no plane field is kept.  The forward propagation's last
// pass (or, for small fields, a pass of its own) takes the FP64 sums
// S_I, S_T, S_II, S_IT, S_TT, from which
//   sum r^2 = S_II / mI^2 - 2 S_IT / (mI mT) + S_TT / mT^2,  sum r I = S_II / mI - S_IT / mT;
// one block adds the block partials in a fixed order, so results are deterministic.
// The adjoint propagation forms dL/df from f and T as its first pass reads them and
// writes dL/dphi from its last pass, so neither dL/df nor A^H dL/df is stored.
#include "olx_common.hpp"

namespace olx {

constexpr int MP_T = 256;   // threads per block
constexpr int MP_B = 1024;  // blocks of the reduction kernels (grid-stride)

__global__ void __launch_bounds__(MP_T) k_mp_expi(const float* __restrict__ phi,
                                                  float2* __restrict__ u, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)MP_T + threadIdx.x; i < n; i += (int64_t)gridDim.x * MP_T) {
        float s, c;
        sincosf(phi[i], &s, &c);
        u[i] = make_float2(c, s);
    }
}

constexpr int MP_NS = 5;  // S_I, S_T, S_II, S_IT, S_TT

// sum over the block of MP_NS FP64 values per thread, in a fixed order
__device__ __forceinline__ void block_sums(double* v) {
    __shared__ double sh[MP_NS][MP_T];
#pragma unroll
    for (int k = 0; k < MP_NS; k++) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = MP_T / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int k = 0; k < MP_NS; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + off];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < MP_NS; k++) v[k] = sh[k][0];
    __syncthreads();
}

// pass 1: the block's S_I, S_T, S_II, S_IT, S_TT (part[k * MP_B + block])
__global__ void __launch_bounds__(MP_T) k_mp_sums(const float2* __restrict__ f,
                                                  const float* __restrict__ t, int64_t n,
                                                  double* __restrict__ part) {
    double v[MP_NS] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int64_t i = blockIdx.x * (int64_t)MP_T + threadIdx.x; i < n; i += (int64_t)MP_B * MP_T) {
        const float2 x = f[i];
        const double I = (double)x.x * x.x + (double)x.y * x.y, T = (double)t[i];
        v[0] += I;
        v[1] += T;
        v[2] += I * I;
        v[3] += I * T;
        v[4] += T * T;
    }
    block_sums(v);
    if (threadIdx.x == 0)
        for (int k = 0; k < MP_NS; k++) part[k * MP_B + blockIdx.x] = v[k];
}

// pass 2 (one block): the totals of `groups` partials (part[k * groups + i]), this
// plane's loss (added to *loss) and the
// constants of dL/df = 2 f dL/dI = (c1 r - c2) f, r = I m_i - T m_t:
// c = {m_i, m_t, c1, c2} with m = 1 / mean, c1 = 4 m_i / N, c2 = 4 m_i^2 sum(r I) / N^2
__global__ void __launch_bounds__(MP_T) k_mp_consts(const double* __restrict__ part, int groups,
                                                    int64_t n, double* __restrict__ c,
                                                    double* __restrict__ loss) {
    double S[MP_NS];
#pragma unroll
    for (int k = 0; k < MP_NS; k++) {
        S[k] = 0.0;
        for (int i = threadIdx.x; i < groups; i += MP_T) S[k] += part[k * groups + i];
    }
    block_sums(S);
    if (threadIdx.x != 0) return;
    const double N = (double)n;
    const double inv_mi = N / S[0], inv_mt = N / S[1];
    const double sr2 = S[2] * inv_mi * inv_mi - 2.0 * S[3] * inv_mi * inv_mt + S[4] * inv_mt * inv_mt;
    const double sri = S[2] * inv_mi - S[3] * inv_mt;
    c[0] = inv_mi;
    c[1] = inv_mt;
    c[2] = 4.0 / N * inv_mi;                    // 2 * (2/N) / mI
    c[3] = 4.0 / N * sri / N * inv_mi * inv_mi;  // 2 * (2/N) * sum rI / (N mI^2)
    *loss += sr2 / N;
}

__global__ void k_mp_zero(double* loss) { *loss = 0.0; }

}  // namespace olx

using namespace olx;

extern "C" int olx_multiplane_mse_c64(olx_ctx* ctx, const float* d_phi, int64_t ny, int64_t nx,
                                      int nplanes, const double* zs, const float* const* d_targets,
                                      double dy, double dx, double lambda, uint32_t flags,
                                      double* d_loss, float* d_grad) {
    if (!ctx || !d_phi || !zs || !d_targets || !d_loss) OLX_FAIL(OLX_EINVAL, "multiplane: null argument");
    if (nplanes < 1) OLX_FAIL(OLX_EINVAL, "multiplane: need at least one plane");
    if (ny <= 0 || nx <= 0) OLX_FAIL(OLX_EINVAL, "multiplane: bad field size");
    if (flags & ~(uint32_t)OLX_BANDLIMIT) OLX_FAIL(OLX_EINVAL, "multiplane: only OLX_BANDLIMIT is accepted");
    for (int p = 0; p < nplanes; p++)
        if (!d_targets[p]) OLX_FAIL(OLX_EINVAL, "multiplane: null target %d", p);
    const int64_t n = ny * nx;
    const size_t fb = sizeof(float2) * (size_t)n;
    const int64_t maxg = std::max<int64_t>(MP_B, ny / 2);  // partials: sums kernel or pass C
    OLX_TRY(ctx->mp_work.reserve(2 * fb + sizeof(double) * (MP_NS * maxg + 4)));
    float2* u = (float2*)ctx->mp_work.p;
    float2* f = (float2*)((char*)ctx->mp_work.p + fb);
    double* part = (double*)((char*)ctx->mp_work.p + 2 * fb);
    double* consts = part + MP_NS * maxg;
    hipStream_t s = ctx->stream;
    const unsigned eb = (unsigned)std::min<int64_t>((n + MP_T - 1) / MP_T, 65536);
    k_mp_expi<<<eb, MP_T, 0, s>>>(d_phi, u, n);
    k_mp_zero<<<1, 1, 0, s>>>(d_loss);
    OLX_TRY(check_launch("multiplane setup"));
    for (int p = 0; p < nplanes; p++) {
        // forward propagation; on the tiled path its last pass also takes the sums
        int groups = 0;
        OLX_TRY(asm_prop_sums(ctx, u, f, ny, nx, dy, dx, lambda, zs[p], flags & OLX_BANDLIMIT,
                              d_targets[p], part, &groups));
        {
            TimedLaunch tl(ctx, "multiplane_loss");
            if (groups == 0) {
                k_mp_sums<<<MP_B, MP_T, 0, s>>>(f, d_targets[p], n, part);
                groups = MP_B;
            }
            k_mp_consts<<<1, MP_T, 0, s>>>(part, groups, n, consts, d_loss);
        }
        OLX_TRY(check_launch("multiplane loss"));
        if (!d_grad) continue;
        // adjoint propagation of dL/df = (c1 r - c2) f, formed as its first pass reads f,
        // whose last pass writes dL/dphi (+)= Im(conj(u) A^H dL/df)
        OLX_TRY(asm_phase_grad(ctx, f, u, d_grad, p > 0 ? 1 : 0, ny, nx, dy, dx, lambda, zs[p],
                               (flags & OLX_BANDLIMIT) | OLX_ADJOINT, d_targets[p], consts));
    }
    return OLX_OK;
}
