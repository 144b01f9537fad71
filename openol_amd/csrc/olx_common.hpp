// olx_common.hpp -- shared host/device definitions of libolx (MI355X / gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/olx.h"

namespace olx {

// ---------------------------------------------------------------- errors
void set_error(const char* fmt, ...);
#define OLX_FAIL(code, ...)            \
    do {                               \
        ::olx::set_error(__VA_ARGS__); \
        return (code);                 \
    } while (0)
#define OLX_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            OLX_FAIL(OLX_EDEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                     __FILE__, __LINE__);                                               \
    } while (0)
#define OLX_TRY(call)           \
    do {                        \
        int rc_ = (call);       \
        if (rc_ != OLX_OK) return rc_; \
    } while (0)

// ---------------------------------------------------------------- device buffers
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    int reserve(size_t b);  // grows (never shrinks); contents not preserved
};

// ---------------------------------------------------------------- kernel timing
struct KernelStat {
    int64_t count = 0;
    double total_ms = 0.0;
};

}  // namespace olx

struct olx_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // twiddle table for the FFT passes (length olx::TWN, float2), built at create
    float2* twiddle = nullptr;
    // propagation workspace (padded intermediate)
    olx::DevBuf work;
    olx::DevBuf scratch2;  // second work buffer (sources aliasing, staging)
    // ordering event: recorded on the previous stream when the bound stream changes
    hipEvent_t order_event = nullptr;
    // cached 1-D tables (chirp spectra, pre/post chirps, twiddles), keyed by parameters,
    // least recently used evicted beyond olx::TABLE_CAP entries
    struct Table {
        std::shared_ptr<olx::DevBuf> buf;
        uint64_t last_use = 0;
    };
    std::map<std::string, Table> tables;
    uint64_t table_tick = 0;
    // WRP / CGH scratch
    olx::DevBuf wrp_counts, wrp_lists, wrp_pts;
    olx::DevBuf lay_pts, lay_counts;
    // point-cloud sort and reduction scratch
    olx::DevBuf obj_keys, obj_tmp;
    olx::DevBuf split_acc;  // split propagation: the summed destination field
    olx::DevBuf mp_work;    // multi-plane loss: exp(i phi), one plane's field, partial sums
    olx::DevBuf cgh_part;   // direct CGH split over point slices: one partial plane per slice
    olx::DevBuf fft64_tmp;  // FP64 1-D FFT above 8192 points: the even and odd halves
    // timing
    bool timing = false;
    int timing_period = 1;   // time every timing_period-th launch of each kernel
    std::map<std::string, uint64_t> timing_seq;  // launches seen per kernel name
    std::map<std::string, olx::KernelStat> stats;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> event_pool;
};

namespace olx {

constexpr int TWN = 16384;  // twiddle table length = largest transform length
// radix-16 stage tables after the twiddle table: for every NS = 2^p <= 1024, entries
// 4 (NS - 1) + 4 k + i = W_{16 NS}^{2^i k} (k < NS, i < 4), so that a stage's w1, w2, w4,
// w8 are two contiguous 16-byte loads per lane (fft_core.hpp)
constexpr int TWS = 8192;
constexpr size_t TABLE_CAP = 64;  // cached tables kept per context

// timing bracket around launches (no-op unless ctx->timing)
struct TimedLaunch {
    olx_ctx* ctx;
    const char* name;
    hipEvent_t a = nullptr, b = nullptr;
    TimedLaunch(olx_ctx* c, const char* n);
    ~TimedLaunch();
};
void flush_timing(olx_ctx* ctx);  // resolves pending events (synchronises them)

int check_launch(const char* what);

// table cache: returns a device buffer for key, creating it with bytes when absent;
// *fresh is set when the caller must fill it
float2* table(olx_ctx* ctx, const std::string& key, size_t count, bool* fresh, int* rc);
// cached natural-order spectrum (scaled by 1/M) of the centred chirp K (qa x^2 + qb x),
// x = (m - M/2) d, zero where |m - M/2| >= lim (olx_prop.hip)
const float2* spectrum_table(olx_ctx* ctx, int M, double d, double K, double qa, int64_t lim,
                             int adj, int* rc, double qb = 0.0);

// forward FP64 DFT (exponent -2 pi i nk/L) of one sequence of length L in place, on the
// context's stream (olx_prop64.hip): O(L log L) in one work-group's LDS, lengths above
// 8192 split once into even and odd halves.  Returns OLX_EUNSUPPORTED, without
// recording an error, when L (or L/2 above 8192) has a prime factor above 13 or L is
// above 16384; the caller then takes another route.
int fft64_one(olx_ctx* ctx, double2* u, int L);

inline bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }

// mixed-radix transforms (olx_prop_mr.hip): whether length L (2 .. 16384) factors
// into 2, 3, 5, 7, 11, 13; the three propagation passes over the row-major
// intermediate W (ny x 2nx); the in-place 2-D FFT (sign of the exponent: inv)
struct PropArgs;
bool mr_supported(int64_t L);
int mr_three_pass(olx_ctx* ctx, const float2* src, float2* dst, PropArgs& a, float2* W);
int mr_fft2d(olx_ctx* ctx, float2* u, int64_t ny, int64_t nx, bool inv);

// fetches (once per device, with the device current) the device address of the
// static workspace of olx_asm_c64_noalloc (olx_prop.hip)
int noalloc_init(int device);

// olx_asm_c64 from src into dst (which may alias src) (olx_prop.hip)
int asm_prop(olx_ctx* ctx, const void* d_src, void* d_dst, int64_t ny, int64_t nx, double dy,
             double dx, double lambda, double z, uint32_t flags);
// asm_prop that also writes, on the tiled path, each pass-C work-group's sums of I, T,
// I^2, I T, T^2 (I = |dst|^2, T = d_t) to d_part[k * G + group] and sets *groups = G
// (at most ny / 2); *groups = 0 when the untiled path ran (olx_prop.hip)
int asm_prop_sums(olx_ctx* ctx, const void* d_src, void* d_dst, int64_t ny, int64_t nx, double dy,
                  double dx, double lambda, double z, uint32_t flags, const float* d_t,
                  double* d_part, int* groups);
// ASM of g (per flags, normally OLX_ADJOINT), consumed as grad (+)= Im(conj(u) A g)
// instead of being stored; g is overwritten.  With d_t and d_c, g is first replaced
// by the loss gradient c1 (|g|^2 m_i - t m_t) - c2 times g, d_c = {m_i, m_t, c1, c2}
// (olx_prop.hip)
int asm_phase_grad(olx_ctx* ctx, void* d_g, const void* d_u, float* d_grad, int accumulate,
                   int64_t ny, int64_t nx, double dy, double dx, double lambda, double z,
                   uint32_t flags, const float* d_t = nullptr, const double* d_c = nullptr);

}  // namespace olx
