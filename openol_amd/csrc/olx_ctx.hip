// olx_ctx.hip -- context, errors, memory helpers and instrumentation of libolx.
#include <cstdarg>
#include <cstring>

#include "olx_common.hpp"

namespace olx {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int DevBuf::reserve(size_t b) {
    if (b <= bytes && p) return OLX_OK;
    if (p) {
        // the old buffer may still be in use by queued work on any stream
        (void)hipDeviceSynchronize();
        (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    if (hipMalloc(&p, b) != hipSuccess) {
        p = nullptr;
        OLX_FAIL(OLX_ENOMEM, "device allocation of %zu bytes failed", b);
    }
    bytes = b;
    return OLX_OK;
}

int check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) OLX_FAIL(OLX_EDEVICE, "%s: launch failed: %s", what, hipGetErrorString(e));
    return OLX_OK;
}

float2* table(olx_ctx* ctx, const std::string& key, size_t count, bool* fresh, int* rc) {
    auto it = ctx->tables.find(key);
    if (it != ctx->tables.end()) {
        *fresh = false;
        it->second.last_use = ++ctx->table_tick;
        return (float2*)it->second.buf->p;
    }
    if (ctx->tables.size() >= TABLE_CAP) {
        // evict the least recently used table; queued work may still read it.  A call
        // fetches at most six tables, so the victim is never one of the current call's.
        auto victim = ctx->tables.begin();
        for (auto i = ctx->tables.begin(); i != ctx->tables.end(); ++i)
            if (i->second.last_use < victim->second.last_use) victim = i;
        (void)hipStreamSynchronize(ctx->stream);
        ctx->tables.erase(victim);
    }
    auto buf = std::make_shared<DevBuf>();
    if ((*rc = buf->reserve(sizeof(float2) * count)) != OLX_OK) return nullptr;
    ctx->tables[key] = {buf, ++ctx->table_tick};
    *fresh = true;
    return (float2*)buf->p;
}

static hipEvent_t get_event(olx_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

TimedLaunch::TimedLaunch(olx_ctx* c, const char* n) : ctx(c), name(n) {
    if (!ctx->timing) return;
    // each kernel name is sampled on every timing_period-th of its own launches
    if (ctx->timing_seq[name]++ % (uint64_t)ctx->timing_period != 0) return;  // a, b stay null
    a = get_event(ctx);
    b = get_event(ctx);
    if (a) (void)hipEventRecord(a, ctx->stream);
}

TimedLaunch::~TimedLaunch() {
    if (!ctx->timing || !a || !b) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back({name, {a, b}});
    if (ctx->pending.size() > 4096) flush_timing(ctx);
}

void flush_timing(olx_ctx* ctx) {
    for (auto& p : ctx->pending) {
        hipEvent_t a = p.second.first, b = p.second.second;
        float ms = 0.f;
        if (hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess) {
            KernelStat& s = ctx->stats[p.first];
            s.count += 1;
            s.total_ms += ms;
        }
        ctx->event_pool.push_back(a);
        ctx->event_pool.push_back(b);
    }
    ctx->pending.clear();
}

// twiddle table W_TWN^m = exp(-2 pi i m / TWN), computed in FP64, rounded once
__global__ void k_twiddle(float2* tw, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    double s, c;
    sincospi(-2.0 * (double)m / (double)n, &s, &c);
    tw[m] = make_float2((float)c, (float)s);
}

// the radix-16 stage tables (olx::TWS): the same FP64 expression as k_twiddle for the
// table index each entry stands for, so the values are bitwise those of the table
__global__ void k_twiddle_stages(float2* st, int n) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 4 * 2047) return;
    const int q = idx / 4;
    const int p = 31 - __clz(q + 1);  // NS = 2^p: q in [NS - 1, 2 NS - 1)
    const int ns = 1 << p, off = idx - 4 * (ns - 1);
    const int k = off / 4, i = off % 4;
    const int m = (k << i) * (n / (16 * ns));
    double s, c;
    sincospi(-2.0 * (double)m / (double)n, &s, &c);
    st[idx] = make_float2((float)c, (float)s);
}

// plain copy, 16 bytes per lane.  Each work-group copies a contiguous span in 16 KiB
// blocks, four independent loads per thread before their stores (a grid-stride
// loop with one load per iteration reached 4.56 TB/s at 1 GiB).
__global__ void k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n) {
    constexpr int U = 4;
    const size_t blk = (size_t)U * blockDim.x;                      // vectors per block
    const size_t nblk = (n + blk - 1) / blk;
    const size_t per = (nblk + gridDim.x - 1) / gridDim.x;          // blocks per work-group
    const size_t b0 = blockIdx.x * per, b1 = b0 + per < nblk ? b0 + per : nblk;
    for (size_t b = b0; b < b1; b++) {
        const size_t base = b * blk + threadIdx.x;
        if (base + (U - 1) * blockDim.x < n) {
            uint4 v[U];
#pragma unroll
            for (int k = 0; k < U; k++) v[k] = src[base + k * blockDim.x];
#pragma unroll
            for (int k = 0; k < U; k++) dst[base + k * blockDim.x] = v[k];
        } else {
            for (int k = 0; k < U; k++)
                if (base + k * blockDim.x < n) dst[base + k * blockDim.x] = src[base + k * blockDim.x];
        }
    }
}

}  // namespace olx

using namespace olx;

extern "C" const char* olx_last_error(void) { return g_err; }
extern "C" int olx_version(void) { return 10000; }

extern "C" int olx_ctx_create(olx_ctx** out, int device) {
    if (!out) OLX_FAIL(OLX_EINVAL, "null output pointer");
    *out = nullptr;
    int n = 0;
    OLX_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) OLX_FAIL(OLX_EINVAL, "device %d out of range (%d)", device, n);
    OLX_HIP(hipSetDevice(device));
    olx_ctx* ctx = new olx_ctx();
    ctx->device = device;
    if (hipMalloc((void**)&ctx->twiddle, sizeof(float2) * (TWN + TWS)) != hipSuccess) {
        delete ctx;
        OLX_FAIL(OLX_ENOMEM, "twiddle table allocation failed");
    }
    if (int rc = noalloc_init(device); rc != OLX_OK) {
        (void)hipFree(ctx->twiddle);
        delete ctx;
        return rc;
    }
    k_twiddle<<<TWN / 256, 256>>>(ctx->twiddle, TWN);
    k_twiddle_stages<<<TWS / 256, 256>>>(ctx->twiddle + TWN, TWN);
    if (hipDeviceSynchronize() != hipSuccess) {
        (void)hipFree(ctx->twiddle);
        delete ctx;
        OLX_FAIL(OLX_EDEVICE, "twiddle table construction failed");
    }
    *out = ctx;
    return OLX_OK;
}

extern "C" int olx_ctx_destroy(olx_ctx* ctx) {
    if (!ctx) return OLX_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto& p : ctx->pending) {
        (void)hipEventDestroy(p.second.first);
        (void)hipEventDestroy(p.second.second);
    }
    ctx->pending.clear();
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->order_event) (void)hipEventDestroy(ctx->order_event);
    if (ctx->twiddle) (void)hipFree(ctx->twiddle);
    delete ctx;
    return OLX_OK;
}

extern "C" int olx_ctx_set_stream(olx_ctx* ctx, void* hip_stream) {
    if (!ctx) OLX_FAIL(OLX_EINVAL, "null context");
    hipStream_t s = (hipStream_t)hip_stream;
    if (s == ctx->stream) return OLX_OK;
    // The workspaces and tables are shared by every stream the context is bound to:
    // work queued from now on waits for what the previous stream already holds.
    if (!ctx->order_event) OLX_HIP(hipEventCreateWithFlags(&ctx->order_event, hipEventDisableTiming));
    OLX_HIP(hipEventRecord(ctx->order_event, ctx->stream));
    OLX_HIP(hipStreamWaitEvent(s, ctx->order_event, 0));
    ctx->stream = s;
    return OLX_OK;
}

extern "C" int olx_ctx_clear_tables(olx_ctx* ctx) {
    if (!ctx) OLX_FAIL(OLX_EINVAL, "null context");
    OLX_HIP(hipStreamSynchronize(ctx->stream));
    ctx->tables.clear();
    return OLX_OK;
}

extern "C" int olx_ctx_synchronize(olx_ctx* ctx) {
    if (!ctx) OLX_FAIL(OLX_EINVAL, "null context");
    OLX_HIP(hipStreamSynchronize(ctx->stream));
    return OLX_OK;
}

extern "C" int olx_dev_alloc(olx_ctx* ctx, void** d_ptr, size_t bytes) {
    if (!ctx || !d_ptr) OLX_FAIL(OLX_EINVAL, "null argument");
    *d_ptr = nullptr;
    if (bytes == 0) return OLX_OK;
    if (hipMalloc(d_ptr, bytes) != hipSuccess) {
        *d_ptr = nullptr;
        OLX_FAIL(OLX_ENOMEM, "device allocation of %zu bytes failed", bytes);
    }
    return OLX_OK;
}

extern "C" int olx_dev_free(olx_ctx* ctx, void* d_ptr) {
    if (!ctx) OLX_FAIL(OLX_EINVAL, "null context");
    if (!d_ptr) return OLX_OK;
    OLX_HIP(hipStreamSynchronize(ctx->stream));
    OLX_HIP(hipFree(d_ptr));
    return OLX_OK;
}

extern "C" int olx_copy_h2d(olx_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
    if (!ctx || (bytes && (!d_dst || !h_src))) OLX_FAIL(OLX_EINVAL, "null argument");
    if (!bytes) return OLX_OK;
    OLX_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    OLX_HIP(hipStreamSynchronize(ctx->stream));
    return OLX_OK;
}

extern "C" int olx_copy_d2h(olx_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
    if (!ctx || (bytes && (!h_dst || !d_src))) OLX_FAIL(OLX_EINVAL, "null argument");
    if (!bytes) return OLX_OK;
    OLX_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    OLX_HIP(hipStreamSynchronize(ctx->stream));
    return OLX_OK;
}

extern "C" int olx_dev_memset(olx_ctx* ctx, void* d_ptr, int value, size_t bytes) {
    if (!ctx || (bytes && !d_ptr)) OLX_FAIL(OLX_EINVAL, "null argument");
    if (!bytes) return OLX_OK;
    OLX_HIP(hipMemsetAsync(d_ptr, value, bytes, ctx->stream));
    return OLX_OK;
}

extern "C" int olx_ctx_enable_timing(olx_ctx* ctx, int on) {
    if (!ctx) OLX_FAIL(OLX_EINVAL, "null context");
    if (on < 0) OLX_FAIL(OLX_EINVAL, "timing: negative sampling period");
    if (!on) flush_timing(ctx);
    ctx->timing = on != 0;
    ctx->timing_period = on > 1 ? on : 1;
    ctx->timing_seq.clear();
    return OLX_OK;
}

extern "C" int olx_ctx_reset_timing(olx_ctx* ctx) {
    if (!ctx) OLX_FAIL(OLX_EINVAL, "null context");
    flush_timing(ctx);
    ctx->stats.clear();
    return OLX_OK;
}

extern "C" int olx_ctx_timing_report(olx_ctx* ctx, char* buf, size_t buflen) {
    if (!ctx || !buf || buflen == 0) OLX_FAIL(OLX_EINVAL, "null argument");
    flush_timing(ctx);
    size_t off = 0;
    buf[0] = 0;
    for (auto& kv : ctx->stats) {
        int w = snprintf(buf + off, buflen - off, "%s %lld %.6f\n", kv.first.c_str(),
                         (long long)kv.second.count, kv.second.total_ms);
        if (w < 0 || (size_t)w >= buflen - off) break;
        off += (size_t)w;
    }
    return OLX_OK;
}

extern "C" int olx_stream_copy(olx_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
    if (!ctx || !d_dst || !d_src) OLX_FAIL(OLX_EINVAL, "null argument");
    if (bytes % 16) OLX_FAIL(OLX_EINVAL, "stream_copy: bytes must be a multiple of 16");
    const size_t n = bytes / 16;
    {
        TimedLaunch tl(ctx, "stream_copy");
        k_copy16<<<4096, 256, 0, ctx->stream>>>((uint4*)d_dst, (const uint4*)d_src, n);
    }
    return check_launch("stream_copy");
}
