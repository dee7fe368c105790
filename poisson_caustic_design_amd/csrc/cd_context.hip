// Context, error reporting, profiling timers and small shared kernels.
#include <hip/hip_runtime.h>

#include <cstring>

#include "cd_internal.h"

namespace {

template <typename T>
__global__ void k_fill(T* __restrict__ p, size_t n, T v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// Sum partials[block][ncomp] over blocks in a fixed order (deterministic).
__global__ void k_reduce_partials(const double* __restrict__ partials, int nblocks, int ncomp,
                                  double* __restrict__ out) {
    __shared__ double smem[16];
    for (int c = 0; c < ncomp; ++c) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partials[(size_t)b * ncomp + c];
        s = cd_block_sum(s, smem);
        if (threadIdx.x == 0) out[c] = s;
    }
}

template <typename T>
int fill_impl(cd_context* ctx, T* p, size_t n, T v) {
    if (n == 0) return CD_OK;
    size_t blocks = (n + CD_BLOCK - 1) / CD_BLOCK;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)blocks), dim3(CD_BLOCK), 0, ctx->stream, p, n, v);
    CD_CHECK_LAUNCH(ctx);
    return CD_OK;
}

}  // namespace

int cd_fill_f32(cd_context* ctx, float* p, size_t n, float v) { return fill_impl(ctx, p, n, v); }
int cd_fill_f64(cd_context* ctx, double* p, size_t n, double v) { return fill_impl(ctx, p, n, v); }
int cd_fill_i32(cd_context* ctx, int* p, size_t n, int v) { return fill_impl(ctx, p, n, v); }

int cd_reduce_partials(cd_context* ctx, const double* partials, int nblocks, int ncomp,
                       double* out_dev) {
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(1024), 0, ctx->stream, partials, nblocks,
                       ncomp, out_dev);
    CD_CHECK_LAUNCH(ctx);
    return CD_OK;
}

int cd_context::ensure(DevBuf& b, size_t bytes) {
    if (b.bytes >= bytes && b.ptr) return CD_OK;
    if (b.ptr) {
        // grow-only: the old buffer may still be in use by queued work
        hipError_t e = hipStreamSynchronize(stream);
        if (e != hipSuccess) { set_error(hipGetErrorString(e)); return CD_ERR_HIP; }
        b.release();
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes ? bytes : 256);
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? CD_ERR_OOM : CD_ERR_HIP;
    }
    b.ptr = p;
    b.bytes = bytes;
    ++gen;  // device addresses changed: captured graphs are stale
    return CD_OK;
}

hipEvent_t cd_context::take_event() {
    if (!event_pool.empty()) {
        hipEvent_t e = event_pool.back();
        event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int cd_context::timer_begin() {
    if (!profiling) return CD_OK;
    open_event = take_event();
    if (!open_event) { set_error("hipEventCreate failed"); return CD_ERR_HIP; }
    hipError_t e = hipEventRecord(open_event, stream);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); return CD_ERR_HIP; }
    return CD_OK;
}

int cd_context::timer_end(int kernel, long long launches, double bytes) {
    if (!profiling || !open_event) return CD_OK;
    hipEvent_t b = take_event();
    if (!b) { set_error("hipEventCreate failed"); return CD_ERR_HIP; }
    hipError_t e = hipEventRecord(b, stream);
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); return CD_ERR_HIP; }
    pending.push_back(PendingTimer{open_event, b, kernel, launches, bytes});
    open_event = nullptr;
    return CD_OK;
}

int cd_context::resolve_timers() {
    for (auto& p : pending) {
        float ms = 0.f;
        hipError_t e = hipEventSynchronize(p.b);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.a, p.b);
        if (e != hipSuccess) { set_error(hipGetErrorString(e)); return CD_ERR_HIP; }
        timers[p.kernel].total_ms += ms;
        timers[p.kernel].launches += p.launches;
        timers[p.kernel].bytes += p.bytes;
        event_pool.push_back(p.a);
        event_pool.push_back(p.b);
    }
    pending.clear();
    return CD_OK;
}

extern "C" {

int cd_get_api_version(void) { return CD_API_VERSION; }

const char* cd_status_string(int s) {
    switch (s) {
        case CD_OK: return "ok";
        case CD_ERR_RASTER_MISS: return "raster sample outside every triangle";
        case CD_ERR_FOLDED: return "folded or degenerate mesh cell";
        case CD_ERR_NOT_CONVERGED: return "sweep cap reached before convergence";
        case CD_ERR_NAN: return "non-finite value in result";
        case CD_ERR_BAD_ARG: return "bad argument";
        case CD_ERR_HIP: return "HIP runtime error";
        case CD_ERR_NO_DEVICE: return "no HIP device";
        case CD_ERR_OOM: return "device out of memory";
        case CD_ERR_STATE: return "call out of order";
        default: return "unknown status";
    }
}

void cd_solve_opts_default(cd_solve_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->max_sweeps = 100000;
    o->stop = CD_STOP_RESIDUAL;
    o->tol = 1e-6;
    o->precision = CD_PREC_MIXED;
    o->check_every = 0;
    o->learned_init = 0;
    o->omega = 0.0;
}

int cd_context_create(int device, cd_context** out) {
    if (!out) return CD_ERR_BAD_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CD_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return CD_ERR_BAD_ARG;
    if (hipSetDevice(device) != hipSuccess) return CD_ERR_HIP;
    auto* ctx = new cd_context();
    ctx->device = device;
    // default-priority, non-blocking stream
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc((void**)&ctx->host_status, sizeof(SolveDev), hipHostMallocCoherent) !=
            hipSuccess) {
        cd_context_destroy(ctx);
        return CD_ERR_HIP;
    }
    *out = ctx;
    return CD_OK;
}

int cd_context_set_graphs(cd_context* ctx, int on) {
    if (!ctx) return CD_ERR_BAD_ARG;
    ctx->graphs = on != 0;
    return CD_OK;
}

void cd_context_destroy(cd_context* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (cd_net* n : std::vector<cd_net*>(ctx->nets)) cd_net_detach(n);
    if (ctx->fcg_exec) (void)hipGraphExecDestroy(ctx->fcg_exec);
    for (cd_design* d : std::vector<cd_design*>(ctx->designs)) cd_design_detach(d);
    ctx->e32.release();
    ctx->e32b.release();
    ctx->r32.release();
    ctx->phi_start.release();
    ctx->fcg_phi.release();
    ctx->e64w.release();
    ctx->phi64w.release();
    ctx->b64w.release();
    ctx->mask.release();
    ctx->partials.release();
    ctx->solve_dev.release();
    ctx->fill.release();
    ctx->lc_r.release();
    ctx->lc_x.release();
    ctx->lc_y.release();
    ctx->lc_stats.release();
    ctx->fcg_p.release();
    ctx->fcg_q.release();
    ctx->fcg_dev.release();
    ctx->fcg_em.release();
    ctx->fcg_emb.release();
    ctx->fcg_rm.release();
    ctx->fcg_zero.release();
    ctx->cap_phi.release();
    if (ctx->host_status) (void)hipHostFree(ctx->host_status);
    for (auto& p : ctx->pending) {
        ctx->event_pool.push_back(p.a);
        ctx->event_pool.push_back(p.b);
    }
    if (ctx->open_event) ctx->event_pool.push_back(ctx->open_event);
    for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int cd_context_synchronize(cd_context* ctx) {
    if (!ctx) return CD_ERR_BAD_ARG;
    CD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return CD_OK;
}

const char* cd_context_last_error(const cd_context* ctx) {
    return ctx ? ctx->last_error.c_str() : "null context";
}

int cd_context_set_profiling(cd_context* ctx, int enable) {
    if (!ctx) return CD_ERR_BAD_ARG;
    ctx->profiling = enable != 0;
    return CD_OK;
}

int cd_context_kernel_stats(cd_context* ctx, int kernel, double* total_ms, long long* launches,
                            double* bytes) {
    if (!ctx || kernel < 0 || kernel >= CD_KERNEL_COUNT) return CD_ERR_BAD_ARG;
    CD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CD_TRY(ctx->resolve_timers());
    if (total_ms) *total_ms = ctx->timers[kernel].total_ms;
    if (launches) *launches = ctx->timers[kernel].launches;
    if (bytes) *bytes = ctx->timers[kernel].bytes;
    return CD_OK;
}

int cd_context_reset_kernel_stats(cd_context* ctx) {
    if (!ctx) return CD_ERR_BAD_ARG;
    CD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    CD_TRY(ctx->resolve_timers());
    for (auto& t : ctx->timers) t = KernelTimer();
    return CD_OK;
}

}  // extern "C"
