// Internal declarations shared by the translation units of libcaustic_amd.so.
// gfx950 only: 64-wide wavefronts, 256-thread workgroups by default.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "cd_api.h"

#define CD_WAVE 64
#define CD_BLOCK 256

// ---------------------------------------------------------------------------
// error plumbing: every HIP call in host code goes through CD_HIP, which
// records the message on the context and returns CD_ERR_HIP from the caller.
// ---------------------------------------------------------------------------
#define CD_HIP(ctx, expr)                                                          \
    do {                                                                           \
        hipError_t _e = (expr);                                                    \
        if (_e != hipSuccess) {                                                    \
            (ctx)->set_error(std::string(#expr) + ": " + hipGetErrorString(_e));   \
            return CD_ERR_HIP;                                                     \
        }                                                                          \
    } while (0)

#define CD_CHECK_LAUNCH(ctx) CD_HIP(ctx, hipGetLastError())

#define CD_TRY(expr)                   \
    do {                               \
        int _s = (expr);               \
        if (_s != CD_OK) return _s;    \
    } while (0)

// Device buffer owned by the library (grow-only).
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Per-kernel-family timing (HIP events, recorded only when profiling is on).
struct KernelTimer {
    double total_ms = 0.0;
    long long launches = 0;
    double bytes = 0.0;
};

// Device-side status of one Poisson solve (written by kernels, read by the
// host through a pinned copy at every convergence check).
struct SolveDev {
    int done;            // set by a check kernel once the stop rule holds
    int done_check;      // check index that set `done` (-1: none)
    int sweeps_at_done;  // max-update rule: sweep count when `done` was set
    int checks;
    double bsum;         // sum of the (non-hole) right-hand side
    double bmean;        // its mean (the zero-mean projection)
    double bnorm;        // ||P b||_2
    double rsum, rnorm;  // last residual sum and ||P b - A phi||_2
    double rel;          // last relative residual
    double rel0;         // first relative residual
    double last_max_update;
    float rmean32;       // mean of the last FP32 residual (projected out in sweeps)
    int cur;             // blocked path: ping-pong buffer holding the current correction
    unsigned int maxupd_bits[16];  // per-sweep max |delta| (float bits) for the max-update rule
    unsigned int seq;    // host mirror: written last by a check (the host waits for it)
    unsigned int pad_;
};

// An event pair recorded around a group of launches; resolved (elapsed time
// read) only after the stream has been synchronised for other reasons, so
// profiling adds no host round trips.
struct PendingTimer {
    hipEvent_t a, b;
    int kernel;
    long long launches;
    double bytes;
};

// detach a net / design from its context (internal; C linkage because the
// design's definition sits in that file's extern "C" block)
extern "C" void cd_net_detach(cd_net* n);
extern "C" void cd_design_detach(cd_design* d);

struct cd_context {
    int device = 0;
    // nets and designs created on this context; destroying the context
    // detaches them (their later calls return CD_ERR_STATE)
    std::vector<cd_net*> nets;
    std::vector<cd_design*> designs;
    // bumped whenever ensure() moves a buffer (captured graphs hold addresses)
    unsigned long long gen = 0;
    // learned FCG iteration body captured as a HIP graph, and what it was
    // captured for
    bool graphs = true;
    hipGraphExec_t fcg_exec = nullptr;
    struct FcgKey {
        int w = 0, h = 0, nsmooth = 0, mid = 0;
        const void* net = nullptr;
        unsigned long long cgen = 0, ngen = 0;
        bool operator==(const FcgKey& o) const {
            return w == o.w && h == o.h && nsmooth == o.nsmooth && mid == o.mid &&
                   net == o.net && cgen == o.cgen && ngen == o.ngen;
        }
    } fcg_key;
    hipStream_t stream = nullptr;
    std::string last_error;
    bool profiling = false;
    KernelTimer timers[CD_KERNEL_COUNT];
    std::vector<hipEvent_t> event_pool;
    std::vector<PendingTimer> pending;
    hipEvent_t open_event = nullptr;

    // solver workspace (grow-only)
    DevBuf e32, e32b, r32, e64w, phi64w, b64w, mask, partials, solve_dev, fill;
    SolveDev* host_status = nullptr;  // pinned, coherent: checks write it directly
    unsigned int status_seq = 0;      // sequence number of the last check enqueued
    int ws_w = 0, ws_h = 0;

    // learned-init network (owned by the caller) and its solver workspace
    cd_net* net = nullptr;
    DevBuf lc_r, lc_x, lc_y, lc_stats;
    DevBuf fcg_p, fcg_q, fcg_dev;  // learned FCG: direction (ping-pong pair), coefficients
    DevBuf phi_start;              // learned FCG: the warm start (FP64)
    DevBuf fcg_phi;                // learned FCG: second buffer of the FP64 iterate
    // learned FCG intermediate grid: correction ping-pong, BF16 residual, a
    // device 0.0f (the mean the intermediate sweeps subtract)
    DevBuf fcg_em, fcg_emb, fcg_rm, fcg_zero;

    // training-pair capture (device buffers owned by the caller)
    float* cap_x = nullptr;
    float* cap_t = nullptr;
    int cap_capacity = 0, cap_count = 0, cap_h = 0, cap_w = 0, cap_coarsen = 0, cap_every = 1;
    DevBuf cap_phi;  // FP64 restricted iterates of the pairs of the current solve

    void set_error(const std::string& s) { last_error = s; }
    int ensure(DevBuf& b, size_t bytes);
    int timer_begin();
    int timer_end(int kernel, long long launches, double bytes);
    int resolve_timers();  // call after the stream has been synchronised
    hipEvent_t take_event();
};

// ---- shared device helpers -------------------------------------------------
__device__ __forceinline__ double cd_wave_sum(double v) {
#pragma unroll
    for (int o = CD_WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, CD_WAVE);
    return v;
}

__device__ __forceinline__ float cd_wave_max(float v) {
#pragma unroll
    for (int o = CD_WAVE / 2; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, CD_WAVE));
    return v;
}

__device__ __forceinline__ double cd_wave_min(double v) {
#pragma unroll
    for (int o = CD_WAVE / 2; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, CD_WAVE));
    return v;
}

__device__ __forceinline__ double cd_wave_max_d(double v) {
#pragma unroll
    for (int o = CD_WAVE / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, CD_WAVE));
    return v;
}

// Block-wide sum of one double per thread (any block size that is a multiple
// of 64, at most 1024).  Result valid in thread 0.
__device__ __forceinline__ double cd_block_sum(double v, double* smem /* >= 16 */) {
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int nthr = blockDim.x * blockDim.y;
    v = cd_wave_sum(v);
    if ((tid & (CD_WAVE - 1)) == 0) smem[tid / CD_WAVE] = v;
    __syncthreads();
    double t = 0.0;
    if (tid < CD_WAVE) {
        t = (tid < nthr / CD_WAVE) ? smem[tid] : 0.0;
        t = cd_wave_sum(t);
    }
    __syncthreads();
    return t;
}

// N block sums at once (each in the order cd_block_sum uses, so the results
// are the same): one pair of barriers instead of N
template <int N>
__device__ __forceinline__ void cd_block_sum_n(double (&v)[N], double* smem /* >= 16 * N */) {
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int nthr = blockDim.x * blockDim.y;
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = cd_wave_sum(v[k]);
    if ((tid & (CD_WAVE - 1)) == 0)
#pragma unroll
        for (int k = 0; k < N; ++k) smem[16 * k + tid / CD_WAVE] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k) {
        double t = 0.0;
        if (tid < CD_WAVE) {
            t = (tid < nthr / CD_WAVE) ? smem[16 * k + tid] : 0.0;
            t = cd_wave_sum(t);
        }
        v[k] = t;
    }
    __syncthreads();
}

// ---- host-side helpers shared across units ---------------------------------
// Deterministic sum of `n` doubles at stride 1 (count values) into out[k] for
// each of `ncomp` interleaved components: partials laid out [block][ncomp].
int cd_reduce_partials(cd_context* ctx, const double* partials, int nblocks, int ncomp,
                       double* out_dev);

// Fill helpers (graph-safe: kernels, not hipMemsetAsync).
int cd_fill_f32(cd_context* ctx, float* p, size_t n, float v);
int cd_fill_f64(cd_context* ctx, double* p, size_t n, double v);
int cd_fill_i32(cd_context* ctx, int* p, size_t n, int v);

// Device buffers of one plain Poisson solve, supplied by the caller instead
// of the context's grow-only workspace (the allocation-free entry point).
// Sizes for a w x h grid: partials 4*nred doubles (nred = ceil(w/64)*ceil(h/4)),
// e32/e32b 2*plane floats each in the bricked layout (mixed precision without
// holes; e32 n floats with holes), r BF16 2*plane (no holes), FP32 2*plane
// (no holes, max-update rule) or FP32 n (holes).
struct SolveWork {
    double* partials = nullptr;
    SolveDev* sd = nullptr;
    float* e32 = nullptr;
    float* e32b = nullptr;
    void* r = nullptr;
};

// Poisson solve on device-resident data.  b is the right-hand side (not
// modified), phi the warm start/result (FP64).  mask: optional per-cell
// neighbour-validity bits (NULL: no holes).  work: NULL takes the context's
// workspace (allocated on demand); otherwise those buffers, and the solve
// allocates nothing (no learned iterations, no training-pair capture).
int cd_solve_device_internal(cd_context* ctx, const double* b, double* phi, const uint8_t* mask,
                             long long valid_cells, int w, int h, const cd_solve_opts* o,
                             cd_solve_stats* st, const SolveWork* work = nullptr);

// Sizes of the solve buffers of a w x h grid: nred reduction blocks (the
// partials hold 4*nred doubles) and npk elements of the bricked layout.
void cd_solve_work_sizes(int w, int h, size_t* nred, size_t* npk);

// Hole mask of a right-hand side (per-cell neighbour-validity bits, NaN
// cells marked) into mask (w*h bytes); the number of holes is reduced into
// fill[0] through partials (nred doubles) and read back through host_buf
// (pinned host memory).
int cd_hole_mask(cd_context* ctx, const double* d_rhs, int w, int h, uint8_t* mask,
                 double* partials, double* fill, double* host_buf, long long* holes);

// Learned network forward on one coarse field (cd_learned.hip).
int cd_net_forward_internal(cd_net* n, const float* d_x, float* d_y, int h, int w);
const cd_net_arch* cd_net_arch_of(const cd_net* n);
// forward on a coarse residual normalised on the fly (stats = sum, sum^2)
int cd_net_forward_coarse(cd_net* n, const float* rc, const double* stats, float* d_y, int h,
                          int w);
// the net's input buffer for an h x w coarse field (BF16, 8 channels per
// cell: the normalised residual, then zeros), and the forward on it
int cd_net_coarse_input(cd_net* n, int h, int w, unsigned short** x8);
int cd_net_forward_prepared(cd_net* n, float* d_y, int h, int w);
// bumped whenever the net's activation buffers move (batch shape changes)
unsigned long long cd_net_gen(const cd_net* n);
