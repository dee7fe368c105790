// Poisson solver for the reference operator (src/solver.cpp:12-61):
//     sum_{valid nbr n} phi_n - cnt * phi_c = D_c
// on the cell-centred grid with Neumann edges (a neighbour outside the grid or
// on a NaN hole of D is not counted).  Replaces poisson_solver
// (src/solver.h:8, src/solver.cpp:70-147).
//
// Method (MI355X-first):
//   * red-black SOR.  Red-black ordering removes the reference's tile data
//     race by construction.  omega defaults to the optimum for this operator
//     (cd_sor_omega); the reference's 2/(1+3.14159265/width) (solver.cpp:71)
//     is kept for the drop-in entry points.
//   * CD_PREC_MIXED: sweeps run in FP32 on a correction e with the right-hand
//     side r = P b - A phi64 (FP64 residual, zero-mean projected); every
//     `check_every` sweeps e is folded into the FP64 base (defect correction)
//     and the FP64 relative residual is evaluated on the device.
//   * CD_PREC_FP64: the same sweeps in FP64 directly on phi.
//   * stop rules are evaluated by device kernels that set a `done` flag; later
//     sweep launches of the same batch see the flag and exit immediately.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <type_traits>

#include "cd_internal.h"

namespace {

constexpr int SWEEP_BX = 64;   // x-pairs per block (128 columns)
constexpr int SWEEP_BY = 4;    // rows per block
constexpr int RED_BX = 64;
constexpr int RED_BY = 4;

// mask bits: neighbour validity (left, up, right, down) and "cell is a hole"
constexpr uint8_t M_L = 1, M_U = 2, M_R = 4, M_D = 8, M_HOLE = 0x80;

// ---------------------------------------------------------------------------
// right-hand-side statistics, hole mask
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_build_mask(const double* __restrict__ b,
                                                    uint8_t* __restrict__ mask, int w, int h,
                                                    double* __restrict__ partials) {
    __shared__ double smem[16];
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y * blockDim.y + threadIdx.y;
    double holes = 0.0;
    if (x < w && y < h) {
        size_t i = (size_t)y * w + x;
        uint8_t m = 0;
        if (isnan(b[i])) { m = M_HOLE; holes = 1.0; }
        if (x > 0 && !isnan(b[i - 1])) m |= M_L;
        if (y > 0 && !isnan(b[i - w])) m |= M_U;
        if (x < w - 1 && !isnan(b[i + 1])) m |= M_R;
        if (y < h - 1 && !isnan(b[i + w])) m |= M_D;
        mask[i] = m;
    }
    double s = cd_block_sum(holes, smem);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// pass 1: sum of b over non-hole cells; pass 2 (mean given): sum of (b-mean)^2
__global__ __launch_bounds__(256) void k_rhs_moments(const double* __restrict__ b, int w, int h,
                                                     const SolveDev* __restrict__ st, int pass,
                                                     double* __restrict__ partials) {
    __shared__ double smem[16];
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y * blockDim.y + threadIdx.y;
    double v = 0.0;
    if (x < w && y < h) {
        double bv = b[(size_t)y * w + x];
        if (!isnan(bv)) {
            if (pass == 0) v = bv;
            else { double d = bv - st->bmean; v = d * d; }
        }
    }
    double s = cd_block_sum(v, smem);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        partials[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// single block: reduce partials and finish the right-hand-side statistics
__global__ void k_rhs_finish(const double* __restrict__ partials, int nblocks, SolveDev* st,
                             int pass, double count) {
    __shared__ double smem[16];
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += partials[b];
    s = cd_block_sum(s, smem);
    if (threadIdx.x == 0) {
        if (pass == 0) {
            st->done = 0;
            st->done_check = -1;
            st->sweeps_at_done = -1;
            st->checks = 0;
            st->bsum = s;
            st->bmean = count > 0 ? s / count : 0.0;
            st->rel = INFINITY;
            st->rel0 = INFINITY;
            st->last_max_update = INFINITY;
            st->rmean32 = 0.f;
            st->cur = 0;
            for (int k = 0; k < 16; ++k) st->maxupd_bits[k] = 0u;
        } else {
            st->bnorm = sqrt(s);
        }
    }
}

// ---------------------------------------------------------------------------
// defect correction: phi64 += e32, e32 = 0 (elementwise)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_flush(double* __restrict__ phi, float* __restrict__ e,
                                               size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        float ev = e[i];
        phi[i] += (double)ev;
        e[i] = 0.f;
    }
}

// Bricked checkerboard layout of the blocked sweeps.  Colour c = (x+y)&1 and
// packed column j = x>>1 split the grid into two planes of h x wp cells
// (wp = (w+1)/2); each plane is stored in bricks of 8 rows x 64 packed columns
// (2 KiB of FP32), bricks row-major.  A wave of the blocked kernel reads
// 8 consecutive rows of one brick as one contiguous 2 KiB stream.
struct Brick {
    int wpb;       // packed columns rounded up to 64
    int nbx;       // bricks per row of bricks
    size_t plane;  // elements per colour plane (rows rounded up to 8)
};

inline Brick make_brick(int w, int h) {
    Brick b;
    const int wp = (w + 1) / 2;
    b.wpb = (wp + 63) / 64 * 64;
    b.nbx = b.wpb / 64;
    b.plane = (size_t)((h + 7) / 8 * 8) * b.wpb;
    return b;
}

__device__ __forceinline__ size_t brick_row(const Brick& b, int c, int Y) {
    return (size_t)c * b.plane + ((size_t)(Y >> 3) * b.nbx) * 512 + (size_t)(Y & 7) * 64;
}
__device__ __forceinline__ unsigned brick_col(int j) { return (unsigned)((j >> 6) * 512 + (j & 63)); }
// bricked index of cell (x, y)
__device__ __forceinline__ size_t brick_at(const Brick& bk, int x, int y) {
    return brick_row(bk, (x + y) & 1, y) + brick_col(x >> 1);
}
// XCD-contiguous workgroup order.  The dispatcher hands consecutive workgroups
// (x fastest, then y, then z) round-robin to the 8 XCDs, each with its own L2;
// this maps the dispatch index to an order in which XCD k takes the k-th
// contiguous eighth of the n workgroups (a bijection on [0, n)), so that
// neighbouring workgroups, whose halos overlap, share an L2.
__device__ __forceinline__ int xcd_order() {
    const int n = gridDim.x * gridDim.y * gridDim.z;
    const int lin = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    const int q = n >> 3, r = n & 7, x = lin & 7;
    return x * q + min(x, r) + (lin >> 3);
}

// Buffer access to a bricked plane pair of at most 2^31 bytes: scalar
// resource + scalar (row) byte offset + the lane's vector byte offset, no
// per-lane 64-bit address arithmetic.  Every access stays in range.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void* p, size_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t rs, unsigned voff, unsigned soff);
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t rs, unsigned voff,
                                                 unsigned soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0));
}
template <>
__device__ __forceinline__ unsigned short buf_load<unsigned short>(__amdgpu_buffer_rsrc_t rs,
                                                                   unsigned voff, unsigned soff) {
    return __builtin_amdgcn_raw_buffer_load_b16(rs, voff, soff, 0);
}

// FP32 -> BF16 (round to nearest even) and back; finite inputs only
__device__ __forceinline__ unsigned short f32_to_bf16(float x) {
    const unsigned u = __float_as_uint(x);
    return (unsigned short)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_to_f32(unsigned short v) {
    return __uint_as_float((unsigned)v << 16);
}
// right-hand side element of the blocked sweeps: BF16 or FP32 storage
__device__ __forceinline__ float rhs_to_f32(unsigned short v) { return bf16_to_f32(v); }
__device__ __forceinline__ float rhs_to_f32(float v) { return v; }

// blocked path: fold the current ping-pong correction into phi64, zero both
__global__ __launch_bounds__(256) void k_flush_packed(double* __restrict__ phi,
                                                      float* __restrict__ ea,
                                                      float* __restrict__ eb,
                                                      const SolveDev* __restrict__ st, int w,
                                                      int h, Brick bk) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    const size_t pi = brick_row(bk, (x + y) & 1, y) + brick_col(x >> 1);
    const float ev = st->cur ? eb[pi] : ea[pi];
    phi[(size_t)y * w + x] += (double)ev;
    ea[pi] = 0.f;
    eb[pi] = 0.f;
}

// ---------------------------------------------------------------------------
// FP64 residual r = (b - bmean) - A phi  (+ FP32 copy for the inner sweeps)
// partials[block] = (sum r, sum r^2)
// ---------------------------------------------------------------------------
// OUT: 0 norms only, 1 FP32 copy of r (natural layout), 2 BF16 copy of r
// (bricked layout, blocked sweeps), 3 FP32 copy of r in the bricked layout
// (blocked sweeps under the max-update rule).  A thread takes ROWS rows (stride
// blockDim.y) of its column before the two block reductions, so a launch
// reduces over few workgroups.  The loads of 4 rows at a time are issued
// before the first is used (clamped, always in-range addresses; cells outside
// the grid contribute nothing): a thread keeps 24 loads in flight instead of
// waiting for each row in turn.
template <int OUT, int ROWS, bool MASK>
__device__ __forceinline__ void residual_rows(const double* __restrict__ phi,
                                              const double* __restrict__ b,
                                              const uint8_t* __restrict__ mask, double bmean,
                                              float* __restrict__ r32,
                                              unsigned short* __restrict__ r16, const Brick& bk,
                                              int w, int h, double& rs, double& rq) {
    constexpr int G = ROWS < 4 ? ROWS : 4;  // rows whose loads are in flight together
    const int ob = xcd_order();
    const int x0 = (ob % gridDim.x) * blockDim.x + threadIdx.x;
    const int y0 = (ob / gridDim.x) * blockDim.y * ROWS + threadIdx.y;
    const int x = min(x0, w - 1);
    const int xl = max(x - 1, 0), xr = min(x + 1, w - 1);
#pragma unroll
    for (int g = 0; g < ROWS; g += G) {
        double c[G], l[G], r[G], u[G], d[G], bb[G];
        uint8_t m[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int y = min(y0 + (g + k) * (int)blockDim.y, h - 1);
            const size_t row = (size_t)y * w;
            c[k] = phi[row + x];
            l[k] = phi[row + xl];
            r[k] = phi[row + xr];
            u[k] = phi[(size_t)max(y - 1, 0) * w + x];
            d[k] = phi[(size_t)min(y + 1, h - 1) * w + x];
            bb[k] = b[row + x];
            if (MASK) m[k] = mask[row + x];
        }
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int y = y0 + (g + k) * (int)blockDim.y;
            if (x0 >= w || y >= h) break;
            double s = 0.0, cnt = 0.0;
            bool hole = false;
            if (MASK) {
                hole = (m[k] & M_HOLE) != 0;
                if (m[k] & M_L) { cnt += 1.0; s += l[k]; }
                if (m[k] & M_U) { cnt += 1.0; s += u[k]; }
                if (m[k] & M_R) { cnt += 1.0; s += r[k]; }
                if (m[k] & M_D) { cnt += 1.0; s += d[k]; }
            } else {
                if (x > 0) { cnt += 1.0; s += l[k]; }
                if (y > 0) { cnt += 1.0; s += u[k]; }
                if (x < w - 1) { cnt += 1.0; s += r[k]; }
                if (y < h - 1) { cnt += 1.0; s += d[k]; }
            }
            double rv = 0.0;
            if (!hole) {
                rv = (bb[k] - bmean) - (s - cnt * c[k]);
                rs += rv;
                rq += rv * rv;
            }
            const size_t i = (size_t)y * w + x;
            if (OUT == 1) r32[i] = (float)rv;
            if (OUT == 2)
                r16[brick_row(bk, (x + y) & 1, y) + brick_col(x >> 1)] = f32_to_bf16((float)rv);
            if (OUT == 3) r32[brick_row(bk, (x + y) & 1, y) + brick_col(x >> 1)] = (float)rv;
        }
    }
}
template <int OUT, int ROWS>
__global__ __launch_bounds__(256) void k_residual(const double* __restrict__ phi,
                                                  const double* __restrict__ b,
                                                  const uint8_t* __restrict__ mask,
                                                  const SolveDev* __restrict__ st,
                                                  float* __restrict__ r32,
                                                  unsigned short* __restrict__ r16, Brick bk,
                                                  int w, int h, double* __restrict__ partials) {
    __shared__ double smem[32];
    const double bmean = st->bmean;
    double rs = 0.0, rq = 0.0;
    if (mask)
        residual_rows<OUT, ROWS, true>(phi, b, mask, bmean, r32, r16, bk, w, h, rs, rq);
    else
        residual_rows<OUT, ROWS, false>(phi, b, mask, bmean, r32, r16, bk, w, h, rs, rq);
    double aq[2] = {rs, rq};
    cd_block_sum_n(aq, smem);
    const double a = aq[0], q = aq[1];
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const size_t bid = (size_t)xcd_order();
        partials[2 * bid] = a;
        partials[2 * bid + 1] = q;
    }
}
// launch of k_residual<OUT, rows> for rows in {1, 2, 4, 8}
template <int OUT>
void launch_residual(int rows, dim3 grid, dim3 blk, hipStream_t s, const double* phi,
                     const double* b, const uint8_t* mask, const SolveDev* st, float* r32,
                     unsigned short* r16, const Brick& bk, int w, int h, double* partials) {
    switch (rows) {
    case 8: hipLaunchKernelGGL((k_residual<OUT, 8>), grid, blk, 0, s, phi, b, mask, st, r32, r16, bk, w, h, partials); break;
    case 4: hipLaunchKernelGGL((k_residual<OUT, 4>), grid, blk, 0, s, phi, b, mask, st, r32, r16, bk, w, h, partials); break;
    case 2: hipLaunchKernelGGL((k_residual<OUT, 2>), grid, blk, 0, s, phi, b, mask, st, r32, r16, bk, w, h, partials); break;
    default: hipLaunchKernelGGL((k_residual<OUT, 1>), grid, blk, 0, s, phi, b, mask, st, r32, r16, bk, w, h, partials); break;
    }
}

// single block: finish a convergence check.  The status also goes to the
// host's pinned coherent mirror, its sequence word `seq` last, after every
// other word has been written: the host waits for that word instead of
// synchronising the stream (a round trip of ~30 us per check).
__global__ void k_check(const double* __restrict__ partials, int nblocks, SolveDev* st,
                        SolveDev* host_st, unsigned seq, double count, int check_id,
                        int stop_rule, double tol) {
    constexpr int kWords = (int)(sizeof(SolveDev) / 4);
    constexpr int kSeq = (int)(offsetof(SolveDev, seq) / 4);
    static_assert(sizeof(SolveDev) % 4 == 0 && kWords <= 1024, "status words");
    __shared__ double smem[32];
    __shared__ unsigned words[kWords];
    double a = 0.0, q = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
        a += partials[2 * b];
        q += partials[2 * b + 1];
    }
    double aq[2] = {a, q};
    cd_block_sum_n(aq, smem);
    a = aq[0];
    q = aq[1];
    if (threadIdx.x == 0) {
        double mean = count > 0 ? a / count : 0.0;
        double rn2 = q - count * mean * mean;
        // rounding can make rn2 slightly negative; a NaN must stay NaN (a
        // non-finite iterate never meets the stop rule)
        double rn = rn2 > 0.0 ? sqrt(rn2) : (rn2 == rn2 ? 0.0 : rn2);
        double rel = st->bnorm > 0 ? rn / st->bnorm : (rn > 0 ? INFINITY : 0.0);
        st->rsum = a;
        st->rnorm = rn;
        st->rel = rel;
        if (st->checks == 0) st->rel0 = rel;
        st->checks += 1;
        st->rmean32 = (float)mean;
        if (stop_rule == CD_STOP_RESIDUAL && !st->done && rel <= tol) {
            st->done = 1;
            st->done_check = check_id;
        }
        st->seq = seq;
        const unsigned* src = reinterpret_cast<const unsigned*>(st);
        for (int k = 0; k < kWords; ++k) words[k] = src[k];
    }
    __syncthreads();
    // one word per thread, then the sequence word as a system-scope release
    // (it makes the other threads' words, ordered before it by the barrier,
    // visible to the host first)
    unsigned* dst = reinterpret_cast<unsigned*>(host_st);
    if (threadIdx.x < kWords && threadIdx.x != kSeq) dst[threadIdx.x] = words[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(dst + kSeq, words[kSeq], __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---------------------------------------------------------------------------
// red-black SOR half-sweep.  `color` 0 updates cells with (x+y) even.
// e: iterate (FP32 correction or FP64 phi); r: right-hand side with its mean
// *rmean projected out; the update is the reference's
//   delta = omega / cnt * (sum - cnt * val - rhs)   (solver.cpp:47, 56)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_rb_sweep(T* __restrict__ e, const T* __restrict__ r,
                                                  const uint8_t* __restrict__ mask, int w, int h,
                                                  int color, T omega,
                                                  const T* __restrict__ rmean, SolveDev* st,
                                                  int track_max, int sweep) {
    if (st->done) return;
    __shared__ float smax[SWEEP_BX * SWEEP_BY / CD_WAVE];
    const int y = blockIdx.y * SWEEP_BY + threadIdx.y;
    const int x = 2 * (blockIdx.x * SWEEP_BX + threadIdx.x) + ((y + color) & 1);
    float m = 0.f;
    if (y < h && x < w) {
        const size_t i = (size_t)y * w + x;
        const T c = e[i];
        T s = 0, cnt = 0;
        bool hole = false;
        if (mask) {
            const uint8_t mk = mask[i];
            hole = (mk & M_HOLE) != 0;
            if (mk & M_L) { cnt += 1; s += e[i - 1]; }
            if (mk & M_U) { cnt += 1; s += e[i - w]; }
            if (mk & M_R) { cnt += 1; s += e[i + 1]; }
            if (mk & M_D) { cnt += 1; s += e[i + w]; }
        } else {
            if (x > 0) { cnt += 1; s += e[i - 1]; }
            if (y > 0) { cnt += 1; s += e[i - w]; }
            if (x < w - 1) { cnt += 1; s += e[i + 1]; }
            if (y < h - 1) { cnt += 1; s += e[i + w]; }
        }
        if (!hole) {
            const T rv = r[i] - *rmean;
            const T d = omega / cnt * (s - cnt * c - rv);
            e[i] = c + d;
            m = (float)fabs(d);
        }
    }
    if (track_max) {
        const int tid = threadIdx.y * SWEEP_BX + threadIdx.x;
        m = cd_wave_max(m);
        if ((tid & (CD_WAVE - 1)) == 0) smax[tid / CD_WAVE] = m;
        __syncthreads();
        if (tid == 0) {
            float t = smax[0];
            for (int k = 1; k < SWEEP_BX * SWEEP_BY / CD_WAVE; ++k) t = fmaxf(t, smax[k]);
            // non-negative floats order like their bit patterns
            atomicMax(&st->maxupd_bits[sweep & 1], __float_as_uint(t));
        }
    }
}

// ---------------------------------------------------------------------------
// Temporally blocked red-black SOR (FP32, packed layout, no holes).
//
// A workgroup of W waves owns a region of W*R rows x 128 columns; the halo is
// H cells on every side.  Lane l of wave v keeps packed column l (cells
// 2l and 2l+1 of every row, one per colour) of rows v*R .. v*R+R-1 in
// registers, with the cells' scaled right-hand side.  Up/down neighbours of a
// cell are the other colour in the rows above/below: registers of the same
// lane, except at a wave's first/last row, which comes through a small LDS
// exchange.  Left/right neighbours are the other colour in the same row at
// packed column l-1+p or l+p (p = column parity of the cell): the same lane or
// the adjacent lane, fetched with a DPP wave shift folded into the first add.
// Half-sweep s (1-based) updates the cells at distance >= s from the region
// border, whose neighbours are exact at that point; after nhalf <= H
// half-sweeps the interior (distance >= H) equals nhalf global half-sweeps
// and is written to the other ping-pong buffer.
// ---------------------------------------------------------------------------
// Geometry is fixed at compile time; CD_BK_W / CD_BK_R / CD_BK_H exist to
// compare variants on the hardware.  The kernel is bound by vector
// ALU issue and by the load phase that precedes the sweeps (a round of
// resident workgroups loads, then sweeps): the interior update takes five
// VALU operations per cell (three adds, the first carrying the DPP shift, and
// two FMAs), border tiles six.  Measured at 2048^2, 5 sweeps per launch:
// 14.3-14.8 us per launch with these updates against 17.6-18.0 us for the
// earlier form (fma(-4, v, sum) - r, then fma(oc, t, v); selects on border
// tiles; 64-bit per-lane addresses); right-hand side in LDS 15.0-15.4 us.
// Halo (= half-sweeps per launch): H = 6 / 8 / 10 / 12 give 3.76 / 3.27 /
// 2.94 / 2.99 us per sweep at 2048^2 and 1.96 / 1.56 / 1.36 / 1.45 at 1024^2;
// 16 waves x 6 rows 3.14 and 1.69.  Earlier geometry comparisons (4 sweeps per
// launch, same 128 x 64 region): 4 waves x 16 rows 22.1 us, 8 x 8 17.0 us,
// 12 x 6 17.7 us, 16 x 4 16.8 us; 8 x 12 and 8 x 16 (taller regions, fewer
// workgroups) 21.7 and 27.0 us.  A persistent variant (one round of
// workgroups walking the tiles, the next tile's loads issued before the
// current tile's sweeps) measured 21.1-21.5 us against 17.6-18.0 us: the
// loop-carried registers cost copies in every half-sweep and the static tile
// split loses the dispatcher's load balance; two tiles per workgroup (the
// second's loads issued before the first's sweeps) 17.2 us against 14.8 us;
// delaying the start of a quarter of the first round's workgroups by 0.9-2.6
// us (to stagger the load and sweep phases) 15.2-16.6 us.
// H = 10 also keeps the learned iteration's 8 smoothing sweeps at two
// launches.
#ifndef CD_BK_R
#define CD_BK_R 4
#endif
#ifndef CD_BK_H
#define CD_BK_H 10
#endif
#ifndef CD_BK_W
#define CD_BK_W 16
#endif
// cache policy bits of the blocked kernel's stores
#ifndef CD_BK_STORE_POLICY
#define CD_BK_STORE_POLICY 16  // sc1: written through, not left dirty in L2
#endif
constexpr int BK_RX = 128;      // region columns (64 packed columns, one per lane)
constexpr int BK_R = CD_BK_R;   // rows per wave
constexpr int BK_NW = CD_BK_W;  // waves per workgroup
constexpr int BK_THREADS = 64 * BK_NW;
constexpr int BK_RY = BK_NW * BK_R;

// Right-hand side of the region in the registers of the owning thread:
// element (colour c, row k of the wave).  (In LDS it measured 4% slower.)
template <int R>
struct RhsRegsT {
    float v[2][R];
    __device__ __forceinline__ float get(int c, int k, int, int) const { return v[c][k]; }
    __device__ __forceinline__ void set(int c, int k, int, int, float x) { v[c][k] = x; }
};
using RhsRegs = RhsRegsT<BK_R>;

// value of lane l-1 (0 in lane 0) / lane l+1 (0 in lane 63): DPP wave shifts
__device__ __forceinline__ float lane_prev(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_next(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130, 0xf, 0xf, true));
}

// Per-lane facts of a region for the sweep helpers: packed column `lane`
// holds cells x0 = X0 + 2 lane and x0 + 1 of rows Yw .. Yw + R - 1.
struct SweepGeom {
    int lane, dx0, dx1;     // region-border distances of the lane's two columns
    int dyw, Yw, h;         // first region row and grid row of the wave, grid height
    bool l0, l1, r0, r1;    // the two cells have a left / right neighbour in the grid
    bool in0, in1;          // the two cells lie in the grid
    float wl0, wl1, wr0, wr1;  // the same as weights 1 / 0
    float om2, om3, om4;    // omega / cnt
    float c1;               // 1 - omega
};

__device__ __forceinline__ SweepGeom sweep_geom(int X0, int RX, int RY, int R, int Yw, int dyw,
                                                int w, int h, float om2, float om3, float om4,
                                                float c1) {
    SweepGeom g;
    g.lane = threadIdx.x & 63;
    const int x0 = X0 + 2 * g.lane;
    g.dx0 = min(2 * g.lane, RX - 1 - 2 * g.lane);
    g.dx1 = min(2 * g.lane + 1, RX - 2 - 2 * g.lane);
    g.dyw = dyw;
    g.Yw = Yw;
    g.h = h;
    g.in0 = (unsigned)x0 < (unsigned)w;
    g.in1 = (unsigned)(x0 + 1) < (unsigned)w;
    g.l0 = x0 > 0;
    g.l1 = x0 + 1 > 0;
    g.r0 = x0 < w - 1;
    g.r1 = x0 + 1 < w - 1;
    g.wl0 = g.l0 ? 1.f : 0.f;
    g.wl1 = g.l1 ? 1.f : 0.f;
    g.wr0 = g.r0 ? 1.f : 0.f;
    g.wr1 = g.r1 ? 1.f : 0.f;
    g.om2 = om2;
    g.om3 = om3;
    g.om4 = om4;
    g.c1 = c1;
    return g;
}

// Rows K0 <= k < K1 of one half-sweep of colour C.  Only rows 0 and R-1 read
// up_row / down_row (the neighbouring waves' rows).
//
// Residual rule (!TRACK): the right-hand side holds nq = -oc r per cell
// (oc = omega / cnt, cnt = in-grid neighbours; 0 outside the grid), and
//     v' = oc S + ((1 - omega) v + nq),   S = sum of the in-grid neighbours,
// which is v + oc (S - cnt v - r), the reference operator's SOR update
// (solver.cpp:29-47), in two FMAs after the neighbour sum.  On border tiles
// (EDGE) S is a weighted sum with weight 0 for a neighbour outside the grid
// and oc comes per cell from `occ`; weights 1 and oc = omega / 4 give the
// interior path's operations, hence its bits.  Cells outside the grid have
// oc = nq = 0 and value 0, and stay 0.
// Max-update rule (TRACK): the right-hand side holds r; the update
// oc (S - cnt v - r) is formed explicitly, as the stop test tracks it.
template <int C, int H, bool EDGE, bool TRACK, int R, int RY, int K0, int K1, typename RV>
__device__ __forceinline__ void blocked_half(float (&ev)[2][R], const RV& rv, const SweepGeom g,
                                             const float (&occ)[2][R], float up_row,
                                             float down_row, float& mx) {
#pragma unroll
    for (int k = K0; k < K1; ++k) {
        const int p = (k + C) & 1;  // column parity of colour-C cells in row k (R even)
        const float o = ev[1 - C][k];
        const float up = k > 0 ? ev[1 - C][k - 1] : up_row;
        const float down = k < R - 1 ? ev[1 - C][k + 1] : down_row;
        const int dy = min(g.dyw + k, RY - 1 - (g.dyw + k));
        const int dx = p ? g.dx1 : g.dx0;
        const float v = ev[C][k];
        if (!TRACK) {
            const float nq = rv.get(C, k, g.dyw, g.lane);
            if (EDGE) {
                const int Y = g.Yw + k;
                const float wu = Y > 0 ? 1.f : 0.f, wd = Y < g.h - 1 ? 1.f : 0.f;  // uniform
                const float wl = p ? g.wl1 : g.wl0, wr = p ? g.wr1 : g.wr0;
                float sum = (p ? o : lane_prev(o)) * wl;
                sum = __builtin_fmaf(wu, up, sum);
                sum = __builtin_fmaf(wr, p ? lane_next(o) : o, sum);
                sum = __builtin_fmaf(wd, down, sum);
                ev[C][k] = __builtin_fmaf(occ[C][k], sum, __builtin_fmaf(g.c1, v, nq));
            } else {
                // Interior region: every cell is updated.  A cell at distance
                // d < s from the region border takes stale neighbours, but
                // after half-sweep s only cells at distance >= s are exact, and
                // those read cells at distance >= s-1 only: the stale values
                // never reach the stored interior (distance >= H >= nhalf) and
                // stay finite (each update is a bounded combination of finite
                // values).
                const float left = p ? o : lane_prev(o);
                const float right = p ? lane_next(o) : o;
                const float sum = left + up + right + down;
                ev[C][k] = __builtin_fmaf(g.om4, sum, __builtin_fmaf(g.c1, v, nq));
            }
        } else {
            const float left = p ? o : lane_prev(o);
            const float right = p ? lane_next(o) : o;
            if (EDGE) {
                // A neighbour outside the grid contributes the cell's own
                // value: the sum is then the sum over the in-grid neighbours
                // plus (4 - cnt) v, so fma(-4, v, sum) = sum_in - cnt v, and
                // omega/cnt (om4 / om3 / om2) scales the update.  Cells outside
                // the grid keep their value (0); they are never read by cells
                // inside it.  The grid border lies at least H cells inside the
                // region, so every in-grid cell is updated: stale values near
                // the region border never reach the stored interior.
                // (Branch-free: uniform per-row branches here keep the
                // compiler from interleaving the rows' dependent chains.)
                const int Y = g.Yw + k;
                const bool lu = Y > 0, ld = Y < g.h - 1;  // uniform over the wave
                const bool ll = p ? g.l1 : g.l0, lr = p ? g.r1 : g.r0;
                const bool in = (p ? g.in1 : g.in0) && Y >= 0 && Y < g.h;
                const float sum = (ll ? left : v) + (lu ? up : v) + (lr ? right : v) + (ld ? down : v);
                const int miss = (int)!ll + (int)!lr + (int)!lu + (int)!ld;
                const float oc = miss == 0 ? g.om4 : (miss == 1 ? g.om3 : g.om2);
                const float t = __builtin_fmaf(-4.f, v, sum) - rv.get(C, k, g.dyw, g.lane);
                const float dl = oc * t;
                ev[C][k] = in ? v + dl : v;
                if (in && dx >= H && dy >= H) mx = fmaxf(mx, fabsf(dl));
            } else {
                const float t = __builtin_fmaf(-4.f, v, left + up + right + down) -
                                rv.get(C, k, g.dyw, g.lane);
                const float dl = g.om4 * t;
                ev[C][k] = v + dl;
                if (dx >= H && dy >= H) mx = fmaxf(mx, fabsf(dl));
            }
        }
    }
}

template <int C, int H, bool EDGE, bool TRACK, int R, int NW, typename RV>
__device__ __forceinline__ void blocked_step(float (&ev)[2][R], const RV& rv,
                                             float (*xch)[NW][2][64], const SweepGeom g,
                                             const float (&occ)[2][R], int wave, float& mx) {
    const int lane = g.lane;
    // publish this wave's first/last row of the colour the update reads
    xch[1 - C][wave][0][lane] = ev[1 - C][0];
    xch[1 - C][wave][1][lane] = ev[1 - C][R - 1];
    // the inner rows need only this wave's registers: update them while the
    // other waves reach the barrier (the order of the updates of one colour
    // does not matter: they read only the other colour)
    blocked_half<C, H, EDGE, TRACK, R, R * NW, 1, R - 1>(ev, rv, g, occ, 0.f, 0.f, mx);
    __syncthreads();
    const float up_row = xch[1 - C][wave > 0 ? wave - 1 : 0][1][lane];
    const float down_row = xch[1 - C][wave < NW - 1 ? wave + 1 : NW - 1][0][lane];
    blocked_half<C, H, EDGE, TRACK, R, R * NW, 0, 1>(ev, rv, g, occ, up_row, down_row, mx);
    blocked_half<C, H, EDGE, TRACK, R, R * NW, R - 1, R>(ev, rv, g, occ, up_row, down_row, mx);
}

// nhalf <= H half-sweeps, colours alternating from 0; mxs[j] receives the
// largest |update| of sweep j over the cells the launch stores (TRACK)
template <int H, bool EDGE, bool TRACK, int R, int NW, typename RV>
__device__ __forceinline__ void blocked_run(float (&ev)[2][R], const RV& rv,
                                            float (*xch)[NW][2][64], const SweepGeom g,
                                            const float (&occ)[2][R], int wave, int nhalf,
                                            float (&mxs)[H / 2]) {
#pragma unroll
    for (int j = 0; j < H / 2; ++j) {
        const int s = 2 * j + 1;
        if (s > nhalf) break;
        float mx = 0.f;
        blocked_step<0, H, EDGE, TRACK, R, NW>(ev, rv, xch, g, occ, wave, mx);
        if (s + 1 <= nhalf) blocked_step<1, H, EDGE, TRACK, R, NW>(ev, rv, xch, g, occ, wave, mx);
        mxs[j] = fmaxf(mxs[j], mx);
    }
}

// Right-hand side of the sweeps, per cell: r - mean (TRACK) or
// nq = (mean - r) oc (residual rule, see blocked_half).  On interior tiles
// every cell lies in the grid with four neighbours (oc = omega / 4); on border
// tiles cells outside the grid get value 0 and nq = oc = 0.
template <bool EDGE, bool TRACK, typename RT, typename RV>
__device__ __forceinline__ void region_setup(float (&ev)[2][BK_R], const RT (&rr)[2][BK_R], RV& rv,
                                             float (&occ)[2][BK_R], const SweepGeom g, float rm) {
#pragma unroll
    for (int k = 0; k < BK_R; ++k) {
        const int Y = g.Yw + k;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int p = (k + c) & 1;
            const float rc = rhs_to_f32(rr[c][k]);
            if (!EDGE) {
                rv.set(c, k, g.dyw, g.lane, TRACK ? rc - rm : (rm - rc) * g.om4);
                continue;
            }
            const bool in = (p ? g.in1 : g.in0) && Y >= 0 && Y < g.h;
            ev[c][k] = in ? ev[c][k] : 0.f;
            if (TRACK) {
                rv.set(c, k, g.dyw, g.lane, in ? rc - rm : 0.f);
            } else {
                const int miss = (int)!(p ? g.l1 : g.l0) + (int)!(p ? g.r1 : g.r0) +
                                 (int)!(Y > 0) + (int)!(Y < g.h - 1);
                const float oc = in ? (miss == 0 ? g.om4 : (miss == 1 ? g.om3 : g.om2)) : 0.f;
                occ[c][k] = oc;
                rv.set(c, k, g.dyw, g.lane, (rm - rc) * oc);
            }
        }
    }
}

// The right-hand side is BF16 under the residual rule (every FP64 check
// re-rounds the current residual, so the rounding never reaches the result)
// and FP32 under the max-update rule (TRACK), which cannot see the smooth
// error a rounded right-hand side leaves behind.
// The intermediate-grid correction em (bricked, wm x hm, fm x fm fine cells
// per cell), which the learned iteration's first fine smoothing launch
// interpolates bilinearly at the fine cell centres as it loads its region.
struct MidGrid {
    const float* em;
    Brick bkm;
    int wm, hm, fm;
};

// Loads of one region: packed column `lane` of the wave's BK_R rows, both
// colours, from the correction and the right-hand side.  Rows and columns are
// clamped into the grid (cells outside it are masked by region_setup).
// PROLONG: the correction is the intermediate-grid correction interpolated
// bilinearly at the cell centres (u, v) = ((x, y) + 0.5) / fm - 0.5, not ein;
// the workgroup first stages the intermediate cells its region needs in
// `patch` (each is read by up to 4 fm^2 fine cells), with the intermediate
// grid's border clamping.  All waves of the workgroup must call it.
constexpr int kPatchMax = (BK_RX / 2 + 2) * (BK_RY / 2 + 2);  // fm >= 2
template <int H, bool PROLONG, typename RT>
__device__ __forceinline__ void region_load(float (&ev)[2][BK_R], RT (&rr)[2][BK_R],
                                            __amdgpu_buffer_rsrc_t ein_rs,
                                            __amdgpu_buffer_rsrc_t r_rs, const Brick& bk,
                                            const MidGrid& mid, float* patch, int tile,
                                            int tiles_x, int w, int h, int wave, int lane) {
    constexpr int IX = BK_RX - 2 * H, IY = BK_RY - 2 * H;
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int X0 = tx * IX - H, Y0 = ty * IY - H, Yw = Y0 + BK_R * wave;
    const int j = min(max(X0 / 2 + lane, 0), (w + 1) / 2 - 1);
    const unsigned col = brick_col(j);
    int pi0[2] = {0, 0}, pj0 = 0, pw = 0;
    float ptu[2] = {0.f, 0.f};
    const float inv = PROLONG ? 1.f / mid.fm : 0.f;
    if (PROLONG) {
        // intermediate cells I0..I1 x J0..J1 cover the region's (clamped) cells
        const int xa = min(max(X0, 0), w - 1), xb = min(max(X0 + BK_RX - 1, 0), w - 1);
        const int ya = min(max(Y0, 0), h - 1), yb = min(max(Y0 + BK_RY - 1, 0), h - 1);
        const int I0 = (int)floorf((xa + 0.5f) * inv - 0.5f);
        const int I1 = (int)floorf((xb + 0.5f) * inv - 0.5f) + 1;
        pj0 = (int)floorf((ya + 0.5f) * inv - 0.5f);
        const int J1 = (int)floorf((yb + 0.5f) * inv - 0.5f) + 1;
        pw = I1 - I0 + 1;
        const int n = pw * (J1 - pj0 + 1);
        for (int q = threadIdx.x; q < n; q += BK_THREADS) {
            const int pr = q / pw, pc = q - pr * pw;
            patch[q] = mid.em[brick_at(mid.bkm, min(max(I0 + pc, 0), mid.wm - 1),
                                       min(max(pj0 + pr, 0), mid.hm - 1))];
        }
        // the lane's two cells 2j + t (t = 0, 1): the same columns in every row
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const int xx = min(2 * j + t, w - 1);
            const float u = (xx + 0.5f) * inv - 0.5f;
            const int i0 = (int)floorf(u);
            ptu[t] = u - i0;
            pi0[t] = i0 - I0;
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < BK_R; ++k) {
        const int Yc = min(max(Yw + k, 0), h - 1);
        int r0 = 0;
        float tv = 0.f;
        if (PROLONG) {
            const float v = (Yc + 0.5f) * inv - 0.5f;
            const int j0 = (int)floorf(v);
            tv = v - j0;
            r0 = (j0 - pj0) * pw;
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const unsigned row = (unsigned)brick_row(bk, c, Yc);
            if (PROLONG) {
                const int t = (c + Yc) & 1;
                const float tu = ptu[t];
                const float* pp = patch + r0 + pi0[t];
                const float a = pp[0], b = pp[1], cc = pp[pw], d = pp[pw + 1];
                ev[c][k] = (a * (1.f - tu) + b * tu) * (1.f - tv) + (cc * (1.f - tu) + d * tu) * tv;
            } else {
                ev[c][k] = buf_load<float>(ein_rs, col * 4u, row * 4u);
            }
            rr[c][k] = buf_load<RT>(r_rs, col * (unsigned)sizeof(RT), row * (unsigned)sizeof(RT));
        }
    }
}

// Setup, nhalf half-sweeps and the store of one loaded region; mxs[j] takes
// the max with the largest |update| of sweep j (TRACK).
template <int H, bool TRACK, typename RT>
__device__ __forceinline__ void blocked_tile(float (&ev)[2][BK_R], const RT (&rr)[2][BK_R],
                                             float (*xch)[BK_NW][2][64], float (&mxs)[H / 2],
                                             __amdgpu_buffer_rsrc_t eout_rs, const Brick& bk,
                                             int tile, int tiles_x, int w, int h, int nhalf,
                                             float om2, float om3, float om4, float rm, int wave,
                                             int lane) {
    constexpr int IX = BK_RX - 2 * H, IY = BK_RY - 2 * H;
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int X0 = tx * IX - H, Y0 = ty * IY - H;
    const int wp = (w + 1) / 2;
    const int Yw = Y0 + BK_R * wave;  // first global row of this wave
    const int dyw = BK_R * wave;      // first region row of this wave
    const unsigned col = brick_col(min(max(X0 / 2 + lane, 0), wp - 1));
    RhsRegs rv;
    const SweepGeom g = sweep_geom(X0, BK_RX, BK_RY, BK_R, Yw, dyw, w, h, om2, om3, om4,
                                   1.f - 4.f * om4);
    const bool edge = X0 < 1 || X0 + BK_RX > w - 1 || Y0 < 1 || Y0 + BK_RY > h - 1;
    float occ[2][BK_R];  // border tiles: omega / cnt per cell, 0 outside the grid
    if (edge) {
        region_setup<true, TRACK>(ev, rr, rv, occ, g, rm);
        blocked_run<H, true, TRACK, BK_R, BK_NW>(ev, rv, xch, g, occ, wave, nhalf, mxs);
    } else {
        region_setup<false, TRACK>(ev, rr, rv, occ, g, rm);
        blocked_run<H, false, TRACK, BK_R, BK_NW>(ev, rv, xch, g, occ, wave, nhalf, mxs);
    }
    // write the interior
#pragma unroll
    for (int k = 0; k < BK_R; ++k) {
        const int Y = Yw + k;
        const int dy = min(dyw + k, BK_RY - 1 - (dyw + k));
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int p = (k + c) & 1;
            const int dx = p ? g.dx1 : g.dx0;
            if (dx >= H && dy >= H && (unsigned)Y < (unsigned)h && (p ? g.in1 : g.in0))
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(ev[c][k]), eout_rs, col * 4u,
                                                      (unsigned)brick_row(bk, c, Y) * 4u,
                                                      CD_BK_STORE_POLICY);
        }
    }
}

// workgroups of a k_rb_blocked launch over ntiles tiles (a multiple of 8)
inline int blocked_grid(int ntiles) { return (ntiles + 7) / 8 * 8; }

// FULL: nhalf == H, known at compile time (every launch but a batch's last),
// which leaves the half-sweeps free of branches.  PROLONG: the sweeps start
// from `mid` interpolated at the cells instead of from ein.
template <int H, bool TRACK, bool FULL, bool PROLONG = false>
__global__ __launch_bounds__(BK_THREADS) void k_rb_blocked(const float* __restrict__ ein,
                                                    float* __restrict__ eout,
                                                    const void* __restrict__ rhs,
                                                    Brick bk, int w, int h,
                                                    int tiles_x, int ntiles, int nhalf,
                                                    float om2, float om3, float om4,
                                                    const float* __restrict__ rmean,
                                                    SolveDev* st, int out_index, MidGrid mid) {
    static_assert(H % 2 == 0 && H >= 2 && 2 * H < BK_RY && BK_R % 2 == 0, "geometry");
    __shared__ float xch[2][BK_NW][2][64];
    __shared__ float smax[BK_NW][H / 2];
    // XCD-aware tile order: gridDim.x is a multiple of 8 and consecutive
    // workgroups go round-robin to the 8 XCDs; give each XCD a contiguous band
    // of tiles so halo re-reads hit its own L2.
    const int per = gridDim.x >> 3;
    const int tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (tile >= ntiles) return;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    using RT = typename std::conditional<TRACK, float, unsigned short>::type;
    const size_t npk = 2 * bk.plane;
    const __amdgpu_buffer_rsrc_t ein_rs = plane_rsrc(ein, npk * sizeof(float));
    const __amdgpu_buffer_rsrc_t r_rs = plane_rsrc(rhs, npk * sizeof(RT));
    const __amdgpu_buffer_rsrc_t eout_rs = plane_rsrc(eout, npk * sizeof(float));
    float ev[2][BK_R];
    RT rr[2][BK_R];
    __shared__ float patch[PROLONG ? kPatchMax : 1];
    region_load<H, PROLONG>(ev, rr, ein_rs, r_rs, bk, mid, patch, tile, tiles_x, w, h, wave,
                            lane);
    // Max-update rule: the stop flag can be set between the launches of one
    // batch (k_maxcheck_blocked), and a finished solve leaves the buffers
    // untouched.  Residual rule: only a check sets it, and no sweeps are
    // enqueued after a check that stopped the solve, so the flag is not read
    // (reading it first would hold every load behind two dependent scalar
    // loads).
    if (TRACK && st->done) return;
    const float rm = *rmean;
    float mxs[H / 2];
#pragma unroll
    for (int j = 0; j < H / 2; ++j) mxs[j] = 0.f;
    blocked_tile<H, TRACK>(ev, rr, xch, mxs, eout_rs, bk, tile, tiles_x, w, h, FULL ? H : nhalf,
                           om2, om3, om4, rm, wave, lane);
    if (TRACK) {
#pragma unroll
        for (int j = 0; j < H / 2; ++j) {
            const float m = cd_wave_max(mxs[j]);
            if (lane == 0) smax[wave][j] = m;
        }
        __syncthreads();
        if (threadIdx.x < H / 2 && 2 * (int)threadIdx.x < nhalf) {
            const int j = threadIdx.x;
            float m = smax[0][j];
            for (int wv = 1; wv < BK_NW; ++wv) m = fmaxf(m, smax[wv][j]);
            atomicMax(&st->maxupd_bits[j], __float_as_uint(m));
        }
    }
    if (threadIdx.x == 0) st->cur = out_index;  // same value from every workgroup
}

// max-update stop rule over the sweeps of one blocked launch (solver.cpp:142).
// The launch has applied all of its nsweeps sweeps, so a stop at any of them
// counts the whole launch.
__global__ void k_maxcheck_blocked(SolveDev* st, int first_sweep, int nsweeps, double tol) {
    if (threadIdx.x != 0) return;
    if (!st->done) {
        for (int j = 0; j < nsweeps; ++j) {
            const float mu = __uint_as_float(st->maxupd_bits[j]);
            st->last_max_update = (double)mu;
            if ((double)mu < tol) {
                st->done = 1;
                st->sweeps_at_done = first_sweep + nsweeps;
                break;
            }
        }
    }
    for (int j = 0; j < 16; ++j) st->maxupd_bits[j] = 0u;
}

// Reopen a solve whose max-update stop is to be confirmed on a fresh
// right-hand side (the sweeps that stopped it read r rounded at the last check).
__global__ void k_reopen(SolveDev* st) {
    if (threadIdx.x != 0) return;
    st->done = 0;
    st->sweeps_at_done = -1;
    for (int j = 0; j < 16; ++j) st->maxupd_bits[j] = 0u;
}

// max-update stop rule after the black half of sweep `sweep` (solver.cpp:142)
__global__ void k_maxcheck(SolveDev* st, int sweep, double tol) {
    if (threadIdx.x != 0) return;
    if (!st->done) {
        float mu = __uint_as_float(st->maxupd_bits[sweep & 1]);
        st->last_max_update = (double)mu;
        if ((double)mu < tol) {
            st->done = 1;
            st->sweeps_at_done = sweep + 1;
        }
    }
    st->maxupd_bits[(sweep + 1) & 1] = 0u;
}

__global__ void k_set_holes_nan(double* __restrict__ phi, const uint8_t* __restrict__ mask,
                                size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && (mask[i] & M_HOLE)) phi[i] = NAN;
}

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// Over-relaxation factor.  > 0: as given.  < 0: the reference's
// 2/(1+3.14159265/width) (solver.cpp:71).  == 0: the optimum for this
// operator: the Jacobi iteration I + A/4 has spectral radius
// rho_J = 1 - sin^2(pi/(2n)) on the zero-mean subspace (n = larger side), and
// red-black ordering is consistent, so omega = 2/(1 + sqrt(1 - rho_J^2)).
double cd_sor_omega(double requested, int w, int h) {
    if (requested > 0) return requested;
    if (requested < 0) return 2.0 / (1.0 + 3.14159265 / w);
    const int n = std::max(w, h);
    const double s = std::sin(M_PI / (2.0 * n));
    const double rj = 1.0 - s * s;
    return 2.0 / (1.0 + std::sqrt(1.0 - rj * rj));
}

// the next check of an adaptive solve comes after this multiple of the sweeps
// the observed contraction rate predicts (0.9 to 1.5 measured no better on the
// 2048^2 and 1024^2 designs: a check also restarts the FP32 correction, so
// where the checks fall moves the sweep count by a few per cent either way)
#ifndef CD_CHECK_MARGIN
#define CD_CHECK_MARGIN 1.05
#endif

int auto_check_every(int w, int h) {
    // SOR with omega near 2 contracts every mode by ~(1 - 2*pi/n) per sweep.
    // The FP32 correction loses accuracy once it has absorbed a ~10x residual
    // reduction, which takes ~ n*ln(10)/(2*pi) sweeps.
    int n = std::max(w, h);
    int k = (int)std::lround(0.36 * n);
    return std::max(8, std::min(k, 2048));
}


// ---------------------------------------------------------------------------
// Learned correction and training-pair capture (no hole mask, positional
// Neumann counts).  Coarse cell (I,J) averages the f x f fine cells
// [f*I, f*I+f) x [f*J, f*J+f).
// ---------------------------------------------------------------------------
// coarse residual rc = block mean of the projected BF16 residual the blocked
// sweeps see (r16 - rmean32).  f < 8: one thread per coarse cell; larger f:
// one wave per coarse cell (lanes stride over its f^2 fine cells, fixed-order
// shuffle tree).  Statistics of rc follow in k_coarse_stats.
// x8 (may be NULL): rc also written as the input of a linear network, BF16
// channel 0 of 8 (a linear bias-free network commutes with the scaling that
// normalises its input, so that input needs no statistics)
__global__ __launch_bounds__(256) void k_restrict_r16(const unsigned short* __restrict__ r16,
                                                      Brick bk, const SolveDev* __restrict__ st,
                                                      int f, int wc, int hc,
                                                      float* __restrict__ rc,
                                                      unsigned short* __restrict__ x8) {
    const float rm = st->rmean32;
    const int ff = f * f;
    if (f < 8) {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= wc * hc) return;
        const int J = i / wc, I = i - J * wc;
        float acc = 0.f;
        for (int yy = J * f; yy < J * f + f; ++yy)
            for (int xx = I * f; xx < I * f + f; ++xx)
                acc += bf16_to_f32(r16[brick_at(bk, xx, yy)]) - rm;
        rc[i] = acc / (float)ff;
        if (x8) *reinterpret_cast<uint4*>(x8 + (size_t)i * 8) =
            make_uint4(f32_to_bf16(acc / (float)ff), 0, 0, 0);
        return;
    }
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (i >= wc * hc) return;  // whole wave
    const int J = i / wc, I = i - J * wc;
    float acc = 0.f;
    for (int k = lane; k < ff; k += 64) {
        const int dy = k / f, dx = k - dy * f;
        acc += bf16_to_f32(r16[brick_at(bk, I * f + dx, J * f + dy)]) - rm;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) {
        rc[i] = acc / (float)ff;
        if (x8) *reinterpret_cast<uint4*>(x8 + (size_t)i * 8) =
            make_uint4(f32_to_bf16(acc / (float)ff), 0, 0, 0);
    }
}

// (mean, rms) of nc coarse values from stats = (sum, sum^2), as every consumer
// of stats computes them
__device__ __forceinline__ void coarse_derived(const double* stats, int nc, double* derived) {
    const double mean = stats[0] / nc;
    const double var = stats[1] / nc - mean * mean;
    derived[0] = mean;
    derived[1] = sqrt(var > 1e-300 ? var : 1e-300);
}

// (sum, sum^2) of the coarse residual in one workgroup (fixed order), for
// coarse grids of up to 64Ki cells; derived (may be NULL) receives (mean, rms);
// x8 (may be NULL) receives the network input: the residual normalised to
// zero mean and unit rms in BF16, channel 0 of 8
__global__ __launch_bounds__(1024) void k_coarse_stats1(const float* __restrict__ rc, int nc,
                                                        double* __restrict__ stats,
                                                        double* __restrict__ derived,
                                                        unsigned short* __restrict__ x8) {
    __shared__ double smem[16];
    __shared__ double ms[2];
    double a = 0.0, q = 0.0;
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        const double v = rc[i];
        a += v;
        q += v * v;
    }
    a = cd_block_sum(a, smem);
    q = cd_block_sum(q, smem);
    if (threadIdx.x == 0) {
        stats[0] = a;
        stats[1] = q;
        if (derived) coarse_derived(stats, nc, derived);
        const double mean = a / nc;
        const double var = q / nc - mean * mean;
        ms[0] = mean;
        ms[1] = sqrt(var > 1e-300 ? var : 1e-300);
    }
    if (!x8) return;
    __syncthreads();
    const double mean = ms[0], sig = ms[1];
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
        uint4 v = make_uint4(0, 0, 0, 0);
        v.x = f32_to_bf16((float)(((double)rc[i] - mean) / sig));
        *reinterpret_cast<uint4*>(x8 + (size_t)i * 8) = v;
    }
}

// (mean, rms) after a multi-workgroup reduction of the coarse statistics
__global__ void k_coarse_derived(const double* __restrict__ stats, int nc,
                                 double* __restrict__ derived) {
    if (threadIdx.x == 0) coarse_derived(stats, nc, derived);
}

// partials (sum, sum^2) of the coarse residual per workgroup of 256 cells
__global__ __launch_bounds__(256) void k_coarse_stats(const float* __restrict__ rc, int nc,
                                                      double* __restrict__ partials) {
    __shared__ double smem[16];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const double v = i < nc ? (double)rc[i] : 0.0;
    const double a = cd_block_sum(v, smem);
    const double q = cd_block_sum(v * v, smem);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = a;
        partials[2 * blockIdx.x + 1] = q;
    }
}

// stats = (sum, sum^2) over n coarse cells -> x = (rc - mean) / rms
__global__ void k_normalize_coarse(const float* __restrict__ rc, const double* __restrict__ stats,
                                   int n, float* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double mean = stats[0] / n;
    const double var = stats[1] / n - mean * mean;
    const double sig = sqrt(var > 1e-300 ? var : 1e-300);
    x[i] = (float)(((double)rc[i] - mean) / sig);
}

// packed-layout launch: grid (ceil(wpb / 256), rows, 2 colour planes); a
// thread owns packed column j of row Y in plane c, cell x = 2j + ((c+Y)&1)
__device__ __forceinline__ bool packed_thread(const Brick& bk, int w, int Y, int& c, int& j,
                                              int& x) {
    c = blockIdx.z;
    j = blockIdx.x * blockDim.x + threadIdx.x;
    x = 2 * j + ((c + Y) & 1);
    return j < bk.wpb && x < w;
}

// e = rms * (nmax/pi)^2 * bilinear(y) at cell centres (rms = derived[1], one
// value for the whole launch), written in the
// bricked layout of the blocked sweeps (the start of the smoothing); one
// thread per packed element, coalesced writes
__device__ __forceinline__ void prolong_bricked_at(const float* __restrict__ y,
                                                   const double* __restrict__ derived, int w,
                                                   int h, int f, int wc, int hc, double nscale,
                                                   const Brick& bk, float* __restrict__ e) {
    const int yy = blockIdx.y;
    int c, j, xx;
    if (!packed_thread(bk, w, yy, c, j, xx)) return;
    const double sig = derived[1];  // rms of the coarse residual
    const float inv = 1.f / f;  // exact for the power-of-two coarsenings in use
    const float u = (xx + 0.5f) * inv - 0.5f, v = (yy + 0.5f) * inv - 0.5f;
    int i0 = (int)floorf(u), j0 = (int)floorf(v);
    const float tu = u - i0, tv = v - j0;
    const int i1 = min(max(i0 + 1, 0), wc - 1), j1 = min(max(j0 + 1, 0), hc - 1);
    i0 = min(max(i0, 0), wc - 1);
    j0 = min(max(j0, 0), hc - 1);
    const float ev = (y[j0 * wc + i0] * (1.f - tu) + y[j0 * wc + i1] * tu) * (1.f - tv) +
                     (y[j1 * wc + i0] * (1.f - tu) + y[j1 * wc + i1] * tu) * tv;
    e[brick_row(bk, c, yy) + brick_col(j)] = (float)(sig * nscale * (double)ev);
}
__global__ __launch_bounds__(256) void k_prolong_bricked(const float* __restrict__ y,
                                                         const double* __restrict__ derived,
                                                         int w, int h, int f, int wc, int hc,
                                                         double nscale, Brick bk,
                                                         float* __restrict__ e) {
    prolong_bricked_at(y, derived, w, h, f, wc, hc, nscale, bk, e);
}

// Intermediate grid of the learned preconditioner (fm x fm fine cells per
// cell): rm = the block sum of the projected fine residual (r16 - rmean32),
// i.e. fm^2 times its block mean, the right-hand side of the unit-spacing
// operator on that grid; BF16, bricked like the fine right-hand side.  One
// thread per packed element of the intermediate grid.
// With the start of the intermediate-grid correction in the same launch (the
// two share the thread layout): em = the coarse output interpolated there.
__global__ __launch_bounds__(256) void k_restrict_to_mid(const unsigned short* __restrict__ r16,
                                                         Brick bk, const SolveDev* __restrict__ st,
                                                         int fm, int wm, int hm, Brick bkm,
                                                         unsigned short* __restrict__ rm,
                                                         const float* __restrict__ y,
                                                         const double* __restrict__ derived,
                                                         int fc, int wc, int hc, double nscale,
                                                         float* __restrict__ em) {
    prolong_bricked_at(y, derived, wm, hm, fc, wc, hc, nscale, bkm, em);
    const int Ym = blockIdx.y;
    int c, j, xm;
    if (!packed_thread(bkm, wm, Ym, c, j, xm)) return;
    const float rmean = st->rmean32;
    float acc = 0.f;
    for (int dy = 0; dy < fm; ++dy)
        for (int dx = 0; dx < fm; ++dx)
            acc += bf16_to_f32(r16[brick_at(bk, xm * fm + dx, Ym * fm + dy)]) - rmean;
    rm[brick_row(bkm, c, Ym) + brick_col(j)] = f32_to_bf16(acc);
}

// ---------------------------------------------------------------------------
// Flexible conjugate gradients (FCG(1)) preconditioned by the network plus
// smoothing sweeps.  With L the operator (negative semi-definite), r the
// projected residual and z = M(r) the preconditioned residual:
//   beta = -(z . Lp_old) / (p_old . Lp_old)   p = z + beta p_old
//   alpha = (p . r) / (p . Lp)                phi += alpha p
// alpha minimises the energy 1/2 phi'(-L)phi + b'phi along p, so every step
// lowers the energy norm of the error.  Lp = Lz + beta Lp_old needs no second
// stencil pass; p . Lp expands into the dot products taken below.
// ---------------------------------------------------------------------------
struct FcgDev {
    double beta, alpha;
    double pq;     // p . Lp of the current direction
    int have;      // a previous direction exists
    int bad;       // the last step had no descent direction
    const float* pprev;  // the previous direction (one of two ping-pong
                         // buffers; the captured iteration body reads it here)
};

// start of the learned iterations of a solve; p0: the zeroed first direction
__global__ void k_fcg_init(FcgDev* fd, const float* p0) {
    if (threadIdx.x != 0) return;
    fd->beta = 0.0;
    fd->alpha = 0.0;
    fd->pq = 0.0;
    fd->have = 0;
    fd->bad = 0;
    fd->pprev = p0;
}

// Lz at packed element (c, Y, j) (cell x): the four neighbours lie in the
// other colour plane, up/down at column j, left/right at (x -+ 1) >> 1
__device__ __forceinline__ float lap_packed(const float* __restrict__ z, const Brick& bk, int c,
                                            int Y, int j, int x, int w, int h, float zc) {
    const int o = 1 - c;
    float s = 0.f, cnt = 0.f;
    if (x > 0) { s += z[brick_row(bk, o, Y) + brick_col((x - 1) >> 1)]; cnt += 1.f; }
    if (x < w - 1) { s += z[brick_row(bk, o, Y) + brick_col((x + 1) >> 1)]; cnt += 1.f; }
    if (Y > 0) { s += z[brick_row(bk, o, Y - 1) + brick_col(j)]; cnt += 1.f; }
    if (Y < h - 1) { s += z[brick_row(bk, o, Y + 1) + brick_col(j)]; cnt += 1.f; }
    return s - cnt * zc;
}

#ifndef CD_DOT_BANDS
#define CD_DOT_BANDS 256
#endif
// partials[4 * block + k]: p_old.Lz (= z.Lp_old: L is symmetric, so Lp_old
// needs neither a stencil pass nor storage), z.r, p_old.r, z.Lz over the
// packed elements of the workgroup's band of consecutive rows of one colour
// plane (fixed order).  Workgroups in XCD-contiguous order: both planes of a
// band, whose stencils read each other, and the neighbouring bands share an
// L2.  Rows go in groups of 4 whose loads are all issued before the first is
// used (clamped, in-range addresses; elements outside the grid contribute
// nothing).
__global__ __launch_bounds__(256) void k_fcg_dots(const float* __restrict__ z,
                                                  const FcgDev* __restrict__ fd,
                                                  const unsigned short* __restrict__ r16,
                                                  const SolveDev* __restrict__ st, Brick bk,
                                                  int w, int h, double* __restrict__ partials) {
    constexpr int G = 4;
    __shared__ double smem[64];
    const float rm = st->rmean32;
    const float* __restrict__ p = fd->pprev;
    // block order (column block fastest, then plane, then band of rows)
    const int ob = xcd_order();
    const int bx = ob % gridDim.x, c = (ob / gridDim.x) & 1, band = ob / (2 * gridDim.x);
    const int o = 1 - c;
    const int rows = (h + gridDim.y - 1) / gridDim.y;  // rows per band
    const int y0 = band * rows, y1 = min(h, y0 + rows);
    const int j0 = bx * blockDim.x + threadIdx.x;
    const int j = min(j0, bk.wpb - 1);
    const unsigned cj = brick_col(j), cjm = brick_col(max(j - 1, 0)),
                   cjp = brick_col(min(j + 1, bk.wpb - 1));
    double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;
    for (int Yb = y0; Yb < y1; Yb += G) {
        float zc[G], zl[G], zr[G], zu[G], zd[G], pv[G];
        unsigned short rv[G];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int Y = min(Yb + k, h - 1);
            const int par = (c + Y) & 1;  // x = 2j + par
            const size_t ro = brick_row(bk, o, Y);
            const size_t i = brick_row(bk, c, Y) + cj;
            zc[k] = z[i];
            zl[k] = z[ro + (par ? cj : cjm)];   // (x - 1) >> 1
            zr[k] = z[ro + (par ? cjp : cj)];   // (x + 1) >> 1
            zu[k] = z[brick_row(bk, o, max(Y - 1, 0)) + cj];
            zd[k] = z[brick_row(bk, o, min(Y + 1, h - 1)) + cj];
            pv[k] = p[i];
            rv[k] = r16[i];
        }
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int Y = Yb + k;
            const int x = 2 * j0 + ((c + Y) & 1);
            if (Y >= y1 || j0 >= bk.wpb || x >= w) continue;
            float sn = 0.f, cnt = 0.f;
            if (x > 0) { sn += zl[k]; cnt += 1.f; }
            if (x < w - 1) { sn += zr[k]; cnt += 1.f; }
            if (Y > 0) { sn += zu[k]; cnt += 1.f; }
            if (Y < h - 1) { sn += zd[k]; cnt += 1.f; }
            const float lz = sn - cnt * zc[k];
            const float rr = bf16_to_f32(rv[k]) - rm;
            d0 += (double)(pv[k] * lz);
            d1 += (double)(zc[k] * rr);
            d2 += (double)(pv[k] * rr);
            d3 += (double)(zc[k] * lz);
        }
    }
    double dd[4] = {d0, d1, d2, d3};
    cd_block_sum_n(dd, smem);
    d0 = dd[0];
    d1 = dd[1];
    d2 = dd[2];
    d3 = dd[3];
    if (threadIdx.x == 0) {
        partials[4 * ob] = d0;
        partials[4 * ob + 1] = d1;
        partials[4 * ob + 2] = d2;
        partials[4 * ob + 3] = d3;
    }
}

// single block: step coefficients from the dot products (fixed summation order)
__global__ void k_fcg_coef(const double* __restrict__ partials, int nblocks, FcgDev* fd) {
    __shared__ double smem[64];
    double d[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int k = 0; k < 4; ++k) d[k] += partials[4 * b + k];
    cd_block_sum_n(d, smem);
    if (threadIdx.x == 0) {
        const double zq = d[0], zr = d[1], pr = d[2], zlz = d[3];  // zq = z.Lp_old
        double beta = 0.0;
        if (fd->have && fd->pq < 0.0) beta = -zq / fd->pq;
        double pq = zlz + 2.0 * beta * zq + beta * beta * (fd->have ? fd->pq : 0.0);
        double prn = zr + beta * (fd->have ? pr : 0.0);
        double alpha = 0.0;
        fd->bad = 0;
        if (pq < 0.0 && isfinite(pq) && isfinite(prn)) {
            alpha = prn / pq;
        } else {
            // no usable direction: restart from z alone next time
            fd->bad = 1;
            beta = 0.0;
            pq = zlz;
        }
        fd->beta = beta;
        fd->alpha = alpha;
        fd->pq = pq;
        fd->have = fd->bad ? 0 : 1;
    }
}

// The learned iteration's FCG step folded into its convergence check:
// p = z + beta p_old (z: the smoothed correction, packed), phi_out = phi_in +
// alpha p, and the FP64 residual of phi_out with its BF16 copy and partial
// norms (as k_residual<2>).  phi and p are ping-pong pairs: the stencil reads
// phi_in and p_old at the four neighbours and forms their new values itself
// (the same FP32 and FP64 operations as at the centre, so every cell's new
// value is one number).  fd->pprev is pointed at p_new for the next iteration.
// x8 (may be NULL): the workgroup also writes the next iteration's network
// input, the lf x lf block means of the new residual (BF16, channel 0 of 8 of
// a coarse grid lwc wide).  The host passes x8 only when the workgroups tile
// the grid exactly and cover whole coarse cells: lf divides 64, and the
// 4*ROWS rows of a workgroup are one or two coarse rows.  The FP64 residual is
// mean-free to rounding, so its block means need no projection.
// rows whose loads a thread issues together, and the most rows a thread takes
// (2 / 8 measured as fast as 4 / 8 and 2 / 4 at 1024^2 and 2048^2)
#ifndef CD_STEP_G
#define CD_STEP_G 2
#endif
#ifndef CD_STEP_ROWS
#define CD_STEP_ROWS 8
#endif
template <int ROWS>
__global__ __launch_bounds__(256) void k_residual_step(const double* __restrict__ phi_in,
                                                       double* __restrict__ phi_out,
                                                       const double* __restrict__ b,
                                                       const SolveDev* __restrict__ st,
                                                       const float* __restrict__ z,
                                                       const float* __restrict__ p_old,
                                                       float* __restrict__ p_new,
                                                       FcgDev* __restrict__ fd,
                                                       unsigned short* __restrict__ r16, Brick bk,
                                                       int w, int h, double* __restrict__ partials,
                                                       unsigned short* __restrict__ x8, int lf,
                                                       int lwc) {
    constexpr int G = ROWS < CD_STEP_G ? ROWS : CD_STEP_G;  // rows whose loads are in flight together
    __shared__ double smem[32];
    __shared__ float csum[2][4][64];  // (coarse row, thread row, column) partial block sums
    float cpart0 = 0.f, cpart1 = 0.f;  // this thread's sums per coarse row
    const double bmean = st->bmean;
    const float alpha = (float)fd->alpha, beta = (float)fd->beta;
    const int ob = xcd_order();
    const int bx = ob % gridDim.x, by = ob / gridDim.x;
    const int x0 = bx * blockDim.x + threadIdx.x;
    const int y0 = by * blockDim.y * ROWS + threadIdx.y;
    const int x = min(x0, w - 1);
    const int xl = max(x - 1, 0), xr = min(x + 1, w - 1);
    // bricked offsets (a plane pair holds < 2^29 elements in blocked mode): the
    // column parts of this cell and of its left / right neighbours are the same
    // on every row; their planes alternate with the row
    const unsigned cj = brick_col(x >> 1), cjl = brick_col(xl >> 1), cjr = brick_col(xr >> 1);
    // (24-bit multiplies, full rate: row and brick indices are < 2^24)
    const unsigned plane = (unsigned)bk.plane, brow = (unsigned)bk.nbx * 512u;
    auto rowo = [&](int c, int y) -> unsigned {
        return (c ? plane : 0u) + __umul24((unsigned)(y >> 3), brow) + (unsigned)(y & 7) * 64u;
    };
    // buffer resources: 32-bit byte offsets, and reads past either end of a
    // buffer return 0 (the host passes grids of < 2^28 cells)
    const size_t n = (size_t)w * h;
    const __amdgpu_buffer_rsrc_t phi_rs = plane_rsrc(phi_in, n * sizeof(double));
    const __amdgpu_buffer_rsrc_t b_rs = plane_rsrc(b, n * sizeof(double));
    const __amdgpu_buffer_rsrc_t z_rs = plane_rsrc(z, 2 * bk.plane * sizeof(float));
    const __amdgpu_buffer_rsrc_t p_rs = plane_rsrc(p_old, 2 * bk.plane * sizeof(float));
    const __amdgpu_buffer_rsrc_t phio_rs = plane_rsrc(phi_out, n * sizeof(double));
    const __amdgpu_buffer_rsrc_t pn_rs = plane_rsrc(p_new, 2 * bk.plane * sizeof(float));
    const __amdgpu_buffer_rsrc_t r16_rs = plane_rsrc(r16, 2 * bk.plane * sizeof(unsigned short));
    auto ld64 = [](__amdgpu_buffer_rsrc_t rs, unsigned off) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, off, 0, 0));
    };
    // new direction and iterate at (xx, yy) from the loaded phi_in, z, p_old
    auto pnew = [&](float zv, float pv) { return zv + beta * pv; };
    auto phinew = [&](double ph, float pn) { return ph + (double)(alpha * pn); };
    double rs = 0.0, rq = 0.0;
#pragma unroll
    for (int g = 0; g < ROWS; g += G) {
        double ph[G][5], bb[G];
        float zv[G][5], pv[G][5];
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int y = min(y0 + (g + k) * (int)blockDim.y, h - 1);
            const int yu = max(y - 1, 0), yd = min(y + 1, h - 1);
            const int c = (x + y) & 1, o = 1 - c;
            // centre, left, up, right, down (a clamped neighbour at the border
            // is some valid cell: its value is not used)
            const unsigned ro = rowo(o, y);
            const unsigned bis[5] = {rowo(c, y) + cj, ro + cjl, rowo(o, yu) + cj, ro + cjr,
                                     rowo(o, yd) + cj};
            // byte offsets of the natural-layout cells: centre, left, up, right,
            // down (the clamped ones at the border are valid cells, unused)
            const unsigned co = (__umul24((unsigned)y, (unsigned)w) + (unsigned)x) * 8u;
            const unsigned ws8 = (unsigned)w * 8u;
            const unsigned pos[5] = {co, co - (unsigned)(x - xl) * 8u, co - (unsigned)(y - yu) * ws8,
                                     co + (unsigned)(xr - x) * 8u, co + (unsigned)(yd - y) * ws8};
#pragma unroll
            for (int t = 0; t < 5; ++t) {
                ph[k][t] = ld64(phi_rs, pos[t]);
                zv[k][t] = buf_load<float>(z_rs, bis[t] * 4u, 0);
                pv[k][t] = buf_load<float>(p_rs, bis[t] * 4u, 0);
            }
            bb[k] = ld64(b_rs, co);
        }
#pragma unroll
        for (int k = 0; k < G; ++k) {
            const int y = y0 + (g + k) * (int)blockDim.y;
            if (x0 >= w || y >= h) break;
            double nv[5];
            float pc = 0.f;
#pragma unroll
            for (int t = 0; t < 5; ++t) {
                const float pn = pnew(zv[k][t], pv[k][t]);
                if (t == 0) pc = pn;
                nv[t] = phinew(ph[k][t], pn);
            }
            // neighbours in the order left, up, right, down
            double s = 0.0, cnt = 0.0;
            if (x > 0) { cnt += 1.0; s += nv[1]; }
            if (y > 0) { cnt += 1.0; s += nv[2]; }
            if (x < w - 1) { cnt += 1.0; s += nv[3]; }
            if (y < h - 1) { cnt += 1.0; s += nv[4]; }
            const double rv = (bb[k] - bmean) - (s - cnt * nv[0]);
            rs += rv;
            rq += rv * rv;
            if (x8) {
                if ((int)threadIdx.y + 4 * (g + k) >= lf) cpart1 += (float)rv;
                else cpart0 += (float)rv;
            }
            const unsigned bi = rowo((x + y) & 1, y) + cj;
            __builtin_amdgcn_raw_buffer_store_b64(
                __builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, nv[0]), phio_rs,
                (__umul24((unsigned)y, (unsigned)w) + (unsigned)x) * 8u, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(pc), pn_rs, bi * 4u, 0, 0);
            __builtin_amdgcn_raw_buffer_store_b16(f32_to_bf16((float)rv), r16_rs, bi * 2u, 0, 0);
        }
    }
    double aq[2] = {rs, rq};
    cd_block_sum_n(aq, smem);
    const double a = aq[0], q = aq[1];
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        partials[2 * ob] = a;
        partials[2 * ob + 1] = q;
        if (ob == 0) fd->pprev = p_new;
    }
    if (x8) {  // uniform: every thread of the launch takes this branch or none
        const int ncr = (4 * ROWS) / lf > 1 ? 2 : 1;  // coarse rows of this workgroup
        csum[0][threadIdx.y][threadIdx.x] = cpart0;
        csum[1][threadIdx.y][threadIdx.x] = cpart1;
        __syncthreads();
        if (threadIdx.y < ncr) {
            // thread row cr sums coarse row cr over the four thread rows, then
            // over groups of lf lanes (fixed order: xor tree)
            const int cr = threadIdx.y;
            float v = ((csum[cr][0][threadIdx.x] + csum[cr][1][threadIdx.x]) +
                       csum[cr][2][threadIdx.x]) + csum[cr][3][threadIdx.x];
            for (int o = 1; o < lf; o <<= 1) v += __shfl_xor(v, o, 64);
            if ((threadIdx.x & (lf - 1)) == 0) {
                const int cx = (bx * 64 + threadIdx.x) / lf;
                const int cy = (by * 4 * ROWS) / lf + cr;
                *reinterpret_cast<uint4*>(x8 + ((size_t)cy * lwc + cx) * 8) =
                    make_uint4(f32_to_bf16(v / (float)(lf * lf)), 0, 0, 0);
            }
        }
    }
}

// launch of k_residual_step<rows> for rows in {1, 2, 4, 8}
void launch_residual_step(int rows, dim3 grid, dim3 blk, hipStream_t s, const double* phi_in,
                          double* phi_out, const double* b, const SolveDev* st, const float* z,
                          const float* p_old, float* p_new, FcgDev* fd, unsigned short* r16,
                          const Brick& bk, int w, int h, double* partials, unsigned short* x8,
                          int lf, int lwc) {
    switch (rows) {
    case 8: hipLaunchKernelGGL(k_residual_step<8>, grid, blk, 0, s, phi_in, phi_out, b, st, z, p_old, p_new, fd, r16, bk, w, h, partials, x8, lf, lwc); break;
    case 4: hipLaunchKernelGGL(k_residual_step<4>, grid, blk, 0, s, phi_in, phi_out, b, st, z, p_old, p_new, fd, r16, bk, w, h, partials, x8, lf, lwc); break;
    case 2: hipLaunchKernelGGL(k_residual_step<2>, grid, blk, 0, s, phi_in, phi_out, b, st, z, p_old, p_new, fd, r16, bk, w, h, partials, x8, lf, lwc); break;
    default: hipLaunchKernelGGL(k_residual_step<1>, grid, blk, 0, s, phi_in, phi_out, b, st, z, p_old, p_new, fd, r16, bk, w, h, partials, x8, lf, lwc); break;
    }
}

// coarse FP64 block mean of phi
__global__ void k_restrict_phi(const double* __restrict__ phi, int w, int f, int wc, int hc,
                               double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= wc * hc) return;
    const int J = i / wc, I = i - J * wc;
    double acc = 0.0;
    for (int yy = J * f; yy < J * f + f; ++yy)
        for (int xx = I * f; xx < I * f + f; ++xx) acc += phi[(size_t)yy * w + xx];
    out[i] = acc / (double)(f * f);
}

// target of one captured pair: (phi*_c - phi_k_c) / (rms * (nmax/pi)^2)
__global__ void k_capture_finish(const double* __restrict__ final_c,
                                 const double* __restrict__ phik_c,
                                 const double* __restrict__ stats, int nc, double nscale,
                                 float* __restrict__ t) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const double mean = stats[0] / nc;
    const double var = stats[1] / nc - mean * mean;
    const double sig = sqrt(var > 1e-300 ? var : 1e-300);
    t[i] = (float)((final_c[i] - phik_c[i]) / (sig * nscale));
}

}  // namespace

// Wait until the check with sequence number seq has written the host's status
// mirror: spin on its sequence word (the status arrives within microseconds of
// the check), then, past a few milliseconds (long sweep batches, or a failed
// launch), synchronise the stream, which also reports errors.
int cd_wait_status(cd_context* ctx, unsigned seq) {
    const volatile unsigned* w = &ctx->host_status->seq;
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0;; ++i) {
        if (*w == seq) {
            std::atomic_thread_fence(std::memory_order_acquire);
            return CD_OK;
        }
        if ((i & 255) == 255 &&
            std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(4))
            break;
    }
    CD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->host_status->seq != seq) {
        ctx->set_error("a check finished without writing the host status");
        return CD_ERR_HIP;
    }
    return CD_OK;
}

int cd_solve_device_internal(cd_context* ctx, const double* b, double* phi, const uint8_t* mask,
                             long long valid_cells, int w, int h, const cd_solve_opts* o,
                             cd_solve_stats* st_out, const SolveWork* work) {
    auto t_start = std::chrono::steady_clock::now();
    if (w < 2 || h < 2 || !b || !phi || !o) return CD_ERR_BAD_ARG;
    if (o->stop != CD_STOP_RESIDUAL && o->stop != CD_STOP_MAX_UPDATE) return CD_ERR_BAD_ARG;
    if (o->precision != CD_PREC_MIXED && o->precision != CD_PREC_FP64) return CD_ERR_BAD_ARG;
    const bool mixed = o->precision == CD_PREC_MIXED;
    const size_t n = (size_t)w * h;
    const double count = (double)valid_cells;

    // workspace
    const dim3 rblk(RED_BX, RED_BY);
    const dim3 rgrid(div_up(w, RED_BX), div_up(h, RED_BY));
    const int nred = rgrid.x * rgrid.y;
    // FP32 sweeps without holes take the temporally blocked kernel on the
    // packed layout; FP64 sweeps and hole masks take the per-half-sweep kernel.
    // (the blocked kernel addresses a plane pair through 32-bit byte offsets)
    const Brick bk = make_brick(w, h);
    const bool blocked = mixed && mask == nullptr && 2 * bk.plane * sizeof(float) < (1ull << 31);
    const size_t npk = 2 * bk.plane;
    SolveWork ws;
    if (work) {
        ws = *work;
    } else {
        // (sum, sum^2) per block for checks, four dot products per block for FCG
        CD_TRY(ctx->ensure(ctx->partials, sizeof(double) * 4 * (size_t)nred));
        CD_TRY(ctx->ensure(ctx->solve_dev, sizeof(SolveDev)));
        if (mixed) {
            CD_TRY(ctx->ensure(ctx->e32, sizeof(float) * (blocked ? npk : n)));
            // FP32 rhs (natural layout), BF16 rhs (bricked layout) or, under
            // the max-update rule, FP32 rhs (bricked layout)
            const size_t rb = blocked ? (o->stop == CD_STOP_MAX_UPDATE ? sizeof(float) : sizeof(unsigned short)) * npk
                                      : sizeof(float) * n;
            CD_TRY(ctx->ensure(ctx->r32, rb));
            if (blocked) CD_TRY(ctx->ensure(ctx->e32b, sizeof(float) * npk));
        }
        ws.partials = ctx->partials.as<double>();
        ws.sd = ctx->solve_dev.as<SolveDev>();
        ws.e32 = mixed ? ctx->e32.as<float>() : nullptr;
        ws.e32b = blocked ? ctx->e32b.as<float>() : nullptr;
        ws.r = mixed ? ctx->r32.ptr : nullptr;
    }
    double* partials = ws.partials;
    SolveDev* sd = ws.sd;
    float* e32 = mixed ? ws.e32 : nullptr;
    float* e32b = blocked ? ws.e32b : nullptr;
    float* r32 = mixed ? static_cast<float*>(ws.r) : nullptr;
    unsigned short* r16 = blocked ? static_cast<unsigned short*>(ws.r) : nullptr;
    const int track = o->stop == CD_STOP_MAX_UPDATE;
    // kernel timers record HIP events (created on demand): not with caller buffers
    auto timer_begin = [&]() -> int { return work ? (int)CD_OK : ctx->timer_begin(); };
    auto timer_end = [&](int kernel, long long launches, double bytes) -> int {
        return work ? (int)CD_OK : ctx->timer_end(kernel, launches, bytes);
    };
    // padding stays finite
    if (mixed && blocked) CD_TRY(cd_fill_f32(ctx, r32, track ? npk : npk / 2, 0.f));

    const double omega = cd_sor_omega(o->omega, w, h);
    const int kmax = o->check_every > 0 ? o->check_every : auto_check_every(w, h);
    const bool adaptive = o->check_every <= 0 && o->stop == CD_STOP_RESIDUAL;
    hipStream_t s = ctx->stream;

    // right-hand-side statistics: mean over non-hole cells, ||P b||
    hipLaunchKernelGGL(k_rhs_moments, rgrid, rblk, 0, s, b, w, h, sd, 0, partials);
    hipLaunchKernelGGL(k_rhs_finish, dim3(1), dim3(1024), 0, s, partials, nred, sd, 0, count);
    hipLaunchKernelGGL(k_rhs_moments, rgrid, rblk, 0, s, b, w, h, sd, 1, partials);
    hipLaunchKernelGGL(k_rhs_finish, dim3(1), dim3(1024), 0, s, partials, nred, sd, 1, count);
    CD_CHECK_LAUNCH(ctx);
    if (mixed) CD_TRY(cd_fill_f32(ctx, e32, blocked ? npk : n, 0.f));
    if (blocked) CD_TRY(cd_fill_f32(ctx, e32b, npk, 0.f));

    const int fblocks = std::min(div_up((long long)n, CD_BLOCK), 65536);
    // residual launch of a check: kResidualRows rows per thread, fewer on
    // smaller grids (more workgroups in flight; the partials stay <= nred)
    const int kResidualRows = nred >= 16384 ? 8 : nred >= 8192 ? 4 : nred >= 4096 ? 2 : 1;
    const dim3 cgrid(div_up(w, RED_BX), div_up(h, RED_BY * kResidualRows));
    const int ncheck = cgrid.x * cgrid.y;  // <= nred
    // flush: fold the current FP32 correction into phi64 first (the learned
    // iterations add theirs to phi64 themselves)
    // the end of every check: k_check on the residual pass's partials, then the
    // host's wait for the status (the residual pass's timer is open)
    auto finish_check = [&](int check_id, int timer_kernels, double bytes,
                            int nparts) -> int {
        const unsigned seq = ++ctx->status_seq;
        hipLaunchKernelGGL(k_check, dim3(1), dim3(1024), 0, s, partials, nparts, sd,
                           ctx->host_status, seq, count, check_id, o->stop, o->tol);
        CD_CHECK_LAUNCH(ctx);
        CD_TRY(timer_end(CD_KERNEL_RESIDUAL, timer_kernels, bytes));
        if (ctx->profiling) {
            CD_HIP(ctx, hipStreamSynchronize(s));
            CD_TRY(ctx->resolve_timers());
        } else {
            CD_TRY(cd_wait_status(ctx, seq));
        }
        return CD_OK;
    };
    auto enqueue_check = [&](int check_id, bool flush = true) -> int {
        CD_TRY(timer_begin());
        if (blocked) {
            if (flush)
                hipLaunchKernelGGL(k_flush_packed, rgrid, rblk, 0, s, phi, e32, e32b, sd, w, h, bk);
            if (track)
                launch_residual<3>(kResidualRows, cgrid, rblk, s, phi, b, mask, sd, r32, r16, bk,
                                   w, h, partials);
            else
                launch_residual<2>(kResidualRows, cgrid, rblk, s, phi, b, mask, sd, r32, r16, bk,
                                   w, h, partials);
        } else if (mixed) {
            hipLaunchKernelGGL(k_flush, dim3(fblocks), dim3(CD_BLOCK), 0, s, phi, e32, n);
            launch_residual<1>(kResidualRows, cgrid, rblk, s, phi, b, mask, sd, r32, r16, bk, w, h,
                               partials);
        } else {
            launch_residual<0>(kResidualRows, cgrid, rblk, s, phi, b, mask, sd, r32, r16, bk, w, h,
                               partials);
        }
        return finish_check(check_id, (mixed && flush) ? 3 : 2,
                            (blocked ? (flush ? 48.0 : 18.0) : (mixed ? 44.0 : 16.0)) * (double)n,
                            ncheck);
    };

    const dim3 sblk(SWEEP_BX, SWEEP_BY);
    const dim3 sgrid(div_up(div_up(w, 2), SWEEP_BX), div_up(h, SWEEP_BY));
    // blocked geometry
    constexpr int BH = CD_BK_H;  // halo = half-sweeps per launch
    const int bix = BK_RX - 2 * BH, biy = BK_RY - 2 * BH;
    const int tiles_x = div_up(w, bix), tiles_y = div_up(h, biy);
    const int ntiles = tiles_x * tiles_y;
    const int bgrid = blocked_grid(ntiles);
    const float om = (float)omega;
    const float om2 = om / 2.f, om3 = om / 3.f, om4 = om / 4.f;
    // pro: the first launch starts from this intermediate-grid correction,
    // interpolated at the cells, instead of from e32
    auto enqueue_blocked_om = [&](int first, int count_sweeps, float om2, float om3, float om4,
                                  const MidGrid* pro = nullptr) -> int {
        CD_TRY(timer_begin());
        int launches = 0;
        const MidGrid none{};
        for (int done_sw = 0; done_sw < count_sweeps;) {
            const int nsw = std::min(BH / 2, count_sweeps - done_sw);
            const bool full = 2 * nsw == BH;
            const int in = launches & 1;
            float* ein = in ? e32b : e32;
            float* eout = in ? e32 : e32b;
            if (track) {
                hipLaunchKernelGGL(full ? (k_rb_blocked<BH, true, true>) : (k_rb_blocked<BH, true, false>),
                                   dim3(bgrid), dim3(BK_THREADS), 0, s, ein, eout, r32, bk, w, h,
                                   tiles_x, ntiles, 2 * nsw, om2, om3, om4, &sd->rmean32, sd,
                                   1 - in, none);
                hipLaunchKernelGGL(k_maxcheck_blocked, dim3(1), dim3(64), 0, s, sd,
                                   first + done_sw, nsw, o->tol);
            } else if (pro && launches == 0) {
                hipLaunchKernelGGL(full ? (k_rb_blocked<BH, false, true, true>)
                                        : (k_rb_blocked<BH, false, false, true>),
                                   dim3(bgrid), dim3(BK_THREADS), 0, s, ein, eout, r16, bk, w, h,
                                   tiles_x, ntiles, 2 * nsw, om2, om3, om4, &sd->rmean32, sd,
                                   1 - in, *pro);
            } else {
                hipLaunchKernelGGL(full ? (k_rb_blocked<BH, false, true>) : (k_rb_blocked<BH, false, false>),
                                   dim3(bgrid), dim3(BK_THREADS), 0, s, ein, eout, r16, bk, w, h,
                                   tiles_x, ntiles, 2 * nsw, om2, om3, om4, &sd->rmean32, sd,
                                   1 - in, none);
            }
            ++launches;
            done_sw += nsw;
        }
        CD_CHECK_LAUNCH(ctx);
        // algorithmic bytes: 12 B per cell per red+black sweep (e read, e write,
        // rhs read), whatever the number of sweeps fused into one launch
        CD_TRY(timer_end(CD_KERNEL_SOR_SWEEP, launches, 12.0 * (double)n * count_sweeps));
        return CD_OK;
    };
    auto enqueue_blocked = [&](int first, int count_sweeps) -> int {
        return enqueue_blocked_om(first, count_sweeps, om2, om3, om4);
    };
    auto enqueue_sweeps = [&](int first, int count_sweeps) -> int {
        if (blocked) return enqueue_blocked(first, count_sweeps);
        CD_TRY(timer_begin());
        for (int k = 0; k < count_sweeps; ++k) {
            const int sw = first + k;
            for (int color = 0; color < 2; ++color) {
                if (mixed) {
                    hipLaunchKernelGGL(k_rb_sweep<float>, sgrid, sblk, 0, s, e32, r32, mask, w,
                                       h, color, (float)omega, &sd->rmean32, sd, track, sw);
                } else {
                    hipLaunchKernelGGL(k_rb_sweep<double>, sgrid, sblk, 0, s, phi, b, mask, w,
                                       h, color, omega, &sd->bmean, sd, track, sw);
                }
            }
            if (track) hipLaunchKernelGGL(k_maxcheck, dim3(1), dim3(64), 0, s, sd, sw, o->tol);
        }
        CD_CHECK_LAUNCH(ctx);
        CD_TRY(timer_end(CD_KERNEL_SOR_SWEEP, 2LL * count_sweeps,
                              (mixed ? 12.0 : 24.0) * (double)n * count_sweeps));
        return CD_OK;
    };

    // ---- learned correction / capture geometry ------------------------------
    const cd_net_arch* arch = ctx->net ? cd_net_arch_of(ctx->net) : nullptr;
    const double nscale = std::pow((double)std::max(w, h) / M_PI, 2.0);
    auto coarse_ok = [&](int f, int lv) {
        const int m = f << (lv - 1);
        return f > 0 && w % f == 0 && h % f == 0 && (w / f) % (1 << (lv - 1)) == 0 &&
               (h / f) % (1 << (lv - 1)) == 0 && m > 0;
    };
    // (the step check addresses the FP64 grids through 32-bit byte offsets)
    bool learned = o->learned_init && arch && blocked && o->stop == CD_STOP_RESIDUAL &&
                   coarse_ok(arch->coarsen, arch->levels) && !work &&
                   (size_t)n * sizeof(double) < (1ull << 31);
    const int lf = arch ? arch->coarsen : 1;
    const int lwc = w / std::max(lf, 1), lhc = h / std::max(lf, 1);
    const bool capture = ctx->cap_x && ctx->cap_coarsen > 0 && w == ctx->cap_coarsen * ctx->cap_w &&
                         h == ctx->cap_coarsen * ctx->cap_h && blocked && !track && !work;
    const int cwc = capture ? ctx->cap_w : 0, chc = capture ? ctx->cap_h : 0;
    constexpr int kMaxPairsPerSolve = 64;
    std::vector<int> cap_slots;
    if (learned) {
        CD_TRY(ctx->ensure(ctx->lc_r, sizeof(float) * (size_t)lwc * lhc));
        CD_TRY(ctx->ensure(ctx->lc_x, sizeof(float) * (size_t)lwc * lhc));
        CD_TRY(ctx->ensure(ctx->lc_y, sizeof(float) * (size_t)lwc * lhc));
        CD_TRY(ctx->ensure(ctx->fcg_p, sizeof(float) * npk));
        CD_TRY(ctx->ensure(ctx->fcg_q, sizeof(float) * npk));
        CD_TRY(ctx->ensure(ctx->fcg_dev, sizeof(FcgDev)));
        CD_TRY(ctx->ensure(ctx->fcg_phi, sizeof(double) * n));
    }
    // intermediate grid of the preconditioner (coarsening a multiple of 4, >= 8)
    const int mid_sweeps = arch ? arch->mid_sweeps : 0;
    const bool mid = learned && mid_sweeps > 0 && lf >= 8 && lf % 4 == 0;
    const int fm = mid ? lf / 4 : 1;
    const int wm = w / fm, hm = h / fm;
    const Brick bkm = make_brick(wm, hm);
    if (mid) {
        const size_t npm = 2 * bkm.plane;
        CD_TRY(ctx->ensure(ctx->fcg_em, sizeof(float) * npm));
        CD_TRY(ctx->ensure(ctx->fcg_emb, sizeof(float) * npm));
        CD_TRY(ctx->ensure(ctx->fcg_rm, sizeof(unsigned short) * npm));
        CD_TRY(ctx->ensure(ctx->fcg_zero, sizeof(float) * 4));
        // padding cells stay finite; the intermediate sweeps subtract a zero mean
        CD_TRY(cd_fill_f32(ctx, ctx->fcg_em.as<float>(), npm, 0.f));
        CD_TRY(cd_fill_f32(ctx, ctx->fcg_emb.as<float>(), npm, 0.f));
        CD_TRY(cd_fill_f32(ctx, ctx->fcg_rm.as<float>(), npm / 2, 0.f));
        CD_TRY(cd_fill_f32(ctx, ctx->fcg_zero.as<float>(), 4, 0.f));
    }
    if (learned || capture)
        CD_TRY(ctx->ensure(ctx->lc_stats, sizeof(double) * 2 * (kMaxPairsPerSolve + 2)));
    if (capture) {
        CD_TRY(ctx->ensure(ctx->cap_phi, sizeof(double) * (size_t)cwc * chc * (kMaxPairsPerSolve + 1)));
        CD_TRY(ctx->ensure(ctx->lc_r, sizeof(float) * (size_t)std::max(cwc * chc, lwc * lhc)));
    }
    // restrict the residual of the last check to (wc x hc) into lc_r; stats
    // slot receives (sum, sum^2) of the coarse residual
    // x8 (may be NULL): also write the network input there (coarse grids of up
    // to 64Ki cells; *x8_done tells whether it was).  direct: x8 receives the
    // coarse residual itself (the input of a linear network) and no statistics
    // are computed
    auto restrict_stats = [&](int f, int wc, int hc, double* stats, double* derived,
                              unsigned short* x8 = nullptr, bool* x8_done = nullptr,
                              bool direct = false) -> int {
        const int nc = wc * hc;
        const int rb = f < 8 ? div_up(nc, 256) : div_up(nc, 4);  // thread / wave per cell
        hipLaunchKernelGGL(k_restrict_r16, dim3(rb), dim3(256), 0, s, r16, bk, sd, f, wc, hc,
                           ctx->lc_r.as<float>(), direct ? x8 : nullptr);
        if (x8_done) *x8_done = direct && x8;
        if (direct && x8) {
            CD_CHECK_LAUNCH(ctx);
            return CD_OK;
        }
        if (nc <= 65536) {
            hipLaunchKernelGGL(k_coarse_stats1, dim3(1), dim3(1024), 0, s, ctx->lc_r.as<float>(),
                               nc, stats, derived, x8);
            if (x8_done) *x8_done = x8 != nullptr;
        } else {
            const int nb = div_up(nc, 256);  // <= nred: the partials buffer is large enough
            hipLaunchKernelGGL(k_coarse_stats, dim3(nb), dim3(256), 0, s, ctx->lc_r.as<float>(),
                               nc, partials);
            CD_TRY(cd_reduce_partials(ctx, partials, nb, 2, stats));
            if (derived)
                hipLaunchKernelGGL(k_coarse_derived, dim3(1), dim3(64), 0, s, stats, nc, derived);
        }
        CD_CHECK_LAUNCH(ctx);
        return CD_OK;
    };
    // restrict the residual of the last check to (wc x hc), normalise into
    // xout; stats slot receives (sum, sum^2) of the coarse residual
    auto restrict_normalize = [&](int f, int wc, int hc, float* xout, double* stats) -> int {
        CD_TRY(restrict_stats(f, wc, hc, stats, nullptr));
        const int nc = wc * hc;
        hipLaunchKernelGGL(k_normalize_coarse, dim3(div_up(nc, 256)), dim3(256), 0, s,
                           ctx->lc_r.as<float>(), stats, nc, xout);
        CD_CHECK_LAUNCH(ctx);
        return CD_OK;
    };
    // the FP64 iterate: phi, or during learned iterations either buffer of the
    // pair (phi, fcg_phi) that the folded FCG steps alternate between
    double* phi_cur = phi;
    auto capture_pair = [&]() -> int {
        if (ctx->cap_count >= ctx->cap_capacity || (int)cap_slots.size() >= kMaxPairsPerSolve)
            return CD_OK;
        const int slot = ctx->cap_count++;
        const int i = (int)cap_slots.size();
        const int nc = cwc * chc;
        double* stats = ctx->lc_stats.as<double>() + 2 * (2 + i);
        CD_TRY(restrict_normalize(ctx->cap_coarsen, cwc, chc, ctx->cap_x + (size_t)slot * nc, stats));
        hipLaunchKernelGGL(k_restrict_phi, dim3(div_up(nc, 256)), dim3(256), 0, s, phi_cur, w,
                           ctx->cap_coarsen, cwc, chc, ctx->cap_phi.as<double>() + (size_t)(i + 1) * nc);
        CD_CHECK_LAUNCH(ctx);
        cap_slots.push_back(slot);
        return CD_OK;
    };
    const float oms = arch ? (float)arch->smooth_omega : om;
    const int nsmooth = arch ? std::max(1, arch->smooth_sweeps) : 4;
    // the smoothing launches start from e32; after `nl` launches the result
    // sits in e32 (nl even) or e32b (nl odd)
    const int smooth_launches = div_up(nsmooth, BH / 2);
    // packed-element kernels: grid (packed width / 256, rows, 2 planes); the
    // dot products stride over rows so that their partials fit nred blocks
    const int pbx = div_up(bk.wpb, 256);
    const dim3 packed_grid(pbx, h, 2);
    // (at most CD_DOT_BANDS bands of rows: each thread accumulates several rows
    // before the block reduction; 1024 measured slower at 1024^2, 128 and 512
    // no faster at 2048^2)
    const int dot_rows = std::max(1, std::min(std::min(h, CD_DOT_BANDS), nred / (2 * pbx)));
    const dim3 dot_grid(pbx, dot_rows, 2);
    const int dot_blocks = 2 * pbx * dot_rows;
    if (dot_blocks > nred) learned = false;  // tiny grids: partials hold nred blocks
    float* zbuf = (smooth_launches & 1) ? e32b : e32;
    int learned_applied = 0, learned_rejected = 0;
    // a linear network maps the unnormalised coarse residual (scaling commutes
    // with it): the restriction writes the network input, no statistics pass
    const bool direct_input = learned && arch->linear;
    // ... and where the check's workgroups cover whole coarse cells, the check
    // that ends an FCG iteration writes the next iteration's network input
    // itself (the iterations after the first start from it)
    // the step check's rows per thread and grid (its partials stay <= nred)
    const int kStepRows = std::min(kResidualRows, CD_STEP_ROWS);
    const dim3 stgrid(div_up(w, RED_BX), div_up(h, RED_BY * kStepRows));
    const int nstep = stgrid.x * stgrid.y;
    const int crows = 4 * kStepRows;  // rows of a step-check workgroup
    const bool fuse_restrict = direct_input && lf >= 2 && lf <= 64 && (lf & (lf - 1)) == 0 &&
                               w % 64 == 0 && h % crows == 0 && crows % lf == 0 &&
                               crows / lf <= 2 && lwc * lf == w && lhc * lf == h;
    unsigned short* net_x8 = nullptr;
    if (fuse_restrict) CD_TRY(cd_net_coarse_input(ctx->net, lhc, lwc, &net_x8));
    bool x8_ready = false;  // the last check wrote the network input

    int check_id = 0;
    int cum = 0;
    CD_TRY(enqueue_check(check_id));
    if (capture && !ctx->host_status->done) CD_TRY(capture_pair());
    double prev_rel = ctx->host_status->rel;
    int k = kmax;
    if (adaptive) k = std::min(kmax, std::max(8, kmax / 4));
    bool done = ctx->host_status->done != 0;
    // the warm start, kept while learned iterations run: if they are abandoned
    // with a residual above the starting one (degenerate right-hand sides of
    // folded meshes can do that: a lower energy norm of the error does not
    // bound the residual), plain SOR restarts from it
    const double start_rel = ctx->host_status->rel;
    if (learned && !done) {
        CD_TRY(cd_fill_f32(ctx, ctx->fcg_p.as<float>(), npk, 0.f));
        CD_TRY(cd_fill_f32(ctx, ctx->fcg_q.as<float>(), npk, 0.f));
        hipLaunchKernelGGL(k_fcg_init, dim3(1), dim3(64), 0, s, ctx->fcg_dev.as<FcgDev>(),
                           ctx->fcg_p.as<float>());
        CD_CHECK_LAUNCH(ctx);
        if (direct_input) {  // (mean, rms) of the unnormalised input: (0, 1)
            double* derived = ctx->lc_stats.as<double>() + 2;
            CD_TRY(cd_fill_f64(ctx, derived, 1, 0.0));
            CD_TRY(cd_fill_f64(ctx, derived + 1, 1, 1.0));
        }
        CD_TRY(ctx->ensure(ctx->phi_start, sizeof(double) * n));
        CD_HIP(ctx, hipMemcpyAsync(ctx->phi_start.ptr, phi, sizeof(double) * n,
                                   hipMemcpyDeviceToDevice, s));
    }
    // learned FCG iterations; abandoned for plain SOR when the residual stops
    // falling (the iterate is kept when its residual is below the starting
    // one: every step lowered the error's energy norm)
    double best_rel = ctx->host_status->rel;
    int stall = 0;
    // one FCG iteration up to the next check: restriction, network, prolongation,
    // smoothing sweeps, FCG update.  Launched eagerly the first time in a solve
    // (which also sizes every buffer), then replayed as one captured graph.
    auto fcg_body = [&](int first_sweep) -> int {
        FcgDev* fd = ctx->fcg_dev.as<FcgDev>();
        CD_TRY(timer_begin());
        // slot 0: (sum, sum^2) of the coarse residual; slot 1: (mean, rms)
        double* stats = ctx->lc_stats.as<double>();
        unsigned short* x8 = nullptr;
        bool prepared = false;
        CD_TRY(cd_net_coarse_input(ctx->net, lhc, lwc, &x8));
        if (x8_ready && x8 == net_x8)
            prepared = true;
        else
            CD_TRY(restrict_stats(lf, lwc, lhc, stats, stats + 2, x8, &prepared, direct_input));
        if (prepared)
            CD_TRY(cd_net_forward_prepared(ctx->net, ctx->lc_y.as<float>(), lhc, lwc));
        else
            CD_TRY(cd_net_forward_coarse(ctx->net, ctx->lc_r.as<float>(), stats,
                                         ctx->lc_y.as<float>(), lhc, lwc));
        if (mid) {
            // coarse -> intermediate grid, smoothing there, -> fine grid
            const int pbm = div_up(bkm.wpb, 256);
            unsigned short* rm = ctx->fcg_rm.as<unsigned short>();
            float* em[2] = {ctx->fcg_em.as<float>(), ctx->fcg_emb.as<float>()};
            hipLaunchKernelGGL(k_restrict_to_mid, dim3(pbm, hm, 2), dim3(256), 0, s, r16, bk, sd,
                               fm, wm, hm, bkm, rm, ctx->lc_y.as<float>(), stats + 2, lf / fm,
                               lwc, lhc, nscale, em[0]);
            const int tx = div_up(wm, bix), ntm = tx * div_up(hm, biy);
            const int gm = blocked_grid(ntm);
            int cur_m = 0;
            for (int done_sw = 0; done_sw < mid_sweeps;) {
                const int nsw = std::min(BH / 2, mid_sweeps - done_sw);
                hipLaunchKernelGGL(2 * nsw == BH ? (k_rb_blocked<BH, false, true>) : (k_rb_blocked<BH, false, false>),
                                   dim3(gm), dim3(BK_THREADS), 0, s, em[cur_m],
                                   em[1 - cur_m], rm, bkm, wm, hm, tx, ntm, 2 * nsw, oms / 2.f,
                                   oms / 3.f, oms / 4.f, ctx->fcg_zero.as<float>(), sd, 0,
                                   MidGrid{});
                cur_m = 1 - cur_m;
                done_sw += nsw;
            }
            CD_CHECK_LAUNCH(ctx);
            CD_TRY(timer_end(CD_KERNEL_LEARNED, 3, 0.0));
            // the fine smoothing starts from em interpolated at the fine cells
            const MidGrid pro{em[cur_m], bkm, wm, hm, fm};
            CD_TRY(enqueue_blocked_om(first_sweep, nsmooth, oms / 2.f, oms / 3.f, oms / 4.f, &pro));
        } else {
            hipLaunchKernelGGL(k_prolong_bricked, packed_grid, dim3(256), 0, s,
                               ctx->lc_y.as<float>(), stats + 2, w, h, lf, lwc, lhc, nscale, bk,
                               e32);
            CD_CHECK_LAUNCH(ctx);
            CD_TRY(timer_end(CD_KERNEL_LEARNED, 3, 0.0));
            CD_TRY(enqueue_blocked_om(first_sweep, nsmooth, oms / 2.f, oms / 3.f, oms / 4.f));
        }
        CD_TRY(timer_begin());
        hipLaunchKernelGGL(k_fcg_dots, dot_grid, dim3(256), 0, s, zbuf, fd, r16, sd, bk, w, h,
                           partials);
        hipLaunchKernelGGL(k_fcg_coef, dim3(1), dim3(1024), 0, s, partials, dot_blocks, fd);
        CD_CHECK_LAUNCH(ctx);
        CD_TRY(timer_end(CD_KERNEL_LEARNED, 2, 10.0 * (double)n));
        return CD_OK;
    };
    // the check that ends an FCG iteration, with the iteration's step folded
    // in: p = z + beta p_old, phi += alpha p, into the other buffer of each
    // pair (the direction pair is fcg_p / fcg_q)
    double* const phis[2] = {phi, ctx->fcg_phi.as<double>()};
    float* const dirs[2] = {ctx->fcg_p.as<float>(), ctx->fcg_q.as<float>()};
    int fk = 0, pk = 0;  // buffers holding the iterate and the last direction
    auto enqueue_step_check = [&](int check_id) -> int {
        CD_TRY(timer_begin());
        launch_residual_step(kStepRows, stgrid, rblk, s, phis[fk], phis[1 - fk], b, sd, zbuf,
                             dirs[pk], dirs[1 - pk], ctx->fcg_dev.as<FcgDev>(), r16, bk, w, h,
                             partials, fuse_restrict ? net_x8 : nullptr, lf, lwc);
        x8_ready = fuse_restrict;
        fk = 1 - fk;
        pk = 1 - pk;
        phi_cur = phis[fk];
        return finish_check(check_id, 2, 38.0 * (double)n, nstep);
    };
    int fcg_iter = 0;
    while (!done && learned && cum < o->max_sweeps) {
        if (o->max_sweeps - cum < nsmooth) { learned = false; break; }  // a cycle no longer fits
        // profiling keeps the launches separate so that every timer records
        const bool replay = ctx->graphs && !ctx->profiling && fcg_iter > 0;
        if (replay) {
            const cd_context::FcgKey key{w,        h,        nsmooth,
                                         mid ? mid_sweeps : 0,   ctx->net, ctx->gen,
                                         cd_net_gen(ctx->net)};
            if (!(ctx->fcg_exec && ctx->fcg_key == key)) {
                if (ctx->fcg_exec) {
                    CD_HIP(ctx, hipGraphExecDestroy(ctx->fcg_exec));
                    ctx->fcg_exec = nullptr;
                }
                hipGraph_t graph = nullptr;
                CD_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
                const int body = fcg_body(cum);
                const hipError_t ce = hipStreamEndCapture(s, &graph);
                if (body != CD_OK) {
                    if (graph) (void)hipGraphDestroy(graph);
                    return body;
                }
                CD_HIP(ctx, ce);
                const hipError_t ie = hipGraphInstantiate(&ctx->fcg_exec, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                CD_HIP(ctx, ie);
                ctx->fcg_key = key;
            }
            CD_HIP(ctx, hipGraphLaunch(ctx->fcg_exec, s));
        } else {
            CD_TRY(fcg_body(cum));
        }
        ++fcg_iter;
        cum += nsmooth;
        CD_TRY(enqueue_step_check(++check_id));
        ++learned_applied;
        const SolveDev& hs = *ctx->host_status;
        done = hs.done != 0;
        if (!std::isfinite(hs.rel)) stall = 7;  // abandoned below, back to the warm start
        if (capture && !done && std::isfinite(hs.rel) &&
            check_id % std::max(ctx->cap_every, 1) == 0)
            CD_TRY(capture_pair());
#ifdef CD_TRACE_LEARNED
        {
            FcgDev hf;
            CD_HIP(ctx, hipMemcpy(&hf, ctx->fcg_dev.ptr, sizeof(hf), hipMemcpyDeviceToHost));
            fprintf(stderr, "fcg %d: rel %.3e alpha %.3e beta %.3e pq %.3e bad %d\n",
                    learned_applied, hs.rel, hf.alpha, hf.beta, hf.pq, hf.bad);
        }
#endif
        // the residual of CG-type iterations may rise for a few steps before
        // it falls; give up only when 8 iterations bring no halving
        if (hs.rel < 0.5 * best_rel) {
            best_rel = hs.rel;
            stall = 0;
        } else if (++stall >= 8) {
            ++learned_rejected;
            learned = false;
            if (!(hs.rel <= start_rel) && !done) {
                // back to the warm start; the next check recomputes its residual
                CD_HIP(ctx, hipMemcpyAsync(phi, ctx->phi_start.ptr, sizeof(double) * n,
                                           hipMemcpyDeviceToDevice, s));
                fk = 0;
                phi_cur = phi;
                CD_TRY(enqueue_check(++check_id, false));
                done = ctx->host_status->done != 0;
            }
        }
        prev_rel = ctx->host_status->rel;
    }
    if (fk) {  // the iterate ended in the second buffer
        CD_HIP(ctx, hipMemcpyAsync(phi, phis[1], sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        fk = 0;
        phi_cur = phi;
    }
    // plain sweeps after learned iterations start from a zero correction (the
    // iterations left their FP32 work in both ping-pong buffers)
    if (fcg_iter > 0 && !done) {
        CD_TRY(cd_fill_f32(ctx, e32, npk, 0.f));
        CD_TRY(cd_fill_f32(ctx, e32b, npk, 0.f));
    }
    // Max-update rule in mixed precision: the sweeps see the right-hand side
    // rounded (BF16 / FP32) at the last check, which perturbs every update by
    // up to omega/4 times that rounding of the check's residual.  A stop is
    // therefore confirmed on a fresh FP64 residual: the check after the
    // stopping batch folds the correction and rounds the new residual, and
    // one more launch of sweeps must meet the rule again; if none of its
    // sweeps does, the solve goes on (and a later stop is confirmed the same
    // way).  FP64 sweeps read b itself and stop at once.
    const bool confirm_stop = track && mixed;
    const int confirm_sweeps = blocked ? BH / 2 : 4;
    bool confirming = false;
    while (!done && cum < o->max_sweeps) {
        const int nsw = std::min(confirming ? confirm_sweeps : k, o->max_sweeps - cum);
        CD_TRY(enqueue_sweeps(cum, nsw));
        cum += nsw;
        ++check_id;
        CD_TRY(enqueue_check(check_id));
        const SolveDev& hs = *ctx->host_status;
        done = hs.done != 0;
        if (!std::isfinite(hs.rel)) break;
        if (confirm_stop) {
            if (done && !confirming && cum < o->max_sweeps) {
                hipLaunchKernelGGL(k_reopen, dim3(1), dim3(64), 0, s, sd);
                CD_CHECK_LAUNCH(ctx);
                confirming = true;
                done = false;
                continue;
            }
            confirming = false;  // confirmed (done) or not: back to full batches
        }
        if (capture && !done && check_id % std::max(ctx->cap_every, 1) == 0)
            CD_TRY(capture_pair());
        if (adaptive && !done) {
            // geometric rate estimate over the last batch -> sweeps still needed
            double rel = hs.rel;
            double rho = (prev_rel > 0 && rel > 0 && rel < prev_rel) ? std::pow(rel / prev_rel, 1.0 / nsw) : 1.0;
            int want = kmax;
            if (rho < 0.999999) {
                double need = std::log(o->tol / rel) / std::log(rho);
                want = (int)std::ceil(need * CD_CHECK_MARGIN) + 2;
            }
            k = std::max(8, std::min(kmax, want));
            prev_rel = rel;
        }
    }
    if (capture && !cap_slots.empty()) {
        const int nc = cwc * chc;
        double* fin = ctx->cap_phi.as<double>();  // slot 0: the final solution
        hipLaunchKernelGGL(k_restrict_phi, dim3(div_up(nc, 256)), dim3(256), 0, s, phi, w,
                           ctx->cap_coarsen, cwc, chc, fin);
        for (size_t i = 0; i < cap_slots.size(); ++i)
            hipLaunchKernelGGL(k_capture_finish, dim3(div_up(nc, 256)), dim3(256), 0, s, fin,
                               fin + (i + 1) * nc, ctx->lc_stats.as<double>() + 2 * (2 + i), nc,
                               nscale, ctx->cap_t + (size_t)cap_slots[i] * nc);
        CD_CHECK_LAUNCH(ctx);
    }
    const SolveDev hs = *ctx->host_status;
    if (mask && o->max_sweeps > 0) {
        hipLaunchKernelGGL(k_set_holes_nan, dim3(div_up((long long)n, CD_BLOCK)), dim3(CD_BLOCK), 0,
                           s, phi, mask, n);
        CD_CHECK_LAUNCH(ctx);
    }
    CD_HIP(ctx, hipStreamSynchronize(s));

    int status = CD_OK;
    if (!std::isfinite(hs.rel)) status = CD_ERR_NAN;
    else if (!hs.done) status = CD_ERR_NOT_CONVERGED;
    if (st_out) {
        std::memset(st_out, 0, sizeof(*st_out));
        st_out->status = status;
        // the residual rule stops at the last check, after all enqueued sweeps
        st_out->sweeps = (hs.done && o->stop == CD_STOP_MAX_UPDATE) ? hs.sweeps_at_done : cum;
        st_out->checks = hs.checks;
        st_out->rel_residual = hs.rel;
        st_out->initial_rel_residual = hs.rel0;
        st_out->last_max_update = hs.last_max_update;
        st_out->learned_applied = learned_applied;
        st_out->learned_rejected = learned_rejected;
        st_out->elapsed_ms = std::chrono::duration<double, std::milli>(
                                 std::chrono::steady_clock::now() - t_start).count();
    }
    return status;
}

// Sizes of the solve buffers of a w x h grid (see SolveWork).
void cd_solve_work_sizes(int w, int h, size_t* nred, size_t* npk) {
    *nred = (size_t)div_up(w, RED_BX) * div_up(h, RED_BY);
    *npk = 2 * make_brick(w, h).plane;
}

// Hole mask of a right-hand side (NaN cells and their neighbours' validity
// bits) and the number of holes, read back through host_buf (pinned).
int cd_hole_mask(cd_context* ctx, const double* d_rhs, int w, int h, uint8_t* mask,
                 double* partials, double* fill, double* host_buf, long long* holes) {
    const dim3 rblk(RED_BX, RED_BY);
    const dim3 rgrid(div_up(w, RED_BX), div_up(h, RED_BY));
    const int nred = rgrid.x * rgrid.y;
    hipLaunchKernelGGL(k_build_mask, rgrid, rblk, 0, ctx->stream, d_rhs, mask, w, h, partials);
    CD_CHECK_LAUNCH(ctx);
    CD_TRY(cd_reduce_partials(ctx, partials, nred, 1, fill));
    CD_HIP(ctx, hipMemcpyAsync(host_buf, fill, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *holes = (long long)host_buf[0];
    return CD_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free solve.  Its buffers (reduction partials, solve state, the
// two ping-pong FP32 corrections, the right-hand side of the sweeps, the hole
// mask and the reduction slots) live in one static device array per GPU,
// sized for grids of up to CD_NOALLOC_MAX_SIDE x CD_NOALLOC_MAX_SIDE cells and
// held until the process ends.  The array's address on each device is taken
// once (hipGetSymbolAddress, with that device current) and kept per device;
// host code passes that address, never the array's name.
// ---------------------------------------------------------------------------
namespace {

constexpr size_t al256(size_t x) { return (x + 255) / 256 * 256; }

struct NoallocLayout {
    size_t partials, sd, e32, e32b, r, mask, fill, total;
};

constexpr NoallocLayout noalloc_layout(int w, int h) {
    // the bricked layout of make_brick and the reduction grid of the checks
    const size_t wp = (size_t)(w + 1) / 2;
    const size_t plane = ((size_t)(h + 7) / 8 * 8) * ((wp + 63) / 64 * 64);
    const size_t npk = 2 * plane;  // >= w*h (natural layout with holes fits)
    const size_t nred = (size_t)((w + RED_BX - 1) / RED_BX) * ((h + RED_BY - 1) / RED_BY);
    NoallocLayout L{};
    size_t o = 0;
    L.partials = o; o += al256(sizeof(double) * 4 * nred);
    L.sd = o;       o += al256(sizeof(SolveDev));
    L.e32 = o;      o += al256(sizeof(float) * npk);
    L.e32b = o;     o += al256(sizeof(float) * npk);
    L.r = o;        o += al256(sizeof(float) * npk);  // FP32 (max-update rule) or BF16
    L.mask = o;     o += al256((size_t)w * h);
    L.fill = o;     o += al256(sizeof(double) * 4);
    L.total = o;
    return L;
}

constexpr size_t kNoallocScratchBytes =
    noalloc_layout(CD_NOALLOC_MAX_SIDE, CD_NOALLOC_MAX_SIDE).total;
static_assert(kNoallocScratchBytes <= ((size_t)CD_NOALLOC_SCRATCH_MIB << 20) &&
                  kNoallocScratchBytes > ((size_t)(CD_NOALLOC_SCRATCH_MIB - 1) << 20),
              "CD_NOALLOC_SCRATCH_MIB states the scratch size");

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch (set up on the first call per
// device; the caller has made ctx->device current).
int noalloc_scratch(cd_context* ctx, unsigned char** out) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return CD_ERR_BAD_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* p = nullptr;
        CD_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            ctx->set_error("hipGetSymbolAddress returned no address for the solve scratch");
            return CD_ERR_HIP;
        }
        g_scratch_ptr[ctx->device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_ptr[ctx->device];
    return CD_OK;
}

}  // namespace

extern "C" {

int cd_poisson_solve_device_noalloc(cd_context* ctx, const double* d_rhs, double* d_phi,
                                    int width, int height, const cd_solve_opts* opts,
                                    cd_solve_stats* stats) {
    if (!ctx || !d_rhs || !d_phi || width < 2 || height < 2) return CD_ERR_BAD_ARG;
    cd_solve_opts def;
    if (!opts) { cd_solve_opts_default(&def); opts = &def; }
    if (opts->learned_init) {
        ctx->set_error("cd_poisson_solve_device_noalloc: learned_init needs the network's "
                       "buffers; use cd_poisson_solve_device");
        return CD_ERR_BAD_ARG;
    }
    const NoallocLayout L = noalloc_layout(width, height);
    if (L.total > kNoallocScratchBytes) {
        ctx->set_error("cd_poisson_solve_device_noalloc: a " + std::to_string(width) + " x " +
                       std::to_string(height) + " grid needs " + std::to_string(L.total) +
                       " bytes of scratch, more than the " +
                       std::to_string(kNoallocScratchBytes) + " held per device (grids up to " +
                       std::to_string(CD_NOALLOC_MAX_SIDE) + " x " +
                       std::to_string(CD_NOALLOC_MAX_SIDE) + ")");
        return CD_ERR_BAD_ARG;
    }
    CD_HIP(ctx, hipSetDevice(ctx->device));
    unsigned char* base = nullptr;
    CD_TRY(noalloc_scratch(ctx, &base));
    SolveWork ws;
    ws.partials = reinterpret_cast<double*>(base + L.partials);
    ws.sd = reinterpret_cast<SolveDev*>(base + L.sd);
    ws.e32 = reinterpret_cast<float*>(base + L.e32);
    ws.e32b = reinterpret_cast<float*>(base + L.e32b);
    ws.r = base + L.r;
    uint8_t* mask = base + L.mask;
    double* fill = reinterpret_cast<double*>(base + L.fill);
    long long holes = 0;
    CD_TRY(cd_hole_mask(ctx, d_rhs, width, height, mask, ws.partials, fill,
                        reinterpret_cast<double*>(ctx->host_status), &holes));
    const long long n = (long long)width * height;
    return cd_solve_device_internal(ctx, d_rhs, d_phi, holes > 0 ? mask : nullptr, n - holes,
                                    width, height, opts, stats, &ws);
}

int cd_poisson_solve_device(cd_context* ctx, const double* d_rhs, double* d_phi, int width,
                            int height, const cd_solve_opts* opts, cd_solve_stats* stats) {
    if (!ctx || !d_rhs || !d_phi || width < 2 || height < 2) return CD_ERR_BAD_ARG;
    cd_solve_opts def;
    if (!opts) { cd_solve_opts_default(&def); opts = &def; }
    CD_HIP(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)width * height;
    size_t nred, npk;
    cd_solve_work_sizes(width, height, &nred, &npk);
    // hole mask: only used by the solve when the right-hand side contains NaN
    CD_TRY(ctx->ensure(ctx->mask, n));
    CD_TRY(ctx->ensure(ctx->partials, sizeof(double) * 4 * nred));  // as the solve
    CD_TRY(ctx->ensure(ctx->fill, sizeof(double) * 4));
    long long holes = 0;
    CD_TRY(cd_hole_mask(ctx, d_rhs, width, height, ctx->mask.as<uint8_t>(),
                        ctx->partials.as<double>(), ctx->fill.as<double>(),
                        reinterpret_cast<double*>(ctx->host_status), &holes));
    const uint8_t* mask = holes > 0 ? ctx->mask.as<uint8_t>() : nullptr;
    return cd_solve_device_internal(ctx, d_rhs, d_phi, mask, (long long)n - holes, width, height,
                                    opts, stats);
}

int cd_context_set_capture(cd_context* ctx, float* d_x, float* d_t, int capacity, int h, int w,
                           int coarsen, int every) {
    if (!ctx) return CD_ERR_BAD_ARG;
    if (d_x && (!d_t || capacity < 1 || h < 1 || w < 1 || coarsen < 1)) return CD_ERR_BAD_ARG;
    ctx->cap_x = d_x;
    ctx->cap_t = d_t;
    ctx->cap_capacity = d_x ? capacity : 0;
    ctx->cap_count = 0;
    ctx->cap_h = h;
    ctx->cap_w = w;
    ctx->cap_coarsen = coarsen;
    ctx->cap_every = every < 1 ? 1 : every;
    return CD_OK;
}

int cd_context_capture_count(const cd_context* ctx, int* count) {
    if (!ctx || !count) return CD_ERR_BAD_ARG;
    *count = ctx->cap_count;
    return CD_OK;
}

int cd_poisson_solve(cd_context* ctx, const double* rhs, double* phi, int width, int height,
                     const cd_solve_opts* opts, cd_solve_stats* stats) {
    if (!ctx || !rhs || !phi || width < 2 || height < 2) return CD_ERR_BAD_ARG;
    CD_HIP(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)width * height;
    CD_TRY(ctx->ensure(ctx->b64w, sizeof(double) * n));
    CD_TRY(ctx->ensure(ctx->phi64w, sizeof(double) * n));
    CD_HIP(ctx, hipMemcpyAsync(ctx->b64w.ptr, rhs, sizeof(double) * n, hipMemcpyHostToDevice,
                               ctx->stream));
    CD_HIP(ctx, hipMemcpyAsync(ctx->phi64w.ptr, phi, sizeof(double) * n, hipMemcpyHostToDevice,
                               ctx->stream));
    int st = cd_poisson_solve_device(ctx, ctx->b64w.as<double>(), ctx->phi64w.as<double>(), width,
                                     height, opts, stats);
    if (st != CD_OK && st != CD_ERR_NOT_CONVERGED && st != CD_ERR_NAN) return st;
    CD_HIP(ctx, hipMemcpyAsync(phi, ctx->phi64w.ptr, sizeof(double) * n, hipMemcpyDeviceToHost,
                               ctx->stream));
    CD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return st;
}

}  // extern "C"
