// Learned initial correction: a bias-free convolutional encoder-decoder with
// hand-written gfx950 kernels (BF16 MFMA, FP32 accumulation) for inference and
// training.  No counterpart in the reference; see cd_api.h for the contract.
//
// Tensors are NHWC.  Activations are BF16 with channel counts that are
// multiples of 8 (16-byte vectors per pixel); gradients are FP32.  A 3x3
// convolution is an implicit GEMM: M = pixels, N = output channels,
// K = 9 * input channels (tap-major: k = tap * Cin + ci, tap = 3*dy + dx),
// computed with v_mfma_f32_16x16x32_bf16.  Lane layouts of that instruction:
//   A (16x32): lane l holds A[l % 16][8*(l/16) + j], j = 0..7
//   B (32x16): lane l holds B[8*(l/16) + j][l % 16]
//   D (16x16): lane l holds D[4*(l/16) + r][l % 16], r = 0..3
#include <hip/hip_runtime.h>

#include <atomic>

#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "cd_internal.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;

namespace {

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

__device__ __forceinline__ u16 to_bf16(float x) {
    const unsigned u = __float_as_uint(x);
    return (u16)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float from_bf16(u16 v) { return __uint_as_float((unsigned)v << 16); }

union V8 {
    bf16x8 v;
    uint4 u;
    u16 h[8];
};

// ---------------------------------------------------------------------------
// 3x3 convolution, implicit GEMM on MFMA.
//   IN 0: BF16 input = channel concatenation of x1 (c1 ch) and x2 (c2 ch)
//   IN 1: FP32 input g (c1 ch), multiplied by (mask > 0) when mask != NULL
//   OUT 0: BF16 output (cy ch), optional ReLU
//   OUT 1: FP32 output split by channel into y1 (cy ch) and y2 (cy2 ch),
//          optionally accumulated
// wt: BF16 weights [nb*16][kpad], kpad a multiple of 32, zero padded.
// One workgroup = 4 waves = 64 consecutive pixels; a wave computes 16 pixels
// by NB*16 output channels.
// ---------------------------------------------------------------------------
template <int NB, int IN, int OUT>
__global__ __launch_bounds__(256) void k_conv3x3(const u16* __restrict__ x1,
                                                 const u16* __restrict__ x2, int c1, int c2,
                                                 const float* __restrict__ g,
                                                 const u16* __restrict__ mask,
                                                 const u16* __restrict__ wt, int kpad, int B,
                                                 int H, int W, u16* __restrict__ ybf,
                                                 float* __restrict__ y1, float* __restrict__ y2,
                                                 int cy, int cy2, int relu, int accumulate,
                                                 int mode1) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long npix = (long long)B * H * W;
    const long long p0 = (long long)blockIdx.x * 64 + wave * 16;
    const int cin = IN == 0 ? c1 + c2 : c1;
    const int K = 9 * cin;
    const long long pa = p0 + (lane & 15);
    const bool pvalid = pa < npix;
    int xb = 0, yb = 0, bb = 0;
    if (pvalid) {
        xb = (int)(pa % W);
        const long long t = pa / W;
        yb = (int)(t % H);
        bb = (int)(t / H);
    }
    floatx4 acc[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) acc[nb] = floatx4{0.f, 0.f, 0.f, 0.f};
    const int kq = 8 * (lane >> 4);
    const u16* wrow = wt + (size_t)(lane & 15) * kpad + kq;
    // A fragment of K chunk kc (8 consecutive k of this lane's pixel)
    auto load_a = [&](int kc) -> V8 {
        V8 a;
        a.u = make_uint4(0, 0, 0, 0);
        const int k = kc + kq;
        if (pvalid && k < K) {
            const int tap = k / cin, ci = k - tap * cin;
            const int yy = yb + tap / 3 - 1, xx = xb + tap % 3 - 1;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                const long long q = ((long long)bb * H + yy) * W + xx;
                if (IN == 0) {
                    if (ci >= c1) {
                        a.u = *reinterpret_cast<const uint4*>(x2 + q * c2 + (ci - c1));
                    } else if (mode1 == 0) {
                        a.u = *reinterpret_cast<const uint4*>(x1 + q * c1 + ci);
                    } else {
                        // x1 holds [B][H/2][W/2][c1]: nearest 2x upsampling
                        const long long q1 =
                            ((long long)bb * (H >> 1) + (yy >> 1)) * (W >> 1) + (xx >> 1);
                        a.u = *reinterpret_cast<const uint4*>(x1 + q1 * c1 + ci);
                    }
                } else {
                    const float4 g0 = *reinterpret_cast<const float4*>(g + q * c1 + ci);
                    const float4 g1 = *reinterpret_cast<const float4*>(g + q * c1 + ci + 4);
                    float gv[8] = {g0.x, g0.y, g0.z, g0.w, g1.x, g1.y, g1.z, g1.w};
                    if (mask) {
                        V8 m;
                        m.u = *reinterpret_cast<const uint4*>(mask + q * c1 + ci);
#pragma unroll
                        for (int j = 0; j < 8; ++j)
                            if (!(from_bf16(m.h[j]) > 0.f)) gv[j] = 0.f;
                    }
#pragma unroll
                    for (int j = 0; j < 8; ++j) a.h[j] = to_bf16(gv[j]);
                }
            }
        }
        return a;
    };
    // K is walked in groups of 4 chunks: all loads of a group are issued
    // before its MFMAs (latency overlap); the accumulation order is unchanged
    constexpr int U = 4;
    for (int kc0 = 0; kc0 < kpad; kc0 += 32 * U) {
        V8 a[U], bw[U][NB];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int kc = kc0 + 32 * u;
            if (kc < kpad) {
                a[u] = load_a(kc);
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    bw[u][nb].u = *reinterpret_cast<const uint4*>(wrow + (size_t)nb * 16 * kpad + kc);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (kc0 + 32 * u < kpad) {
#pragma unroll
                for (int nb = 0; nb < NB; ++nb)
                    acc[nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[u].v, bw[u][nb].v, acc[nb],
                                                                      0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const long long p = p0 + 4 * (lane >> 4) + r;
            const int ch = nb * 16 + (lane & 15);
            if (p >= npix) continue;
            float v = acc[nb][r];
            if (OUT == 0) {
                if (ch < cy) {
                    if (relu) v = fmaxf(v, 0.f);
                    ybf[p * cy + ch] = to_bf16(v);
                }
            } else {
                float* dst = nullptr;
                if (ch < cy) dst = y1 + p * cy + ch;
                else if (ch - cy < cy2) dst = y2 + p * cy2 + (ch - cy);
                if (dst) *dst = accumulate ? *dst + v : v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Forward 3x3 convolution on one 8x8 output tile per workgroup (4 waves, two
// tile rows each).  The tile's input (10x10 pixels with the zero-padded halo,
// x1 read through nearest 2x upsampling when mode1, then x2 concatenated) and
// the weights of this workgroup's NB*16 output channels are staged in LDS by
// one round of coalesced loads; the MFMA loop then reads only LDS.  The A/B
// fragments and the order of the K chunks are those of k_conv3x3<NB, 0, OUT>,
// so every output is bit-identical to it; the tiles cut the latency chain of
// the per-chunk global gathers, which sets the time of these small layers.
//   OUT 0: BF16 output ybf [B][H][W][cy] (ReLU optional);
//   OUT 1: FP32 output yf [B][H][W][cy].
// grid: (ceil(W/8), ceil(H/8), B * cblocks), output channels
// [16*NB*cblk, 16*NB*(cblk+1)) with cblk = blockIdx.z % cblocks.
// LDS: 100 * (cin + 8) + NB*16 * (kpad + 8) u16 (row paddings spread banks).
// ---------------------------------------------------------------------------
constexpr int CT = 8;  // output tile side

// Stage a P x P patch of the (concatenated) input, image origin (oy, ox), into
// LDS at xs (u16 pixel stride xstride): thread t takes vector t % ng of pixels
// t / ng, t / ng + 256 / ng, ... (one division per thread, none in the loop).
// The loads go in groups of SB without branches (each from a clamped, valid
// pixel; out-of-image pixels are zeroed after the load), so that a group waits
// for one L2 round trip, not one per vector.  mode1 as for k_conv3x3_tile.
template <int P>
__device__ __forceinline__ void stage_patch(u16* __restrict__ xs, int xstride,
                                            const u16* __restrict__ x1,
                                            const u16* __restrict__ x2, int c1, int c2, int H,
                                            int W, int b, int oy, int ox, int mode1) {
    constexpr int SB = 8;
    constexpr int NP = P * P;
    const int ng = (c1 + c2) >> 3;
    const int pstride = 256 / ng;
    const int gi = threadIdx.x % ng;
    const int ci = gi * 8;
    // this thread's source: channels ci.. of x2 (after the c1 of x1) or x1;
    // mode1 applies to x1 only (a pooled input has no x2)
    const bool use2 = ci >= c1;
    const u16* __restrict__ src = use2 ? x2 + (ci - c1) : x1 + ci;
    const int cs = use2 ? c2 : c1;
    const int up = (!use2 && mode1 == 1) ? 1 : 0;  // nearest upsampling: source at half size
    if (threadIdx.x >= pstride * ng) return;
    for (int pb = threadIdx.x / ng; pb < NP; pb += SB * pstride) {
        uint4 v[SB];
        bool ok[SB];
        if (mode1 == 2) {
            // x1 at twice the resolution, 2x2 average pooled as k_avgpool2 does
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int px = pb + u * pstride;
                const int yy = oy + px / P, xx = ox + px % P;
                ok[u] = px < NP && yy >= 0 && yy < H && xx >= 0 && xx < W;
                const int yc = min(max(yy, 0), H - 1), xc = min(max(xx, 0), W - 1);
                const long long r0 = ((long long)b * (2 * H) + 2 * yc) * (2 * W) + 2 * xc;
                const long long r1 = r0 + 2 * W;
                V8 p00, p01, p10, p11, o;
                p00.u = *reinterpret_cast<const uint4*>(src + r0 * cs);
                p01.u = *reinterpret_cast<const uint4*>(src + (r0 + 1) * cs);
                p10.u = *reinterpret_cast<const uint4*>(src + r1 * cs);
                p11.u = *reinterpret_cast<const uint4*>(src + (r1 + 1) * cs);
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float sm = from_bf16(p00.h[e]) + from_bf16(p01.h[e]) +
                                     from_bf16(p10.h[e]) + from_bf16(p11.h[e]);
                    o.h[e] = to_bf16(0.25f * sm);
                }
                v[u] = o.u;
            }
        } else {
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int px = pb + u * pstride;
                const int yy = oy + px / P, xx = ox + px % P;
                ok[u] = px < NP && yy >= 0 && yy < H && xx >= 0 && xx < W;
                const int yc = min(max(yy, 0), H - 1) >> up, xc = min(max(xx, 0), W - 1) >> up;
                const long long q = ((long long)b * (H >> up) + yc) * (W >> up) + xc;
                v[u] = *reinterpret_cast<const uint4*>(src + q * cs);
            }
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const int px = pb + u * pstride;
            if (px < NP)
                *reinterpret_cast<uint4*>(xs + px * xstride + ci) =
                    ok[u] ? v[u] : make_uint4(0, 0, 0, 0);
        }
    }
}

// Stage `rows` weight rows of kpad (u16) from wsrc into ws (row stride kpad + 8):
// vector it = row * kv + cv for it = threadIdx.x, +256, ... with (row, cv)
// advanced incrementally, in branch-free groups of 8 loads (a vector past the
// last row is loaded from that row and stored there: the value it has)
__device__ __forceinline__ void stage_weights(u16* __restrict__ ws,
                                              const u16* __restrict__ wsrc, int rows,
                                              int kpad) {
    constexpr int SB = 8;
    const int kv = kpad >> 3, wstride = kpad + 8;
    const int dq = 256 / kv, dr = 256 - dq * kv;
    int row = threadIdx.x / kv, cv = threadIdx.x - row * kv;
    while (row < rows) {
        uint4 v[SB];
        int rr[SB], cc[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            rr[u] = min(row, rows - 1);
            cc[u] = cv;
            v[u] = *reinterpret_cast<const uint4*>(wsrc + (size_t)rr[u] * kpad + cv * 8);
            row += dq;
            cv += dr;
            if (cv >= kv) { cv -= kv; ++row; }
        }
#pragma unroll
        for (int u = 0; u < SB; ++u)
            *reinterpret_cast<uint4*>(ws + rr[u] * wstride + cc[u] * 8) = v[u];
    }
}

// The K loop of a 3x3 convolution read from LDS: an input patch of side P
// (pixel stride xstride, cin channels) and NB*16 weight rows (stride kpad + 8).
// Lane l's A pixel of M-tile m has patch index pbase[m] (its 3x3 neighbourhood
// starts there); acc[m][nb] accumulates output channels nb*16 + (l % 16) in
// the order of the K chunks, the order k_conv3x3 uses.  k = kc + kq = tap *
// cin + cc is tracked incrementally: cin is a multiple of 8, so the lane's 8
// consecutive k lie in one tap.
template <int NB, int MT, int P>
__device__ __forceinline__ void conv_lds(floatx4 (&acc)[MT][NB], const u16* __restrict__ xs,
                                         int xstride, int cin, const u16* __restrict__ ws,
                                         int kpad, const int (&pbase)[MT], int lane) {
    const int K = 9 * cin;
    const int kq = 8 * (lane >> 4);
    const int wstride = kpad + 8;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[m][nb] = floatx4{0.f, 0.f, 0.f, 0.f};
    const u16* wrow = ws + (lane & 15) * wstride + kq;
    int cc = kq, tdy = 0, tdx = 0;
    while (cc >= cin) {
        cc -= cin;
        if (++tdx == 3) { tdx = 0; ++tdy; }
    }
    for (int kc = 0; kc < kpad; kc += 32) {
        V8 a[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            a[m].u = make_uint4(0, 0, 0, 0);
            if (kc + kq < K)
                a[m].u = *reinterpret_cast<const uint4*>(
                    xs + (pbase[m] + tdy * P + tdx) * xstride + cc);
        }
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            V8 bw;
            bw.u = *reinterpret_cast<const uint4*>(wrow + nb * 16 * wstride + kc);
#pragma unroll
            for (int m = 0; m < MT; ++m)
                acc[m][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[m].v, bw.v, acc[m][nb],
                                                                     0, 0, 0);
        }
        cc += 32;
        while (cc >= cin) {
            cc -= cin;
            if (++tdx == 3) { tdx = 0; ++tdy; }
        }
    }
}

template <int NB, int OUT>
__global__ __launch_bounds__(256) void k_conv3x3_tile(const u16* __restrict__ x1,
                                                      const u16* __restrict__ x2, int c1, int c2,
                                                      const u16* __restrict__ wt, int kpad, int H,
                                                      int W, int cblocks, u16* __restrict__ ybf,
                                                      float* __restrict__ yf, int cy, int relu,
                                                      int mode1) {
    extern __shared__ uint4 tile_lds[];
    u16* xs = reinterpret_cast<u16*>(tile_lds);
    const int cin = c1 + c2;
    const int xstride = cin + 8;                      // u16 per tile pixel
    u16* ws = xs + 100 * xstride;                     // 16-byte aligned: xstride % 8 == 0
    const int cblk = blockIdx.z % cblocks, b = blockIdx.z / cblocks;
    const int tx0 = blockIdx.x * CT, ty0 = blockIdx.y * CT;
    stage_patch<10>(xs, xstride, x1, x2, c1, c2, H, W, b, ty0 - 1, tx0 - 1, mode1);
    stage_weights(ws, wt + (size_t)cblk * NB * 16 * kpad, NB * 16, kpad);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // this lane's A pixel: tile row 2*wave + (l%16)/8, column (l%16)%8
    const int ar = 2 * wave + ((lane & 15) >> 3), ac = lane & 7;
    const int pbase[1] = {ar * 10 + ac};
    floatx4 acc2[1][NB];
    conv_lds<NB, 1, 10>(acc2, xs, xstride, cin, ws, kpad, pbase, lane);
    floatx4(&acc)[NB] = acc2[0];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // D row 4*(l/16) + r is tile pixel (2*wave + row/8, row % 8)
            const int row = 4 * (lane >> 4) + r;
            const int yy = ty0 + 2 * wave + (row >> 3), xx = tx0 + (row & 7);
            const int ch = (cblk * NB + nb) * 16 + (lane & 15);
            if (yy >= H || xx >= W || ch >= cy) continue;
            const long long p = ((long long)b * H + yy) * W + xx;
            float v = acc[nb][r];
            if (OUT == 0) {
                if (relu) v = fmaxf(v, 0.f);
                ybf[p * cy + ch] = to_bf16(v);
            } else {
                yf[p * cy + ch] = v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Two consecutive 3x3 convolutions of one resolution level in one launch
// (inference): conv a reads its inputs as k_conv3x3_tile does (mode1) on the
// 12x12 patch around an 8x8 output tile and computes the 10x10 region that
// conv b needs (7 M-tiles of 16 pixels, two per wave), which stays in LDS as
// BF16 (zero outside the image: conv b's padding); conv b then computes the
// tile from it.  Every workgroup computes all output channels of both convs
// (NBA*16 and NBB*16).  The A/B fragments and the K order of both are those of
// k_conv3x3_tile, and the intermediate is rounded to BF16 as the stored tensor
// is, so the output is bit-identical to the two launches; conv a's output is
// not written (inference only).  wb_early: conv b's weights are staged with
// conv a's at the start (separate LDS); else after conv a, over conv a's.
// grid: (ceil(W/8), ceil(H/8), B).  LDS (u16): 144 (cin + 8) + 100 (cya + 8) +
// NBA*16 (kpada + 8) [+ NBB*16 (kpadb + 8)], or max of the two weight blocks.
// ---------------------------------------------------------------------------
template <int NBA, int NBB>
__global__ __launch_bounds__(256) void k_conv_pair_tile(
    const u16* __restrict__ x1, const u16* __restrict__ x2, int c1, int c2,
    const u16* __restrict__ wa, int kpada, int cya, int relua, const u16* __restrict__ wb,
    int kpadb, int cyb, int relub, u16* __restrict__ ybf, int H, int W, int mode1,
    int wb_early) {
    extern __shared__ uint4 tile_lds[];
    const int cin = c1 + c2;
    const int xstride = cin + 8, istride = cya + 8;
    u16* xs = reinterpret_cast<u16*>(tile_lds);
    u16* ism = xs + 144 * xstride;
    u16* wsa = ism + 100 * istride;
    u16* wsb = wb_early ? wsa + NBA * 16 * (kpada + 8) : wsa;
    const int b = blockIdx.z;
    const int tx0 = blockIdx.x * CT, ty0 = blockIdx.y * CT;
    stage_patch<12>(xs, xstride, x1, x2, c1, c2, H, W, b, ty0 - 2, tx0 - 2, mode1);
    stage_weights(wsa, wa, NBA * 16, kpada);
    if (wb_early) stage_weights(wsb, wb, NBB * 16, kpadb);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    {
        // conv a: M-tiles wave and wave + 4 of the 10x10 region (pixels past
        // 100 compute on pixel 0 and are dropped)
        int pbase[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const int q = (wave + 4 * m) * 16 + (lane & 15);
            pbase[m] = q < 100 ? (q / 10) * 12 + q % 10 : 0;
        }
        floatx4 acc[2][NBA];
        conv_lds<NBA, 2, 12>(acc, xs, xstride, cin, wsa, kpada, pbase, lane);
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int nb = 0; nb < NBA; ++nb)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int q = (wave + 4 * m) * 16 + 4 * (lane >> 4) + r;
                    const int ch = nb * 16 + (lane & 15);
                    if (q >= 100 || ch >= cya) continue;
                    const int yy = ty0 - 1 + q / 10, xx = tx0 - 1 + q % 10;
                    float v = acc[m][nb][r];
                    if (relua) v = fmaxf(v, 0.f);
                    ism[q * istride + ch] =
                        (yy >= 0 && yy < H && xx >= 0 && xx < W) ? to_bf16(v) : (u16)0;
                }
    }
    __syncthreads();
    if (!wb_early) {
        stage_weights(wsb, wb, NBB * 16, kpadb);
        __syncthreads();
    }
    // conv b on the 8x8 tile: lane's A pixel tile row 2*wave + (l%16)/8,
    // column (l%16)%8, i.e. region pixel (row + 1, column + 1) whose 3x3
    // neighbourhood starts at region index row*10 + column
    const int ar = 2 * wave + ((lane & 15) >> 3), ac = lane & 7;
    const int pbase[1] = {ar * 10 + ac};
    floatx4 acc[1][NBB];
    conv_lds<NBB, 1, 10>(acc, ism, istride, cya, wsb, kpadb, pbase, lane);
#pragma unroll
    for (int nb = 0; nb < NBB; ++nb)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = 4 * (lane >> 4) + r;
            const int yy = ty0 + 2 * wave + (row >> 3), xx = tx0 + (row & 7);
            const int ch = nb * 16 + (lane & 15);
            if (yy >= H || xx >= W || ch >= cyb) continue;
            const long long p = ((long long)b * H + yy) * W + xx;
            float v = acc[0][nb][r];
            if (relub) v = fmaxf(v, 0.f);
            ybf[p * cyb + ch] = to_bf16(v);
        }
}

// ---------------------------------------------------------------------------
// Weight gradient dW[co][k] += sum_p gm[p][co] * xs[p][k], gm = g*(mask>0),
// xs the 3x3-shifted (concatenated) input: MFMA with M = output channels,
// N = k, reduction over pixels in steps of 32.  grid: (co blocks, k blocks,
// pixel chunks); each wave reduces `ppw` pixels and adds its 16x16 tile with
// FP32 atomics.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_wgrad(const u16* __restrict__ x1,
                                               const u16* __restrict__ x2, int c1, int c2,
                                               const float* __restrict__ g,
                                               const u16* __restrict__ mask, int cg, int B,
                                               int H, int W, float* __restrict__ dw, int kpad,
                                               int ppw) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int co = blockIdx.x * 16 + (lane & 15);
    const int kcol = blockIdx.y * 16 + (lane & 15);
    const int cin = c1 + c2;
    const int K = 9 * cin;
    const long long npix = (long long)B * H * W;
    const long long pbeg = ((long long)blockIdx.z * 4 + wave) * ppw;
    int tap = 0, ci = 0, dy = 0, dx = 0;
    const bool kvalid = kcol < K;
    if (kvalid) {
        tap = kcol / cin;
        ci = kcol - tap * cin;
        dy = tap / 3 - 1;
        dx = tap % 3 - 1;
    }
    const bool cvalid = co < cg;
    floatx4 acc = {0.f, 0.f, 0.f, 0.f};
    for (long long ps = pbeg; ps < pbeg + ppw && ps < npix; ps += 32) {
        V8 a, b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const long long p = ps + 8 * (lane >> 4) + j;
            float gv = 0.f;
            u16 xv = 0;
            if (p < npix) {
                if (cvalid) {
                    gv = g[p * cg + co];
                    if (mask && !(from_bf16(mask[p * cg + co]) > 0.f)) gv = 0.f;
                }
                if (kvalid) {
                    const int xx = (int)(p % W) + dx;
                    const long long t = p / W;
                    const int yy = (int)(t % H) + dy;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                        const long long q = (t / H * H + yy) * W + xx;
                        xv = ci < c1 ? x1[q * c1 + ci] : x2[q * c2 + (ci - c1)];
                    }
                }
            }
            a.h[j] = to_bf16(gv);
            b.h[j] = xv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int m = blockIdx.x * 16 + 4 * (lane >> 4) + r;  // output channel row of D
        if (m < cg && kvalid) atomicAdd(dw + (size_t)m * kpad + kcol, acc[r]);
    }
}

// ---- elementwise / pooling ---------------------------------------------------
__global__ void k_prep_input(const float* __restrict__ x, u16* __restrict__ x8, long long npix) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    V8 v;
    v.u = make_uint4(0, 0, 0, 0);
    v.h[0] = to_bf16(x[p]);
    *reinterpret_cast<uint4*>(x8 + p * 8) = v.u;
}

// normalised coarse residual -> 8-channel BF16 network input: x = (rc - mean)
// / rms from stats = (sum, sum^2) over n cells (the solver's coarse residual)
__global__ void k_prep_normalized(const float* __restrict__ rc, const double* __restrict__ stats,
                                  int n, u16* __restrict__ x8) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double mean = stats[0] / n;
    const double var = stats[1] / n - mean * mean;
    const double sig = sqrt(var > 1e-300 ? var : 1e-300);
    V8 v;
    v.u = make_uint4(0, 0, 0, 0);
    v.h[0] = to_bf16((float)(((double)rc[p] - mean) / sig));
    *reinterpret_cast<uint4*>(x8 + (long long)p * 8) = v.u;
}

__global__ void k_avgpool2(const u16* __restrict__ x, u16* __restrict__ y, int B, int H, int W,
                           int C) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    const long long n = (long long)B * Ho * Wo * C;
    if (i >= n) return;
    const int c = (int)(i % C);
    long long t = i / C;
    const int xo = (int)(t % Wo);
    t /= Wo;
    const int yo = (int)(t % Ho);
    const int b = (int)(t / Ho);
    const long long base = (((long long)b * H + 2 * yo) * W + 2 * xo) * C + c;
    const float s = from_bf16(x[base]) + from_bf16(x[base + C]) + from_bf16(x[base + (long long)W * C]) +
                    from_bf16(x[base + (long long)W * C + C]);
    y[i] = to_bf16(0.25f * s);
}

// dx[fine] += dy[coarse] / 4
__global__ void k_avgpool2_bwd(const float* __restrict__ dy, float* __restrict__ dx, int B, int H,
                               int W, int C) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)B * H * W * C;
    if (i >= n) return;
    const int c = (int)(i % C);
    long long t = i / C;
    const int x = (int)(t % W);
    t /= W;
    const int y = (int)(t % H);
    const int b = (int)(t / H);
    dx[i] += 0.25f * dy[(((long long)b * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
}

// nearest 2x upsampling of x [B][H/2][W/2][C] to y [B][H][W][C]
__global__ void k_upsample2(const u16* __restrict__ x, u16* __restrict__ y, int B, int H, int W,
                            int C) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = (long long)B * H * W * C;
    if (i >= n) return;
    const int c = (int)(i % C);
    long long t = i / C;
    const int xx = (int)(t % W);
    t /= W;
    const int yy = (int)(t % H);
    const int b = (int)(t / H);
    y[i] = x[(((long long)b * (H / 2) + yy / 2) * (W / 2) + xx / 2) * C + c];
}

// dx[coarse] += sum of the 4 fine gradients
__global__ void k_upsample2_bwd(const float* __restrict__ dy, float* __restrict__ dx, int B, int H,
                                int W, int C) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int Ho = H / 2, Wo = W / 2;
    const long long n = (long long)B * Ho * Wo * C;
    if (i >= n) return;
    const int c = (int)(i % C);
    long long t = i / C;
    const int xo = (int)(t % Wo);
    t /= Wo;
    const int yo = (int)(t % Ho);
    const int b = (int)(t / Ho);
    const long long base = (((long long)b * H + 2 * yo) * W + 2 * xo) * C + c;
    dx[i] += dy[base] + dy[base + C] + dy[base + (long long)W * C] + dy[base + (long long)W * C + C];
}

// supervised loss on the output y (FP32 [npix]): partial sums of (y - t)^2
// and dL/dy = 2 (y - t) / npix
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ y, const float* __restrict__ t,
                                              float* __restrict__ gy, long long npix,
                                              double* __restrict__ partials) {
    __shared__ double smem[16];
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (p < npix) {
        const float d = y[p] - t[p];
        e2 = (double)d * d;
        if (gy) gy[p] = 2.f * d / (float)npix;
    }
    const double s = cd_block_sum(e2, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Neumann 5-point operator (sum of in-grid neighbours - count * centre) of a
// [B][H][W] field with element stride `stride`
__device__ __forceinline__ float lap_field(const float* __restrict__ v, int stride, long long p,
                                           int x, int y, int H, int W) {
    const float c = v[p * stride];
    float s = 0.f, cnt = 0.f;
    if (x > 0) { s += v[(p - 1) * stride]; cnt += 1.f; }
    if (x < W - 1) { s += v[(p + 1) * stride]; cnt += 1.f; }
    if (y > 0) { s += v[(p - W) * stride]; cnt += 1.f; }
    if (y < H - 1) { s += v[(p + W) * stride]; cnt += 1.f; }
    return s - cnt * c;
}

// out = in + Lc(in) / 8: one symmetric smoothing step (gain 1 - lambda/8 on
// the operator's eigenvalue lambda in [0, 8]); used on the network output and
// to weight the operator loss towards the frequencies the fine smoother
// cannot reach
__global__ __launch_bounds__(256) void k_smooth_step(const float* __restrict__ in,
                                                     float* __restrict__ out, int B, int H, int W) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)B * H * W) return;
    const int xx = (int)(p % W), yy = (int)((p / W) % H);
    out[p] = in[p] + 0.125f * lap_field(in, 1, p, xx, yy, H, W);
}

// k smoothing steps in one launch: a 16x16 output tile with a k-cell halo
// lives in LDS (two buffers); step s updates the cells at distance >= s from
// the halo border, so the tile interior equals k global steps.  Same
// arithmetic and summation order as k_smooth_step (bit-identical).  The input
// has element stride `istride` (8: head channel 0).
// 8x8 tiles (256 workgroups at 128x128, halo regions of (8 + 2k)^2 cells):
// 6.2 us for k = 6 against 10.2 us with 16x16 tiles (64 workgroups), whose
// waves carry three to four cells each through the steps
#ifndef CD_SMOOTH_TILE
#define CD_SMOOTH_TILE 8
#endif
constexpr int kSmoothTile = CD_SMOOTH_TILE;
constexpr int kSmoothMaxK = 12;  // halo <= 12: 2 (tile + 2k)^2 floats of LDS
// cells of a region per thread, at most
constexpr int kSmoothCells =
    ((kSmoothTile + 2 * kSmoothMaxK) * (kSmoothTile + 2 * kSmoothMaxK) + 255) / 256;
// The k smoothing steps on an R x R region (R = kSmoothTile + 2k, image origin
// (X0, Y0)) held in LDS at a (b: scratch of the same size); returns the buffer
// holding the result.  Each thread keeps the region coordinates of its (at
// most kSmoothCells) cells for every step.
__device__ __forceinline__ float* smooth_lds(float* a, float* b, int R, int k, int X0, int Y0,
                                             int H, int W) {
    const int RR = R * R;
    int rxs[kSmoothCells], rys[kSmoothCells];
#pragma unroll
    for (int u = 0; u < kSmoothCells; ++u) {
        const int idx = threadIdx.x + 256 * u;
        rys[u] = idx / R;
        rxs[u] = idx - rys[u] * R;
    }
    for (int st = 1; st <= k; ++st) {
#pragma unroll
        for (int u = 0; u < kSmoothCells; ++u) {
            const int idx = threadIdx.x + 256 * u;
            if (256 * u >= RR) break;  // no thread has cells from here on
            if (idx >= RR) continue;
            const int rx = rxs[u], ry = rys[u];
            const int gx = X0 + rx, gy = Y0 + ry;
            float c = a[idx];
            if (rx >= st && rx < R - st && ry >= st && ry < R - st && gx >= 0 && gx < W &&
                gy >= 0 && gy < H) {
                float sum = 0.f, cnt = 0.f;
                if (gx > 0) { sum += a[idx - 1]; cnt += 1.f; }
                if (gx < W - 1) { sum += a[idx + 1]; cnt += 1.f; }
                if (gy > 0) { sum += a[idx - R]; cnt += 1.f; }
                if (gy < H - 1) { sum += a[idx + R]; cnt += 1.f; }
                c = c + 0.125f * (sum - cnt * c);
            }
            b[idx] = c;
        }
        __syncthreads();
        float* t = a;
        a = b;
        b = t;
    }
    return a;
}

// the tile's interior of the smoothed region to out [B][H][W]
__device__ __forceinline__ void smooth_store(const float* a, int R, int k, int X0, int Y0,
                                             int H, int W, long long base,
                                             float* __restrict__ out) {
    for (int idx = threadIdx.x; idx < kSmoothTile * kSmoothTile; idx += blockDim.x) {
        const int ty = idx / kSmoothTile, tx = idx - ty * kSmoothTile;
        const int gx = X0 + k + tx, gy = Y0 + k + ty;
        if (gx < W && gy < H) out[base + (long long)gy * W + gx] = a[(ty + k) * R + tx + k];
    }
}

// k smoothing steps of a field read with element stride istride (8: head
// channel 0), one tile plus its k-cell halo per workgroup; the loads of all of
// a thread's cells are issued before the first LDS store
__global__ __launch_bounds__(256) void k_smooth_tiled(const float* __restrict__ in, int istride,
                                                      float* __restrict__ out, int B, int H,
                                                      int W, int k) {
    extern __shared__ float lds[];
    const int R = kSmoothTile + 2 * k, RR = R * R;
    float* a = lds;
    float* b = lds + RR;
    const int bz = blockIdx.z;
    const int X0 = blockIdx.x * kSmoothTile - k, Y0 = blockIdx.y * kSmoothTile - k;
    const long long base = (long long)bz * H * W;
    float v[kSmoothCells];
#pragma unroll
    for (int u = 0; u < kSmoothCells; ++u) {
        const int idx = threadIdx.x + 256 * u;
        const int ry = idx / R, rx = idx - ry * R;
        const int gx = X0 + rx, gy = Y0 + ry;
        v[u] = (idx < RR && gx >= 0 && gx < W && gy >= 0 && gy < H)
                   ? in[(base + (long long)gy * W + gx) * istride]
                   : 0.f;
    }
#pragma unroll
    for (int u = 0; u < kSmoothCells; ++u)
        if (threadIdx.x + 256 * u < RR) a[threadIdx.x + 256 * u] = v[u];
    __syncthreads();
    a = smooth_lds(a, b, R, k, X0, Y0, H, W);
    smooth_store(a, R, k, X0, Y0, H, W, base, out);
}

// Head convolution and output smoothing in one launch (inference, K steps):
// the head (16 output channels of which channel 0 is the output, no ReLU) on
// the (8 + 2K)^2 region around an 8x8 tile, read from a (10 + 2K)^2 input
// patch in LDS with the A/B fragments and K order of k_conv3x3_tile, so its
// values are those k_conv3x3_tile<1, 1> stores; outside the image the region
// holds 0, as k_smooth_tiled loads it; then the K steps as k_smooth_tiled
// takes them.  The output is bit-identical to the two launches.  grid
// (ceil(W/8), ceil(H/8), B); LDS (bytes): 2 P^2 (C + 8) + 2*16 (kpad + 8)
// + 8 R^2, P = R + 2, R = 8 + 2K.
template <int K>
__global__ __launch_bounds__(256) void k_head_smooth(const u16* __restrict__ x, int C,
                                                     const u16* __restrict__ wt, int kpad, int H,
                                                     int W, float* __restrict__ out) {
    constexpr int R = kSmoothTile + 2 * K, P = R + 2, NQ = R * R;
    constexpr int MT = (NQ + 63) / 64;  // M-tiles of 16 region pixels per wave
    static_assert(kSmoothTile == CT, "the head region is centred on the conv tile");
    extern __shared__ uint4 tile_lds[];
    const int xstride = C + 8;
    u16* xs = reinterpret_cast<u16*>(tile_lds);
    u16* ws = xs + P * P * xstride;
    float* a = reinterpret_cast<float*>(ws + 16 * (kpad + 8));
    float* b2 = a + NQ;
    const int b = blockIdx.z;
    const int tx0 = blockIdx.x * CT, ty0 = blockIdx.y * CT;
    const int X0 = tx0 - K, Y0 = ty0 - K;  // region origin
    // (x2 = x with no channels: never read; a null x2 here crashed the compiler)
    stage_patch<P>(xs, xstride, x, x, C, 0, H, W, b, Y0 - 1, X0 - 1, 0);
    stage_weights(ws, wt, 16, kpad);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int pbase[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const int q = (wave + 4 * m) * 16 + (lane & 15);
        pbase[m] = q < NQ ? (q / R) * P + q % R : 0;
    }
    floatx4 acc[MT][1];
    conv_lds<1, MT, P>(acc, xs, xstride, C, ws, kpad, pbase, lane);
    if ((lane & 15) == 0) {  // output channel 0
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = (wave + 4 * m) * 16 + 4 * (lane >> 4) + r;
                if (q >= NQ) continue;
                const int gy = Y0 + q / R, gx = X0 + q % R;
                a[q] = (gy >= 0 && gy < H && gx >= 0 && gx < W) ? acc[m][0][r] : 0.f;
            }
    }
    __syncthreads();
    const float* res = smooth_lds(a, b2, R, K, X0, Y0, H, W);
    smooth_store(res, R, K, X0, Y0, H, W, (long long)b * H * W, out);
}

// operator residual u = cs * Lc y - x
__global__ __launch_bounds__(256) void k_op_residual(const float* __restrict__ y,
                                                     const float* __restrict__ x,
                                                     float* __restrict__ u, int B, int H, int W,
                                                     float cs) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)B * H * W) return;
    const int xx = (int)(p % W), yy = (int)((p / W) % H);
    u[p] = cs * lap_field(y, 1, p, xx, yy, H, W) - x[p];
}

// partial sums of v^2; dv = (2 / npix) v in place when want_grad
__global__ __launch_bounds__(256) void k_sumsq(float* __restrict__ v, long long npix, int want_grad,
                                               double* __restrict__ partials) {
    __shared__ double smem[16];
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double e2 = 0.0;
    if (p < npix) {
        e2 = (double)v[p] * v[p];
        if (want_grad) v[p] = 2.f * v[p] / (float)npix;
    }
    const double s = cd_block_sum(e2, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// energy loss per pixel: with u = cs Lc y - x, J = y'(-cs Lc)y/2 + x'y has
// pixel contributions y (x - u) / 2 and dJ/dy = -u; partial sums of the
// contributions, gy = -u / npix when want_grad
__global__ __launch_bounds__(256) void k_energy(const float* __restrict__ y,
                                                const float* __restrict__ x,
                                                const float* __restrict__ u,
                                                float* __restrict__ gy, long long npix,
                                                int want_grad, double* __restrict__ partials) {
    __shared__ double smem[16];
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (p < npix) {
        e = 0.5 * (double)y[p] * ((double)x[p] - (double)u[p]);
        if (want_grad) gy[p] = -u[p] / (float)npix;
    }
    const double s = cd_block_sum(e, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// gy = cs * Lc du (Lc symmetric)
__global__ __launch_bounds__(256) void k_op_grad(const float* __restrict__ du,
                                                 float* __restrict__ gy, int B, int H, int W,
                                                 float cs) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (long long)B * H * W) return;
    const int xx = (int)(p % W), yy = (int)((p / W) % H);
    gy[p] = cs * lap_field(du, 1, p, xx, yy, H, W);
}

// head channel 0 -> y (stride 8 -> 1), and dL/dy -> head gradient
__global__ void k_head_out(const float* __restrict__ head, float* __restrict__ y, long long npix) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < npix) y[p] = head[p * 8];
}
__global__ void k_head_grad(const float* __restrict__ gy, float* __restrict__ ghead,
                            long long npix) {
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    ghead[p * 8] = gy[p];
#pragma unroll
    for (int c = 1; c < 8; ++c) ghead[p * 8 + c] = 0.f;
}

// FP32 master -> BF16 forward weights [coutpad][kpad] (same layout)
__global__ void k_cast_bf16(const float* __restrict__ w, u16* __restrict__ o, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) o[i] = to_bf16(w[i]);
}

// transposed, tap-flipped BF16 weights for the input gradient:
// wt[ci][tap' * coutA + co] = w[co][(8 - tap') * cin + ci]
__global__ void k_transpose_w(const float* __restrict__ w, u16* __restrict__ wt, int cout, int coutA,
                              int cin, int kpad, int cinpad, int kpadT) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)cinpad * kpadT) return;
    const int ci = (int)(i / kpadT);
    const int kk = (int)(i - (long long)ci * kpadT);
    float v = 0.f;
    if (ci < cin && kk < 9 * coutA) {
        const int tp = kk / coutA, co = kk - tp * coutA;
        if (co < cout) v = w[(size_t)co * kpad + (8 - tp) * cin + ci];
    }
    wt[i] = to_bf16(v);
}

__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, long long n, float lr, float b1, float b2,
                       float c1, float c2, float eps) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

template <typename T>
__global__ void k_zero(T* __restrict__ p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = T(0);
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================
struct NetConv {
    int c1, c2, cout, coutA;  // inputs (concatenated), logical and allocated outputs
    int relu;
    int K, kpad, coutpad;     // forward GEMM dims
    int cinpad, kpadT;        // input-gradient GEMM dims
    size_t woff;              // offset in the flat parameter arrays
    size_t nparam;            // coutpad * kpad
    int level;                // resolution level
    // tensors (indices into NetState::acts)
    int in1, in2, out;
};

struct NetTensor {
    int level, C;             // channels (multiple of 8)
    DevBuf act;               // BF16 (or FP32 for the head)
    DevBuf grad;              // FP32
};

struct cd_net {
    cd_context* ctx = nullptr;
    cd_net_arch arch;
    std::vector<NetConv> convs;
    size_t nparam = 0;
    DevBuf w, m, v, gw, wbf, wtbf, partials, head, ghead, x8;
    DevBuf yout, ytmp, gy, ures, utmp;  // FP32 [B][H][W]: output, dL/dy, loss work
    unsigned long long gen = 0;          // bumped when the activation buffers move
    unsigned id = 0;                     // unique per net created in the process
    size_t wtoff_total = 0;
    std::vector<size_t> wtoff;
    long long step = 0;
    // activation tensors for the current (batch, h, w)
    std::vector<NetTensor> t;
    int batch = 0, bh = 0, bw = 0;
    bool weights_dirty = true;
    bool fused_forward = true;  // inference forwards fold the decoder upsampling and the
                                // encoder pooling into the convs
    bool tiled_forward = true;  // forward convolutions on LDS-staged 8x8 tiles (bit-identical)
    bool head_pending = false;  // fused forward left the head to finish_output (k_head_smooth)
};

namespace {

// Dynamic LDS available to k_conv3x3_tile on this context's device (the
// kernels' limit raised to the device's per-workgroup maximum once).
size_t tiled_lds_limit(cd_context* c) {
    static std::atomic<long long> limit{-1};
    long long v = limit.load();
    if (v >= 0) return (size_t)v;
    int bytes = 0;
    if (hipDeviceGetAttribute(&bytes, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) !=
        hipSuccess)
        bytes = 0;
    const void* fns[] = {reinterpret_cast<const void*>(&k_conv3x3_tile<1, 0>),
                         reinterpret_cast<const void*>(&k_conv3x3_tile<2, 0>),
                         reinterpret_cast<const void*>(&k_conv3x3_tile<1, 1>),
                         reinterpret_cast<const void*>(&k_conv_pair_tile<1, 1>),
                         reinterpret_cast<const void*>(&k_conv_pair_tile<2, 2>),
                         reinterpret_cast<const void*>(&k_head_smooth<6>)};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) !=
            hipSuccess) {
            (void)hipGetLastError();
            bytes = std::min(bytes, 64 * 1024);
        }
    limit.store(bytes);
    return (size_t)bytes;
}

#ifndef CD_CONV_WIDE_MIN_BLOCKS
#define CD_CONV_WIDE_MIN_BLOCKS 512
#endif
constexpr long long kConvWideMinBlocks = CD_CONV_WIDE_MIN_BLOCKS;

// mode1: 0 reads in1 as stored; 1 reads `x1src` (the half-resolution tensor
// in1 is upsampled from) through nearest 2x upsampling instead; 2 reads
// `x1src` (the double-resolution tensor in1 is pooled from) through 2x2
// average pooling (the tiled kernel; otherwise the pooled tensor is written
// first)
int launch_conv(cd_net* n, const NetConv& cv, int B, int H, int W, bool grad_input,
                int mode1 = 0, const u16* x1src = nullptr) {
    cd_context* c = n->ctx;
    hipStream_t s = c->stream;
    const long long npix = (long long)B * H * W;
    const int blocks = div_up(npix, 64);
    // LDS the tiled kernel needs for this convolution (8x8 tile + halo, the
    // weights of up to 32 output channels)
    // 32 output channels per workgroup where the image has enough 8x8 tiles to
    // fill the GPU, else 16 (twice the workgroups, half the weights staged by
    // each: the deep levels' convolutions are bound by that staging)
    const long long tiles_t = (long long)B * div_up(H, CT) * div_up(W, CT);
    const int nbw_t = ((cv.coutpad / 16) % 2 == 0 && tiles_t * (cv.coutpad / 32) >= kConvWideMinBlocks) ? 2 : 1;
    const size_t lds_t = sizeof(u16) * (100 * (size_t)(cv.c1 + cv.c2 + 8) +
                                        (size_t)nbw_t * 16 * (cv.kpad + 8));
    const bool tiled = !grad_input && n->tiled_forward && lds_t <= tiled_lds_limit(c);
    if (mode1 == 2 && cv.c2 > 0) return CD_ERR_BAD_ARG;  // a pooled input has no skip part
    if (mode1 == 2 && !tiled) {
        const long long cnt = npix * cv.c1;
        hipLaunchKernelGGL(k_avgpool2, dim3(div_up(cnt, 256)), dim3(256), 0, s, x1src,
                           n->t[cv.in1].act.as<u16>(), B, 2 * H, 2 * W, cv.c1);
        mode1 = 0;
        x1src = nullptr;
    }
    if (tiled) {
        const u16* x1 = mode1 ? x1src
                              : (cv.in1 < 0 ? n->x8.as<u16>() : n->t[cv.in1].act.as<u16>());
        const u16* x2 = cv.in2 < 0 ? nullptr : n->t[cv.in2].act.as<u16>();
        const u16* wt = n->wbf.as<u16>() + cv.woff;
        const int nb = cv.coutpad / 16;
        // up to 32 output channels per workgroup (LDS), the rest in more blocks
        const int nbw = nbw_t;
        const int cblocks = nb / nbw;
        const size_t lds = lds_t;
        const dim3 grid(div_up(W, CT), div_up(H, CT), B * cblocks);
        if (cv.out < 0) {  // head: FP32 output [npix][8]
            if (nbw != 1 || cblocks != 1) return CD_ERR_BAD_ARG;
            hipLaunchKernelGGL((k_conv3x3_tile<1, 1>), grid, dim3(256), lds, s, x1, x2, cv.c1,
                               cv.c2, wt, cv.kpad, H, W, cblocks, (u16*)nullptr,
                               n->head.as<float>(), 8, 0, mode1);
        } else {
            u16* y = n->t[cv.out].act.as<u16>();
            if (nbw == 2)
                hipLaunchKernelGGL((k_conv3x3_tile<2, 0>), grid, dim3(256), lds, s, x1, x2,
                                   cv.c1, cv.c2, wt, cv.kpad, H, W, cblocks, y, (float*)nullptr,
                                   cv.coutA, cv.relu, mode1);
            else
                hipLaunchKernelGGL((k_conv3x3_tile<1, 0>), grid, dim3(256), lds, s, x1, x2,
                                   cv.c1, cv.c2, wt, cv.kpad, H, W, cblocks, y, (float*)nullptr,
                                   cv.coutA, cv.relu, mode1);
        }
    } else if (!grad_input) {
        const u16* x1 = mode1 ? x1src
                              : (cv.in1 < 0 ? n->x8.as<u16>() : n->t[cv.in1].act.as<u16>());
        const u16* x2 = cv.in2 < 0 ? nullptr : n->t[cv.in2].act.as<u16>();
        const u16* wt = n->wbf.as<u16>() + cv.woff;
        const int nb = cv.coutpad / 16;
        if (cv.out < 0) {  // head: FP32 output [npix][8]
#define CONV_F(NB) hipLaunchKernelGGL((k_conv3x3<NB, 0, 1>), dim3(blocks), dim3(256), 0, s, x1, x2, cv.c1, cv.c2, (const float*)nullptr, (const u16*)nullptr, wt, cv.kpad, B, H, W, (u16*)nullptr, n->head.as<float>(), (float*)nullptr, 8, 0, 0, 0, mode1)
            if (nb == 1) CONV_F(1);
            else return CD_ERR_BAD_ARG;
#undef CONV_F
        } else {
            u16* y = n->t[cv.out].act.as<u16>();
#define CONV_B(NB) hipLaunchKernelGGL((k_conv3x3<NB, 0, 0>), dim3(blocks), dim3(256), 0, s, x1, x2, cv.c1, cv.c2, (const float*)nullptr, (const u16*)nullptr, wt, cv.kpad, B, H, W, y, (float*)nullptr, (float*)nullptr, cv.coutA, 0, cv.relu, 0, mode1)
            switch (nb) {
                case 1: CONV_B(1); break;
                case 2: CONV_B(2); break;
                case 3: CONV_B(3); break;
                case 4: CONV_B(4); break;
                case 6: CONV_B(6); break;
                case 8: CONV_B(8); break;
                default: return CD_ERR_BAD_ARG;
            }
#undef CONV_B
        }
    } else {
        // input gradient: conv of the masked output gradient with transposed weights
        const float* g = cv.out < 0 ? n->ghead.as<float>() : n->t[cv.out].grad.as<float>();
        const u16* mask = (cv.out >= 0 && cv.relu) ? n->t[cv.out].act.as<u16>() : nullptr;
        const u16* wt = n->wtbf.as<u16>() + n->wtoff[&cv - &n->convs[0]];
        float* d1 = n->t[cv.in1].grad.as<float>();
        float* d2 = cv.in2 < 0 ? nullptr : n->t[cv.in2].grad.as<float>();
        const int nb = cv.cinpad / 16;
#define CONV_G(NB) hipLaunchKernelGGL((k_conv3x3<NB, 1, 1>), dim3(blocks), dim3(256), 0, s, (const u16*)nullptr, (const u16*)nullptr, cv.coutA, 0, g, mask, wt, cv.kpadT, B, H, W, (u16*)nullptr, d1, d2, cv.c1, cv.c2, 0, 1, 0)
        switch (nb) {
            case 1: CONV_G(1); break;
            case 2: CONV_G(2); break;
            case 3: CONV_G(3); break;
            case 4: CONV_G(4); break;
            case 6: CONV_G(6); break;
            case 8: CONV_G(8); break;
            default: return CD_ERR_BAD_ARG;
        }
#undef CONV_G
    }
    CD_CHECK_LAUNCH(c);
    return CD_OK;
}

int launch_wgrad(cd_net* n, const NetConv& cv, int B, int H, int W) {
    cd_context* c = n->ctx;
    const u16* x1 = cv.in1 < 0 ? n->x8.as<u16>() : n->t[cv.in1].act.as<u16>();
    const u16* x2 = cv.in2 < 0 ? nullptr : n->t[cv.in2].act.as<u16>();
    const float* g = cv.out < 0 ? n->ghead.as<float>() : n->t[cv.out].grad.as<float>();
    const u16* mask = (cv.out >= 0 && cv.relu) ? n->t[cv.out].act.as<u16>() : nullptr;
    const long long npix = (long long)B * H * W;
    const int ppw = 1024;
    dim3 grid(div_up(cv.coutA, 16), div_up(cv.K, 16), div_up(npix, 4LL * ppw));
    hipLaunchKernelGGL(k_wgrad, grid, dim3(256), 0, c->stream, x1, x2, cv.c1, cv.c2, g, mask,
                       cv.coutA, B, H, W, n->gw.as<float>() + cv.woff, cv.kpad, ppw);
    CD_CHECK_LAUNCH(c);
    return CD_OK;
}

template <typename T>
int zero(cd_context* c, T* p, long long n) {
    if (n <= 0) return CD_OK;
    hipLaunchKernelGGL(k_zero<T>, dim3(div_up(n, 256)), dim3(256), 0, c->stream, p, n);
    CD_CHECK_LAUNCH(c);
    return CD_OK;
}

// Build the layer list (independent of batch size).
void build_layers(cd_net* n) {
    const int L = n->arch.levels;
    const int* C = n->arch.channels;
    n->convs.clear();
    n->t.clear();
    auto tensor = [&](int level, int ch) {
        NetTensor t;
        t.level = level;
        t.C = ch;
        n->t.push_back(std::move(t));
        return (int)n->t.size() - 1;
    };
    auto conv = [&](int in1, int c1, int in2, int c2, int out, int cout, int coutA, int relu,
                    int level) {
        NetConv cv;
        cv.c1 = c1; cv.c2 = c2; cv.cout = cout; cv.coutA = coutA; cv.relu = relu;
        cv.K = 9 * (c1 + c2);
        cv.kpad = (cv.K + 31) / 32 * 32;
        cv.coutpad = (coutA + 15) / 16 * 16;
        cv.cinpad = (c1 + c2 + 15) / 16 * 16;
        cv.kpadT = (9 * coutA + 31) / 32 * 32;
        cv.woff = n->nparam;
        cv.nparam = (size_t)cv.coutpad * cv.kpad;
        n->nparam += cv.nparam;
        cv.level = level;
        cv.in1 = in1; cv.in2 = in2; cv.out = out;
        n->convs.push_back(cv);
    };
    const int act = n->arch.linear ? 0 : 1;  // ReLU on hidden convolutions
    std::vector<int> skip(L);
    int prev = -1, prevC = 8;  // -1: the 8-channel input tensor
    for (int l = 0; l < L; ++l) {
        int in = prev, inC = prevC;
        if (l > 0) {
            const int p = tensor(l, prevC);  // pooled
            n->t[p].level = l;
            in = p;
        }
        const int a = tensor(l, C[l]);
        conv(in, inC, -1, 0, a, C[l], C[l], act, l);
        const int s = tensor(l, C[l]);
        conv(a, C[l], -1, 0, s, C[l], C[l], act, l);
        skip[l] = s;
        prev = s;
        prevC = C[l];
    }
    int d = skip[L - 1], dC = C[L - 1];
    for (int l = L - 2; l >= 0; --l) {
        const int u = tensor(l, dC);  // upsampled
        const int b = tensor(l, C[l]);
        conv(u, dC, skip[l], C[l], b, C[l], C[l], act, l);
        const int e = tensor(l, C[l]);
        conv(b, C[l], -1, 0, e, C[l], C[l], act, l);
        d = e;
        dC = C[l];
    }
    conv(d, dC, -1, 0, -1, 1, 8, 0, 0);  // head: 1 output channel stored with stride 8
    n->wtoff.clear();
    size_t off = 0;
    for (auto& cv : n->convs) {
        n->wtoff.push_back(off);
        off += (size_t)cv.cinpad * cv.kpadT;
    }
    n->wtoff_total = off;
}

int ensure_batch(cd_net* n, int B, int H, int W) {
    cd_context* c = n->ctx;
    const int L = n->arch.levels;
    if (H % (1 << (L - 1)) || W % (1 << (L - 1))) {
        c->set_error("net input size must be divisible by 2^(levels-1)");
        return CD_ERR_BAD_ARG;
    }
    if (n->batch == B && n->bh == H && n->bw == W) return CD_OK;
    ++n->gen;
    for (auto& t : n->t) {
        const long long px = (long long)B * (H >> t.level) * (W >> t.level);
        CD_TRY(c->ensure(t.act, sizeof(u16) * px * t.C));
        CD_TRY(c->ensure(t.grad, sizeof(float) * px * t.C));
    }
    const long long px = (long long)B * H * W;
    CD_TRY(c->ensure(n->x8, sizeof(u16) * px * 8));
    CD_TRY(c->ensure(n->head, sizeof(float) * px * 8));
    CD_TRY(c->ensure(n->ghead, sizeof(float) * px * 8));
    for (DevBuf* b : {&n->yout, &n->ytmp, &n->gy, &n->ures, &n->utmp})
        CD_TRY(c->ensure(*b, sizeof(float) * px));
    CD_TRY(c->ensure(n->partials, sizeof(double) * (size_t)div_up(px, 256)));
    n->batch = B;
    n->bh = H;
    n->bw = W;
    return CD_OK;
}

int refresh_weights(cd_net* n) {
    if (!n->weights_dirty) return CD_OK;
    cd_context* c = n->ctx;
    hipLaunchKernelGGL(k_cast_bf16, dim3(div_up((long long)n->nparam, 256)), dim3(256), 0,
                       c->stream, n->w.as<float>(), n->wbf.as<u16>(), (long long)n->nparam);
    for (size_t i = 0; i < n->convs.size(); ++i) {
        const NetConv& cv = n->convs[i];
        const long long cnt = (long long)cv.cinpad * cv.kpadT;
        hipLaunchKernelGGL(k_transpose_w, dim3(div_up(cnt, 256)), dim3(256), 0, c->stream,
                           n->w.as<float>() + cv.woff, n->wtbf.as<u16>() + n->wtoff[i], cv.cout,
                           cv.coutA, cv.c1 + cv.c2, cv.kpad, cv.cinpad, cv.kpadT);
    }
    CD_CHECK_LAUNCH(c);
    n->weights_dirty = false;
    return CD_OK;
}

// forward pass from the 8-channel input already in n->x8; head output in n->head
// fused (inference): the upsampled decoder inputs are read through the
// convolutions that consume them (a pure index remap) instead of being
// materialised; the output is bit-identical.  Training needs the materialised
// tensors (weight gradients read them).  Pooling stays a separate pass: read
// through the convolution (mode 2) it costs more than it saves.
// Convolutions a and b (b reads only a's output, at a's resolution) in one
// k_conv_pair_tile launch where the LDS allows; returns false (nothing
// launched) where it does not apply.
bool launch_conv_pair(cd_net* n, const NetConv& a, const NetConv& b, int B, int H, int W,
                      int mode1, const u16* x1src, int* err) {
    *err = CD_OK;
    // up to 32 output channels: one workgroup computing all 64 channels of
    // both convs of a 32x32 or 16x16 level (4-16 workgroups) measured 3.6-4.6 us
    // slower than the two launches that split the channels over 4 workgroups
    // per tile
    const int nba = a.coutpad / 16, nbb = b.coutpad / 16;
    if (nba != nbb || !(nba == 1 || nba == 2)) return false;
    if (b.in1 != a.out || b.in2 >= 0 || b.out < 0 || a.out < 0 || b.c1 != a.coutA) return false;
    if (mode1 == 2 && a.c2 > 0) return false;
    const size_t fixed = 144 * (size_t)(a.c1 + a.c2 + 8) + 100 * (size_t)(a.coutA + 8);
    const size_t wa = (size_t)nba * 16 * (a.kpad + 8), wb = (size_t)nbb * 16 * (b.kpad + 8);
    const size_t limit = tiled_lds_limit(n->ctx);
    int early = 1;
    size_t lds = sizeof(u16) * (fixed + wa + wb);
    if (lds > limit) {
        early = 0;
        lds = sizeof(u16) * (fixed + std::max(wa, wb));
        if (lds > limit) return false;
    }
    const u16* x1 = mode1 ? x1src : (a.in1 < 0 ? n->x8.as<u16>() : n->t[a.in1].act.as<u16>());
    const u16* x2 = a.in2 < 0 ? nullptr : n->t[a.in2].act.as<u16>();
    const u16* wta = n->wbf.as<u16>() + a.woff;
    const u16* wtb = n->wbf.as<u16>() + b.woff;
    u16* y = n->t[b.out].act.as<u16>();
    const dim3 grid(div_up(W, CT), div_up(H, CT), B);
    hipStream_t s = n->ctx->stream;
#define PAIR(NB) hipLaunchKernelGGL((k_conv_pair_tile<NB, NB>), grid, dim3(256), lds, s, x1, x2, a.c1, a.c2, wta, a.kpad, a.coutA, a.relu, wtb, b.kpad, b.coutA, b.relu, y, H, W, mode1, early)
    if (nba == 1) PAIR(1);
    else PAIR(2);
#undef PAIR
    return true;
}

// the head + output smoothing kernel applies: 6 smoothing steps (the kernel's
// instantiation), one 16-channel block, no skip input, and the LDS fits
constexpr int kHeadSmoothK = 6;
bool head_smooth_ok(cd_net* n, const NetConv& hd) {
    if (n->arch.out_smooth != kHeadSmoothK || hd.coutpad != 16 || hd.in2 >= 0 || hd.in1 < 0)
        return false;
    constexpr int R = kSmoothTile + 2 * kHeadSmoothK, P = R + 2;
    const size_t lds = sizeof(u16) * ((size_t)P * P * (hd.c1 + 8) + 16 * (size_t)(hd.kpad + 8)) +
                       sizeof(float) * 2 * R * R;
    return lds <= tiled_lds_limit(n->ctx);
}

int forward(cd_net* n, int B, int H, int W, bool fused = false) {
    cd_context* c = n->ctx;
    CD_TRY(refresh_weights(n));
    n->head_pending = false;
    const int L = n->arch.levels;
    const u16* pooled_from = nullptr;  // fused: the next conv pools this tensor as it reads
    for (int idx = 0; idx < (int)n->convs.size(); ++idx) {
        const NetConv& cv0 = n->convs[idx];
        const int Hl = H >> cv0.level, Wl = W >> cv0.level;
        const int mode1 = (fused && cv0.in2 >= 0) ? 1 : pooled_from ? 2 : 0;
        const u16* x1src = mode1 == 1 ? n->t[n->convs[idx - 1].out].act.as<u16>()
                                      : mode1 == 2 ? pooled_from : nullptr;
        // fused: the first conv of a level's pair (encoder, or the decoder's
        // concatenating one) and the next one in one launch where possible
        bool paired = false;
        if (fused && n->tiled_forward && idx + 1 < (int)n->convs.size() &&
            (cv0.in2 >= 0 || (idx < 2 * L && idx % 2 == 0))) {
            int err = CD_OK;
            paired = launch_conv_pair(n, cv0, n->convs[idx + 1], B, Hl, Wl, mode1, x1src, &err);
            CD_TRY(err);
            if (paired) ++idx;  // the pair's second conv is done
        }
        const NetConv& cv = n->convs[idx];
        // fused: the head and the output smoothing run as one launch in
        // finish_output
        if (!paired && fused && n->tiled_forward && cv.out < 0 && head_smooth_ok(n, cv)) {
            n->head_pending = true;
            continue;
        }
        if (!paired) CD_TRY(launch_conv(n, cv, B, Hl, Wl, false, mode1, x1src));
        pooled_from = nullptr;
        // after the second conv of an encoder level below the bottleneck: pool
        if (idx < 2 * L && idx % 2 == 1 && cv.level < L - 1) {
            const NetConv& nx = n->convs[idx + 1];  // next level's first conv reads the pooled tensor
            if (fused) {
                pooled_from = n->t[cv.out].act.as<u16>();
            } else {
                const long long cnt = (long long)B * (Hl / 2) * (Wl / 2) * n->t[nx.in1].C;
                hipLaunchKernelGGL(k_avgpool2, dim3(div_up(cnt, 256)), dim3(256), 0, c->stream,
                                   n->t[cv.out].act.as<u16>(), n->t[nx.in1].act.as<u16>(), B, Hl,
                                   Wl, n->t[nx.in1].C);
            }
        }
        // before a decoder concat conv: upsample the previous decoder output
        if (!fused && idx + 1 < (int)n->convs.size()) {
            const NetConv& nx = n->convs[idx + 1];
            if (nx.in2 >= 0) {
                const int Hn = H >> nx.level, Wn = W >> nx.level;
                const long long cnt = (long long)B * Hn * Wn * n->t[nx.in1].C;
                hipLaunchKernelGGL(k_upsample2, dim3(div_up(cnt, 256)), dim3(256), 0, c->stream,
                                   n->t[cv.out].act.as<u16>(), n->t[nx.in1].act.as<u16>(), B, Hn,
                                   Wn, n->t[nx.in1].C);
            }
        }
        CD_CHECK_LAUNCH(c);
    }
    return CD_OK;
}

// backward pass: n->ghead holds dL/dhead; accumulates n->gw
int backward(cd_net* n, int B, int H, int W) {
    cd_context* c = n->ctx;
    for (auto& t : n->t) {
        const long long px = (long long)B * (H >> t.level) * (W >> t.level);
        CD_TRY(zero(c, t.grad.as<float>(), px * t.C));
    }
    const int L = n->arch.levels;
    for (int idx = (int)n->convs.size() - 1; idx >= 0; --idx) {
        const NetConv& cv = n->convs[idx];
        const int Hl = H >> cv.level, Wl = W >> cv.level;
        // gradient into this conv's output is complete here: its consumers
        // (later convs, pooling, upsampling) were processed before it
        CD_TRY(launch_wgrad(n, cv, B, Hl, Wl));
        if (cv.in1 >= 0) CD_TRY(launch_conv(n, cv, B, Hl, Wl, true));
        // route input gradients through upsampling / pooling
        if (cv.in2 >= 0) {
            // in1 is the upsampled previous decoder output (level+1)
            const NetConv& pv = n->convs[idx - 1];
            const long long cnt = (long long)B * (Hl / 2) * (Wl / 2) * n->t[cv.in1].C;
            hipLaunchKernelGGL(k_upsample2_bwd, dim3(div_up(cnt, 256)), dim3(256), 0, c->stream,
                               n->t[cv.in1].grad.as<float>(), n->t[pv.out].grad.as<float>(), B, Hl,
                               Wl, n->t[cv.in1].C);
        }
        if (idx < 2 * L && idx % 2 == 0 && cv.level > 0) {
            // in1 is the pooled skip of the level above
            const NetConv& pv = n->convs[idx - 1];
            const long long cnt = (long long)B * (Hl * 2) * (Wl * 2) * n->t[cv.in1].C;
            hipLaunchKernelGGL(k_avgpool2_bwd, dim3(div_up(cnt, 256)), dim3(256), 0, c->stream,
                               n->t[cv.in1].grad.as<float>(), n->t[pv.out].grad.as<float>(), B,
                               Hl * 2, Wl * 2, n->t[cv.in1].C);
        }
        CD_CHECK_LAUNCH(c);
    }
    return CD_OK;
}

// k symmetric smoothing steps on a [B][H][W] field held in *a (b is scratch);
// the result ends in *a
int smooth_steps(cd_net* n, DevBuf& a, DevBuf& b, int k, int B, int H, int W) {
    if (k <= 0) return CD_OK;
    const long long px = (long long)B * H * W;
    if (k <= kSmoothMaxK) {
        const int R = kSmoothTile + 2 * k;
        const dim3 grid(div_up(W, kSmoothTile), div_up(H, kSmoothTile), B);
        hipLaunchKernelGGL(k_smooth_tiled, grid, dim3(256), sizeof(float) * 2 * R * R,
                           n->ctx->stream, a.as<float>(), 1, b.as<float>(), B, H, W, k);
        std::swap(a.ptr, b.ptr);
        std::swap(a.bytes, b.bytes);
    } else {
        for (int i = 0; i < k; ++i) {
            hipLaunchKernelGGL(k_smooth_step, dim3(div_up(px, 256)), dim3(256), 0,
                               n->ctx->stream, a.as<float>(), b.as<float>(), B, H, W);
            std::swap(a.ptr, b.ptr);
            std::swap(a.bytes, b.bytes);
        }
    }
    CD_CHECK_LAUNCH(n->ctx);
    return CD_OK;
}

// network output y = S^out_smooth (head channel 0) into dst (n->yout when null)
int finish_output(cd_net* n, int B, int H, int W, float* dst = nullptr) {
    const long long px = (long long)B * H * W;
    float* y = dst ? dst : n->yout.as<float>();
    const int k = n->arch.out_smooth;
    if (n->head_pending) {  // the fused forward's head with the smoothing
        n->head_pending = false;
        const NetConv& hd = n->convs.back();
        constexpr int R = kSmoothTile + 2 * kHeadSmoothK, P = R + 2;
        const size_t lds = sizeof(u16) * ((size_t)P * P * (hd.c1 + 8) + 16 * (size_t)(hd.kpad + 8)) +
                           sizeof(float) * 2 * R * R;
        const dim3 grid(div_up(W, CT), div_up(H, CT), B);
        hipLaunchKernelGGL(k_head_smooth<kHeadSmoothK>, grid, dim3(256), lds, n->ctx->stream,
                           n->t[hd.in1].act.as<u16>(), hd.c1, n->wbf.as<u16>() + hd.woff, hd.kpad,
                           H, W, y);
        CD_CHECK_LAUNCH(n->ctx);
        return CD_OK;
    }
    if (k > 0 && k <= kSmoothMaxK) {
        // head channel 0 -> k steps -> y in one launch
        const int R = kSmoothTile + 2 * k;
        const dim3 grid(div_up(W, kSmoothTile), div_up(H, kSmoothTile), B);
        hipLaunchKernelGGL(k_smooth_tiled, grid, dim3(256), sizeof(float) * 2 * R * R,
                           n->ctx->stream, n->head.as<float>(), 8, y, B, H, W, k);
        CD_CHECK_LAUNCH(n->ctx);
        return CD_OK;
    }
    hipLaunchKernelGGL(k_head_out, dim3(div_up(px, 256)), dim3(256), 0, n->ctx->stream,
                       n->head.as<float>(), n->yout.as<float>(), px);
    CD_CHECK_LAUNCH(n->ctx);
    CD_TRY(smooth_steps(n, n->yout, n->ytmp, k, B, H, W));
    if (dst)
        CD_HIP(n->ctx, hipMemcpyAsync(dst, n->yout.ptr, sizeof(float) * px,
                                      hipMemcpyDeviceToDevice, n->ctx->stream));
    return CD_OK;
}

// dL/dy in n->gy -> dL/dhead (S is symmetric, so its transpose is itself)
int head_grad(cd_net* n, int B, int H, int W) {
    const long long px = (long long)B * H * W;
    CD_TRY(smooth_steps(n, n->gy, n->ytmp, n->arch.out_smooth, B, H, W));
    hipLaunchKernelGGL(k_head_grad, dim3(div_up(px, 256)), dim3(256), 0, n->ctx->stream,
                       n->gy.as<float>(), n->ghead.as<float>(), px);
    CD_CHECK_LAUNCH(n->ctx);
    return CD_OK;
}

int read_loss(cd_net* n, int nb, long long px, double* loss) {
    cd_context* c = n->ctx;
    CD_TRY(c->ensure(c->fill, sizeof(double) * 4));
    CD_TRY(cd_reduce_partials(c, n->partials.as<double>(), nb, 1, c->fill.as<double>()));
    double s = 0.0;
    CD_HIP(c, hipMemcpyAsync(&s, c->fill.ptr, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    CD_HIP(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = s / (double)px;
    return CD_OK;
}

int loss_and_head_grad(cd_net* n, const float* d_t, int B, int H, int W, double* loss,
                       bool want_grad) {
    const long long px = (long long)B * H * W;
    const int nb = div_up(px, 256);
    CD_TRY(finish_output(n, B, H, W));
    hipLaunchKernelGGL(k_loss, dim3(nb), dim3(256), 0, n->ctx->stream, n->yout.as<float>(), d_t,
                       want_grad ? n->gy.as<float>() : nullptr, px, n->partials.as<double>());
    CD_CHECK_LAUNCH(n->ctx);
    if (want_grad) CD_TRY(head_grad(n, B, H, W));
    return read_loss(n, nb, px, loss);
}

// coarse operator scale (n/pi)^2, n the larger side: makes outputs O(1)
inline float op_scale(int H, int W) {
    const double m = (double)std::max(H, W) / M_PI;
    return (float)(m * m);
}

// operator loss: u = cs Lc y - x, v = S^loss_smooth u, loss = mean(v^2);
// dL/dy = cs Lc S^loss_smooth (2 v / npix).  With arch.loss_energy the loss is
// the energy mean(y (x - u) / 2) and dL/dy = -u / npix.
int op_loss_and_head_grad(cd_net* n, const float* d_x, int B, int H, int W, double* loss,
                          bool want_grad) {
    cd_context* c = n->ctx;
    const long long px = (long long)B * H * W;
    const int nb = div_up(px, 256);
    const float cs = op_scale(H, W);
    CD_TRY(finish_output(n, B, H, W));
    hipLaunchKernelGGL(k_op_residual, dim3(nb), dim3(256), 0, c->stream, n->yout.as<float>(), d_x,
                       n->ures.as<float>(), B, H, W, cs);
    CD_CHECK_LAUNCH(c);
    if (n->arch.loss_energy) {
        hipLaunchKernelGGL(k_energy, dim3(nb), dim3(256), 0, c->stream, n->yout.as<float>(), d_x,
                           n->ures.as<float>(), n->gy.as<float>(), px, want_grad ? 1 : 0,
                           n->partials.as<double>());
        CD_CHECK_LAUNCH(c);
        if (want_grad) CD_TRY(head_grad(n, B, H, W));
        return read_loss(n, nb, px, loss);
    }
    CD_TRY(smooth_steps(n, n->ures, n->utmp, n->arch.loss_smooth, B, H, W));
    hipLaunchKernelGGL(k_sumsq, dim3(nb), dim3(256), 0, c->stream, n->ures.as<float>(), px,
                       want_grad ? 1 : 0, n->partials.as<double>());
    CD_CHECK_LAUNCH(c);
    if (want_grad) {
        CD_TRY(smooth_steps(n, n->ures, n->utmp, n->arch.loss_smooth, B, H, W));
        hipLaunchKernelGGL(k_op_grad, dim3(nb), dim3(256), 0, c->stream, n->ures.as<float>(),
                           n->gy.as<float>(), B, H, W, cs);
        CD_CHECK_LAUNCH(c);
        CD_TRY(head_grad(n, B, H, W));
    }
    return read_loss(n, nb, px, loss);
}

int prep(cd_net* n, const float* d_x, int B, int H, int W) {
    CD_TRY(ensure_batch(n, B, H, W));
    const long long px = (long long)B * H * W;
    hipLaunchKernelGGL(k_prep_input, dim3(div_up(px, 256)), dim3(256), 0, n->ctx->stream, d_x,
                       n->x8.as<u16>(), px);
    CD_CHECK_LAUNCH(n->ctx);
    return CD_OK;
}

// one Adam step (beta 0.9/0.999, eps 1e-8) from the accumulated gradient
int adam_step(cd_net* n, float lr) {
    cd_context* c = n->ctx;
    n->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
    hipLaunchKernelGGL(k_adam, dim3(div_up((long long)n->nparam, 256)), dim3(256), 0, c->stream,
                       n->w.as<float>(), n->m.as<float>(), n->v.as<float>(), n->gw.as<float>(),
                       (long long)n->nparam, lr, b1, b2, c1, c2, 1e-8f);
    CD_CHECK_LAUNCH(c);
    n->weights_dirty = true;
    return CD_OK;
}

}  // namespace

// solver-side entry: forward on the coarse residual rc (batch 1), normalised
// on the fly with stats = (sum, sum^2); output FP32 [h][w]
int cd_net_forward_coarse(cd_net* n, const float* rc, const double* stats, float* d_y, int h,
                          int w) {
    CD_TRY(ensure_batch(n, 1, h, w));
    hipLaunchKernelGGL(k_prep_normalized, dim3(div_up(h * w, 256)), dim3(256), 0, n->ctx->stream,
                       rc, stats, h * w, n->x8.as<u16>());
    CD_CHECK_LAUNCH(n->ctx);
    CD_TRY(forward(n, 1, h, w, n->fused_forward));
    return finish_output(n, 1, h, w, d_y);
}

// The network's BF16 input buffer for one h x w coarse field (sized here),
// for a producer that writes the normalised input itself (channel 0, the
// other 7 zero), and the forward on that input
int cd_net_coarse_input(cd_net* n, int h, int w, unsigned short** x8) {
    CD_TRY(ensure_batch(n, 1, h, w));
    *x8 = n->x8.as<u16>();
    return CD_OK;
}
int cd_net_forward_prepared(cd_net* n, float* d_y, int h, int w) {
    CD_TRY(ensure_batch(n, 1, h, w));
    CD_TRY(forward(n, 1, h, w, n->fused_forward));
    return finish_output(n, 1, h, w, d_y);
}

// solver-side entry: forward on one coarse field (batch 1), output FP32 [h][w]
int cd_net_forward_internal(cd_net* n, const float* d_x, float* d_y, int h, int w) {
    CD_TRY(prep(n, d_x, 1, h, w));
    CD_TRY(forward(n, 1, h, w, n->fused_forward));
    return finish_output(n, 1, h, w, d_y);
}

const cd_net_arch* cd_net_arch_of(const cd_net* n) { return &n->arch; }
unsigned long long cd_net_gen(const cd_net* n) {
    return ((unsigned long long)n->id << 32) | (n->gen & 0xffffffffULL);
}

extern "C" void cd_net_detach(cd_net* n) {
    cd_context* c = n->ctx;
    if (!c) return;
    if (c->net == n) c->net = nullptr;
    for (size_t i = 0; i < c->nets.size(); ++i)
        if (c->nets[i] == n) { c->nets.erase(c->nets.begin() + i); break; }
    n->ctx = nullptr;
}

extern "C" {

void cd_net_arch_default(cd_net_arch* a) {
    if (!a) return;
    std::memset(a, 0, sizeof(*a));
    a->levels = 4;
    a->channels[0] = 16;
    a->channels[1] = 32;
    a->channels[2] = 64;
    a->channels[3] = 64;
    a->coarsen = 4;
    a->smooth_sweeps = 4;
    a->smooth_omega = 1.2;
    a->linear = 1;
    a->out_smooth = 0;
    a->loss_smooth = 0;
    a->loss_energy = 0;
    a->mid_sweeps = 0;
}

int cd_net_create(cd_context* ctx, const cd_net_arch* arch, cd_net** out) {
    if (!ctx || !arch || !out) return CD_ERR_BAD_ARG;
    *out = nullptr;
    if (arch->levels < 2 || arch->levels > 6 || arch->coarsen < 1) return CD_ERR_BAD_ARG;
    if (arch->out_smooth < 0 || arch->out_smooth > 64 || arch->loss_smooth < 0 ||
        arch->loss_smooth > 64 || arch->loss_energy < 0 || arch->loss_energy > 1 ||
        arch->mid_sweeps < 0 || arch->mid_sweeps > 64)
        return CD_ERR_BAD_ARG;
    for (int l = 0; l < arch->levels; ++l)
        if (arch->channels[l] <= 0 || arch->channels[l] % 8 || arch->channels[l] > 128)
            return CD_ERR_BAD_ARG;
    CD_HIP(ctx, hipSetDevice(ctx->device));
    auto* n = new cd_net();
    n->ctx = ctx;
    n->arch = *arch;
    build_layers(n);
    int st = CD_OK;
    if ((st = ctx->ensure(n->w, sizeof(float) * n->nparam)) != CD_OK ||
        (st = ctx->ensure(n->m, sizeof(float) * n->nparam)) != CD_OK ||
        (st = ctx->ensure(n->v, sizeof(float) * n->nparam)) != CD_OK ||
        (st = ctx->ensure(n->gw, sizeof(float) * n->nparam)) != CD_OK ||
        (st = ctx->ensure(n->wbf, sizeof(u16) * n->nparam)) != CD_OK ||
        (st = ctx->ensure(n->wtbf, sizeof(u16) * n->wtoff_total)) != CD_OK ||
        (st = zero(ctx, n->w.as<float>(), (long long)n->nparam)) != CD_OK ||
        (st = zero(ctx, n->m.as<float>(), (long long)n->nparam)) != CD_OK ||
        (st = zero(ctx, n->v.as<float>(), (long long)n->nparam)) != CD_OK) {
        delete n;
        return st;
    }
    static std::atomic<unsigned> next_id{0};
    n->id = ++next_id;
    ctx->nets.push_back(n);
    *out = n;
    return CD_OK;
}

void cd_net_destroy(cd_net* n) {
    if (!n) return;
    if (n->ctx) {
        (void)hipStreamSynchronize(n->ctx->stream);
        cd_net_detach(n);
    }
    delete n;
}

size_t cd_net_num_params(const cd_net* n) { return n ? n->nparam : 0; }

int cd_net_set_params(cd_net* n, const float* host, size_t count) {
    if (!n || !host || count != n->nparam) return CD_ERR_BAD_ARG;
    if (!n->ctx) return CD_ERR_STATE;
    cd_context* c = n->ctx;
    CD_HIP(c, hipMemcpyAsync(n->w.ptr, host, sizeof(float) * count, hipMemcpyHostToDevice,
                             c->stream));
    CD_TRY(zero(c, n->m.as<float>(), (long long)count));
    CD_TRY(zero(c, n->v.as<float>(), (long long)count));
    n->step = 0;
    n->weights_dirty = true;
    CD_HIP(c, hipStreamSynchronize(c->stream));
    return CD_OK;
}

int cd_net_get_params(cd_net* n, float* host, size_t count) {
    if (!n || !host || count != n->nparam) return CD_ERR_BAD_ARG;
    if (!n->ctx) return CD_ERR_STATE;
    cd_context* c = n->ctx;
    CD_HIP(c, hipMemcpyAsync(host, n->w.ptr, sizeof(float) * count, hipMemcpyDeviceToHost,
                             c->stream));
    CD_HIP(c, hipStreamSynchronize(c->stream));
    return CD_OK;
}

int cd_net_init_params(cd_net* n, unsigned long long seed) {
    if (!n) return CD_ERR_BAD_ARG;
    if (!n->ctx) return CD_ERR_STATE;
    std::vector<float> w(n->nparam, 0.f);
    std::mt19937_64 rng(seed);
    for (const NetConv& cv : n->convs) {
        const bool first = cv.in1 < 0 && &cv == &n->convs[0];  // sees input channel 0 only
        const int fan_in = 9 * (first ? 1 : cv.c1 + cv.c2);
        // ReLU: He-normal with a small head.  Linear: variance 1/(3 fan_in)
        // per layer, so the initial operator is close to zero and the first
        // steps of the operator loss are well conditioned.
        const float sd = n->arch.linear ? std::sqrt(1.f / (3.f * fan_in))
                                        : std::sqrt(2.f / (float)fan_in);
        std::normal_distribution<float> nd(0.f, sd);
        // head: small for ReLU nets; zero for linear ones (the zero operator,
        // operator loss exactly 1 at the start: no amplified initial noise)
        const float scale = cv.out < 0 ? (n->arch.linear ? 0.f : 0.1f) : 1.f;
        for (int co = 0; co < cv.cout; ++co)
            for (int k = 0; k < cv.K; ++k) {
                const int ci = k % (cv.c1 + cv.c2);
                if (first && ci != 0) continue;
                w[cv.woff + (size_t)co * cv.kpad + k] = scale * nd(rng);
            }
    }
    return cd_net_set_params(n, w.data(), w.size());
}

int cd_net_forward(cd_net* n, const float* d_x, float* d_y, int batch, int h, int w) {
    if (!n || !d_x || !d_y || batch < 1) return CD_ERR_BAD_ARG;
    if (!n->ctx) return CD_ERR_STATE;
    CD_HIP(n->ctx, hipSetDevice(n->ctx->device));
    CD_TRY(prep(n, d_x, batch, h, w));
    CD_TRY(forward(n, batch, h, w, n->fused_forward));
    CD_TRY(finish_output(n, batch, h, w, d_y));
    CD_HIP(n->ctx, hipStreamSynchronize(n->ctx->stream));
    return CD_OK;
}

int cd_net_set_fused_forward(cd_net* n, int on) {
    if (!n) return CD_ERR_BAD_ARG;
    if (!n->ctx) return CD_ERR_STATE;
    n->fused_forward = on != 0;
    return CD_OK;
}

int cd_net_set_tiled_forward(cd_net* n, int on) {
    if (!n) return CD_ERR_BAD_ARG;
    if (!n->ctx) return CD_ERR_STATE;
    n->tiled_forward = on != 0;
    ++n->gen;  // captured solver graphs hold the previous kernels
    return CD_OK;
}

int cd_net_loss_grad(cd_net* n, const float* d_x, const float* d_t, int batch, int h, int w,
                     double* loss, float* host_grad, size_t count) {
    if (!n || !d_x || !d_t || batch < 1) return CD_ERR_BAD_ARG;
    if (!n->ctx) return CD_ERR_STATE;
    if (host_grad && count != n->nparam) return CD_ERR_BAD_ARG;
    cd_context* c = n->ctx;
    CD_HIP(c, hipSetDevice(c->device));
    CD_TRY(prep(n, d_x, batch, h, w));
    CD_TRY(forward(n, batch, h, w));
    CD_TRY(loss_and_head_grad(n, d_t, batch, h, w, loss, true));
    CD_TRY(zero(c, n->gw.as<float>(), (long long)n->nparam));
    CD_TRY(backward(n, batch, h, w));
    if (host_grad) {
        CD_HIP(c, hipMemcpyAsync(host_grad, n->gw.ptr, sizeof(float) * n->nparam,
                                 hipMemcpyDeviceToHost, c->stream));
    }
    CD_HIP(c, hipStreamSynchronize(c->stream));
    return CD_OK;
}

int cd_net_loss_grad_operator(cd_net* n, const float* d_x, int batch, int h, int w,
                              double* loss, float* host_grad, size_t count) {
    if (!n || !d_x || batch < 1) return CD_ERR_BAD_ARG;
    if (!n->ctx) return CD_ERR_STATE;
    if (host_grad && count != n->nparam) return CD_ERR_BAD_ARG;
    cd_context* c = n->ctx;
    CD_HIP(c, hipSetDevice(c->device));
    CD_TRY(prep(n, d_x, batch, h, w));
    CD_TRY(forward(n, batch, h, w));
    CD_TRY(op_loss_and_head_grad(n, d_x, batch, h, w, loss, true));
    CD_TRY(zero(c, n->gw.as<float>(), (long long)n->nparam));
    CD_TRY(backward(n, batch, h, w));
    if (host_grad) {
        CD_HIP(c, hipMemcpyAsync(host_grad, n->gw.ptr, sizeof(float) * n->nparam,
                                 hipMemcpyDeviceToHost, c->stream));
    }
    CD_HIP(c, hipStreamSynchronize(c->stream));
    return CD_OK;
}

int cd_net_train_step(cd_net* n, const float* d_x, const float* d_t, int batch, int h, int w,
                      float lr, double* loss) {
    CD_TRY(cd_net_loss_grad(n, d_x, d_t, batch, h, w, loss, nullptr, 0));
    return adam_step(n, lr);
}

int cd_net_train_step_operator(cd_net* n, const float* d_x, int batch, int h, int w, float lr,
                               double* loss) {
    CD_TRY(cd_net_loss_grad_operator(n, d_x, batch, h, w, loss, nullptr, 0));
    return adam_step(n, lr);
}

int cd_context_set_net(cd_context* ctx, cd_net* net) {
    if (!ctx) return CD_ERR_BAD_ARG;
    if (net && net->ctx != ctx) return CD_ERR_BAD_ARG;
    ctx->net = net;
    return CD_OK;
}

}  // extern "C"
