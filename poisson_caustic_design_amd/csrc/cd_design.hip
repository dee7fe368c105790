// Caustic_design on the device: the transport (OT) step, the height step and
// the initialisation of src/caustic_design.cpp:190-364, as gfx950 kernels over
// device-resident state.  All geometry is FP64 (single pass, bandwidth-
// trivial); the Poisson solves use cd_solve_device_internal.
#include <hip/hip_runtime.h>

#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <system_error>
#include <thread>
#include <vector>

#include "cd_geometry.h"
#include "cd_internal.h"

using namespace cdgeo;

namespace {

enum ReduceOp { R_SUM = 0, R_MIN = 1, R_MAX = 2 };

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// generic deterministic finish of per-block partials [block][ncomp]
// ---------------------------------------------------------------------------
__global__ void k_finish(const double* __restrict__ partials, int nblocks, int ncomp, int op0,
                         int op1, int op2, double* __restrict__ out) {
    __shared__ double smem[16];
    const int ops[3] = {op0, op1, op2};
    for (int c = 0; c < ncomp; ++c) {
        const int op = ops[c];
        double acc = op == R_SUM ? 0.0 : (op == R_MIN ? INFINITY : -INFINITY);
        for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
            const double v = partials[(size_t)b * ncomp + c];
            acc = op == R_SUM ? acc + v : (op == R_MIN ? fmin(acc, v) : fmax(acc, v));
        }
        if (op == R_SUM) {
            acc = cd_block_sum(acc, smem);
        } else {
            acc = op == R_MIN ? cd_wave_min(acc) : cd_wave_max_d(acc);
            if ((threadIdx.x & 63) == 0) smem[threadIdx.x / 64] = acc;
            __syncthreads();
            if (threadIdx.x == 0) {
                for (int k = 1; k < (int)blockDim.x / 64; ++k)
                    acc = op == R_MIN ? fmin(acc, smem[k]) : fmax(acc, smem[k]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) out[c] = acc;
    }
}

__device__ __forceinline__ double block_min(double v, double* smem) {
    v = cd_wave_min(v);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) smem[tid / 64] = v;
    __syncthreads();
    double r = smem[0];
    for (int k = 1; k < (int)blockDim.x / 64; ++k) r = fmin(r, smem[k]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_max(double v, double* smem) {
    v = cd_wave_max_d(v);
    const int tid = threadIdx.x;
    if ((tid & 63) == 0) smem[tid / 64] = v;
    __syncthreads();
    double r = smem[0];
    for (int k = 1; k < (int)blockDim.x / 64; ++k) r = fmax(r, smem[k]);
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// initialisation (caustic_design.cpp:334-364)
// ---------------------------------------------------------------------------
// min/max over non-NaN pixels (utils.cpp:104-115)
// image ingest (main.cpp:13-27, 95-98, 222): gray of the RGBA8 source pixel
// a nearest resize samples, src = (x * old_w / new_w, y * old_h / new_h) in
// integer arithmetic; gray = 0.299 r + 0.587 g + 0.114 b of the bytes / 255
// (same operation order, no contraction: bit-equal to the reference)
__global__ __launch_bounds__(256) void k_gray_resize(const unsigned char* __restrict__ rgba,
                                                     int w, int h, int ow, int oh,
                                                     double* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)ow * oh) return;
    const int y = (int)(i / ow), x = (int)(i - (long long)y * ow);
    const int sx = (int)((long long)x * w / ow), sy = (int)((long long)y * h / oh);
    const unsigned char* px = rgba + ((size_t)sy * w + sx) * 4;
    const double r = px[0] / 255.0, g = px[1] / 255.0, b = px[2] / 255.0;
    out[i] = (0.299 * r) + (0.587 * g) + (0.114 * b);
}

__global__ __launch_bounds__(256) void k_minmax(const double* __restrict__ img, size_t n,
                                                double* __restrict__ partials) {
    __shared__ double smem[16];
    double mn = INFINITY, mx = -INFINITY;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const double v = img[i];
        if (!isnan(v)) { mn = fmin(mn, v); mx = fmax(mx, v); }
    }
    mn = block_min(mn, smem);
    mx = block_max(mx, smem);
    if (threadIdx.x == 0) { partials[2 * blockIdx.x] = mn; partials[2 * blockIdx.x + 1] = mx; }
}

// scale to [0,1] (utils.cpp:117-126); NaN pixels become 0 as the reference's
// default-initialised output; a NaN first pixel seeds min/max with NaN there.
__global__ __launch_bounds__(256) void k_scale(const double* __restrict__ img,
                                               double* __restrict__ out, size_t n,
                                               const double* __restrict__ mm) {
    double mn = mm[0], mx = mm[1];
    if (isnan(img[0])) { mn = NAN; mx = NAN; }
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const double v = img[i];
        out[i] = isnan(v) ? 0.0 : 0.0 + (1.0 - 0.0) * (v - mn) / (mx - mn);
    }
}

// regular mesh (mesh.cpp:45-54), source heights zero
__global__ __launch_bounds__(256) void k_mesh_init(double* __restrict__ txy,
                                                   double* __restrict__ sxy,
                                                   double* __restrict__ sz, int mx, int my,
                                                   double W, double H) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mx * my) return;
    const int y = i / mx, x = i - y * mx;
    const double px = (double)x * W / (mx - 1), py = (double)y * H / (my - 1);
    txy[2 * i] = px;
    txy[2 * i + 1] = py;
    sxy[2 * i] = px;
    sxy[2 * i + 1] = py;
    sz[i] = 0.0;
}

// image integral over quad slot k of vertex v (polygon_utils.cpp:311-366):
// one thread per (vertex, adjacent-triangle slot)
// One thread per (vertex, adjacent-quad slot), 64 threads per workgroup; the
// clipping work polygons live in LDS (one column per lane) instead of
// per-thread scratch memory.
constexpr int kQuadBlock = 64;
__global__ __launch_bounds__(kQuadBlock) void k_quad_integrals(const double* __restrict__ pixels,
                                                               int nx, int ny, double W,
                                                               const double* __restrict__ xy,
                                                               int mx, int my,
                                                               double* __restrict__ quadacc) {
    __shared__ double clipbuf[4 * CLIP_MAX * kQuadBlock];
    double* buf = clipbuf + threadIdx.x;
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long V = (long long)mx * my;
    if (gid >= 6 * V) return;
    const int v = (int)(gid / 6), slot = (int)(gid - 6LL * v);
    const int y = v / mx, x = v - y * mx;
    int adj[6];
    const int n = adjacent_triangles(mx, my, x, y, adj);
    double acc = 0.0;
    if (slot < n) {
        double q[8];
        triangle_quad(mx, xy, v, adj[slot], q);
        double xmin = q[0], xmax = q[0], ymin = q[1], ymax = q[1];
        for (int i = 1; i < 4; ++i) {
            xmin = fmin(xmin, q[2 * i]); xmax = fmax(xmax, q[2 * i]);
            ymin = fmin(ymin, q[2 * i + 1]); ymax = fmax(ymax, q[2 * i + 1]);
        }
        const double px = W / (double)nx;
        const double ylo = fmax(floor(ymin / px), 0.0), yhi = fmin(ceil(ymax / px), (double)ny);
        const double xlo = fmax(floor(xmin / px), 0.0), xhi = fmin(ceil(xmax / px), (double)nx);
        for (int py = (int)ylo; py < yhi; ++py)
            for (int pxi = (int)xlo; pxi < xhi; ++pxi) {
                const double cx = ((double)pxi + 0.5) * px, cy = ((double)py + 0.5) * px;
                acc += quad_pixel_area_s<kQuadBlock>(q, cx, cy, px, buf) *
                       pixels[(size_t)py * nx + pxi];
            }
    }
    quadacc[gid] = acc;
}

// per-vertex sum of the quad integrals in slot order + block partial sums
__global__ __launch_bounds__(256) void k_cell_sums(const double* __restrict__ quadacc, int V,
                                                   double* __restrict__ out,
                                                   double* __restrict__ partials) {
    __shared__ double smem[16];
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    double a = 0.0;
    if (v < V) {
        for (int k = 0; k < 6; ++k) a += quadacc[6 * (size_t)v + k];
        out[v] = a;
    }
    const double s = cd_block_sum(v < V ? a : 0.0, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// normalise so the target areas sum to W*H (polygon_utils.cpp:382-386)
__global__ __launch_bounds__(256) void k_scale_areas(double* __restrict__ a, int V, double WH,
                                                     const double* __restrict__ total) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < V) a[v] *= WH / total[0];
}

// ---------------------------------------------------------------------------
// transport step (caustic_design.cpp:190-266)
// ---------------------------------------------------------------------------
// per-vertex partitioned dual-cell area and relative mismatch
// (mesh.cpp:121-231, polygon_utils.cpp:216-236, caustic_design.cpp:199-209)
__global__ __launch_bounds__(256) void k_cell_errors(const double* __restrict__ xy,
                                                     const double* __restrict__ tareas, int mx,
                                                     int my, double* __restrict__ area,
                                                     double* __restrict__ errors,
                                                     double* __restrict__ partials) {
    __shared__ double smem[16];
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    double folded = 0.0;
    if (v < mx * my) {
        const int y = v / mx, x = v - y * mx;
        int adj[6];
        const int n = adjacent_triangles(mx, my, x, y, adj);
        double a = 0.0;
        for (int k = 0; k < n; ++k) {
            double q[8];
            triangle_quad(mx, xy, v, adj[k], q);
            a += shoelace4(q);
        }
        area[v] = a;
        errors[v] = (tareas[v] - a) / a;
        if (!(a > 0.0)) folded = 1.0;
    }
    const double s = cd_block_sum(folded, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// Pass 1 of the rasterization (mesh.cpp:241-254) and of the inverse map
// (mesh.cpp:354-364): every triangle claims the samples it contains; a sample
// keeps the lowest claiming triangle index.
//
// Sample grids of the two scatters.  px/py are the exact sample coordinates
// the reference uses; jf/if_ map a coordinate to a fractional sample index
// (used only to narrow the candidates, every candidate is still tested).
struct RasterGrid {  // mesh.cpp:245,252: node-centred samples of the N x N grid
    int nx, ny;
    double W, H, eps;
    __device__ double px(int j) const { return (double)j * (W - eps) / (nx - 1) + 0.5 * eps; }
    __device__ double py(int i) const { return (double)i * (H - eps) / (ny - 1) + 0.5 * eps; }
    __device__ double jf(double x) const { return (x - 0.5 * eps) / ((W - eps) / (nx - 1)); }
    __device__ double if_(double y) const { return (y - 0.5 * eps) / ((H - eps) / (ny - 1)); }
    static constexpr bool kSkipRows = true;  // rows outside [ymin, ymax] +- 1e-9 are skipped
};
struct InverseGrid {  // mesh.cpp:355-362: source vertices pulled eps inside the domain
    int nx, ny;
    double W, H, eps;
    __device__ double px(int j) const {
        return eps + ((double)j * W / (nx - 1)) * ((W - 2 * eps) / W);
    }
    __device__ double py(int i) const {
        return eps + ((double)i * H / (ny - 1)) * ((H - 2 * eps) / H);
    }
    __device__ double jf(double x) const { return (x - eps) / ((W - 2 * eps) / W) / (W / (nx - 1)); }
    __device__ double if_(double y) const { return (y - eps) / ((H - 2 * eps) / H) / (H / (ny - 1)); }
    static constexpr bool kSkipRows = false;
};

struct Tri2 {
    double ax, ay, bx, by, cx, cy, xmin, xmax, ymin, ymax;
};

__device__ __forceinline__ bool load_tri(const double* __restrict__ xy, int mx, int t, Tri2& T) {
    int a, b, c;
    tri_vertices(mx, t, a, b, c);
    T.ax = xy[2 * a]; T.ay = xy[2 * a + 1];
    T.bx = xy[2 * b]; T.by = xy[2 * b + 1];
    T.cx = xy[2 * c]; T.cy = xy[2 * c + 1];
    T.xmin = fmin(T.ax, fmin(T.bx, T.cx)); T.xmax = fmax(T.ax, fmax(T.bx, T.cx));
    T.ymin = fmin(T.ay, fmin(T.by, T.cy)); T.ymax = fmax(T.ay, fmax(T.by, T.cy));
    return isfinite(T.xmin) && isfinite(T.xmax) && isfinite(T.ymin) && isfinite(T.ymax);
}

// x-extent of the triangle on the horizontal line y clamped into [ymin, ymax]
__device__ __forceinline__ void row_extent(const Tri2& T, double y, double& xl, double& xr) {
    const double yc = fmin(fmax(y, T.ymin), T.ymax);
    xl = INFINITY;
    xr = -INFINITY;
    const double P[3][2] = {{T.ax, T.ay}, {T.bx, T.by}, {T.cx, T.cy}};
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const double px = P[e][0], py = P[e][1], qx = P[(e + 1) % 3][0], qy = P[(e + 1) % 3][1];
        if (yc < fmin(py, qy) || yc > fmax(py, qy)) continue;
        if (py == qy) {
            xl = fmin(xl, fmin(px, qx));
            xr = fmax(xr, fmax(px, qx));
        } else {
            const double s = fmin(fmax((yc - py) / (qy - py), 0.0), 1.0);
            const double x = px + s * (qx - px);
            xl = fmin(xl, x);
            xr = fmax(xr, x);
        }
    }
}

// candidate rows [i0, i1] and, per row, columns [j0, j1] (one sample of
// margin each side); every candidate is tested with the reference's
// barycentric predicate, so the owners are those of a full bounding-box scan
template <typename G>
__device__ __forceinline__ void tri_rows(const G& g, const Tri2& T, int& i0, int& i1) {
    // clamped in floating point first: a run-away vertex cannot overflow the
    // integer conversion
    i0 = (int)fmax(0.0, floor(g.if_(T.ymin)) - 1.0);
    i1 = (int)fmin((double)(g.ny - 1), ceil(g.if_(T.ymax)) + 1.0);
}

template <typename G>
__device__ __forceinline__ void scan_row(const G& g, const Tri2& T, int key, int i,
                                         int* __restrict__ owner) {
    const double py = g.py(i);
    if (G::kSkipRows && (py < T.ymin - 1e-9 || py > T.ymax + 1e-9)) return;
    double xl, xr;
    row_extent(T, py, xl, xr);
    const int j0 = (int)fmax(0.0, floor(g.jf(xl)) - 1.0);
    const int j1 = (int)fmin((double)(g.nx - 1), ceil(g.jf(xr)) + 1.0);
    for (int j = j0; j <= j1; ++j) {
        double u, v, w;
        barycentric(T.ax, T.ay, T.bx, T.by, T.cx, T.cy, g.px(j), py, u, v, w);
        if (inside(u, v)) atomicMin(&owner[(size_t)i * g.nx + j], key);
    }
}

// one thread per triangle; triangles spanning more than kBigRows candidate
// rows go to a list that one workgroup per triangle scans row-parallel.  A
// sample keeps the lowest key of the triangles that hold it: the triangle
// index, or with `rank` the triangle's position in the reference's BVH
// first-hit order (the gathers map it back through that order).
constexpr int kBigRows = 16;
constexpr int kBigBlocks = 1024;  // workgroups striding over the big-triangle list

template <typename G>
__global__ __launch_bounds__(256) void k_scatter_small(const double* __restrict__ xy, int mx,
                                                       int my, G g, int* __restrict__ owner,
                                                       const int* __restrict__ rank,
                                                       int* __restrict__ big, int* __restrict__ nbig) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * (mx - 1) * (my - 1)) return;
    Tri2 T;
    if (!load_tri(xy, mx, t, T)) return;
    int i0, i1;
    tri_rows(g, T, i0, i1);
    if (i1 - i0 + 1 > kBigRows) {
        big[atomicAdd(nbig, 1)] = t;
        return;
    }
    const int key = rank ? rank[t] : t;
    for (int i = i0; i <= i1; ++i) scan_row(g, T, key, i, owner);
}

template <typename G>
__global__ __launch_bounds__(256) void k_scatter_big(const double* __restrict__ xy, int mx, G g,
                                                     int* __restrict__ owner,
                                                     const int* __restrict__ rank,
                                                     const int* __restrict__ big,
                                                     const int* __restrict__ nbig) {
    const int n = *nbig;
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const int t = big[k];
        Tri2 T;
        if (!load_tri(xy, mx, t, T)) continue;
        int i0, i1;
        tri_rows(g, T, i0, i1);
        const int key = rank ? rank[t] : t;
        for (int i = i0 + (int)threadIdx.x; i <= i1; i += blockDim.x) scan_row(g, T, key, i, owner);
    }
}

// Rasterization, pass 2: piecewise-linear interpolation of per-vertex values
// in the owning triangle (mesh.cpp:264-274).  partials: (sum, misses).
__global__ __launch_bounds__(256) void k_raster_gather(const double* __restrict__ xy, int mx,
                                                       const double* __restrict__ vals0,
                                                       const double* __restrict__ vals1,
                                                       int nx, int ny, double W, double H,
                                                       const int* __restrict__ owner,
                                                       const int* __restrict__ order,
                                                       double* __restrict__ out0,
                                                       double* __restrict__ out1,
                                                       double* __restrict__ partials) {
    __shared__ double smem[16];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double sum = 0.0, miss = 0.0;
    if (s < (long long)nx * ny) {
        const int i = (int)(s / nx), j = (int)(s - (long long)i * nx);
        int t = owner[s];
        if (t == INT_MAX) {
            out0[s] = NAN;
            if (out1) out1[s] = NAN;
            miss = 1.0;
        } else {
            if (order) t = order[t];
            const double eps = 1e-8;
            const double py = (double)i * (H - eps) / (ny - 1) + 0.5 * eps;
            const double px = (double)j * (W - eps) / (nx - 1) + 0.5 * eps;
            int a, b, c;
            tri_vertices(mx, t, a, b, c);
            double u, v, w;
            barycentric(xy[2 * a], xy[2 * a + 1], xy[2 * b], xy[2 * b + 1], xy[2 * c],
                        xy[2 * c + 1], px, py, u, v, w);
            const double r0 = vals0[a] * u + vals0[b] * v + vals0[c] * w;
            out0[s] = r0;
            sum = r0;
            if (out1) out1[s] = vals1[a] * u + vals1[b] * v + vals1[c] * w;
        }
    }
    const double a0 = cd_block_sum(sum, smem);
    const double a1 = cd_block_sum(miss, smem);
    if (threadIdx.x == 0) { partials[2 * blockIdx.x] = a0; partials[2 * blockIdx.x + 1] = a1; }
}

// subtractAverage (utils.cpp:60-86) with the sum already reduced
__global__ __launch_bounds__(256) void k_subtract_mean(double* __restrict__ a, size_t n,
                                                       const double* __restrict__ sum,
                                                       double count) {
    const double mean = sum[0] / count;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const double v = a[i];
        if (!isnan(v)) a[i] = v - mean;
    }
}

// Gradient of phi at a cell by clamped central differences (utils.cpp:3-20).
struct GradX {
    const double* p; int w, h;
    __device__ double operator()(int x, int y) const {
        const int xp = min(x + 1, w - 1), xm = max(x - 1, 0);
        return (p[(size_t)y * w + xp] - p[(size_t)y * w + xm]) / 2.0;
    }
};
struct GradY {
    const double* p; int w, h;
    __device__ double operator()(int x, int y) const {
        const int yp = min(y + 1, h - 1), ym = max(y - 1, 0);
        return (p[(size_t)yp * w + x] - p[(size_t)ym * w + x]) / 2.0;
    }
};
struct Grid {
    const double* p; int w;
    __device__ double operator()(int x, int y) const { return p[(size_t)y * w + x]; }
};

// grad(phi) at the cell-centred vertex lookup + boundary-constrained step
// (caustic_design.cpp:224-265, mesh.cpp:485-532).  partials: max displacement.
__global__ __launch_bounds__(256) void k_vertex_step(const double* __restrict__ phi, int nx,
                                                     int ny, double W, double H, int mx, int my,
                                                     double step, double* __restrict__ xy,
                                                     double* __restrict__ vgrad,
                                                     double* __restrict__ partials) {
    __shared__ double smem[16];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double dist = 0.0;
    if (i < mx * my) {
        const int y = i / mx, x = i - y * mx;
        const double px = xy[2 * i], py = xy[2 * i + 1];
        const double fx = (px / W) * nx - 0.5, fy = (py / H) * ny - 0.5;
        double vx = bilinear(GradX{phi, nx, ny}, nx, ny, fx, fy);
        double vy = bilinear(GradY{phi, nx, ny}, nx, ny, fx, fy);
        vgrad[2 * i] = vx;
        vgrad[2 * i + 1] = vy;
        if (x == 0 || x == mx - 1) vx = 0.0;
        if (y == 0 || y == my - 1) vy = 0.0;
        const double dt = W / mx;
        const double nxp = px + vx * dt * step, nyp = py + vy * dt * step;
        const double dx = px - nxp, dy = py - nyp;
        dist = sqrt(dx * dx + dy * dy + 0.0 * 0.0);
        xy[2 * i] = nxp;
        xy[2 * i + 1] = nyp;
    }
    const double m = block_max(dist, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

// Folded triangles of a deformed mesh: the structured triangulation
// (mesh.cpp:57-63) has every triangle counter-clockwise in the undeformed
// mesh, (b-a) x (c-a) = +dx*dy > 0; a fold reverses (or collapses) at least
// one, so its orientation is <= 0 (or not finite).  The reference neither
// detects nor reports folds; it keeps stepping.  partials: count per block.
__global__ __launch_bounds__(256) void k_fold_count(const double* __restrict__ xy, int mx,
                                                    int my, double* __restrict__ partials) {
    __shared__ double smem[16];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    double folded = 0.0;
    if (t < 2 * (mx - 1) * (my - 1)) {
        int a, b, c;
        tri_vertices(mx, t, a, b, c);
        const double ax = xy[2 * a], ay = xy[2 * a + 1];
        const double cr = (xy[2 * b] - ax) * (xy[2 * c + 1] - ay) -
                          (xy[2 * b + 1] - ay) * (xy[2 * c] - ax);
        if (!(cr > 0.0)) folded = 1.0;
    }
    const double s = cd_block_sum(folded, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// height step (caustic_design.cpp:269-332)
// ---------------------------------------------------------------------------
// Inverse map, pass 2 + refractive normals (mesh.cpp:365-402, 677-722).
// partials: misses.
__global__ __launch_bounds__(256) void k_inverse_normals(const double* __restrict__ xy, int mx,
                                                         int my, double W, double H,
                                                         const int* __restrict__ owner,
                                                         const int* __restrict__ order,
                                                         const double* __restrict__ sz,
                                                         double focal, double ior,
                                                         double* __restrict__ inv,
                                                         double* __restrict__ nrm_x,
                                                         double* __restrict__ nrm_y,
                                                         double* __restrict__ partials) {
    __shared__ double smem[16];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double miss = 0.0;
    if (i < mx * my) {
        const int y = i / mx, x = i - y * mx;
        int t = owner[i];
        const double sx = (double)x * W / (mx - 1), sy = (double)y * H / (my - 1);
        if (t == INT_MAX) {
            miss = 1.0;
            inv[2 * i] = NAN; inv[2 * i + 1] = NAN;
            nrm_x[i] = NAN; nrm_y[i] = NAN;
        } else {
            if (order) t = order[t];
            const double eps = 1e-8;
            const double px = eps + sx * ((W - 2 * eps) / W);
            const double py = eps + sy * ((H - 2 * eps) / H);
            int a, b, c;
            tri_vertices(mx, t, a, b, c);
            double u, v, w;
            barycentric(xy[2 * a], xy[2 * a + 1], xy[2 * b], xy[2 * b + 1], xy[2 * c],
                        xy[2 * c + 1], px, py, u, v, w);
            const int ay_ = a / mx, ax_ = a - ay_ * mx;
            const int by_ = b / mx, bx_ = b - by_ * mx;
            const int cy_ = c / mx, cx_ = c - cy_ * mx;
            const double sax = (double)ax_ * W / (mx - 1), say = (double)ay_ * H / (my - 1);
            const double sbx = (double)bx_ * W / (mx - 1), sby = (double)by_ * H / (my - 1);
            const double scx = (double)cx_ * W / (mx - 1), scy = (double)cy_ * H / (my - 1);
            double ix = sax * u + sbx * v + scx * w;
            double iy = say * u + sby * v + scy * w;
            if (x == 0) ix = 0.0;
            if (x == mx - 1) ix = W;
            if (y == 0) iy = 0.0;
            if (y == my - 1) iy = H;
            inv[2 * i] = ix;
            inv[2 * i + 1] = iy;
            // transmitted direction, normalised (utils.cpp:370-384)
            double t0 = ix - sx, t1 = iy - sy, t2 = 0 - sz[i] + focal;
            const double len = sqrt(0.0 + t0 * t0 + t1 * t1 + t2 * t2);
            t0 /= len; t1 /= len; t2 /= len;
            const double n0 = t0 - 0.0 * ior, n1 = t1 - 0.0 * ior, n2 = t2 - 1.0 * ior;
            nrm_x[i] = n0 / n2;
            nrm_y[i] = n1 / n2;
        }
    }
    const double m = cd_block_sum(miss, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

// divergence with a zero outer ring (utils.cpp:22-39); partials: sum
__global__ __launch_bounds__(256) void k_divergence(const double* __restrict__ nx_,
                                                    const double* __restrict__ ny_, int w, int h,
                                                    double* __restrict__ out,
                                                    double* __restrict__ partials) {
    __shared__ double smem[16];
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (s < (long long)w * h) {
        const int y = (int)(s / w), x = (int)(s - (long long)y * w);
        if (x == 0 || x == w - 1 || y == 0 || y == h - 1) {
            v = 0.0;
        } else {
            const double dx = (nx_[s + 1] - nx_[s - 1]) / 2.0;
            const double dy = (ny_[s + w] - ny_[s - w]) / 2.0;
            v = dx + dy;
        }
        out[s] = v;
    }
    const double a = cd_block_sum(v, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = a;
}

// vertex heights from h (caustic_design.cpp:322-329); partials: min
__global__ __launch_bounds__(256) void k_height_sample(const double* __restrict__ hg, int nx,
                                                       int ny, double W, double H, int mx, int my,
                                                       double* __restrict__ hv,
                                                       double* __restrict__ partials) {
    __shared__ double smem[16];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double mn = INFINITY;
    if (i < mx * my) {
        const int y = i / mx, x = i - y * mx;
        const double sx = (double)x * W / (mx - 1), sy = (double)y * H / (my - 1);
        const double v = bilinear(Grid{hg, nx}, nx, ny, (sx / W) * nx - 0.5, (sy / H) * ny - 0.5);
        hv[i] = v;
        mn = v;
    }
    const double m = block_min(mn, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = m;
}

// shift by min(0, min h) and store (mesh.cpp:724-742); partials: update sum
__global__ __launch_bounds__(256) void k_height_apply(const double* __restrict__ hv, int V,
                                                      const double* __restrict__ mn,
                                                      double* __restrict__ sz,
                                                      double* __restrict__ partials) {
    __shared__ double smem[16];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double u = 0.0;
    if (i < V) {
        const double shift = fmin(0.0, mn[0]);
        const double hh = hv[i] - shift;
        u = (hh - sz[i]) * (hh - sz[i]);
        sz[i] = hh;
    }
    const double a = cd_block_sum(u, smem);
    if (threadIdx.x == 0) partials[blockIdx.x] = a;
}

// Laplacian smoothing of the target mesh (mesh.cpp:534-606): neighbour
// average from the unmodified points (interior: x and y over 4 neighbours;
// edges: the free coordinate over 3 in the reference's summation order;
// corners kept), then p = avg * f + p * (1 - f) for every vertex
__global__ __launch_bounds__(256) void k_laplacian_smooth(const double* __restrict__ src,
                                                          double* __restrict__ dst, int mx,
                                                          int my, double f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= mx * my) return;
    const int y = i / mx, x = i - y * mx;
    auto X = [&](int xx, int yy) { return src[2 * (yy * mx + xx)]; };
    auto Y = [&](int xx, int yy) { return src[2 * (yy * mx + xx) + 1]; };
    double ax = src[2 * i], ay = src[2 * i + 1];
    const bool left = x == 0, right = x == mx - 1, top = y == 0, bottom = y == my - 1;
    if ((left || right) && (top || bottom)) {
        // corner: unchanged
    } else if (left) {
        ay = ((0.0 + Y(x + 1, y)) + Y(x, y + 1)) + Y(x, y - 1);
        ay /= 3.0;
    } else if (right) {
        ay = ((0.0 + Y(x - 1, y)) + Y(x, y - 1)) + Y(x, y + 1);
        ay /= 3.0;
    } else if (top) {
        ax = ((0.0 + X(x - 1, y)) + X(x + 1, y)) + X(x, y + 1);
        ax /= 3.0;
    } else if (bottom) {
        ax = ((0.0 + X(x - 1, y)) + X(x, y - 1)) + X(x + 1, y);
        ax /= 3.0;
    } else {
        ax = (((0.0 + X(x - 1, y)) + X(x + 1, y)) + X(x, y - 1)) + X(x, y + 1);
        ay = (((0.0 + Y(x - 1, y)) + Y(x + 1, y)) + Y(x, y - 1)) + Y(x, y + 1);
        ax /= 4.0;
        ay /= 4.0;
    }
    dst[2 * i] = ax * f + src[2 * i] * (1.0 - f);
    dst[2 * i + 1] = ay * f + src[2 * i + 1] * (1.0 - f);
}

// gradient of a grid by central differences, clamped at the borders
// (calculate_gradient, utils.cpp:3-20): out[0..G) = d/dx, out[G..2G) = d/dy
__global__ __launch_bounds__(256) void k_gradient_grid(const double* __restrict__ g, int w,
                                                       int h, double* __restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w) return;
    const size_t row = (size_t)y * w;
    out[row + x] = (g[row + min(x + 1, w - 1)] - g[row + max(x - 1, 0)]) / 2.0;
    out[(size_t)w * h + row + x] =
        (g[(size_t)min(y + 1, h - 1) * w + x] - g[(size_t)max(y - 1, 0) * w + x]) / 2.0;
}

// Partitioned dual cells (Mesh::build_target_partitioned_dual_cells,
// mesh.cpp:201-231): for vertex v, slot k < 6 holds its k-th adjacent
// triangle in ascending index (tri[6v+k], -1 past the last) and that
// triangle's quad (v_i, mid(ij), centroid(ijk), mid(ik)) as 4 points of
// {x, y, 0} (mesh.cpp:174-199; target points have z = 0).
__global__ __launch_bounds__(256) void k_dual_cells(const double* __restrict__ xy, int mx, int my,
                                                    int* __restrict__ tri,
                                                    double* __restrict__ quads) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= mx * my) return;
    const int y = v / mx, x = v - y * mx;
    int adj[6];
    const int n = adjacent_triangles(mx, my, x, y, adj);
    for (int k = 0; k < 6; ++k) {
        double* o = quads + ((size_t)v * 6 + k) * 12;
        if (k < n) {
            double q[8];
            triangle_quad(mx, xy, v, adj[k], q);
            for (int c = 0; c < 4; ++c) {
                o[3 * c] = q[2 * c];
                o[3 * c + 1] = q[2 * c + 1];
                o[3 * c + 2] = 0.0;
            }
        } else {
            for (int c = 0; c < 12; ++c) o[c] = 0.0;
        }
        tri[(size_t)v * 6 + k] = k < n ? adj[k] : -1;
    }
}

__global__ void k_points3(const double* __restrict__ xy, const double* __restrict__ z, int V,
                          double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= V) return;
    out[3 * i] = xy[2 * i];
    out[3 * i + 1] = xy[2 * i + 1];
    out[3 * i + 2] = z ? z[i] : 0.0;
}

}  // namespace

// ===========================================================================
// host side
// ===========================================================================
struct cd_design {
    cd_context* ctx = nullptr;
    cd_design_params p;
    int mx = 0, my = 0, nx = 0, ny = 0;
    size_t V = 0, G = 0;
    int T = 0;
    bool initialized = false;
    bool miss_recovery = false;  // smooth and continue on a raster miss
    int recoveries = 0;          // misses recovered since the last init
    DevBuf smooth_tmp;           // V x 2 copy of the target points
    DevBuf txy, sxy, sz, tareas, errors, area, vgrad, inv, nrm_x, nrm_y, hv, quadacc;
    DevBuf big, nbig;  // scatter: triangles spanning many sample rows, and their count
    DevBuf pixels, image, raster, phi, h, div, normx, normy, owner, owner_src, owner_v;
    DevBuf partials, scalars, tmp3;
    DevBuf grad, cells_t, cells_q;  // on demand: gradient grid, dual-cell export
    cd_solve_stats last_phi{}, last_h{};
    long long folded_cells = 0;      // dual cells of area <= 0 at the last transport step's start
    long long folded_triangles = 0;  // inverted triangles after the last transport step
    bool fold_fail = false;          // a transport step that folds the mesh fails (CD_ERR_FOLDED)
    int first_hit = 0;               // 1: samples take the reference BVH's first hit
    DevBuf bvh_rank, order_t, order_s;   // first-hit position per triangle; orders of both meshes
    std::vector<double> h_xy;            // host copies for the order's build
    std::vector<int> h_order, h_rank;
};

namespace {

int alloc_design(cd_design* d) {
    cd_context* c = d->ctx;
    const size_t V = d->V, G = d->G;
    CD_TRY(c->ensure(d->txy, 16 * V));
    CD_TRY(c->ensure(d->sxy, 16 * V));
    CD_TRY(c->ensure(d->sz, 8 * V));
    CD_TRY(c->ensure(d->tareas, 8 * V));
    CD_TRY(c->ensure(d->errors, 8 * V));
    CD_TRY(c->ensure(d->area, 8 * V));
    CD_TRY(c->ensure(d->vgrad, 16 * V));
    CD_TRY(c->ensure(d->inv, 16 * V));
    CD_TRY(c->ensure(d->nrm_x, 8 * V));
    CD_TRY(c->ensure(d->nrm_y, 8 * V));
    CD_TRY(c->ensure(d->hv, 8 * V));
    CD_TRY(c->ensure(d->quadacc, 6 * 8 * V));
    CD_TRY(c->ensure(d->owner_v, 4 * V));
    CD_TRY(c->ensure(d->big, 4 * (size_t)d->T));
    CD_TRY(c->ensure(d->nbig, sizeof(int)));
    CD_TRY(c->ensure(d->pixels, 8 * G));
    CD_TRY(c->ensure(d->image, 8 * G));
    CD_TRY(c->ensure(d->raster, 8 * G));
    CD_TRY(c->ensure(d->phi, 8 * G));
    CD_TRY(c->ensure(d->h, 8 * G));
    CD_TRY(c->ensure(d->div, 8 * G));
    CD_TRY(c->ensure(d->normx, 8 * G));
    CD_TRY(c->ensure(d->normy, 8 * G));
    CD_TRY(c->ensure(d->owner, 4 * G));
    CD_TRY(c->ensure(d->owner_src, 4 * G));
    const size_t nb = (std::max(G, 6 * V) + CD_BLOCK - 1) / CD_BLOCK;
    CD_TRY(c->ensure(d->partials, 8 * 3 * nb));
    CD_TRY(c->ensure(d->scalars, 8 * 16));
    CD_TRY(c->ensure(d->tmp3, 24 * V));
    return CD_OK;
}

// folded-triangle count of the target mesh into scalars[slot]
int count_folds(cd_design* d, int slot) {
    const int tb = div_up(d->T, CD_BLOCK);
    hipLaunchKernelGGL(k_fold_count, dim3(tb), dim3(CD_BLOCK), 0, d->ctx->stream,
                       d->txy.as<double>(), d->mx, d->my, d->partials.as<double>());
    CD_CHECK_LAUNCH(d->ctx);
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(1024), 0, d->ctx->stream, d->partials.as<double>(),
                       tb, 1, (int)R_SUM, (int)R_SUM, (int)R_SUM, d->scalars.as<double>() + slot);
    CD_CHECK_LAUNCH(d->ctx);
    return CD_OK;
}

int finish(cd_design* d, int nblocks, int ncomp, int op0, int op1, int op2, int slot) {
    hipLaunchKernelGGL(k_finish, dim3(1), dim3(1024), 0, d->ctx->stream, d->partials.as<double>(),
                       nblocks, ncomp, op0, op1, op2, d->scalars.as<double>() + slot);
    CD_CHECK_LAUNCH(d->ctx);
    return CD_OK;
}

int read_scalars(cd_design* d, double* out, int n) {
    cd_context* c = d->ctx;
    CD_HIP(c, hipMemcpyAsync(c->host_status, d->scalars.ptr, sizeof(double) * n,
                             hipMemcpyDeviceToHost, c->stream));
    CD_HIP(c, hipStreamSynchronize(c->stream));
    CD_TRY(c->resolve_timers());
    std::memcpy(out, c->host_status, sizeof(double) * n);
    return CD_OK;
}

// rasterize per-vertex values (one or two fields) over triangulation xy
// first-hit order of mesh xy (reference BVH, bvh.cpp:20-155) into `order`,
// positions into d->bvh_rank; both null in the default lowest-index mode
int first_hit_order(cd_design* d, const double* xy, DevBuf& order, const int** rank_out,
                    const int** order_out) {
    *rank_out = nullptr;
    *order_out = nullptr;
    if (!d->first_hit) return CD_OK;
    cd_context* c = d->ctx;
    const size_t V = d->V, T = d->T;
    CD_TRY(c->ensure(order, 4 * T));
    CD_TRY(c->ensure(d->bvh_rank, 4 * T));
    d->h_xy.resize(2 * V);
    d->h_order.resize(T);
    d->h_rank.resize(T);
    CD_HIP(c, hipMemcpyAsync(d->h_xy.data(), xy, 16 * V, hipMemcpyDeviceToHost, c->stream));
    CD_HIP(c, hipStreamSynchronize(c->stream));
    CD_TRY(cd_bvh_first_hit_order(d->h_xy.data(), d->mx, d->my, d->h_order.data()));
    for (size_t k = 0; k < T; ++k) d->h_rank[d->h_order[k]] = (int)k;
    CD_HIP(c, hipMemcpyAsync(order.ptr, d->h_order.data(), 4 * T, hipMemcpyHostToDevice,
                             c->stream));
    CD_HIP(c, hipMemcpyAsync(d->bvh_rank.ptr, d->h_rank.data(), 4 * T, hipMemcpyHostToDevice,
                             c->stream));
    CD_HIP(c, hipStreamSynchronize(c->stream));  // the host vectors are reused
    *rank_out = d->bvh_rank.as<int>();
    *order_out = order.as<int>();
    return CD_OK;
}

int raster(cd_design* d, const double* xy, int* owner, bool rebuild_owner, const double* v0,
           const double* v1, double* out0, double* out1, double* misses_out) {
    cd_context* c = d->ctx;
    hipStream_t s = c->stream;
    // owners of the target mesh are rebuilt here; those of the source mesh
    // were built at initialisation, in the order kept in order_s
    const int* rank = nullptr;
    const int* order = nullptr;
    if (rebuild_owner) CD_TRY(first_hit_order(d, xy, d->order_t, &rank, &order));
    else if (d->first_hit) order = d->order_s.as<int>();
    const int tb = div_up(d->T, CD_BLOCK);
    const int gb = div_up((long long)d->G, CD_BLOCK);
    CD_TRY(c->timer_begin());
    if (rebuild_owner) {
        CD_TRY(cd_fill_i32(c, owner, d->G, INT_MAX));
        CD_TRY(cd_fill_i32(c, d->nbig.as<int>(), 1, 0));
        const RasterGrid g{d->nx, d->ny, d->p.width, d->p.height, 1e-8};
        hipLaunchKernelGGL(k_scatter_small<RasterGrid>, dim3(tb), dim3(CD_BLOCK), 0, s, xy, d->mx,
                           d->my, g, owner, rank, d->big.as<int>(), d->nbig.as<int>());
        hipLaunchKernelGGL(k_scatter_big<RasterGrid>, dim3(kBigBlocks), dim3(CD_BLOCK), 0, s, xy,
                           d->mx, g, owner, rank, d->big.as<int>(), d->nbig.as<int>());
    }
    hipLaunchKernelGGL(k_raster_gather, dim3(gb), dim3(CD_BLOCK), 0, s, xy, d->mx, v0, v1, d->nx,
                       d->ny, d->p.width, d->p.height, owner, order, out0, out1,
                       d->partials.as<double>());
    CD_CHECK_LAUNCH(c);
    CD_TRY(finish(d, gb, 2, R_SUM, R_SUM, R_SUM, 0));  // scalars[0]=sum, [1]=misses
    CD_TRY(c->timer_end(CD_KERNEL_RASTER, rebuild_owner ? 3 : 1,
                        (rebuild_owner ? 8.0 : 0.0) * d->G + (out1 ? 20.0 : 12.0) * d->G));
    if (misses_out) {
        double sc[2];
        CD_TRY(read_scalars(d, sc, 2));
        *misses_out = sc[1];
    }
    return CD_OK;
}

// owners of the static source triangulation (mesh.cpp:291-345 locates
// samples in the undeformed source mesh, whose xy never change)
int source_owners(cd_design* d) {
    cd_context* c = d->ctx;
    hipStream_t s = c->stream;
    const int* rank = nullptr;
    const int* order = nullptr;
    CD_TRY(first_hit_order(d, d->sxy.as<double>(), d->order_s, &rank, &order));
    CD_TRY(cd_fill_i32(c, d->owner_src.as<int>(), d->G, INT_MAX));
    CD_TRY(cd_fill_i32(c, d->nbig.as<int>(), 1, 0));
    const RasterGrid g{d->nx, d->ny, d->p.width, d->p.height, 1e-8};
    hipLaunchKernelGGL(k_scatter_small<RasterGrid>, dim3(div_up(d->T, CD_BLOCK)), dim3(CD_BLOCK),
                       0, s, d->sxy.as<double>(), d->mx, d->my, g, d->owner_src.as<int>(), rank,
                       d->big.as<int>(), d->nbig.as<int>());
    hipLaunchKernelGGL(k_scatter_big<RasterGrid>, dim3(kBigBlocks), dim3(CD_BLOCK), 0, s,
                       d->sxy.as<double>(), d->mx, g, d->owner_src.as<int>(), rank,
                       d->big.as<int>(), d->nbig.as<int>());
    CD_CHECK_LAUNCH(c);
    return CD_OK;
}

int subtract_mean(cd_design* d, double* a, int slot_sum) {
    hipLaunchKernelGGL(k_subtract_mean, dim3(std::min(div_up((long long)d->G, CD_BLOCK), 65536)),
                       dim3(CD_BLOCK), 0, d->ctx->stream, a, d->G,
                       d->scalars.as<double>() + slot_sum, (double)d->G);
    CD_CHECK_LAUNCH(d->ctx);
    return CD_OK;
}

int design_init_from_device_image(cd_design* d, const double* img) {
    cd_context* c = d->ctx;
    hipStream_t s = c->stream;
    const int V = (int)d->V;
    // min-max normalisation (caustic_design.cpp:335, utils.cpp:88-129)
    const int nbm = std::min(div_up((long long)d->G, CD_BLOCK), 4096);
    hipLaunchKernelGGL(k_minmax, dim3(nbm), dim3(CD_BLOCK), 0, s, img, d->G,
                       d->partials.as<double>());
    CD_CHECK_LAUNCH(c);
    CD_TRY(finish(d, nbm, 2, R_MIN, R_MAX, R_SUM, 0));
    hipLaunchKernelGGL(k_scale, dim3(std::min(div_up((long long)d->G, CD_BLOCK), 65536)),
                       dim3(CD_BLOCK), 0, s, img, d->pixels.as<double>(), d->G,
                       d->scalars.as<double>());
    // mesh (caustic_design.cpp:339; mesh.cpp:3-27)
    hipLaunchKernelGGL(k_mesh_init, dim3(div_up(V, CD_BLOCK)), dim3(CD_BLOCK), 0, s,
                       d->txy.as<double>(), d->sxy.as<double>(), d->sz.as<double>(), d->mx,
                       d->my, d->p.width, d->p.height);
    // target areas over the initial dual cells (caustic_design.cpp:341-348)
    hipLaunchKernelGGL(k_quad_integrals, dim3(div_up(6LL * V, kQuadBlock)), dim3(kQuadBlock), 0, s,
                       d->pixels.as<double>(), d->nx, d->ny, d->p.width, d->txy.as<double>(),
                       d->mx, d->my, d->quadacc.as<double>());
    hipLaunchKernelGGL(k_cell_sums, dim3(div_up(V, CD_BLOCK)), dim3(CD_BLOCK), 0, s,
                       d->quadacc.as<double>(), V, d->tareas.as<double>(),
                       d->partials.as<double>());
    CD_CHECK_LAUNCH(c);
    CD_TRY(finish(d, div_up(V, CD_BLOCK), 1, R_SUM, R_SUM, R_SUM, 2));
    hipLaunchKernelGGL(k_scale_areas, dim3(div_up(V, CD_BLOCK)), dim3(CD_BLOCK), 0, s,
                       d->tareas.as<double>(), V, d->p.width * d->p.height,
                       d->scalars.as<double>() + 2);
    // phi = h = 0 (caustic_design.cpp:354-363)
    CD_TRY(cd_fill_f64(c, d->phi.as<double>(), d->G, 0.0));
    CD_TRY(cd_fill_f64(c, d->h.as<double>(), d->G, 0.0));
    CD_TRY(source_owners(d));
    d->recoveries = 0;
    d->initialized = true;
    d->folded_cells = 0;
    d->folded_triangles = 0;
    return CD_OK;
}

int check_params(const cd_design_params* p) {
    if (!p) return CD_ERR_BAD_ARG;
    if (p->mesh_res_x < 2 || p->mesh_res_y < 2 || p->res_x < 2 || p->res_y < 2) return CD_ERR_BAD_ARG;
    if (!(p->width > 0) || !(p->height > 0)) return CD_ERR_BAD_ARG;
    return CD_OK;
}

}  // namespace

extern "C" {

void cd_design_params_default(cd_design_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->ior = 1.49;
    p->step_size = (double)0.05f;
    cd_solve_opts_default(&p->phi_solve);
    p->phi_solve.stop = CD_STOP_MAX_UPDATE;
    p->phi_solve.tol = 0.0000001;
    p->phi_solve.omega = -1.0;
    cd_solve_opts_default(&p->h_solve);
    p->h_solve.stop = CD_STOP_MAX_UPDATE;
    p->h_solve.tol = 0.00000001;
    p->h_solve.omega = -1.0;
}

int cd_design_create(cd_context* ctx, const cd_design_params* p, cd_design** out) {
    if (!ctx || !out) return CD_ERR_BAD_ARG;
    *out = nullptr;
    CD_TRY(check_params(p));
    CD_HIP(ctx, hipSetDevice(ctx->device));
    auto* d = new cd_design();
    d->ctx = ctx;
    d->p = *p;
    d->mx = p->mesh_res_x; d->my = p->mesh_res_y;
    d->nx = p->res_x; d->ny = p->res_y;
    d->V = (size_t)d->mx * d->my;
    d->G = (size_t)d->nx * d->ny;
    d->T = 2 * (d->mx - 1) * (d->my - 1);
    int st = alloc_design(d);
    if (st != CD_OK) { delete d; return st; }
    ctx->designs.push_back(d);
    *out = d;
    return CD_OK;
}

void cd_design_destroy(cd_design* d) {
    if (!d) return;
    if (d->ctx) {
        (void)hipStreamSynchronize(d->ctx->stream);
        cd_design_detach(d);
    }
    delete d;
}

void cd_design_detach(cd_design* d) {
    cd_context* c = d->ctx;
    if (!c) return;
    for (size_t i = 0; i < c->designs.size(); ++i)
        if (c->designs[i] == d) { c->designs.erase(c->designs.begin() + i); break; }
    d->ctx = nullptr;
}

int cd_design_set_solve_opts(cd_design* d, const cd_solve_opts* phi, const cd_solve_opts* h) {
    if (!d) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    if (phi) d->p.phi_solve = *phi;
    if (h) d->p.h_solve = *h;
    return CD_OK;
}

int cd_design_init(cd_design* d, const double* image) {
    if (!d || !image) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    cd_context* c = d->ctx;
    CD_HIP(c, hipSetDevice(c->device));
    CD_HIP(c, hipMemcpyAsync(d->image.ptr, image, 8 * d->G, hipMemcpyHostToDevice, c->stream));
    CD_TRY(design_init_from_device_image(d, d->image.as<double>()));
    CD_HIP(c, hipStreamSynchronize(c->stream));
    return CD_OK;
}

int cd_design_init_device(cd_design* d, const double* d_image) {
    if (!d || !d_image) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    CD_HIP(d->ctx, hipSetDevice(d->ctx->device));
    return design_init_from_device_image(d, d_image);
}

namespace {
int smooth_target(cd_design* d, double factor) {
    cd_context* c = d->ctx;
    const int V = (int)d->V;
    CD_TRY(c->ensure(d->smooth_tmp, sizeof(double) * 2 * (size_t)V));
    CD_HIP(c, hipMemcpyAsync(d->smooth_tmp.ptr, d->txy.ptr, sizeof(double) * 2 * (size_t)V,
                             hipMemcpyDeviceToDevice, c->stream));
    hipLaunchKernelGGL(k_laplacian_smooth, dim3(div_up(V, CD_BLOCK)), dim3(CD_BLOCK), 0,
                       c->stream, d->smooth_tmp.as<double>(), d->txy.as<double>(), d->mx, d->my,
                       factor);
    CD_CHECK_LAUNCH(c);
    return CD_OK;
}
}  // namespace

int cd_design_laplacian_smoothing(cd_design* d, double factor) {
    if (!d) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    if (!d->initialized) return CD_ERR_STATE;
    CD_HIP(d->ctx, hipSetDevice(d->ctx->device));
    CD_TRY(smooth_target(d, factor));
    CD_HIP(d->ctx, hipStreamSynchronize(d->ctx->stream));
    return CD_OK;
}

int cd_design_set_fold_check(cd_design* d, int fail) {
    if (!d) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    d->fold_fail = fail != 0;
    return CD_OK;
}

int cd_design_count_folds(cd_design* d, long long* folded_triangles) {
    if (!d || !folded_triangles) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    if (!d->initialized) return CD_ERR_STATE;
    CD_HIP(d->ctx, hipSetDevice(d->ctx->device));
    CD_TRY(count_folds(d, 9));
    double sc[10];
    CD_TRY(read_scalars(d, sc, 10));
    *folded_triangles = (long long)sc[9];
    return CD_OK;
}

int cd_design_set_first_hit(cd_design* d, int mode) {
    if (!d || (mode != 0 && mode != 1)) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    d->first_hit = mode;
    if (d->initialized) {  // the source owners were built in the previous mode
        CD_HIP(d->ctx, hipSetDevice(d->ctx->device));
        CD_TRY(source_owners(d));
        CD_HIP(d->ctx, hipStreamSynchronize(d->ctx->stream));
    }
    return CD_OK;
}

int cd_design_set_miss_recovery(cd_design* d, int on) {
    if (!d) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    d->miss_recovery = on != 0;
    return CD_OK;
}

int cd_design_transport_step(cd_design* d, double* step_out) {
    if (!d) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    if (!d->initialized) return CD_ERR_STATE;
    cd_context* c = d->ctx;
    hipStream_t s = c->stream;
    CD_HIP(c, hipSetDevice(c->device));
    const int V = (int)d->V;
    const int vb = div_up(V, CD_BLOCK);
    // dual-cell areas and mismatch (caustic_design.cpp:194-209)
    CD_TRY(c->timer_begin());
    hipLaunchKernelGGL(k_cell_errors, dim3(vb), dim3(CD_BLOCK), 0, s, d->txy.as<double>(),
                       d->tareas.as<double>(), d->mx, d->my, d->area.as<double>(),
                       d->errors.as<double>(), d->partials.as<double>());
    CD_CHECK_LAUNCH(c);
    CD_TRY(finish(d, vb, 1, R_SUM, R_SUM, R_SUM, 4));  // scalars[4] = folded cells
    CD_TRY(c->timer_end(CD_KERNEL_VERTEX, 1, 64.0 * V));
    // rasterize onto the grid (caustic_design.cpp:213, mesh.cpp:234-288)
    double misses = 0.0;
    CD_TRY(raster(d, d->txy.as<double>(), d->owner.as<int>(), true, d->errors.as<double>(),
                  nullptr, d->raster.as<double>(), nullptr, &misses));
    if (misses > 0) {
        if (step_out) *step_out = NAN;
        if (d->miss_recovery) {
            // caustic_design.cpp:215-218: smooth the target mesh and report a
            // NaN step (the driver's `step < conv` test keeps iterating)
            CD_TRY(smooth_target(d, (double)0.1f));
            d->recoveries += 1;
            CD_HIP(c, hipStreamSynchronize(s));
            return CD_OK;
        }
        c->set_error("raster miss: " + std::to_string((long long)misses) + " samples");
        return CD_ERR_RASTER_MISS;
    }
    // zero-mean projection (caustic_design.cpp:221) and the phi solve (:222)
    CD_TRY(subtract_mean(d, d->raster.as<double>(), 0));
    int st = cd_solve_device_internal(c, d->raster.as<double>(), d->phi.as<double>(), nullptr,
                                      (long long)d->G, d->nx, d->ny, &d->p.phi_solve,
                                      &d->last_phi);
    if (st != CD_OK && st != CD_ERR_NOT_CONVERGED) return st;
    // gradient, bilinear vertex lookup, boundary-constrained step (:224-265)
    CD_TRY(c->timer_begin());
    hipLaunchKernelGGL(k_vertex_step, dim3(vb), dim3(CD_BLOCK), 0, s, d->phi.as<double>(), d->nx,
                       d->ny, d->p.width, d->p.height, d->mx, d->my, d->p.step_size,
                       d->txy.as<double>(), d->vgrad.as<double>(), d->partials.as<double>());
    CD_CHECK_LAUNCH(c);
    CD_TRY(finish(d, vb, 1, R_MAX, R_SUM, R_SUM, 3));
    // orientation of every triangle of the stepped mesh
    CD_TRY(count_folds(d, 8));
    CD_TRY(c->timer_end(CD_KERNEL_VERTEX, 2, 64.0 * V + 48.0 * d->T));
    double sc[9];
    CD_TRY(read_scalars(d, sc, 9));
    d->folded_cells = (long long)sc[4];
    d->folded_triangles = (long long)sc[8];
    if (step_out) *step_out = sc[3] / d->p.width;
    if (d->fold_fail && d->folded_triangles > 0) {
        c->set_error("transport step folded the mesh: " + std::to_string(d->folded_triangles) +
                     " inverted triangles");
        return CD_ERR_FOLDED;
    }
    return CD_OK;
}

int cd_design_height_step(cd_design* d, double* update_out) {
    if (!d) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    if (!d->initialized) return CD_ERR_STATE;
    cd_context* c = d->ctx;
    hipStream_t s = c->stream;
    CD_HIP(c, hipSetDevice(c->device));
    const int V = (int)d->V;
    const int vb = div_up(V, CD_BLOCK);
    // inverse transport map + refractive normals (caustic_design.cpp:271, mesh.cpp:348-409, 677-722)
    const double focal = d->nx / d->p.width * d->p.focal_l;
    const int* rank = nullptr;
    const int* order = nullptr;
    CD_TRY(first_hit_order(d, d->txy.as<double>(), d->order_t, &rank, &order));
    CD_TRY(c->timer_begin());
    CD_TRY(cd_fill_i32(c, d->owner_v.as<int>(), d->V, INT_MAX));
    CD_TRY(cd_fill_i32(c, d->nbig.as<int>(), 1, 0));
    const InverseGrid ig{d->mx, d->my, d->p.width, d->p.height, 1e-8};
    hipLaunchKernelGGL(k_scatter_small<InverseGrid>, dim3(div_up(d->T, CD_BLOCK)), dim3(CD_BLOCK),
                       0, s, d->txy.as<double>(), d->mx, d->my, ig, d->owner_v.as<int>(), rank,
                       d->big.as<int>(), d->nbig.as<int>());
    hipLaunchKernelGGL(k_scatter_big<InverseGrid>, dim3(kBigBlocks), dim3(CD_BLOCK), 0, s,
                       d->txy.as<double>(), d->mx, ig, d->owner_v.as<int>(), rank, d->big.as<int>(),
                       d->nbig.as<int>());
    hipLaunchKernelGGL(k_inverse_normals, dim3(vb), dim3(CD_BLOCK), 0, s, d->txy.as<double>(),
                       d->mx, d->my, d->p.width, d->p.height, d->owner_v.as<int>(), order,
                       d->sz.as<double>(), focal, d->p.ior, d->inv.as<double>(),
                       d->nrm_x.as<double>(), d->nrm_y.as<double>(), d->partials.as<double>());
    CD_CHECK_LAUNCH(c);
    CD_TRY(finish(d, vb, 1, R_SUM, R_SUM, R_SUM, 5));
    CD_TRY(c->timer_end(CD_KERNEL_VERTEX, 3, 72.0 * V));
    double sc6[6];
    CD_TRY(read_scalars(d, sc6, 6));
    if (sc6[5] > 0) {
        c->set_error("inverse map: " + std::to_string((long long)sc6[5]) +
                     " source vertices outside the target mesh");
        return CD_ERR_RASTER_MISS;
    }
    // normals rasterized over the static source mesh (caustic_design.cpp:274-277)
    double misses = 0.0;
    CD_TRY(raster(d, d->sxy.as<double>(), d->owner_src.as<int>(), false, d->nrm_x.as<double>(),
                  d->nrm_y.as<double>(), d->normx.as<double>(), d->normy.as<double>(), &misses));
    if (misses > 0) {
        c->set_error("source raster miss");
        return CD_ERR_RASTER_MISS;
    }
    // divergence, zero-mean projection, height solve (caustic_design.cpp:307-311)
    const int gb = div_up((long long)d->G, CD_BLOCK);
    hipLaunchKernelGGL(k_divergence, dim3(gb), dim3(CD_BLOCK), 0, s, d->normx.as<double>(),
                       d->normy.as<double>(), d->nx, d->ny, d->div.as<double>(),
                       d->partials.as<double>());
    CD_CHECK_LAUNCH(c);
    CD_TRY(finish(d, gb, 1, R_SUM, R_SUM, R_SUM, 0));
    CD_TRY(subtract_mean(d, d->div.as<double>(), 0));
    int st = cd_solve_device_internal(c, d->div.as<double>(), d->h.as<double>(), nullptr,
                                      (long long)d->G, d->nx, d->ny, &d->p.h_solve, &d->last_h);
    if (st != CD_OK && st != CD_ERR_NOT_CONVERGED) return st;
    // vertex heights (caustic_design.cpp:322-329, mesh.cpp:724-742)
    CD_TRY(c->timer_begin());
    hipLaunchKernelGGL(k_height_sample, dim3(vb), dim3(CD_BLOCK), 0, s, d->h.as<double>(), d->nx,
                       d->ny, d->p.width, d->p.height, d->mx, d->my, d->hv.as<double>(),
                       d->partials.as<double>());
    CD_CHECK_LAUNCH(c);
    CD_TRY(finish(d, vb, 1, R_MIN, R_SUM, R_SUM, 6));
    hipLaunchKernelGGL(k_height_apply, dim3(vb), dim3(CD_BLOCK), 0, s, d->hv.as<double>(), V,
                       d->scalars.as<double>() + 6, d->sz.as<double>(), d->partials.as<double>());
    CD_CHECK_LAUNCH(c);
    CD_TRY(finish(d, vb, 1, R_SUM, R_SUM, R_SUM, 7));
    CD_TRY(c->timer_end(CD_KERNEL_VERTEX, 2, 48.0 * V));
    double sc[8];
    CD_TRY(read_scalars(d, sc, 8));
    if (update_out) *update_out = sc[7];
    return CD_OK;
}

int cd_design_run(cd_design* d, int max_transport_steps, double conv_tres, int height_steps,
                  cd_run_stats* rs) {
    if (!d) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    if (!d->initialized) return CD_ERR_STATE;
    auto t0 = std::chrono::steady_clock::now();
    cd_run_stats r;
    std::memset(&r, 0, sizeof(r));
    r.first_folded_step = -1;
    int status = CD_OK;
    const int recoveries0 = d->recoveries;
    // main.cpp:243-256: OT steps, stop after the first step below conv_tres
    for (int it = 0; it < max_transport_steps; ++it) {
        double step = 0.0;
        int st = cd_design_transport_step(d, &step);
        r.transport_steps++;
        r.last_step = step;
        if (st == CD_OK || st == CD_ERR_FOLDED) {
            r.folded_triangles = d->folded_triangles;
            if (d->folded_triangles > 0) {
                r.folded_steps++;
                if (r.first_folded_step < 0) r.first_folded_step = it;
            }
        }
        r.phi_sweeps += d->last_phi.sweeps;
        r.checks += d->last_phi.checks;
        r.learned_applied += d->last_phi.learned_applied;
        r.learned_rejected += d->last_phi.learned_rejected;
        if (d->last_phi.rel_residual > r.max_phi_residual) r.max_phi_residual = d->last_phi.rel_residual;
        if (d->last_phi.status == CD_ERR_NOT_CONVERGED) r.unconverged_solves++;
        if (st != CD_OK) { status = st; break; }
        if (step < conv_tres) break;
    }
    // main.cpp:260-262: height steps
    for (int it = 0; status == CD_OK && it < height_steps; ++it) {
        double upd = 0.0;
        int st = cd_design_height_step(d, &upd);
        r.height_steps++;
        r.h_sweeps += d->last_h.sweeps;
        r.checks += d->last_h.checks;
        r.learned_applied += d->last_h.learned_applied;
        r.learned_rejected += d->last_h.learned_rejected;
        if (d->last_h.rel_residual > r.max_h_residual) r.max_h_residual = d->last_h.rel_residual;
        if (d->last_h.status == CD_ERR_NOT_CONVERGED) r.unconverged_solves++;
        if (st != CD_OK) { status = st; break; }
    }
    r.status = status;
    r.recoveries = d->recoveries - recoveries0;
    r.elapsed_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (rs) *rs = r;
    return status;
}

int cd_designs_run_batch(cd_design* const* designs, const double* const* d_images, int n,
                         int max_transport_steps, double conv_tres, int height_steps,
                         cd_run_stats* stats) {
    if (n < 0 || (n > 0 && (!designs || !stats)) || n > CD_BATCH_MAX) return CD_ERR_BAD_ARG;
    // one design per context: a context's stream and workspace serve one
    // design at a time
    for (int i = 0; i < n; ++i) {
        if (!designs[i]) return CD_ERR_BAD_ARG;
        if (!designs[i]->ctx) return CD_ERR_STATE;
        for (int j = 0; j < i; ++j)
            if (designs[j]->ctx == designs[i]->ctx || designs[j] == designs[i]) return CD_ERR_BAD_ARG;
    }
    auto one = [&](int i) {
        cd_run_stats r;
        std::memset(&r, 0, sizeof(r));
        int st = CD_OK;
        if (d_images && d_images[i]) st = cd_design_init_device(designs[i], d_images[i]);
        if (st == CD_OK) st = cd_design_run(designs[i], max_transport_steps, conv_tres, height_steps, &r);
        r.status = st;
        stats[i] = r;
    };
    std::vector<std::thread> workers;
    int spawned = 0;
    try {
        workers.reserve(n > 0 ? n - 1 : 0);
        for (int i = 1; i < n; ++i) {
            workers.emplace_back(one, i);
            ++spawned;
        }
    } catch (const std::system_error&) {
        // threads that could not start run inline below
    }
    if (n > 0) one(0);
    for (int i = 1 + spawned; i < n; ++i) one(i);
    for (auto& t : workers) t.join();
    for (int i = 0; i < n; ++i)
        if (stats[i].status != CD_OK) return stats[i].status;
    return CD_OK;
}

int cd_image_gray_resize_device(cd_context* c, const unsigned char* d_rgba, int width,
                                int height, int out_w, int out_h, double* d_out) {
    if (!c || !d_rgba || !d_out || width < 1 || height < 1 || out_w < 1 || out_h < 1)
        return CD_ERR_BAD_ARG;
    CD_HIP(c, hipSetDevice(c->device));
    const long long n = (long long)out_w * out_h;
    hipLaunchKernelGGL(k_gray_resize, dim3(div_up(n, 256)), dim3(256), 0, c->stream, d_rgba,
                       width, height, out_w, out_h, d_out);
    CD_CHECK_LAUNCH(c);
    return CD_OK;
}

int cd_image_gray_resize(cd_context* c, const unsigned char* rgba, int width, int height,
                         int out_w, int out_h, double* out) {
    if (!c || !rgba || !out || width < 1 || height < 1 || out_w < 1 || out_h < 1)
        return CD_ERR_BAD_ARG;
    CD_HIP(c, hipSetDevice(c->device));
    const size_t in_bytes = (size_t)width * height * 4;
    const size_t n = (size_t)out_w * out_h;
    DevBuf din, dout;
    CD_TRY(c->ensure(din, in_bytes));
    CD_TRY(c->ensure(dout, sizeof(double) * n));
    CD_HIP(c, hipMemcpyAsync(din.ptr, rgba, in_bytes, hipMemcpyHostToDevice, c->stream));
    CD_TRY(cd_image_gray_resize_device(c, din.as<unsigned char>(), width, height, out_w, out_h,
                                       dout.as<double>()));
    CD_HIP(c, hipMemcpyAsync(out, dout.ptr, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    CD_HIP(c, hipStreamSynchronize(c->stream));
    return CD_OK;
}

int cd_design_get_params(const cd_design* d, cd_design_params* p) {
    if (!d || !p) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    *p = d->p;
    return CD_OK;
}

size_t cd_design_field_size(const cd_design* d, int field) {
    if (!d) return 0;
    if (!d->ctx) return 0;
    switch (field) {
        case CD_FIELD_TARGET_POINTS:
        case CD_FIELD_SOURCE_POINTS: return 3 * d->V;
        case CD_FIELD_NORMALS:
        case CD_FIELD_VERTEX_GRADIENT:
        case CD_FIELD_INVERSE_MAP: return 2 * d->V;
        case CD_FIELD_ERRORS:
        case CD_FIELD_TARGET_AREAS: return d->V;
        case CD_FIELD_PHI:
        case CD_FIELD_H:
        case CD_FIELD_RASTER:
        case CD_FIELD_DIVERGENCE:
        case CD_FIELD_NORM_X:
        case CD_FIELD_NORM_Y:
        case CD_FIELD_PIXELS:
        case CD_FIELD_IMAGE: return d->G;
        case CD_FIELD_GRADIENT: return 2 * d->G;
        default: return 0;
    }
}

const void* cd_design_device_ptr(cd_design* d, int field) {
    if (!d) return nullptr;
    if (!d->ctx) return nullptr;
    switch (field) {
        case CD_FIELD_PHI: return d->phi.ptr;
        case CD_FIELD_H: return d->h.ptr;
        case CD_FIELD_RASTER: return d->raster.ptr;
        case CD_FIELD_ERRORS: return d->errors.ptr;
        case CD_FIELD_TARGET_AREAS: return d->tareas.ptr;
        case CD_FIELD_DIVERGENCE: return d->div.ptr;
        case CD_FIELD_NORM_X: return d->normx.ptr;
        case CD_FIELD_NORM_Y: return d->normy.ptr;
        case CD_FIELD_PIXELS: return d->pixels.ptr;
        case CD_FIELD_IMAGE: return d->image.ptr;
        case CD_FIELD_INVERSE_MAP: return d->inv.ptr;
        case CD_FIELD_VERTEX_GRADIENT: return d->vgrad.ptr;
        default: return nullptr;
    }
}

int cd_design_get(cd_design* d, int field, double* out, size_t count) {
    if (!d || !out) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    const size_t n = cd_design_field_size(d, field);
    if (n == 0 || count < n) return CD_ERR_BAD_ARG;
    cd_context* c = d->ctx;
    hipStream_t s = c->stream;
    CD_HIP(c, hipSetDevice(c->device));
    const void* src = nullptr;
    const int V = (int)d->V;
    switch (field) {
        case CD_FIELD_TARGET_POINTS:
            hipLaunchKernelGGL(k_points3, dim3(div_up(V, CD_BLOCK)), dim3(CD_BLOCK), 0, s,
                               d->txy.as<double>(), (const double*)nullptr, V, d->tmp3.as<double>());
            CD_CHECK_LAUNCH(c);
            src = d->tmp3.ptr;
            break;
        case CD_FIELD_SOURCE_POINTS:
            hipLaunchKernelGGL(k_points3, dim3(div_up(V, CD_BLOCK)), dim3(CD_BLOCK), 0, s,
                               d->sxy.as<double>(), d->sz.as<double>(), V, d->tmp3.as<double>());
            CD_CHECK_LAUNCH(c);
            src = d->tmp3.ptr;
            break;
        case CD_FIELD_GRADIENT:
            CD_TRY(c->ensure(d->grad, 16 * d->G));
            hipLaunchKernelGGL(k_gradient_grid, dim3(div_up(d->nx, 256), d->ny), dim3(256), 0, s,
                               d->phi.as<double>(), d->nx, d->ny, d->grad.as<double>());
            CD_CHECK_LAUNCH(c);
            src = d->grad.ptr;
            break;
        case CD_FIELD_NORMALS:
            CD_HIP(c, hipMemcpyAsync(out, d->nrm_x.ptr, 8 * d->V, hipMemcpyDeviceToHost, s));
            CD_HIP(c, hipMemcpyAsync(out + d->V, d->nrm_y.ptr, 8 * d->V, hipMemcpyDeviceToHost, s));
            CD_HIP(c, hipStreamSynchronize(s));
            return CD_OK;
        default:
            src = cd_design_device_ptr(d, field);
    }
    if (!src) return CD_ERR_BAD_ARG;
    CD_HIP(c, hipMemcpyAsync(out, src, 8 * n, hipMemcpyDeviceToHost, s));
    CD_HIP(c, hipStreamSynchronize(s));
    return CD_OK;
}

int cd_design_get_dual_cells(cd_design* d, int initial, int* triangles, double* quads,
                             size_t vertices) {
    if (!d || !triangles || !quads) return CD_ERR_BAD_ARG;
    if (!d->ctx || !d->initialized) return CD_ERR_STATE;
    if (vertices < d->V) return CD_ERR_BAD_ARG;
    cd_context* c = d->ctx;
    hipStream_t s = c->stream;
    CD_HIP(c, hipSetDevice(c->device));
    CD_TRY(c->ensure(d->cells_t, sizeof(int) * 6 * d->V));
    CD_TRY(c->ensure(d->cells_q, sizeof(double) * 72 * d->V));
    // the initial target mesh is the regular grid the source points keep
    const double* xy = initial ? d->sxy.as<double>() : d->txy.as<double>();
    hipLaunchKernelGGL(k_dual_cells, dim3(div_up((int)d->V, 256)), dim3(256), 0, s, xy, d->mx,
                       d->my, d->cells_t.as<int>(), d->cells_q.as<double>());
    CD_CHECK_LAUNCH(c);
    CD_HIP(c, hipMemcpyAsync(triangles, d->cells_t.ptr, sizeof(int) * 6 * d->V,
                             hipMemcpyDeviceToHost, s));
    CD_HIP(c, hipMemcpyAsync(quads, d->cells_q.ptr, sizeof(double) * 72 * d->V,
                             hipMemcpyDeviceToHost, s));
    CD_HIP(c, hipStreamSynchronize(s));
    return CD_OK;
}

int cd_design_set(cd_design* d, int field, const double* in, size_t count) {
    if (!d || !in) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    const size_t n = cd_design_field_size(d, field);
    if (n == 0 || count < n) return CD_ERR_BAD_ARG;
    cd_context* c = d->ctx;
    hipStream_t s = c->stream;
    CD_HIP(c, hipSetDevice(c->device));
    if (field == CD_FIELD_TARGET_POINTS || field == CD_FIELD_SOURCE_POINTS) {
        // accept {x,y,z} triples; target z and source xy are fixed by construction
        std::vector<double> xy(2 * d->V), z(d->V);
        for (size_t i = 0; i < d->V; ++i) {
            xy[2 * i] = in[3 * i];
            xy[2 * i + 1] = in[3 * i + 1];
            z[i] = in[3 * i + 2];
        }
        if (field == CD_FIELD_TARGET_POINTS) {
            CD_HIP(c, hipMemcpyAsync(d->txy.ptr, xy.data(), 16 * d->V, hipMemcpyHostToDevice, s));
        } else {
            CD_HIP(c, hipMemcpyAsync(d->sz.ptr, z.data(), 8 * d->V, hipMemcpyHostToDevice, s));
        }
        CD_HIP(c, hipStreamSynchronize(s));
        return CD_OK;
    }
    void* dst = const_cast<void*>(cd_design_device_ptr(d, field));
    if (!dst || field == CD_FIELD_NORMALS) return CD_ERR_BAD_ARG;
    CD_HIP(c, hipMemcpyAsync(dst, in, 8 * n, hipMemcpyHostToDevice, s));
    CD_HIP(c, hipStreamSynchronize(s));
    return CD_OK;
}

int cd_design_last_solve_stats(const cd_design* d, cd_solve_stats* phi, cd_solve_stats* h) {
    if (!d) return CD_ERR_BAD_ARG;
    if (!d->ctx) return CD_ERR_STATE;
    if (phi) *phi = d->last_phi;
    if (h) *h = d->last_h;
    return CD_OK;
}

}  // extern "C"
