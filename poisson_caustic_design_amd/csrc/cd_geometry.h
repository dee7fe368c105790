// Device geometry primitives of the caustic-design path (FP64).  Operation
// order follows the reference expressions so that, compiled without FP
// contraction, results match the FP64 CPU restatement bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

namespace cdgeo {

// Structured triangulation of src/mesh.cpp:57-63: cell (i,j) with base vertex
// idx = i*mx + j holds triangle 2k = {idx, idx+1, idx+mx} and
// 2k+1 = {idx+mx, idx+1, idx+mx+1}, k = i*(mx-1) + j.
__device__ __forceinline__ void tri_vertices(int mx, int t, int& a, int& b, int& c) {
    const int cell = t >> 1;
    const int cw = mx - 1;
    const int i = cell / cw, j = cell - i * cw;
    const int idx = i * mx + j;
    if ((t & 1) == 0) { a = idx; b = idx + 1; c = idx + mx; }
    else { a = idx + mx; b = idx + 1; c = idx + mx + 1; }
}

// Triangles adjacent to vertex (x,y), ascending triangle index; returns count.
__device__ __forceinline__ int adjacent_triangles(int mx, int my, int x, int y, int* out) {
    int n = 0;
    const int cw = mx - 1;
    if (y > 0 && x > 0) out[n++] = 2 * ((y - 1) * cw + (x - 1)) + 1;
    if (y > 0 && x < mx - 1) { out[n++] = 2 * ((y - 1) * cw + x); out[n++] = 2 * ((y - 1) * cw + x) + 1; }
    if (y < my - 1 && x > 0) { out[n++] = 2 * (y * cw + (x - 1)); out[n++] = 2 * (y * cw + (x - 1)) + 1; }
    if (y < my - 1 && x < mx - 1) out[n++] = 2 * (y * cw + x);
    return n;
}

// Quad (v_i, mid(ij), centroid(ijk), mid(ik)) of vertex vi in triangle t
// (src/mesh.cpp:174-199); q receives x0,y0,...,x3,y3.
__device__ __forceinline__ void triangle_quad(int mx, const double* __restrict__ xy, int vi, int t,
                                              double* q) {
    int v[3];
    tri_vertices(mx, t, v[0], v[1], v[2]);
    const int p = (v[0] == vi) ? 0 : ((v[1] == vi) ? 1 : 2);
    const int vj = v[(p + 1) % 3], vk = v[(p + 2) % 3];
    const double ax = xy[2 * vi], ay = xy[2 * vi + 1];
    const double bx = xy[2 * vj], by = xy[2 * vj + 1];
    const double cx = xy[2 * vk], cy = xy[2 * vk + 1];
    q[0] = ax; q[1] = ay;
    q[2] = (ax + bx) / 2.0; q[3] = (ay + by) / 2.0;
    q[4] = (ax + bx + cx) / 3.0; q[5] = (ay + by + cy) / 3.0;
    q[6] = (ax + cx) / 2.0; q[7] = (ay + cy) / 2.0;
}

// Signed shoelace area, summation order of polygon_utils.cpp:216-236.
__device__ __forceinline__ double shoelace4(const double* p) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int k = (i + 1) & 3;
        s += p[2 * i] * p[2 * k + 1] - p[2 * k] * p[2 * i + 1];
    }
    return 0.5 * s;
}

// Barycentric coordinates (u weights a, v weights b, w = 1-u-v weights c) of
// point p in triangle (a,b,c): the formula of bvh.cpp:162-190 applied with
// (t0,t1,t2) = (c,b,a) as at bvh.cpp:262.
__device__ __forceinline__ void barycentric(double ax, double ay, double bx, double by, double cx,
                                            double cy, double px, double py, double& u, double& v,
                                            double& w) {
    const double v0x = ax - cx, v0y = ay - cy;
    const double v1x = bx - cx, v1y = by - cy;
    const double v2x = px - cx, v2y = py - cy;
    const double d00 = v0x * v0x + v0y * v0y;
    const double d01 = v0x * v1x + v0y * v1y;
    const double d02 = v0x * v2x + v0y * v2y;
    const double d11 = v1x * v1x + v1y * v1y;
    const double d12 = v1x * v2x + v1y * v2y;
    const double den = d00 * d11 - d01 * d01;
    const double inv = 1 / den;
    u = (d11 * d02 - d01 * d12) * inv;
    v = (d00 * d12 - d01 * d02) * inv;
    w = 1.0 - u - v;
}

// inside test of bvh.cpp:264 (eps 1e-12)
__device__ __forceinline__ bool inside(double u, double v) {
    const double eps = 1e-12;
    return u >= -eps && v >= -eps && (u + v) <= 1.0 + eps;
}

// Bilinear lookup with floor/ceil and clamp-to-edge (caustic_design.cpp:156-188).
template <typename F>
__device__ __forceinline__ double bilinear(F at, int w, int h, double x, double y) {
    int x0 = (int)floor(x), y0 = (int)floor(y);
    int x1 = (int)ceil(x), y1 = (int)ceil(y);
    x0 = min(max(x0, 0), w - 1);
    x1 = min(max(x1, 0), w - 1);
    y0 = min(max(y0, 0), h - 1);
    y1 = min(max(y1, 0), h - 1);
    const double fx1 = x - x0, fx0 = 1.0 - fx1;
    const double fy1 = y - y0, fy0 = 1.0 - fy1;
    const double top = fx0 * at(x0, y0) + fx1 * at(x1, y0);
    const double bot = fx0 * at(x0, y1) + fx1 * at(x1, y1);
    return fy0 * top + fy1 * bot;
}

// ---- convex clipping (Sutherland-Hodgman) ---------------------------------
// Side of c relative to the directed line a->b: sign of (b-a) x (c-b)
// (polygon_utils.cpp:33-41).  A subject vertex is kept unless strictly on the
// outer side; a crossing is emitted only for strictly opposite consecutive
// vertices with crossing parameter strictly inside (0,1)
// (polygon_utils.cpp:43-60, 94-116).
__device__ __forceinline__ int side_of(double ax, double ay, double bx, double by, double cx,
                                       double cy) {
    const double x = (bx - ax) * (cy - by) - (by - ay) * (cx - bx);
    return x < 0 ? -1 : (x > 0 ? 1 : 0);
}

__device__ __forceinline__ bool crossing(double ax, double ay, double bx, double by, double sx,
                                         double sy, double ex, double ey, double& ox, double& oy) {
    const double cex = bx - ax, cey = by - ay;
    const double dsx = ex - sx, dsy = ey - sy;
    const double dx = ax - sx, dy = ay - sy;
    const double den = dsx * cey - dsy * cex;
    if (den == 0.0) return false;
    const double t = (dx * cey - dy * cex) / den;
    if (t <= 0 || t >= 1) return false;
    ox = sx + t * dsx;
    oy = sy + t * dsy;
    return true;
}

constexpr int CLIP_MAX = 12;

// Clip polygon (px,py; n vertices) against the directed edge a->b keeping the
// side `dir`; output into (qx,qy); returns the output count.  Vertex i of a
// polygon is stored at index i * S (S = 1: a private array; S = 64: one
// column per lane of a 64-wide LDS block, bank-conflict free).
template <int S>
__device__ __forceinline__ int clip_edge_s(const double* px, const double* py, int n, double ax,
                                           double ay, double bx, double by, int dir, double* qx,
                                           double* qy) {
    int m = 0;
    if (n == 0) return 0;
    double sx = px[(n - 1) * S], sy = py[(n - 1) * S];
    int ss = side_of(ax, ay, bx, by, sx, sy);
    for (int i = 0; i < n; ++i) {
        const double ex = px[i * S], ey = py[i * S];
        const int se = side_of(ax, ay, bx, by, ex, ey);
        if (ss != -dir && m < CLIP_MAX) { qx[m * S] = sx; qy[m * S] = sy; ++m; }
        if (ss + se == 0 && ss != 0 && m < CLIP_MAX) {
            double ox, oy;
            if (crossing(ax, ay, bx, by, sx, sy, ex, ey, ox, oy)) {
                qx[m * S] = ox;
                qy[m * S] = oy;
                ++m;
            }
        }
        sx = ex; sy = ey; ss = se;
    }
    return m;
}

template <int S>
__device__ __forceinline__ double shoelace_s(const double* x, const double* y, int n) {
    if (n < 3) return 0.0;
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        const int k = (i + 1 == n) ? 0 : i + 1;
        s += x[i * S] * y[k * S] - x[k * S] * y[i * S];
    }
    return 0.5 * s;
}

__device__ __forceinline__ int clip_edge(const double* px, const double* py, int n, double ax,
                                         double ay, double bx, double by, int dir, double* qx,
                                         double* qy) {
    return clip_edge_s<1>(px, py, n, ax, ay, bx, by, dir, qx, qy);
}

__device__ __forceinline__ double shoelace_n(const double* x, const double* y, int n) {
    return shoelace_s<1>(x, y, n);
}

// Signed area of quad q (4 vertices) intersected with the pixel square of
// side px centred at (cx, cy), square vertices ordered as
// polygon_utils.cpp:336-346.  buf holds the four work polygons (ax, ay, bx,
// by; CLIP_MAX vertices each, stride S).
template <int S>
__device__ __forceinline__ double quad_pixel_area_s(const double* q, double cx, double cy,
                                                    double px, double* buf) {
    const double sx[4] = {cx - px / 2.0, cx - px / 2.0, cx + px / 2.0, cx + px / 2.0};
    const double sy[4] = {cy - px / 2.0, cy + px / 2.0, cy + px / 2.0, cy - px / 2.0};
    const int dir = side_of(sx[0], sy[0], sx[1], sy[1], sx[2], sy[2]);
    double* ax = buf;
    double* ay = buf + CLIP_MAX * S;
    double* bx = buf + 2 * CLIP_MAX * S;
    double* by = buf + 3 * CLIP_MAX * S;
    for (int i = 0; i < 4; ++i) { ax[i * S] = q[2 * i]; ay[i * S] = q[2 * i + 1]; }
    int n = clip_edge_s<S>(ax, ay, 4, sx[3], sy[3], sx[0], sy[0], dir, bx, by);
    if (n == 0) return 0.0;
    n = clip_edge_s<S>(bx, by, n, sx[0], sy[0], sx[1], sy[1], dir, ax, ay);
    if (n == 0) return 0.0;
    n = clip_edge_s<S>(ax, ay, n, sx[1], sy[1], sx[2], sy[2], dir, bx, by);
    if (n == 0) return 0.0;
    n = clip_edge_s<S>(bx, by, n, sx[2], sy[2], sx[3], sy[3], dir, ax, ay);
    return shoelace_s<S>(ax, ay, n);
}

__device__ __forceinline__ double quad_pixel_area(const double* q, double cx, double cy,
                                                  double px) {
    double buf[4 * CLIP_MAX];
    return quad_pixel_area_s<1>(q, cx, cy, px, buf);
}

}  // namespace cdgeo
