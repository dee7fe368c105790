// First-hit order of the reference's triangle BVH, on the host.
//
// The reference locates raster samples and grid vertices in a mesh with a
// bounding-volume hierarchy (bvh.cpp): Bvh::build (bvh.cpp:20-48) partitions
// an array of triangle ids recursively (bvh.cpp:78-155: bounding box of the
// range's vertices, widened by half an epsilon, split at the middle of its
// longer side, leaves of at most 5 triangles or depth 30, mesh.cpp:235,292,349)
// and Bvh::query (bvh.cpp:196-266) walks the tree depth first, first child
// first, and returns the first triangle of a leaf that holds the point.  The
// leaves cover consecutive ranges of the partitioned array in walk order, and
// every box on the way to a triangle's leaf holds the triangle's own box, so
// the triangle returned for a point is the containing triangle with the
// lowest position in the partitioned array.  This file computes that array;
// the rasterization kernels then keep, per sample, the containing triangle of
// lowest position instead of lowest index (the two differ only where
// triangles overlap, i.e. on folded meshes, and on shared edges).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>
#include <utility>
#include <vector>

#include "cd_api.h"

namespace {

constexpr int kLeafSize = 5;   // build_target_bvh(5, 30), mesh.cpp:235
constexpr int kMaxDepth = 30;

struct OrderBuilder {
    const double* xy;
    int mx, T;
    std::vector<double> cx, cy;  // centroids, permuted with the ids
    int* ids;

    void vertices(int t, int v[3]) const {  // mesh.cpp:57-63
        const int cell = t >> 1, cw = mx - 1;
        const int i = cell / cw, j = cell - i * cw, idx = i * mx + j;
        if ((t & 1) == 0) { v[0] = idx; v[1] = idx + 1; v[2] = idx + mx; }
        else { v[0] = idx + mx; v[1] = idx + 1; v[2] = idx + mx + 1; }
    }

    // area-weighted polygon centroid, polygon_utils.cpp:238-263 (a triangle
    // of zero area yields non-finite coordinates there as here)
    void centroid(int t, double& ox, double& oy) const {
        int v[3];
        vertices(t, v);
        double sa = 0.0, sx = 0.0, sy = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double x0 = xy[2 * v[k]], y0 = xy[2 * v[k] + 1];
            const double x1 = xy[2 * v[(k + 1) % 3]], y1 = xy[2 * v[(k + 1) % 3] + 1];
            const double a = x0 * y1 - x1 * y0;
            sa += a;
            sx += (x0 + x1) * a;
            sy += (y0 + y1) * a;
        }
        sa *= 0.5;
        ox = sx / (6 * sa);
        oy = sy / (6 * sa);
    }

    // two-sided partition of bvh.cpp:50-76; the split value arrives as float
    int split(int start, int end, int dim, float split_value) {
        const double s = split_value;
        std::vector<double>& c = dim ? cy : cx;
        int left = start, right = end - 1;
        while (left < right) {
            while (left < end && c[left] < s) ++left;
            while (right >= start && c[right] >= s) --right;
            if (left >= right) break;
            std::swap(cx[left], cx[right]);
            std::swap(cy[left], cy[right]);
            std::swap(ids[left], ids[right]);
            ++left;
            --right;
        }
        // the reference reads the centroid at `left` even when left == end:
        // inside the array that is the next range's first entry; past its
        // end (undefined there) it is taken as not below the split value
        const double cl = left < T ? c[left] : std::numeric_limits<double>::quiet_NaN();
        return cl <= s ? std::min(end, left + 1) : left;
    }

    void node(int start, int end, int level) {
        double mnx = std::numeric_limits<double>::infinity(), mny = mnx;
        double mxx = -mnx, mxy = -mnx;
        for (int i = start; i < end; ++i) {
            int v[3];
            vertices(ids[i], v);
            for (int k = 0; k < 3; ++k) {
                const double x = xy[2 * v[k]], y = xy[2 * v[k] + 1];
                mnx = std::fmin(mnx, x);
                mny = std::fmin(mny, y);
                mxx = std::fmax(mxx, x);
                mxy = std::fmax(mxy, y);
            }
        }
        const double eps = DBL_EPSILON;
        mnx -= 0.5 * eps;
        mny -= 0.5 * eps;
        mxx += 0.5 * eps;
        mxy += 0.5 * eps;
        if (end - start <= kLeafSize || level >= kMaxDepth) return;
        const int dim = (mxx - mnx) < (mxy - mny) ? 1 : 0;
        const double mid = dim == 0 ? 0.5 * (mxx + mnx) : 0.5 * (mxy + mny);
        const int m = split(start, end, dim, (float)mid);
        if (m == start || m == end) return;
        node(start, m, level + 1);
        node(m, end, level + 1);
    }
};

}  // namespace

extern "C" int cd_bvh_first_hit_order(const double* xy, int mx, int my, int* order) {
    if (!xy || !order || mx < 2 || my < 2) return CD_ERR_BAD_ARG;
    const long long T = 2LL * (mx - 1) * (my - 1);
    if (T > (long long)std::numeric_limits<int>::max() / 2) return CD_ERR_BAD_ARG;
    OrderBuilder b{xy, mx, (int)T, {}, {}, order};
    b.cx.resize(T);
    b.cy.resize(T);
    for (int t = 0; t < (int)T; ++t) {
        order[t] = t;
        b.centroid(t, b.cx[t], b.cy[t]);
    }
    b.node(0, (int)T, 0);
    return CD_OK;
}
