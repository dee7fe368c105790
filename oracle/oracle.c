/*
 * FP64 CPU restatement of the caustic-design hot path (the oracle).
 * TEST INFRASTRUCTURE ONLY -- see oracle.h.  Single-threaded, plain C.
 *
 * Pinning (each function is checked against something other than itself in
 * tests/test_oracle_*.py):
 *   orc_poisson_sor    bit-exact vs the reference's poisson_solver built from
 *                      its own sources (oracle/_ref), threads=1; converged
 *                      solution vs a 2-D DCT-II direct solve (scipy.fft);
 *                      eigenvector known-answer test; red-black ordering vs
 *                      the same DCT solution.
 *   orc_residual       closed form on an eigenvector (A v = lambda v).
 *   orc_subtract_average, orc_gradient, orc_divergence, orc_scale_minmax,
 *   orc_bilinear       vs oracle/_ref on random inputs; linear-field closed
 *                      forms; committed golden vectors (tests/golden/).
 *   orc_cell_areas     sum = W*H on any unfolded mesh; 1/3 of adjacent
 *                      triangle areas; vs oracle/_ref errors/areas.
 *   orc_clip_area, orc_integrate_polygon, orc_target_areas
 *                      exact overlap of axis-aligned rectangles; uniform image
 *                      => area-proportional; vs oracle/_ref.
 *   orc_raster         linear fields reproduced exactly; vs oracle/_ref raster.
 *   orc_vertex_step, orc_inverse_map, orc_normals, orc_set_heights
 *                      closed forms (identity map, zero displacement => (0,0),
 *                      min-shift) and the full design trajectory vs oracle/_ref.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * Poisson SOR
 * ====================================================================== */

/* One cell update of the reference operator (solver.cpp:20-58): neighbours
 * outside the domain or whose D is NaN are not counted. */
static inline double sor_cell(const double* D, double* phi, int w, int h, int x, int y,
                              double omega) {
    size_t i = (size_t)y * w + x;
    double val = phi[i];
    double s = 0.0, c = 0.0;
    if (x != 0 && !isnan(D[i - 1])) { c += 1.0; s += phi[i - 1]; }
    if (y != 0 && !isnan(D[i - w])) { c += 1.0; s += phi[i - w]; }
    if (x != w - 1 && !isnan(D[i + 1])) { c += 1.0; s += phi[i + 1]; }
    if (y != h - 1 && !isnan(D[i + w])) { c += 1.0; s += phi[i + w]; }
    double delta = omega / c * (s - c * val - D[i]);
    phi[i] += delta;
    return delta;
}

double orc_residual(const double* D, const double* phi, int w, int h, double* r,
                    double* pd_norm) {
    /* mean of the non-NaN right-hand side (the projection P of utils.cpp:60-86) */
    double sum = 0.0;
    long cnt = 0;
    for (size_t i = 0; i < (size_t)w * h; ++i)
        if (!isnan(D[i])) { sum += D[i]; cnt++; }
    double mean = cnt ? sum / (double)cnt : 0.0;
    double rr = 0.0, bb = 0.0;
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            if (isnan(D[i])) { if (r) r[i] = 0.0; continue; }
            double s = 0.0, c = 0.0;
            if (x != 0 && !isnan(D[i - 1])) { c += 1.0; s += phi[i - 1]; }
            if (y != 0 && !isnan(D[i - w])) { c += 1.0; s += phi[i - w]; }
            if (x != w - 1 && !isnan(D[i + 1])) { c += 1.0; s += phi[i + 1]; }
            if (y != h - 1 && !isnan(D[i + w])) { c += 1.0; s += phi[i + w]; }
            double b = D[i] - mean;
            double ri = b - (s - c * phi[i]);
            if (r) r[i] = ri;
            rr += ri * ri;
            bb += b * b;
        }
    }
    if (pd_norm) *pd_norm = sqrt(bb);
    return sqrt(rr);
}

int orc_poisson_sor(const double* D, double* phi, int w, int h, const orc_sor_opts* o,
                    orc_sor_stats* st) {
    double omega = o->omega > 0 ? o->omega : 2.0 / (1.0 + 3.14159265 / w); /* solver.cpp:71 */
    int check = o->check_every > 0 ? o->check_every : 1;
    int sweeps = 0;
    double last = 0.0;
    double rel = INFINITY;
    for (int it = 0; it < o->max_iterations; ++it) {
        double mu = 0.0;
        if (o->ordering == ORC_ORDER_RED_BLACK) {
            for (int color = 0; color < 2; ++color)
                for (int y = 0; y < h; ++y)
                    for (int x = (y + color) & 1; x < w; x += 2) {
                        double d = fabs(sor_cell(D, phi, w, h, x, y, omega));
                        if (d > mu) mu = d;
                    }
        } else {
            for (int y = 0; y < h; ++y)
                for (int x = 0; x < w; ++x) {
                    double d = fabs(sor_cell(D, phi, w, h, x, y, omega));
                    if (d > mu) mu = d;
                }
        }
        sweeps++;
        last = mu;
        if (o->max_update_tol > 0 && mu < o->max_update_tol) break; /* solver.cpp:142 */
        if (o->rel_residual_tol > 0 && sweeps % check == 0) {
            double bn;
            double rn = orc_residual(D, phi, w, h, NULL, &bn);
            rel = bn > 0 ? rn / bn : rn;
            if (rel <= o->rel_residual_tol) break;
        }
    }
    if (st) {
        double bn;
        double rn = orc_residual(D, phi, w, h, NULL, &bn);
        st->sweeps = sweeps;
        st->last_max_update = last;
        st->rel_residual = bn > 0 ? rn / bn : rn;
    }
    (void)rel;
    return 0;
}

/* ======================================================================
 * grid utilities
 * ====================================================================== */

void orc_subtract_average(double* a, int w, int h) {
    double sum = 0.0;
    int count = 0;
    for (size_t i = 0; i < (size_t)w * h; ++i)
        if (!isnan(a[i])) { sum += a[i]; count++; }
    double avg = sum / count;
    for (size_t i = 0; i < (size_t)w * h; ++i)
        if (!isnan(a[i])) a[i] = a[i] - avg;
}

void orc_gradient(const double* a, int w, int h, double* gx, double* gy) {
    for (int y = 0; y < h; ++y) {
        int yp = y + 1 < h ? y + 1 : h - 1, ym = y > 0 ? y - 1 : 0;
        for (int x = 0; x < w; ++x) {
            int xp = x + 1 < w ? x + 1 : w - 1, xm = x > 0 ? x - 1 : 0;
            gx[(size_t)y * w + x] = (a[(size_t)y * w + xp] - a[(size_t)y * w + xm]) / 2.0;
            gy[(size_t)y * w + x] = (a[(size_t)yp * w + x] - a[(size_t)ym * w + x]) / 2.0;
        }
    }
}

void orc_divergence(const double* nx, const double* ny, int w, int h, double* out) {
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            size_t i = (size_t)y * w + x;
            if (x == 0 || x == w - 1 || y == 0 || y == h - 1) { out[i] = 0.0; continue; }
            /* x-derivative of nx plus y-derivative of ny, interior central differences */
            double dx = (nx[i + 1] - nx[i - 1]) / 2.0;
            double dy = (ny[i + w] - ny[i - w]) / 2.0;
            out[i] = dx + dy;
        }
    }
}

int orc_scale_minmax(const double* a, int w, int h, double lo, double hi, double* out) {
    if (w <= 0 || h <= 0) return -1;
    /* std::min/std::max semantics: m = (b < m) ? b : m, seeded by a[0] */
    double mn = a[0], mx = a[0];
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        if (isnan(a[i])) continue;
        if (a[i] < mn) mn = a[i];
        if (mx < a[i]) mx = a[i];
    }
    for (size_t i = 0; i < (size_t)w * h; ++i)
        out[i] = isnan(a[i]) ? 0.0 : lo + (hi - lo) * (a[i] - mn) / (mx - mn);
    return 0;
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

double orc_bilinear(const double* img, int w, int h, double x, double y) {
    int x0 = (int)floor(x), y0 = (int)floor(y);
    int x1 = (int)ceil(x), y1 = (int)ceil(y);
    x0 = clampi(x0, 0, w - 1);
    x1 = clampi(x1, 0, w - 1);
    y0 = clampi(y0, 0, h - 1);
    y1 = clampi(y1, 0, h - 1);
    double fx1 = x - x0, fx0 = 1.0 - fx1;
    double fy1 = y - y0, fy0 = 1.0 - fy1;
    double top = fx0 * img[(size_t)y0 * w + x0] + fx1 * img[(size_t)y0 * w + x1];
    double bot = fx0 * img[(size_t)y1 * w + x0] + fx1 * img[(size_t)y1 * w + x1];
    return fy0 * top + fy1 * bot;
}

/* ======================================================================
 * mesh
 * ====================================================================== */

void orc_mesh_points(int mx, int my, double W, double H, double* xy) {
    for (int i = 0; i < my; ++i)
        for (int j = 0; j < mx; ++j) {
            xy[2 * ((size_t)i * mx + j)] = (double)j * W / (mx - 1);
            xy[2 * ((size_t)i * mx + j) + 1] = (double)i * H / (my - 1);
        }
}

void orc_mesh_triangles(int mx, int my, int* tri) {
    size_t t = 0;
    for (int i = 0; i < my - 1; ++i)
        for (int j = 0; j < mx - 1; ++j) {
            int idx = i * mx + j;
            tri[3 * t + 0] = idx; tri[3 * t + 1] = idx + 1; tri[3 * t + 2] = idx + mx; t++;
            tri[3 * t + 0] = idx + mx; tri[3 * t + 1] = idx + 1; tri[3 * t + 2] = idx + mx + 1; t++;
        }
}

/* Triangles adjacent to vertex (x,y) of the structured mesh, ascending index. */
static int adjacent_triangles(int mx, int my, int x, int y, int* out) {
    int n = 0;
    int cw = mx - 1;
    if (y > 0 && x > 0) out[n++] = 2 * ((y - 1) * cw + (x - 1)) + 1;
    if (y > 0 && x < mx - 1) { out[n++] = 2 * ((y - 1) * cw + x); out[n++] = 2 * ((y - 1) * cw + x) + 1; }
    if (y < my - 1 && x > 0) { out[n++] = 2 * (y * cw + (x - 1)); out[n++] = 2 * (y * cw + (x - 1)) + 1; }
    if (y < my - 1 && x < mx - 1) out[n++] = 2 * (y * cw + x);
    return n;
}

static void triangle_vertices(int mx, int t, int* v) {
    int cell = t >> 1;
    int cw = mx - 1;
    int i = cell / cw, j = cell % cw;
    int idx = i * mx + j;
    if ((t & 1) == 0) { v[0] = idx; v[1] = idx + 1; v[2] = idx + mx; }
    else { v[0] = idx + mx; v[1] = idx + 1; v[2] = idx + mx + 1; }
}

/* Quad (v_i, mid(ij), centroid(ijk), mid(ik)) of vertex vi in triangle t,
 * with (j,k) the next two vertices in the triangle's cyclic order
 * (mesh.cpp:174-199). q receives 8 doubles. */
static void triangle_quad(int mx, const double* xy, int vi, int t, double* q) {
    int v[3];
    triangle_vertices(mx, t, v);
    int p = 0;
    while (v[p] != vi) p++;
    int vj = v[(p + 1) % 3], vk = v[(p + 2) % 3];
    const double* a = xy + 2 * (size_t)vi;
    const double* b = xy + 2 * (size_t)vj;
    const double* c = xy + 2 * (size_t)vk;
    for (int d = 0; d < 2; ++d) {
        q[0 + d] = a[d];
        q[2 + d] = (a[d] + b[d]) / 2.0;
        q[4 + d] = (a[d] + b[d] + c[d]) / 3.0;
        q[6 + d] = (a[d] + c[d]) / 2.0;
    }
}

static double shoelace(const double* p, int n) {
    if (n < 3) return 0.0;
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
        int k = (i + 1) % n;
        s += p[2 * i] * p[2 * k + 1] - p[2 * k] * p[2 * i + 1];
    }
    return 0.5 * s;
}

void orc_cell_areas(int mx, int my, const double* xy, double* area) {
    int adj[6];
    double q[8];
    for (int y = 0; y < my; ++y)
        for (int x = 0; x < mx; ++x) {
            int vi = y * mx + x;
            int n = adjacent_triangles(mx, my, x, y, adj);
            double a = 0.0;
            for (int k = 0; k < n; ++k) {
                triangle_quad(mx, xy, vi, adj[k], q);
                a += shoelace(q, 4);
            }
            area[vi] = a;
        }
}

/* ---- Sutherland-Hodgman clipping against a convex polygon --------------
 * Side of c relative to the directed line a->b: sign of (b-a) x (c-b)
 * (polygon_utils.cpp:33-41).  A subject vertex is kept unless it lies strictly
 * on the outer side; a crossing point is emitted only when two consecutive
 * vertices lie strictly on opposite sides and the crossing parameter along the
 * subject edge is strictly inside (0,1) (polygon_utils.cpp:43-60, 94-116).
 * Output order per clip edge: for each subject edge S->E starting with
 * (last->first): S if kept, then the crossing. */
#define ORC_MAXPOLY 64

static int side_of(const double* a, const double* b, const double* c) {
    double x = (b[0] - a[0]) * (c[1] - b[1]) - (b[1] - a[1]) * (c[0] - b[0]);
    return x < 0 ? -1 : (x > 0 ? 1 : 0);
}

static int crossing(const double* a, const double* b, const double* s, const double* e,
                    double* out) {
    double ex = b[0] - a[0], ey = b[1] - a[1];      /* clip edge direction     */
    double sx = e[0] - s[0], sy = e[1] - s[1];      /* subject edge direction  */
    double dx = a[0] - s[0], dy = a[1] - s[1];
    double den = sx * ey - sy * ex;
    if (den == 0.0) return 0;
    double t = (dx * ey - dy * ex) / den;
    if (t <= 0 || t >= 1) return 0;
    out[0] = s[0] + t * sx;
    out[1] = s[1] + t * sy;
    return 1;
}

static int clip_edge(const double* in, int n, const double* a, const double* b, int dir,
                     double* out) {
    int m = 0;
    if (n == 0) return 0;
    const double* s = in + 2 * (n - 1);
    int ss = side_of(a, b, s);
    for (int i = 0; i < n; ++i) {
        const double* e = in + 2 * i;
        int se = side_of(a, b, e);
        if (ss != -dir && m < ORC_MAXPOLY) { out[2 * m] = s[0]; out[2 * m + 1] = s[1]; m++; }
        if (ss + se == 0 && ss != 0 && m < ORC_MAXPOLY) {
            if (crossing(a, b, s, e, out + 2 * m)) m++;
        }
        s = e;
        ss = se;
    }
    return m;
}

double orc_clip_area(const double* sub, int ns, const double* clip, int nc) {
    double bufA[2 * ORC_MAXPOLY], bufB[2 * ORC_MAXPOLY];
    int dir = side_of(clip, clip + 2, clip + 4);
    int n = clip_edge(sub, ns, clip + 2 * (nc - 1), clip, dir, bufB);
    double* cur = bufB;
    double* nxt = bufA;
    for (int i = 0; i < nc - 1 && n > 0; ++i) {
        n = clip_edge(cur, n, clip + 2 * i, clip + 2 * (i + 1), dir, nxt);
        double* t = cur; cur = nxt; nxt = t;
    }
    return shoelace(cur, n);
}

double orc_integrate_polygon(const double* image, int nx, int ny, double W, const double* xy,
                             int n) {
    double xmin = xy[0], xmax = xy[0], ymin = xy[1], ymax = xy[1];
    for (int i = 1; i < n; ++i) {
        if (xy[2 * i] < xmin) xmin = xy[2 * i];
        if (xy[2 * i] > xmax) xmax = xy[2 * i];
        if (xy[2 * i + 1] < ymin) ymin = xy[2 * i + 1];
        if (xy[2 * i + 1] > ymax) ymax = xy[2 * i + 1];
    }
    double px = W / (double)nx;                                   /* polygon_utils.cpp:327 */
    double s = 0.0;
    double ylo = fmax(floor(ymin / px), 0.0), yhi = fmin(ceil(ymax / px), (double)ny);
    double xlo = fmax(floor(xmin / px), 0.0), xhi = fmin(ceil(xmax / px), (double)nx);
    for (int y = (int)ylo; y < yhi; ++y)
        for (int x = (int)xlo; x < xhi; ++x) {
            double cx = ((double)x + 0.5) * px, cy = ((double)y + 0.5) * px;
            double sq[8] = {cx - px / 2.0, cy - px / 2.0, cx - px / 2.0, cy + px / 2.0,
                            cx + px / 2.0, cy + px / 2.0, cx + px / 2.0, cy - px / 2.0};
            s += orc_clip_area(xy, n, sq, 4) * image[(size_t)y * nx + x];
        }
    return s;
}

void orc_target_areas(const double* image, int nx, int ny, double W, double H, int mx, int my,
                      double* out) {
    double* xy = (double*)malloc(sizeof(double) * 2 * (size_t)mx * my);
    orc_mesh_points(mx, my, W, H, xy);
    int adj[6];
    double q[8];
    double total = 0.0;
    for (int y = 0; y < my; ++y)
        for (int x = 0; x < mx; ++x) {
            int vi = y * mx + x;
            int n = adjacent_triangles(mx, my, x, y, adj);
            double a = 0.0;
            for (int k = 0; k < n; ++k) {
                triangle_quad(mx, xy, vi, adj[k], q);
                a += orc_integrate_polygon(image, nx, ny, W, q, 4);
            }
            out[vi] = a;
            total += a;
        }
    double scale = (W * H) / total;                                /* polygon_utils.cpp:382 */
    for (size_t i = 0; i < (size_t)mx * my; ++i) out[i] *= scale;
    free(xy);
}

/* ---- point location ------------------------------------------------- */

/* Barycentric coordinates as bvh.cpp:162-190 called with (c, b, a, p) for the
 * triangle (a, b, c) (bvh.cpp:262): u weights a, v weights b, w = 1-u-v weights c. */
static void barycentric(const double* a, const double* b, const double* c, double px, double py,
                        double* u, double* v, double* w) {
    double v0x = a[0] - c[0], v0y = a[1] - c[1];
    double v1x = b[0] - c[0], v1y = b[1] - c[1];
    double v2x = px - c[0], v2y = py - c[1];
    double d00 = v0x * v0x + v0y * v0y;
    double d01 = v0x * v1x + v0y * v1y;
    double d02 = v0x * v2x + v0y * v2y;
    double d11 = v1x * v1x + v1y * v1y;
    double d12 = v1x * v2x + v1y * v2y;
    double den = d00 * d11 - d01 * d01;
    double inv = 1 / den;
    *u = (d11 * d02 - d01 * d12) * inv;
    *v = (d00 * d12 - d01 * d02) * inv;
    *w = 1.0 - *u - *v;
}

static int inside(double u, double v) {
    const double eps = 1e-12;
    return u >= -eps && v >= -eps && (u + v) <= 1.0 + eps;
}

typedef struct {
    int gx, gy;
    double x0, y0, sx, sy;
    int* start;  /* gx*gy+1 */
    int* items;
} tri_bins;

static void bins_build(tri_bins* b, int mx, int my, const double* xy) {
    int T = 2 * (mx - 1) * (my - 1);
    size_t V = (size_t)mx * my;
    double xmin = xy[0], xmax = xy[0], ymin = xy[1], ymax = xy[1];
    for (size_t i = 0; i < V; ++i) {
        if (xy[2 * i] < xmin) xmin = xy[2 * i];
        if (xy[2 * i] > xmax) xmax = xy[2 * i];
        if (xy[2 * i + 1] < ymin) ymin = xy[2 * i + 1];
        if (xy[2 * i + 1] > ymax) ymax = xy[2 * i + 1];
    }
    b->gx = mx; b->gy = my;
    b->x0 = xmin; b->y0 = ymin;
    b->sx = (xmax - xmin) / b->gx * (1 + 1e-9) + 1e-300;
    b->sy = (ymax - ymin) / b->gy * (1 + 1e-9) + 1e-300;
    int nb = b->gx * b->gy;
    b->start = (int*)calloc((size_t)nb + 1, sizeof(int));
    int v[3];
    for (int pass = 0; pass < 2; ++pass) {
        int* fill = NULL;
        if (pass == 1) {
            for (int k = 0; k < nb; ++k) b->start[k + 1] += b->start[k];
            b->items = (int*)malloc(sizeof(int) * (size_t)(b->start[nb] + 1));
            fill = (int*)malloc(sizeof(int) * (size_t)nb);
            memcpy(fill, b->start, sizeof(int) * (size_t)nb);
        }
        for (int t = 0; t < T; ++t) {
            triangle_vertices(mx, t, v);
            double tx0 = xy[2 * v[0]], tx1 = tx0, ty0 = xy[2 * v[0] + 1], ty1 = ty0;
            for (int k = 1; k < 3; ++k) {
                double X = xy[2 * v[k]], Y = xy[2 * v[k] + 1];
                if (X < tx0) tx0 = X;
                if (X > tx1) tx1 = X;
                if (Y < ty0) ty0 = Y;
                if (Y > ty1) ty1 = Y;
            }
            int bx0 = clampi((int)floor((tx0 - b->x0) / b->sx) - 1, 0, b->gx - 1);
            int bx1 = clampi((int)floor((tx1 - b->x0) / b->sx) + 1, 0, b->gx - 1);
            int by0 = clampi((int)floor((ty0 - b->y0) / b->sy) - 1, 0, b->gy - 1);
            int by1 = clampi((int)floor((ty1 - b->y0) / b->sy) + 1, 0, b->gy - 1);
            for (int yy = by0; yy <= by1; ++yy)
                for (int xx = bx0; xx <= bx1; ++xx) {
                    int k = yy * b->gx + xx;
                    if (pass == 0) b->start[k + 1]++;
                    else b->items[fill[k]++] = t;
                }
        }
        free(fill);
    }
}

static void bins_free(tri_bins* b) { free(b->start); free(b->items); }

/* lowest-index containing triangle, or -1 */
static int locate(const tri_bins* b, int mx, const double* xy, double px, double py, double* u,
                  double* v, double* w, int* tv) {
    int bx = (int)floor((px - b->x0) / b->sx), by = (int)floor((py - b->y0) / b->sy);
    if (bx < 0 || by < 0 || bx >= b->gx || by >= b->gy) return -1;
    int k = by * b->gx + bx;
    for (int s = b->start[k]; s < b->start[k + 1]; ++s) {
        int t = b->items[s];
        triangle_vertices(mx, t, tv);
        barycentric(xy + 2 * (size_t)tv[0], xy + 2 * (size_t)tv[1], xy + 2 * (size_t)tv[2], px, py,
                    u, v, w);
        if (inside(*u, *v)) return t;
    }
    return -1;
}

int orc_raster(int mx, int my, const double* xy, const double* values, int nx, int ny, double W,
               double H, double* out) {
    const double eps = 1e-8;                                       /* mesh.cpp:241 */
    tri_bins b;
    bins_build(&b, mx, my, xy);
    int misses = 0;
    int tv[3];
    for (int i = 0; i < ny; ++i) {
        double py = (double)i * (H - eps) / (ny - 1) + 0.5 * eps;    /* mesh.cpp:252 */
        for (int j = 0; j < nx; ++j) {
            double px = (double)j * (W - eps) / (nx - 1) + 0.5 * eps; /* mesh.cpp:245 */
            double u, v, w;
            int t = locate(&b, mx, xy, px, py, &u, &v, &w, tv);
            if (t < 0) { out[(size_t)i * nx + j] = NAN; misses++; continue; }
            out[(size_t)i * nx + j] = values[tv[0]] * u + values[tv[1]] * v + values[tv[2]] * w;
        }
    }
    bins_free(&b);
    return misses;
}

double orc_vertex_step(const double* phi, int nx, int ny, double W, double H, int mx, int my,
                       double* xy) {
    double* gx = (double*)malloc(sizeof(double) * (size_t)nx * ny);
    double* gy = (double*)malloc(sizeof(double) * (size_t)nx * ny);
    orc_gradient(phi, nx, ny, gx, gy);                              /* caustic_design.cpp:225 */
    const double step = (double)0.05f;                              /* caustic_design.cpp:250 */
    const double dt = W / mx;                                       /* mesh.cpp:522          */
    double maxd = 0.0;
    for (int i = 0; i < mx * my; ++i) {
        int y = i / mx, x = i % mx;
        double px = xy[2 * i], py = xy[2 * i + 1];
        double fx = (px / W) * nx - 0.5, fy = (py / H) * ny - 0.5; /* caustic_design.cpp:231-238 */
        double vx = orc_bilinear(gx, nx, ny, fx, fy);
        double vy = orc_bilinear(gy, nx, ny, fx, fy);
        /* boundary rules of mesh.cpp:489-512 */
        int xe = (x == 0 || x == mx - 1), ye = (y == 0 || y == my - 1);
        if (xe) vx = 0.0;
        if (ye) vy = 0.0;
        double nxp = px + vx * dt * step, nyp = py + vy * dt * step;
        double dx = px - nxp, dy = py - nyp;
        double dist = sqrt(dx * dx + dy * dy + 0.0 * 0.0);
        if (maxd < dist) maxd = dist;
        xy[2 * i] = nxp;
        xy[2 * i + 1] = nyp;
    }
    free(gx);
    free(gy);
    return maxd / W;
}

void orc_laplacian_smoothing(double* xy, int mx, int my, double f) {
    const int V = mx * my;
    double* avg = (double*)malloc(sizeof(double) * 2 * (size_t)V);
#define PX(xx, yy) xy[2 * ((yy) * mx + (xx))]
#define PY(xx, yy) xy[2 * ((yy) * mx + (xx)) + 1]
    for (int i = 0; i < V; ++i) {
        const int y = i / mx, x = i % mx;
        double ax = xy[2 * i], ay = xy[2 * i + 1];
        const int left = x == 0, right = x == mx - 1, top = y == 0, bottom = y == my - 1;
        if ((left || right) && (top || bottom)) {
            /* corner: unchanged */
        } else if (left) {
            ay = ((0.0 + PY(x + 1, y)) + PY(x, y + 1)) + PY(x, y - 1);
            ay /= 3.0;
        } else if (right) {
            ay = ((0.0 + PY(x - 1, y)) + PY(x, y - 1)) + PY(x, y + 1);
            ay /= 3.0;
        } else if (top) {
            ax = ((0.0 + PX(x - 1, y)) + PX(x + 1, y)) + PX(x, y + 1);
            ax /= 3.0;
        } else if (bottom) {
            ax = ((0.0 + PX(x - 1, y)) + PX(x, y - 1)) + PX(x + 1, y);
            ax /= 3.0;
        } else {
            ax = (((0.0 + PX(x - 1, y)) + PX(x + 1, y)) + PX(x, y - 1)) + PX(x, y + 1);
            ay = (((0.0 + PY(x - 1, y)) + PY(x + 1, y)) + PY(x, y - 1)) + PY(x, y + 1);
            ax /= 4.0;
            ay /= 4.0;
        }
        avg[2 * i] = ax;
        avg[2 * i + 1] = ay;
    }
#undef PX
#undef PY
    for (int i = 0; i < 2 * V; ++i) xy[i] = avg[i] * f + xy[i] * (1.0 - f);
    free(avg);
}

int orc_inverse_map(int mx, int my, double W, double H, const double* target_xy, double* inv_xy) {
    const double eps = 1e-8;                                       /* mesh.cpp:351 */
    tri_bins b;
    bins_build(&b, mx, my, target_xy);
    int misses = 0;
    int tv[3];
    for (int i = 0; i < mx * my; ++i) {
        int y = i / mx, x = i % mx;
        double sx = (double)x * W / (mx - 1), sy = (double)y * H / (my - 1);
        double px = eps + sx * ((W - 2 * eps) / W), py = eps + sy * ((H - 2 * eps) / H);
        double u, v, w;
        int t = locate(&b, mx, target_xy, px, py, &u, &v, &w, tv);
        if (t < 0) { inv_xy[2 * i] = NAN; inv_xy[2 * i + 1] = NAN; misses++; continue; }
        /* interpolate the regular source positions of the triangle's vertices */
        double ax = (double)(tv[0] % mx) * W / (mx - 1), ay = (double)(tv[0] / mx) * H / (my - 1);
        double bx = (double)(tv[1] % mx) * W / (mx - 1), by = (double)(tv[1] / mx) * H / (my - 1);
        double cx = (double)(tv[2] % mx) * W / (mx - 1), cy = (double)(tv[2] / mx) * H / (my - 1);
        double ix = ax * u + bx * v + cx * w;
        double iy = ay * u + by * v + cy * w;
        /* boundary overrides of mesh.cpp:384-402 */
        if (x == 0) ix = 0.0;
        if (x == mx - 1) ix = W;
        if (y == 0) iy = 0.0;
        if (y == my - 1) iy = H;
        inv_xy[2 * i] = ix;
        inv_xy[2 * i + 1] = iy;
    }
    bins_free(&b);
    return misses;
}

void orc_normals(int mx, int my, double W, double H, const double* inv_xy, const double* src_z,
                 double focal, double ior, double* nx_out, double* ny_out) {
    for (int i = 0; i < mx * my; ++i) {
        int y = i / mx, x = i % mx;
        double sx = (double)x * W / (mx - 1), sy = (double)y * H / (my - 1);
        double t0 = inv_xy[2 * i] - sx, t1 = inv_xy[2 * i + 1] - sy;
        double t2 = 0 - src_z[i] + focal;                           /* mesh.cpp:699 */
        double len = sqrt(0.0 + t0 * t0 + t1 * t1 + t2 * t2);       /* utils.cpp:370-384 */
        t0 /= len; t1 /= len; t2 /= len;
        /* n = t - ior * (0,0,1); output (n_x/n_z, n_y/n_z) (mesh.cpp:711-718) */
        double n0 = t0 - 0.0 * ior, n1 = t1 - 0.0 * ior, n2 = t2 - 1.0 * ior;
        nx_out[i] = n0 / n2;
        ny_out[i] = n1 / n2;
    }
}

double orc_set_heights(const double* hgrid, int nx, int ny, double W, double H, int mx, int my,
                       double* z) {
    size_t V = (size_t)mx * my;
    double* hv = (double*)malloc(sizeof(double) * V);
    double mn = 0.0;
    for (size_t i = 0; i < V; ++i) {
        int y = (int)(i / mx), x = (int)(i % mx);
        double sx = (double)x * W / (mx - 1), sy = (double)y * H / (my - 1);
        hv[i] = orc_bilinear(hgrid, nx, ny, (sx / W) * nx - 0.5, (sy / H) * ny - 0.5);
        if (mn > hv[i]) mn = hv[i];                                 /* mesh.cpp:726-733 */
    }
    double upd = 0.0;
    for (size_t i = 0; i < V; ++i) {
        double hh = hv[i] - mn;
        upd += (hh - z[i]) * (hh - z[i]);
        z[i] = hh;
    }
    free(hv);
    return upd;
}

/* ======================================================================
 * design orchestration
 * ====================================================================== */

struct orc_design {
    int mx, my, nx, ny;
    double W, H, f;
    int stop, max_sweeps, ordering, check_every;
    double phi_tol, h_tol;
    int last_sweeps;
    double last_residual;
    double *phi, *h, *raster, *div, *normx, *normy, *pixels;
    double *txy, *sz, *tareas, *errors, *nrmx, *nrmy, *inv, *area;
};

orc_design* orc_design_create(int mx, int my, int nx, int ny, double W, double H, double f) {
    orc_design* d = (orc_design*)calloc(1, sizeof(orc_design));
    d->mx = mx; d->my = my; d->nx = nx; d->ny = ny;
    d->W = W; d->H = H; d->f = f;
    d->stop = ORC_STOP_REFERENCE;
    d->phi_tol = 1e-7;                                             /* caustic_design.cpp:222 */
    d->h_tol = 1e-8;                                               /* caustic_design.cpp:311 */
    d->max_sweeps = 100000;
    d->ordering = ORC_ORDER_LEXICOGRAPHIC;
    d->check_every = 1;
    size_t G = (size_t)nx * ny, V = (size_t)mx * my;
    d->phi = (double*)calloc(G, sizeof(double));
    d->h = (double*)calloc(G, sizeof(double));
    d->raster = (double*)calloc(G, sizeof(double));
    d->div = (double*)calloc(G, sizeof(double));
    d->normx = (double*)calloc(G, sizeof(double));
    d->normy = (double*)calloc(G, sizeof(double));
    d->pixels = (double*)calloc(G, sizeof(double));
    d->txy = (double*)calloc(2 * V, sizeof(double));
    d->sz = (double*)calloc(V, sizeof(double));
    d->tareas = (double*)calloc(V, sizeof(double));
    d->errors = (double*)calloc(V, sizeof(double));
    d->nrmx = (double*)calloc(V, sizeof(double));
    d->nrmy = (double*)calloc(V, sizeof(double));
    d->inv = (double*)calloc(2 * V, sizeof(double));
    d->area = (double*)calloc(V, sizeof(double));
    return d;
}

void orc_design_destroy(orc_design* d) {
    if (!d) return;
    free(d->phi); free(d->h); free(d->raster); free(d->div); free(d->normx); free(d->normy);
    free(d->pixels); free(d->txy); free(d->sz); free(d->tareas); free(d->errors);
    free(d->nrmx); free(d->nrmy); free(d->inv); free(d->area);
    free(d);
}

void orc_design_set_solver(orc_design* d, int stop, double phi_tol, double h_tol, int max_sweeps,
                           int ordering, int check_every) {
    d->stop = stop;
    d->phi_tol = phi_tol;
    d->h_tol = h_tol;
    d->max_sweeps = max_sweeps;
    d->ordering = ordering;
    d->check_every = check_every;
}

int orc_design_init(orc_design* d, const double* image) {
    orc_scale_minmax(image, d->nx, d->ny, 0.0, 1.0, d->pixels);      /* caustic_design.cpp:335 */
    orc_mesh_points(d->mx, d->my, d->W, d->H, d->txy);               /* caustic_design.cpp:339 */
    memset(d->sz, 0, sizeof(double) * (size_t)d->mx * d->my);
    orc_target_areas(d->pixels, d->nx, d->ny, d->W, d->H, d->mx, d->my, d->tareas); /* :348 */
    memset(d->phi, 0, sizeof(double) * (size_t)d->nx * d->ny);       /* :354-363 */
    memset(d->h, 0, sizeof(double) * (size_t)d->nx * d->ny);
    return 0;
}

static void design_solve(orc_design* d, const double* rhs, double* x, double tol) {
    orc_sor_opts o;
    memset(&o, 0, sizeof(o));
    o.max_iterations = d->max_sweeps;
    o.ordering = d->ordering;
    o.check_every = d->check_every;
    if (d->stop == ORC_STOP_REFERENCE) o.max_update_tol = tol;
    else o.rel_residual_tol = tol;
    orc_sor_stats st;
    orc_poisson_sor(rhs, x, d->nx, d->ny, &o, &st);
    d->last_sweeps = st.sweeps;
    d->last_residual = st.rel_residual;
}

int orc_design_transport_step(orc_design* d, double* step) {
    size_t V = (size_t)d->mx * d->my;
    orc_cell_areas(d->mx, d->my, d->txy, d->area);                  /* caustic_design.cpp:195-197 */
    for (size_t i = 0; i < V; ++i)
        d->errors[i] = (d->tareas[i] - d->area[i]) / d->area[i];    /* :199-209 */
    int miss = orc_raster(d->mx, d->my, d->txy, d->errors, d->nx, d->ny, d->W, d->H, d->raster);
    if (miss) { if (step) *step = NAN; return 1; }
    orc_subtract_average(d->raster, d->nx, d->ny);                  /* :221 */
    design_solve(d, d->raster, d->phi, d->phi_tol);                 /* :222 */
    double s = orc_vertex_step(d->phi, d->nx, d->ny, d->W, d->H, d->mx, d->my, d->txy);
    if (step) *step = s;
    return 0;
}

int orc_design_height_step(orc_design* d, double* update) {
    size_t V = (size_t)d->mx * d->my;
    int miss = orc_inverse_map(d->mx, d->my, d->W, d->H, d->txy, d->inv);
    if (miss) return 1;
    double focal = d->nx / d->W * d->f;                             /* caustic_design.cpp:271 */
    orc_normals(d->mx, d->my, d->W, d->H, d->inv, d->sz, focal, 1.49, d->nrmx, d->nrmy);
    double* sxy = (double*)malloc(sizeof(double) * 2 * V);
    orc_mesh_points(d->mx, d->my, d->W, d->H, sxy);
    int m1 = orc_raster(d->mx, d->my, sxy, d->nrmx, d->nx, d->ny, d->W, d->H, d->normx);
    int m2 = orc_raster(d->mx, d->my, sxy, d->nrmy, d->nx, d->ny, d->W, d->H, d->normy);
    free(sxy);
    if (m1 || m2) return 1;
    orc_divergence(d->normx, d->normy, d->nx, d->ny, d->div);       /* :307 */
    orc_subtract_average(d->div, d->nx, d->ny);                     /* :308 */
    design_solve(d, d->div, d->h, d->h_tol);                        /* :311 */
    double u = orc_set_heights(d->h, d->nx, d->ny, d->W, d->H, d->mx, d->my, d->sz);
    if (update) *update = u;
    return 0;
}

int orc_design_last_sweeps(const orc_design* d) { return d->last_sweeps; }
double orc_design_last_residual(const orc_design* d) { return d->last_residual; }

const double* orc_design_grid(const orc_design* d, int which) {
    return orc_design_grid_mut((orc_design*)d, which);
}

double* orc_design_grid_mut(orc_design* d, int which) {
    switch (which) {
        case 0: return d->phi;
        case 1: return d->h;
        case 2: return d->raster;
        case 3: return d->div;
        case 4: return d->normx;
        case 5: return d->normy;
        case 6: return d->pixels;
        default: return NULL;
    }
}

const double* orc_design_vec(const orc_design* d, int which) {
    return orc_design_vec_mut((orc_design*)d, which);
}

double* orc_design_vec_mut(orc_design* d, int which) {
    switch (which) {
        case 0: return d->txy;
        case 1: return d->sz;
        case 2: return d->tareas;
        case 3: return d->errors;
        case 4: return d->nrmx;
        case 5: return d->nrmy;
        case 6: return d->inv;
        default: return NULL;
    }
}
