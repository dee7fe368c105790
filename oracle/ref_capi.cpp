// C entry points over the reference's own C++ classes, compiled together with
// the reference sources where they lie (see oracle/Makefile).  This file is
// test infrastructure: it lets the test suite and the CPU-baseline leg of the
// benchmark drive the unmodified reference code through ctypes.
//
// Every wrapper flattens/unflattens the reference's nested std::vector layout
// ([row][col], row-major, points as {x,y,z}) and silences the reference's
// stdout progress prints so that callers that print machine-readable output
// are not polluted.
//
// Reference interfaces driven here:
//   poisson_solver              src/solver.h:8, src/solver.cpp:70-147
//   class Caustic_design        src/caustic_design.h:7-66
//   subtractAverage             src/utils.cpp:60-86
//   calculate_gradient          src/utils.cpp:3-20
//   calculate_divergence        src/utils.cpp:22-39
//   scale_matrix_proportional   src/utils.cpp:88-129
//   bilinearInterpolation       src/caustic_design.cpp:156-188
//   integrate_cell_intensities  src/polygon_utils.cpp:311-366
//   Mesh::calculate_inverted_transport_map  src/mesh.cpp:348-409
//   Mesh::laplacian_smoothing   src/mesh.cpp:534-606
//   save_solid_obj              src/utils.cpp:198-260
//   export_grid_to_svg          src/utils.cpp:262-307
//   export_triangles_to_svg     src/utils.cpp:309-344

#include <unistd.h>
#include <fcntl.h>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bvh.h"
#include "caustic_design.h"

// Defined (non-static) in src/caustic_design.cpp but not declared in a header.
double bilinearInterpolation(const std::vector<std::vector<double>>& image, double x, double y);
// Defined with C++ linkage in src/polygon_utils.cpp:311; the header's extern "C"
// declaration (src/polygon_utils.h:41) has a different parameter type.
double integrate_cell_intensities(std::vector<std::vector<double>>& image,
                                  std::vector<std::vector<double>>& input_polygon, int image_w,
                                  int image_h, double width);

namespace {

struct Quiet {
    int saved = -1;
    Quiet() {
        std::fflush(stdout);
        std::cout.flush();
        saved = dup(1);
        int devnull = open("/dev/null", O_WRONLY);
        if (devnull >= 0) { dup2(devnull, 1); close(devnull); }
    }
    ~Quiet() {
        std::fflush(stdout);
        std::cout.flush();
        if (saved >= 0) { dup2(saved, 1); close(saved); }
    }
};

std::vector<std::vector<double>> to_grid(const double* a, int w, int h) {
    std::vector<std::vector<double>> g(h, std::vector<double>(w));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) g[y][x] = a[(size_t)y * w + x];
    return g;
}

void from_grid(const std::vector<std::vector<double>>& g, double* a) {
    size_t k = 0;
    for (const auto& row : g)
        for (double v : row) a[k++] = v;
}

}  // namespace

extern "C" {

// Builds the reference's Bvh (bvh.cpp) over an mx x my mesh with vertices xy
// (x,y pairs; triangulation of mesh.cpp:57-63, as Mesh builds it) with the
// parameters of mesh.cpp:235, and queries n points; face[k] = -1 on a miss.
void ref_bvh_query(const double* xy, int mx, int my, const double* pts, int n, int* face) {
    std::vector<std::vector<double>> points((size_t)mx * my);
    for (size_t v = 0; v < points.size(); ++v) points[v] = {xy[2 * v], xy[2 * v + 1], 0.0};
    std::vector<std::vector<int>> tris;
    for (int i = 0; i < my - 1; ++i)
        for (int j = 0; j < mx - 1; ++j) {
            const int idx = i * mx + j;
            tris.push_back({idx, idx + 1, idx + mx});
            tris.push_back({idx + mx, idx + 1, idx + mx + 1});
        }
    Bvh bvh(tris, points);
    bvh.build(5, 30);
    for (int k = 0; k < n; ++k) {
        Hit hit;
        bool found = false;
        bvh.query({pts[2 * k], pts[2 * k + 1]}, hit, found);
        face[k] = found ? hit.face_id : -1;
    }
}


int ref_poisson_solver(const double* D, double* phi, int w, int h, int max_iterations,
                       double threshold, int max_threads) {
    auto d = to_grid(D, w, h);
    auto p = to_grid(phi, w, h);
    {
        Quiet q;
        poisson_solver(d, p, w, h, max_iterations, threshold, max_threads);
    }
    from_grid(p, phi);
    return 0;
}

void ref_subtract_average(double* a, int w, int h) {
    auto g = to_grid(a, w, h);
    subtractAverage(g);
    from_grid(g, a);
}

void ref_gradient(const double* a, int w, int h, double* gx, double* gy) {
    auto g = calculate_gradient(to_grid(a, w, h));
    from_grid(g[0], gx);
    from_grid(g[1], gy);
}

void ref_divergence(const double* nx, const double* ny, int w, int h, double* out) {
    auto d = calculate_divergence(to_grid(nx, w, h), to_grid(ny, w, h), w, h);
    from_grid(d, out);
}

void ref_scale_minmax(const double* a, int w, int h, double lo, double hi, double* out) {
    auto s = scale_matrix_proportional(to_grid(a, w, h), lo, hi);
    from_grid(s, out);
}

double ref_bilinear(const double* a, int w, int h, double x, double y) {
    auto g = to_grid(a, w, h);
    Quiet q;
    return bilinearInterpolation(g, x, y);
}

// Image integral over one polygon (n vertices, xy interleaved).
double ref_integrate_polygon(const double* image, int w, int h, double width,
                             const double* xy, int n) {
    auto g = to_grid(image, w, h);
    std::vector<std::vector<double>> poly(n, std::vector<double>(2));
    for (int i = 0; i < n; ++i) { poly[i][0] = xy[2 * i]; poly[i][1] = xy[2 * i + 1]; }
    return integrate_cell_intensities(g, poly, w, h, width);
}

void* ref_design_create(int mesh_x, int mesh_y, int res_x, int res_y, double width,
                        double height, double focal_l, double thickness, int threads) {
    auto* cd = new Caustic_design();
    cd->set_mesh_resolution(mesh_x, mesh_y);
    cd->set_domain_resolution(res_x, res_y);
    cd->set_mesh_size(width, height);
    cd->set_lens_focal_length(focal_l);
    cd->set_lens_thickness(thickness);
    cd->set_solver_max_threads(threads);
    cd->mesh = nullptr;
    return cd;
}

// nthreads == 0 makes poisson_solver run zero tiles and return after one
// empty iteration (solver.cpp:73-83, 142): the step then costs everything but
// the Poisson solve.
void ref_design_set_threads(void* h, int threads) {
    static_cast<Caustic_design*>(h)->set_solver_max_threads(threads);
}

void ref_design_destroy(void* h) {
    auto* cd = static_cast<Caustic_design*>(h);
    delete cd->mesh;
    delete cd;
}

void ref_design_init(void* h, const double* image) {
    auto* cd = static_cast<Caustic_design*>(h);
    auto img = to_grid(image, cd->resolution_x, cd->resolution_y);
    Quiet q;
    cd->initialize_solvers(img);
}

double ref_transport_step(void* h) {
    auto* cd = static_cast<Caustic_design*>(h);
    Quiet q;
    return cd->perform_transport_iteration();
}

void ref_height_step(void* h) {
    auto* cd = static_cast<Caustic_design*>(h);
    Quiet q;
    cd->perform_height_map_iteration(0);
}

static void copy_points(const std::vector<std::vector<double>>& pts, double* out, int stride) {
    for (size_t i = 0; i < pts.size(); ++i)
        for (int k = 0; k < stride; ++k) out[i * stride + k] = pts[i][k];
}

void ref_get_target_points(void* h, double* xyz) {
    copy_points(static_cast<Caustic_design*>(h)->mesh->target_points, xyz, 3);
}

void ref_get_source_points(void* h, double* xyz) {
    copy_points(static_cast<Caustic_design*>(h)->mesh->source_points, xyz, 3);
}

void ref_set_target_points(void* h, const double* xyz) {
    auto& pts = static_cast<Caustic_design*>(h)->mesh->target_points;
    for (size_t i = 0; i < pts.size(); ++i)
        for (int k = 0; k < 3; ++k) pts[i][k] = xyz[i * 3 + k];
}

void ref_set_source_z(void* h, const double* z) {
    auto& pts = static_cast<Caustic_design*>(h)->mesh->source_points;
    for (size_t i = 0; i < pts.size(); ++i) pts[i][2] = z[i];
}

// Which grid: 0 phi, 1 h, 2 raster, 3 divergence, 4 norm_x, 5 norm_y, 6 pixels
int ref_get_grid(void* h, int which, double* out) {
    auto* cd = static_cast<Caustic_design*>(h);
    const std::vector<std::vector<double>>* g = nullptr;
    switch (which) {
        case 0: g = &cd->phi; break;
        case 1: g = &cd->h; break;
        case 2: g = &cd->raster; break;
        case 3: g = &cd->divergence; break;
        case 4: g = &cd->norm_x; break;
        case 5: g = &cd->norm_y; break;
        case 6: g = &cd->pixels; break;
        default: return -1;
    }
    if (g->empty()) return -2;
    from_grid(*g, out);
    return 0;
}

int ref_set_grid(void* h, int which, const double* in) {
    auto* cd = static_cast<Caustic_design*>(h);
    std::vector<std::vector<double>>* g = nullptr;
    switch (which) {
        case 0: g = &cd->phi; break;
        case 1: g = &cd->h; break;
        default: return -1;
    }
    *g = to_grid(in, cd->resolution_x, cd->resolution_y);
    return 0;
}

// Which vector: 0 errors, 1 target_areas, 2 vertex_gradient x, 3 vertex_gradient y,
// 4 normals x, 5 normals y
int ref_get_vector(void* h, int which, double* out) {
    auto* cd = static_cast<Caustic_design*>(h);
    const std::vector<double>* v = nullptr;
    switch (which) {
        case 0: v = &cd->errors; break;
        case 1: v = &cd->target_areas; break;
        case 2: if (cd->vertex_gradient.size() < 2) return -2; v = &cd->vertex_gradient[0]; break;
        case 3: if (cd->vertex_gradient.size() < 2) return -2; v = &cd->vertex_gradient[1]; break;
        case 4: if (cd->normals.size() < 2) return -2; v = &cd->normals[0]; break;
        case 5: if (cd->normals.size() < 2) return -2; v = &cd->normals[1]; break;
        default: return -1;
    }
    std::memcpy(out, v->data(), v->size() * sizeof(double));
    return (int)v->size();
}

// Inverse transport map of the current target mesh; returns the number of
// points written (the reference pushes nothing for a source vertex it cannot
// locate, so a count below V means misaligned output).
int ref_inverted_map(void* h, double* xy) {
    auto* cd = static_cast<Caustic_design*>(h);
    std::vector<std::vector<double>> inv;
    {
        Quiet q;
        inv = cd->mesh->calculate_inverted_transport_map();
    }
    for (size_t i = 0; i < inv.size(); ++i) { xy[2 * i] = inv[i][0]; xy[2 * i + 1] = inv[i][1]; }
    return (int)inv.size();
}

// points as V x `stride` doubles -> the reference's {x, y, z} vectors
static std::vector<std::vector<double>> to_points(const double* a, int n, int stride) {
    std::vector<std::vector<double>> p(n, std::vector<double>(3, 0.0));
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < stride && k < 3; ++k) p[i][k] = a[(size_t)i * stride + k];
    return p;
}

// Mesh::laplacian_smoothing (mesh.cpp:534-606) on the design's target mesh
// first_target_cells (which 1) or target_cells (which 0), flattened: for each
// vertex its quad count, then that many quads of 4 points {x, y, z}, in the
// reference's own listing order; returns the number of doubles written.
long ref_get_target_cells(void* h, int which, double* out) {
    auto* cd = static_cast<Caustic_design*>(h);
    const auto& cells = which ? cd->first_target_cells : cd->target_cells;
    long o = 0;
    for (const auto& cell : cells) {
        out[o++] = (double)cell.size();
        for (const auto& quad : cell)
            for (const auto& pt : quad)
                for (int k = 0; k < 3; ++k) out[o++] = pt[k];
    }
    return o;
}

// Caustic_design::calculate_vertex_normal on caller points (V x 3)
void ref_vertex_normal(void* h, const double* xyz, int vertex, double* out) {
    auto* cd = static_cast<Caustic_design*>(h);
    std::vector<std::vector<double>> pts(cd->mesh->source_points.size(), std::vector<double>(3));
    for (size_t i = 0; i < pts.size(); ++i)
        for (int k = 0; k < 3; ++k) pts[i][k] = xyz[i * 3 + k];
    const std::vector<double> n = cd->calculate_vertex_normal(pts, vertex);
    for (int k = 0; k < 3; ++k) out[k] = n[k];
}

void ref_laplacian_smoothing(void* h, double factor) {
    auto* cd = static_cast<Caustic_design*>(h);
    cd->mesh->laplacian_smoothing(cd->mesh->target_points, factor);
}

// The writers run through the public Mesh methods, which supply the
// reference's own (private) structured triangulation (mesh.cpp:57-63).
void ref_save_solid_obj(const char* path, const double* front, const double* back, int res_x,
                        int res_y, double thickness, double width, double height) {
    Quiet q;
    Mesh m(width, height, res_x, res_y);
    m.target_points = to_points(front, res_x * res_y, 3);
    m.source_points = to_points(back, res_x * res_y, 3);
    m.save_solid_obj_target(thickness, path);  // mesh.cpp:757-759
}

void ref_export_grid_svg(const char* path, const double* pts, int stride, int res_x, int res_y,
                         double width, double height, double stroke) {
    Quiet q;
    Mesh m(width, height, res_x, res_y);
    m.target_points = to_points(pts, res_x * res_y, stride);
    m.export_paramererization_to_svg(path, stroke);  // mesh.cpp:608-610
}

void ref_export_triangles_svg(const char* path, const double* pts, int stride, int res_x,
                              int res_y, double width, double height, double stroke) {
    Quiet q;
    Mesh m(width, height, res_x, res_y);
    m.target_points = to_points(pts, res_x * res_y, stride);
    m.export_to_svg(path, stroke);  // mesh.cpp:101-103
}

}  // extern "C"
