// Caustic_design drop-in (src/caustic_design.h:7-66, src/caustic_design.cpp)
// over the C ABI.  All state lives on the device; public members are host
// mirrors refreshed after each call.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "caustic_design.h"
#include "cd_api.h"

namespace {

void check(cd_context* ctx, int rc, const char* what) {
    if (rc == CD_OK) return;
    throw std::runtime_error(std::string(what) + ": " + cd_status_string(rc) + ": " +
                             (ctx ? cd_context_last_error(ctx) : ""));
}

std::vector<std::vector<double>> to_grid(const std::vector<double>& a, int w, int h) {
    std::vector<std::vector<double>> g(h, std::vector<double>(w));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) g[y][x] = a[(size_t)y * w + x];
    return g;
}

}  // namespace

// constructor defaults as caustic_design.cpp:3-14
Caustic_design::Caustic_design()
    : mesh(nullptr), mesh_res_x(0), mesh_res_y(0), resolution_x(0), resolution_y(0), width(0.0),
      height(0.0), focal_l(0.0), thickness(0.0), nthreads(0) {}

Caustic_design::~Caustic_design() {
    cd_design_destroy(design_);
    cd_context_destroy(ctx_);
    delete mesh;
}

void Caustic_design::set_mesh_resolution(int w, int h) { mesh_res_x = w; mesh_res_y = h; }
void Caustic_design::set_domain_resolution(int w, int h) { resolution_x = w; resolution_y = h; }
void Caustic_design::set_mesh_size(double w, double h) { width = w; height = h; }
void Caustic_design::set_lens_focal_length(double f) { focal_l = f; }
void Caustic_design::set_lens_thickness(double t) { thickness = t; }
void Caustic_design::set_solver_max_threads(int n) { nthreads = n; }
void Caustic_design::set_host_mirroring(bool on) { mirror_ = on; }
void Caustic_design::set_exit_on_miss(bool on) { exit_on_miss_ = on; }
void Caustic_design::set_residual_stop(bool on, double rel_tol) {
    residual_stop_ = on;
    rel_tol_ = rel_tol;
}

void Caustic_design::initialize_solvers(std::vector<std::vector<double>> image) {
    if (!ctx_) {
        const char* env = std::getenv("CD_DEVICE");
        check(nullptr, cd_context_create(env ? std::atoi(env) : 0, &ctx_), "cd_context_create");
    }
    cd_design_destroy(design_);
    design_ = nullptr;
    cd_design_params p;
    cd_design_params_default(&p);
    p.mesh_res_x = mesh_res_x;
    p.mesh_res_y = mesh_res_y;
    p.res_x = resolution_x;
    p.res_y = resolution_y;
    p.width = width;
    p.height = height;
    p.focal_l = focal_l;
    p.thickness = thickness;
    if (residual_stop_) {
        p.phi_solve.stop = CD_STOP_RESIDUAL;
        p.phi_solve.tol = rel_tol_;
        p.phi_solve.omega = 0.0;
        p.h_solve.stop = CD_STOP_RESIDUAL;
        p.h_solve.tol = rel_tol_;
        p.h_solve.omega = 0.0;
    }
    check(ctx_, cd_design_create(ctx_, &p, &design_), "cd_design_create");
    // samples in overlapping triangles of a folded mesh take the triangle the
    // reference's BVH query returns (bvh.cpp:196-266), as in Mesh
    check(ctx_, cd_design_set_first_hit(design_, 1), "cd_design_set_first_hit");
    if ((int)image.size() < resolution_y) throw std::invalid_argument("image rows < resolution_y");
    std::vector<double> img((size_t)resolution_x * resolution_y);
    for (int y = 0; y < resolution_y; ++y) {
        if ((int)image[y].size() < resolution_x) throw std::invalid_argument("image row too short");
        for (int x = 0; x < resolution_x; ++x) img[(size_t)y * resolution_x + x] = image[y][x];
    }
    check(ctx_, cd_design_init(design_, img.data()), "initialize_solvers");
    delete mesh;
    mesh = new Mesh();
    mesh->width = width;
    mesh->height = height;
    mesh->res_x = mesh_res_x;
    mesh->res_y = mesh_res_y;
    for (int i = 0; i < mesh_res_y - 1; ++i)
        for (int j = 0; j < mesh_res_x - 1; ++j) {
            int idx = i * mesh_res_x + j;
            mesh->triangles.push_back({idx, idx + 1, idx + mesh_res_x});
            mesh->triangles.push_back({idx + mesh_res_x, idx + 1, idx + mesh_res_x + 1});
        }
    refresh_vertices();
    cell_order_.clear();
    gradient.clear();
    if (mirror_) {
        refresh_cells(1, first_target_cells);
        target_cells = first_target_cells;
    }
    std::vector<double> ta(cd_design_field_size(design_, CD_FIELD_TARGET_AREAS));
    check(ctx_, cd_design_get(design_, CD_FIELD_TARGET_AREAS, ta.data(), ta.size()), "get");
    target_areas = ta;
    if (mirror_) {
        std::vector<double> g(cd_design_field_size(design_, CD_FIELD_PIXELS));
        check(ctx_, cd_design_get(design_, CD_FIELD_PIXELS, g.data(), g.size()), "get");
        pixels = to_grid(g, resolution_x, resolution_y);
        refresh_grids();
    }
}

void Caustic_design::refresh_vertices() {
    const size_t V = (size_t)mesh_res_x * mesh_res_y;
    std::vector<double> t(3 * V), s(3 * V);
    check(ctx_, cd_design_get(design_, CD_FIELD_TARGET_POINTS, t.data(), t.size()), "get");
    check(ctx_, cd_design_get(design_, CD_FIELD_SOURCE_POINTS, s.data(), s.size()), "get");
    mesh->target_points.assign(V, point_t(3));
    mesh->source_points.assign(V, point_t(3));
    for (size_t i = 0; i < V; ++i)
        for (int k = 0; k < 3; ++k) {
            mesh->target_points[i][k] = t[3 * i + k];
            mesh->source_points[i][k] = s[3 * i + k];
        }
}

// Dual cells from the device, each vertex's quads listed as the reference
// lists them: Mesh::find_adjacent_elements (mesh.cpp:121-147) inserts the
// vertex's triangle indices, ascending, into a std::unordered_set<int> and
// copies it out in iteration order.  The same container on the same
// insertion sequence gives that order here; it depends on the topology only,
// so it is worked out once per design.
void Caustic_design::refresh_cells(int initial, std::vector<std::vector<polygon_t>> &cells) {
    const size_t V = (size_t)mesh_res_x * mesh_res_y;
    std::vector<int> tri(6 * V);
    std::vector<double> q(72 * V);
    check(ctx_, cd_design_get_dual_cells(design_, initial, tri.data(), q.data(), V),
          "cd_design_get_dual_cells");
    if (cell_order_.size() != 6 * V) {
        cell_order_.assign(6 * V, 0xFF);
        for (size_t v = 0; v < V; ++v) {
            std::unordered_set<int> adjacent;
            for (int k = 0; k < 6 && tri[6 * v + k] >= 0; ++k) adjacent.insert(tri[6 * v + k]);
            int pos = 0;
            for (int t : adjacent)
                for (int k = 0; k < 6; ++k)
                    if (tri[6 * v + k] == t) cell_order_[6 * v + pos++] = (unsigned char)k;
        }
    }
    cells.assign(V, std::vector<polygon_t>());
    for (size_t v = 0; v < V; ++v) {
        for (int pos = 0; pos < 6 && cell_order_[6 * v + pos] != 0xFF; ++pos) {
            const double* o = &q[(6 * v + cell_order_[6 * v + pos]) * 12];
            polygon_t quad(4, point_t(3));
            for (int c = 0; c < 4; ++c)
                for (int k = 0; k < 3; ++k) quad[c][k] = o[3 * c + k];
            cells[v].push_back(quad);
        }
    }
}

// Mesh::get_vertex_neighbor_ids (mesh.cpp:761-788)
void Mesh::get_vertex_neighbor_ids(int vertex_id, int &left_vertex, int &right_vertex,
                                   int &top_vertex, int &bottom_vertex) const {
    const int y = vertex_id / res_x, x = vertex_id % res_x;
    left_vertex = x != 0 ? y * res_x + (x - 1) : -1;
    top_vertex = y != 0 ? (y - 1) * res_x + x : -1;
    right_vertex = x != res_x - 1 ? y * res_x + (x + 1) : -1;
    bottom_vertex = y != res_y - 1 ? (y + 1) * res_x + x : -1;
}

namespace {

// unit vector of a 3-vector, each component divided by the length
// (normalize, utils.cpp:370-384)
void unit3(const double a[3], double out[3]) {
    const double len = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    for (int i = 0; i < 3; ++i) out[i] = a[i] / len;
}

// normal of triangle (p1, p2, p3) and its angle at p1
// (calculate_angle_and_normal_from_triangle, utils.cpp:386-405)
void angle_and_normal(const point_t &p1, const point_t &p2, const point_t &p3, double n[3],
                      double &angle) {
    const double u[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    const double v[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
    const double c[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2],
                         u[0] * v[1] - u[1] * v[0]};
    unit3(c, n);
    double nu[3], nv[3];
    unit3(u, nu);
    unit3(v, nv);
    const double r = nu[0] * nv[0] + nu[1] * nv[1] + nu[2] * nv[2];
    angle = r <= -1.0 ? M_PI : (r >= 1.0 ? 0.0 : std::acos(r));
}

}  // namespace

// Caustic_design::calculate_vertex_normal (caustic_design.cpp:63-143): the
// angle-weighted mean of the normals of the four corner triangles around the
// vertex that exist, with x and y negated after normalising.
std::vector<double> Caustic_design::calculate_vertex_normal(std::vector<point_t> &points,
                                                            int vertex_index) {
    if (!mesh) throw std::logic_error("calculate_vertex_normal before initialize_solvers");
    int l = 0, r = 0, t = 0, b = 0;
    mesh->get_vertex_neighbor_ids(vertex_index, l, r, t, b);
    double acc[3] = {0.0, 0.0, 0.0};
    const int pairs[4][2] = {{l, t}, {b, l}, {r, b}, {t, r}};
    const bool have[4] = {l != -1 && t != -1, l != -1 && b != -1, r != -1 && b != -1,
                          r != -1 && t != -1};
    for (int k = 0; k < 4; ++k) {
        if (!have[k]) continue;
        double n[3], angle;
        angle_and_normal(points[vertex_index], points[pairs[k][0]], points[pairs[k][1]], n, angle);
        for (int i = 0; i < 3; ++i) acc[i] += n[i] * angle;
    }
    std::vector<double> out = {acc[0], acc[1], acc[2]};
    const double mag = std::sqrt(acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2]);
    if (mag > 1e-12) {
        out[0] /= -mag;
        out[1] /= -mag;
        out[2] /= mag;
    }
    return out;
}

void Caustic_design::refresh_grids() {
    auto grid = [&](int field) {
        std::vector<double> g(cd_design_field_size(design_, field));
        check(ctx_, cd_design_get(design_, field, g.data(), g.size()), "get");
        return to_grid(g, resolution_x, resolution_y);
    };
    phi = grid(CD_FIELD_PHI);
    h = grid(CD_FIELD_H);
    raster = grid(CD_FIELD_RASTER);
    divergence = grid(CD_FIELD_DIVERGENCE);
    norm_x = grid(CD_FIELD_NORM_X);
    norm_y = grid(CD_FIELD_NORM_Y);
    const size_t V = (size_t)mesh_res_x * mesh_res_y;
    std::vector<double> v(V), v2(2 * V);
    check(ctx_, cd_design_get(design_, CD_FIELD_ERRORS, v.data(), V), "get");
    errors = v;
    check(ctx_, cd_design_get(design_, CD_FIELD_VERTEX_GRADIENT, v2.data(), 2 * V), "get");
    vertex_gradient.assign(2, std::vector<double>(V));
    for (size_t i = 0; i < V; ++i) {
        vertex_gradient[0][i] = v2[2 * i];
        vertex_gradient[1][i] = v2[2 * i + 1];
    }
    check(ctx_, cd_design_get(design_, CD_FIELD_NORMALS, v2.data(), 2 * V), "get");
    normals.assign(2, std::vector<double>(V));
    for (size_t i = 0; i < V; ++i) {
        normals[0][i] = v2[i];
        normals[1][i] = v2[V + i];
    }
}

double Caustic_design::perform_transport_iteration() {
    if (!design_) throw std::logic_error("perform_transport_iteration before initialize_solvers");
    // the step's dual cells are those of the mesh it starts from (caustic_design.cpp:195-196)
    if (mirror_) refresh_cells(0, target_cells);
    double step = 0.0;
    int rc = cd_design_transport_step(design_, &step);
    if (rc == CD_ERR_RASTER_MISS && exit_on_miss_) {
        std::printf("interpolation miss!\r\n");
        std::exit(0);
    }
    check(ctx_, rc, "perform_transport_iteration");
    refresh_vertices();
    if (mirror_) {
        refresh_grids();
        const size_t G = (size_t)resolution_x * resolution_y;
        std::vector<double> g(2 * G);
        check(ctx_, cd_design_get(design_, CD_FIELD_GRADIENT, g.data(), g.size()), "get");
        std::vector<double> gx(g.begin(), g.begin() + G), gy(g.begin() + G, g.end());
        gradient = {to_grid(gx, resolution_x, resolution_y), to_grid(gy, resolution_x, resolution_y)};
    }
    return step;
}

void Caustic_design::perform_height_map_iteration(int /*itr*/) {
    if (!design_) throw std::logic_error("perform_height_map_iteration before initialize_solvers");
    double upd = 0.0;
    int rc = cd_design_height_step(design_, &upd);
    if (rc == CD_ERR_RASTER_MISS && exit_on_miss_) {
        std::printf("interpolation miss!\r\n");
        std::exit(0);
    }
    check(ctx_, rc, "perform_height_map_iteration");
    std::printf("height max update %.5e\r\n", upd);
    refresh_vertices();
    if (mirror_) refresh_grids();
}

// file outputs (caustic_design.cpp:20-26,55-61): the reference's writers'
// formats, with the thickness member as it is at call time
void Caustic_design::save_solid_obj_target(const std::string& filename) {
    if (!design_) throw std::logic_error("save_solid_obj_target before initialize_solvers");
    const size_t V = (size_t)mesh_res_x * mesh_res_y;
    std::vector<double> tgt(3 * V), src(3 * V);
    check(ctx_, cd_design_get(design_, CD_FIELD_TARGET_POINTS, tgt.data(), 3 * V), "get");
    check(ctx_, cd_design_get(design_, CD_FIELD_SOURCE_POINTS, src.data(), 3 * V), "get");
    check(ctx_, cd_obj_write_solid(filename.c_str(), tgt.data(), src.data(), mesh_res_x,
                                   mesh_res_y, thickness, width, height), "save_solid_obj_target");
}
void Caustic_design::save_solid_obj_source(const std::string& filename) {
    if (!design_) throw std::logic_error("save_solid_obj_source before initialize_solvers");
    const size_t V = (size_t)mesh_res_x * mesh_res_y;
    std::vector<double> src(3 * V);
    check(ctx_, cd_design_get(design_, CD_FIELD_SOURCE_POINTS, src.data(), 3 * V), "get");
    check(ctx_, cd_obj_write_solid(filename.c_str(), src.data(), src.data(), mesh_res_x,
                                   mesh_res_y, thickness, width, height), "save_solid_obj_source");
}
void Caustic_design::export_paramererization_to_svg(const std::string& filename,
                                                      double line_width) {
    if (!design_) throw std::logic_error("export_paramererization_to_svg before initialize_solvers");
    check(ctx_, cd_design_export_svg(design_, filename.c_str(), CD_SVG_PARAMETERIZATION,
                                     line_width), "export_paramererization_to_svg");
}
void Caustic_design::export_inverted_transport_map(std::string filename, double stroke_width) {
    if (!design_) throw std::logic_error("export_inverted_transport_map before initialize_solvers");
    check(ctx_, cd_design_export_svg(design_, filename.c_str(), CD_SVG_INVERTED_MAP, stroke_width),
          "export_inverted_transport_map");
}
