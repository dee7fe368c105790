/*
 * caustic_design.h -- drop-in replacement of the reference's orchestrator
 * class (dylanmsu/poisson_caustic_design src/caustic_design.h:7-66).
 *
 * Same class name, setters, step methods and public data members that
 * callers read (main.cpp:224-269 reads h; tools read phi, errors, raster,
 * mesh->target_points, mesh->source_points).  State lives on the GPU (C ABI,
 * cd_api.h); after every call the public members are refreshed from the
 * device ("host mirrors").  set_host_mirroring(false) keeps only the scalar
 * results and the vertex arrays current, for callers that do not read the
 * grids.
 *
 * Errors: a raster miss, which the reference answers with exit(0)
 * (src/mesh.cpp:277-279), throws std::runtime_error here unless
 * set_exit_on_miss(true) restores the reference's exit(0).
 */
#ifndef CAUSTIC_DESIGN_H
#define CAUSTIC_DESIGN_H

#include <string>
#include <vector>

#include "solver.h"

typedef std::vector<double> point_t;   /* src/mesh.h:17 */
typedef std::vector<point_t> polygon_t; /* src/mesh.h:18 */

struct cd_context;
struct cd_design;

/* The parts of class Mesh (src/mesh.h:31-104) that callers read: the point
 * sets, sizes, the structured triangulation and the index helper.  The
 * reference's other Mesh methods are the internals of the two steps
 * (dual cells, BVH rasterization, step_grid, inverse map, normals, height
 * setting), which run on the device inside perform_transport_iteration and
 * perform_height_map_iteration; their results are the Caustic_design members. */
class Mesh {
public:
    std::vector<point_t> source_points;
    std::vector<point_t> target_points;
    double width = 0.0;
    double height = 0.0;
    int res_x = 0;
    int res_y = 0;
    std::vector<std::vector<int>> triangles; /* structured, src/mesh.cpp:57-63 */

    /* left/right/top/bottom neighbour of a vertex, -1 at the border (mesh.cpp:761-788) */
    void get_vertex_neighbor_ids(int vertex_id, int &left_vertex, int &right_vertex,
                                 int &top_vertex, int &bottom_vertex) const;
};

class Caustic_design {
public:
    Caustic_design();
    ~Caustic_design();

    Mesh *mesh;
    std::vector<std::vector<double>> phi;
    std::vector<double> errors;
    /* partitioned dual cells of the target mesh as the last transport step
     * found it (caustic_design.cpp:195-196), and of the initial mesh
     * (caustic_design.cpp:343-346); filled while host mirroring is on */
    std::vector<std::vector<polygon_t>> target_cells;
    std::vector<std::vector<polygon_t>> first_target_cells;
    std::vector<double> target_areas;
    std::vector<std::vector<double>> pixels;
    std::vector<std::vector<double>> raster;
    /* gradient[0], gradient[1] of phi (caustic_design.cpp:225) */
    std::vector<std::vector<std::vector<double>>> gradient;
    std::vector<std::vector<double>> h;
    std::vector<std::vector<double>> divergence;
    std::vector<std::vector<double>> norm_x;
    std::vector<std::vector<double>> norm_y;
    std::vector<std::vector<double>> vertex_gradient;
    std::vector<std::vector<double>> normals;

    int mesh_res_x;
    int mesh_res_y;
    int resolution_x;
    int resolution_y;
    double width;
    double height;
    double focal_l;
    double thickness;
    int nthreads;

    /* Angle-weighted normal of vertex vertex_index of a res_x x res_y grid
     * of points (caustic_design.cpp:63-143); computed on the host from the
     * caller's points (one vertex; no device state involved). */
    std::vector<double> calculate_vertex_normal(std::vector<point_t> &points, int vertex_index);

    double perform_transport_iteration();
    void perform_height_map_iteration(int itr);
    void initialize_solvers(std::vector<std::vector<double>> image);

    void set_mesh_resolution(int width, int heigth);
    void set_domain_resolution(int width, int heigth);
    void set_mesh_size(double width, double heigth);
    void set_lens_focal_length(double focal_length);
    void set_lens_thickness(double thickness);
    void set_solver_max_threads(int n_threads);

    void save_solid_obj_target(const std::string &filename);
    void save_solid_obj_source(const std::string &filename);
    void export_paramererization_to_svg(const std::string &filename, double line_width);
    void export_inverted_transport_map(std::string filename, double stroke_width);

    /* --- additions (no reference counterpart) --- */
    void set_host_mirroring(bool on);
    void set_exit_on_miss(bool on);
    /* Stop rule of the two Poisson solves: false (default) keeps the
     * reference's max-update tolerances 1e-7 / 1e-8; true switches both to
     * the FP64 relative residual tolerance rel_tol. */
    void set_residual_stop(bool on, double rel_tol);
    cd_design *device_design() { return design_; }

private:
    cd_context *ctx_ = nullptr;
    cd_design *design_ = nullptr;
    bool mirror_ = true;
    bool exit_on_miss_ = false;
    bool residual_stop_ = false;
    double rel_tol_ = 1e-6;
    std::vector<unsigned char> cell_order_;  /* per vertex: quad slots in the
                                               reference's listing order */
    void refresh_vertices();
    void refresh_grids();
    void refresh_cells(int initial, std::vector<std::vector<polygon_t>> &cells);
};

#endif  // CAUSTIC_DESIGN_H
