// A caller written against the reference's own headers and call sequence
// (main.cpp:218-269, src/solver.h), compiled unchanged against the drop-in
// headers in include/ and linked with libcaustic_amd.so.  Used by
// tests/test_gpu_dropin.py.
//
//   dropin_main <mesh_res> <out_dir/>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "caustic_design.h"
#include "solver.h"

int main(int argc, char** argv) {
    if (argc < 3) return 1;
    const int mesh_resolution = std::atoi(argv[1]);
    const std::string output_path = argv[2];
    const int N = 4 * mesh_resolution;
    // a deterministic target: two Gaussian blobs on a dim background
    std::vector<std::vector<double>> pixels(N, std::vector<double>(N));
    for (int y = 0; y < N; ++y)
        for (int x = 0; x < N; ++x) {
            const double u = (x + 0.5) / N, v = (y + 0.5) / N;
            pixels[y][x] = 0.1 + std::exp(-((u - 0.3) * (u - 0.3) + (v - 0.4) * (v - 0.4)) / 0.01) +
                           0.5 * std::exp(-((u - 0.7) * (u - 0.7) + (v - 0.6) * (v - 0.6)) / 0.02);
        }

    {   // the exact input, for the comparison made by the test
        FILE* f = std::fopen((output_path + "image.f64").c_str(), "wb");
        if (!f) return 1;
        for (auto& row : pixels) std::fwrite(row.data(), sizeof(double), row.size(), f);
        std::fclose(f);
    }

    Caustic_design caustic_design;
    caustic_design.set_mesh_resolution(mesh_resolution, mesh_resolution);
    caustic_design.set_domain_resolution(N, N);
    caustic_design.set_mesh_size(1.0, 1.0);
    caustic_design.set_lens_focal_length(1.5);
    caustic_design.set_lens_thickness(0.2);
    caustic_design.set_solver_max_threads(1);
    caustic_design.initialize_solvers(pixels);

    // the dual cells of the initial mesh, flattened as: per vertex its quad
    // count, then the quads' 4 points {x, y, z}
    auto dump_cells = [&](const std::vector<std::vector<polygon_t>>& cells, const char* name) {
        FILE* f = std::fopen((output_path + name).c_str(), "wb");
        if (!f) std::exit(1);
        for (const auto& cell : cells) {
            const double n = (double)cell.size();
            std::fwrite(&n, sizeof(double), 1, f);
            for (const auto& quad : cell)
                for (const auto& pt : quad) std::fwrite(pt.data(), sizeof(double), 3, f);
        }
        std::fclose(f);
    };
    auto dump_grid = [&](const std::vector<std::vector<double>>& g, FILE* f) {
        for (const auto& row : g) std::fwrite(row.data(), sizeof(double), row.size(), f);
    };
    dump_cells(caustic_design.first_target_cells, "cells_first.f64");

    int steps = 0;
    for (int itr = 0; itr < 50; itr++) {
        double step_size = caustic_design.perform_transport_iteration();
        ++steps;
        if (itr == 0) {  // the first step's cells, phi and gradient
            dump_cells(caustic_design.target_cells, "cells_step1.f64");
            FILE* f = std::fopen((output_path + "phi_grad1.f64").c_str(), "wb");
            if (!f) return 1;
            dump_grid(caustic_design.phi, f);
            dump_grid(caustic_design.gradient[0], f);
            dump_grid(caustic_design.gradient[1], f);
            std::fclose(f);
        }
        if (step_size < (double)0.01f) break;
    }
    for (int itr = 0; itr < 3; itr++) caustic_design.perform_height_map_iteration(itr);
    caustic_design.save_solid_obj_source(output_path + "output.obj");
    {   // vertex normals of the final lens surface at every 7th vertex
        FILE* f = std::fopen((output_path + "normals.f64").c_str(), "wb");
        if (!f) return 1;
        auto& pts = caustic_design.mesh->source_points;
        for (int v = 0; v < (int)pts.size(); v += 7) {
            const std::vector<double> n = caustic_design.calculate_vertex_normal(pts, v);
            std::fwrite(n.data(), sizeof(double), 3, f);
        }
        std::fclose(f);
        FILE* g = std::fopen((output_path + "source_points.f64").c_str(), "wb");
        if (!g) return 1;
        for (auto& p : pts) std::fwrite(p.data(), sizeof(double), 3, g);
        std::fclose(g);
    }

    double hs = 0.0;
    for (auto& row : caustic_design.h)
        for (double v : row) hs += std::fabs(v);
    std::printf("steps %d\nh_abs_sum %.17g\n", steps, hs);
    std::printf("mesh_points %zu\n", caustic_design.mesh->source_points.size());

    // the free solver entry point (src/solver.h:8) on a small problem
    const int w = 24, h = 16;
    std::vector<std::vector<double>> D(h, std::vector<double>(w)), phi(h, std::vector<double>(w, 0.0));
    double mean = 0.0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) mean += (D[y][x] = std::cos(3.14159265358979 * (x + 0.5) / w) +
                                                       0.5 * std::cos(2 * 3.14159265358979 * (y + 0.5) / h));
    mean /= w * h;
    for (auto& row : D)
        for (double& v : row) v -= mean;
    poisson_solver(D, phi, w, h, 100000, 1e-10, 1);
    double rmax = 0.0;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            double s = 0.0, cnt = 0.0;
            if (x > 0) { s += phi[y][x - 1]; cnt += 1; }
            if (x < w - 1) { s += phi[y][x + 1]; cnt += 1; }
            if (y > 0) { s += phi[y - 1][x]; cnt += 1; }
            if (y < h - 1) { s += phi[y + 1][x]; cnt += 1; }
            rmax = std::fmax(rmax, std::fabs(s - cnt * phi[y][x] - D[y][x]));
        }
    std::printf("poisson_max_residual %.3e\n", rmax);
    return 0;
}
