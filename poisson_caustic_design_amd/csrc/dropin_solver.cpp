// poisson_solver drop-in (src/solver.h:8, src/solver.cpp:70-147) over the C ABI.
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "cd_api.h"
#include "solver.h"

double solver_progress = 0.0;

namespace {

cd_context* shared_context() {
    static std::once_flag once;
    static cd_context* ctx = nullptr;
    static int status = CD_OK;
    std::call_once(once, [] {
        const char* env = std::getenv("CD_DEVICE");
        status = cd_context_create(env ? std::atoi(env) : 0, &ctx);
    });
    if (!ctx) throw std::runtime_error(std::string("poisson_solver: ") + cd_status_string(status));
    return ctx;
}

}  // namespace

void poisson_solver(std::vector<std::vector<double>>& D, std::vector<std::vector<double>>& phi,
                    int width, int height, int max_iterations, double convergence_threshold,
                    int max_threads) {
    // zero tiles when max_threads <= 0: the reference sweeps nothing (solver.cpp:73-83)
    if (max_threads <= 0 || max_iterations <= 0) return;
    if ((int)D.size() < height || (int)phi.size() < height)
        throw std::invalid_argument("poisson_solver: grid smaller than height");
    std::vector<double> d((size_t)width * height), p((size_t)width * height);
    for (int y = 0; y < height; ++y) {
        if ((int)D[y].size() < width || (int)phi[y].size() < width)
            throw std::invalid_argument("poisson_solver: row shorter than width");
        for (int x = 0; x < width; ++x) {
            d[(size_t)y * width + x] = D[y][x];
            p[(size_t)y * width + x] = phi[y][x];
        }
    }
    cd_solve_opts o;
    cd_solve_opts_default(&o);
    o.max_sweeps = max_iterations;
    o.stop = CD_STOP_MAX_UPDATE;
    o.tol = convergence_threshold;
    o.omega = -1.0;  // the reference's over-relaxation (solver.cpp:71)
    cd_solve_stats st;
    cd_context* ctx = shared_context();
    int rc = cd_poisson_solve(ctx, d.data(), p.data(), width, height, &o, &st);
    if (rc != CD_OK && rc != CD_ERR_NOT_CONVERGED && rc != CD_ERR_NAN)
        throw std::runtime_error(std::string("poisson_solver: ") + cd_status_string(rc) + ": " +
                                 cd_context_last_error(ctx));
    solver_progress = (double)st.sweeps / (double)max_iterations;
    for (int y = 0; y < height; ++y)
        for (int x = 0; x < width; ++x) phi[y][x] = p[(size_t)y * width + x];
}
