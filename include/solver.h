/*
 * solver.h -- drop-in replacement of the reference's Poisson entry point
 * (dylanmsu/poisson_caustic_design src/solver.h:1-11).
 *
 * Same declaration, same argument meaning; the sweeps run on the GPU through
 * the C ABI (cd_api.h: cd_poisson_solve) with the reference's stop rule
 * "max |delta| over one sweep < convergence_threshold" (src/solver.cpp:142)
 * and its sweep cap `max_iterations`.  `max_threads` keeps the reference's
 * corner case: a value <= 0 runs zero tiles and leaves phi untouched
 * (src/solver.cpp:73-83).  The device is taken from the environment variable
 * CD_DEVICE (default 0).  HIP failures throw std::runtime_error; reaching the
 * sweep cap is silent, as in the reference.
 */
#ifndef SOLVER_H
#define SOLVER_H

#include <cmath>
#include <string>
#include <thread>
#include <vector>

void poisson_solver(std::vector<std::vector<double>> &D, std::vector<std::vector<double>> &phi,
                    int width, int height, int max_iterations, double convergence_threshold,
                    int max_threads);

/* Progress of the last solve in [0,1] (the reference declares it at
 * src/solver.h:10 but never updates it; here it is sweeps / max_iterations). */
extern double solver_progress;

#endif  // SOLVER_H
