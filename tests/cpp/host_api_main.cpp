// Drives the C++ mirror of the reference host API (include/ndg_sem.hpp) the way the reference's
// src/main.cpp does: tables, mesh, initial conditions, solver, solve, solution. Writes the final
// packed p, u, v as raw doubles to argv[1] and prints the step count and final time.
//   host_api_main <out.bin> <resolution> <N> <t_end>
// The mesh is the square generator's [-1,1]^2, periodic in x and y, with one Gaussian pulse.
#include <cstdio>
#include <cstdlib>

#include "ndg_sem.hpp"

using namespace SEM::Device;

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s out.bin resolution N t_end\n", argv[0]);
        return 2;
    }
    const int res = std::atoi(argv[2]), N = std::atoi(argv[3]);
    const double t_end = std::atof(argv[4]);
    try {
        Entities::NDG_t<Polynomials::LegendrePolynomial_t> NDG(N, 5);
        ndg_square_mesh_params params{};
        params.resolution = res;
        params.x_min = -1.0; params.x_max = 1.0; params.y_min = -1.0; params.y_max = 1.0;
        for (int k = 0; k < 4; ++k) params.bc[k] = NDG_BC_WALL;
        params.x_periodic = 1;
        params.y_periodic = 1;
        Meshes::Mesh2D_t mesh(Meshes::square_mesh(params), N, N, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, NDG);
        mesh.initial_conditions({{0.1, -0.2, 0.2, 1.0}});
        Solvers::Solver2D_t solver(0.5, {t_end}, 0.0);
        solver.solve(NDG, mesh);
        std::vector<double> p, u, v;
        mesh.get_solution(p, u, v);
        FILE *f = std::fopen(argv[1], "wb");
        if (!f) return 3;
        std::fwrite(p.data(), sizeof(double), p.size(), f);
        std::fwrite(u.data(), sizeof(double), u.size(), f);
        std::fwrite(v.data(), sizeof(double), v.size(), f);
        std::fclose(f);
        std::printf("steps %lld time %.17g dofs %zu\n", (long long)solver.n_steps_, solver.time_, 3 * p.size());
    } catch (const Error &e) {
        std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
