// Drives the C++ mirror of the reference host API (include/ndg_sem.hpp) the way the reference's
// src/main.cpp does: tables, mesh, initial conditions, solver, solve, solution. Writes the final
// packed p, u, v as raw doubles to argv[1], the output-point x, y, p, u, v to argv[1].out, VTK
// files to argv[1]_vtk/ and argv[1]_vtk2/, and prints the step count and final time.
//   host_api_main <out.bin> <resolution> <N> <t_end> [mesh.su2 [TAG_A:TAG_B ...]]
// The mesh is the square generator's [-1,1]^2, periodic in x and y, or the given SU2 file (the
// reference's Mesh2D_t(path, ...) constructor; with marker pairs, read_su2 joins them
// periodically), with one Gaussian pulse.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "ndg_sem.hpp"

using namespace SEM::Device;

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s out.bin resolution N t_end [mesh.su2 [TAG_A:TAG_B ...]]\n", argv[0]);
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
        std::vector<std::pair<std::string, std::string>> periodic;  // SU2 marker pairs joined periodically
        for (int i = 6; i < argc; ++i) {
            const std::string spec = argv[i];
            const size_t colon = spec.find(':');
            if (colon == std::string::npos) return 2;
            periodic.emplace_back(spec.substr(0, colon), spec.substr(colon + 1));
        }
        Meshes::Mesh2D_t mesh =
            argc == 6   ? Meshes::Mesh2D_t(std::filesystem::path(argv[5]), N, N, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, NDG)
            : argc > 6  ? Meshes::Mesh2D_t(Meshes::read_su2(argv[5], periodic), N, N, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, NDG)
                        : Meshes::Mesh2D_t(Meshes::square_mesh(params), N, N, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, NDG);
        mesh.initial_conditions({{0.1, -0.2, 0.2, 1.0}});
        // output events at t_end / 2 and t_end (the reference's write_complete_data calls)
        Solvers::Solver2D_t solver(0.5, {t_end / 2, t_end}, 0.0);
        std::vector<double> events;
        solver.solve(NDG, mesh, [&](double time, Meshes::Mesh2D_t &) { events.push_back(time); });
        std::printf("events");
        for (double e : events) std::printf(" %.17g", e);
        std::printf("\n");
        std::vector<double> p, u, v;
        mesh.get_solution(p, u, v);
        FILE *f = std::fopen(argv[1], "wb");
        if (!f) return 3;
        std::fwrite(p.data(), sizeof(double), p.size(), f);
        std::fwrite(u.data(), sizeof(double), u.size(), f);
        std::fwrite(v.data(), sizeof(double), v.size(), f);
        std::fclose(f);
        // Mesh2D_t::get_solution: the output-point values, written after the state as x, y, p, u, v
        mesh.get_solution(NDG);
        const std::string out2 = std::string(argv[1]) + ".out";
        FILE *g = std::fopen(out2.c_str(), "wb");
        if (!g) return 3;
        for (const std::vector<double> *a : {&mesh.x_output_, &mesh.y_output_, &mesh.p_output_, &mesh.u_output_,
                                             &mesh.v_output_})
            std::fwrite(a->data(), sizeof(double), a->size(), g);
        std::fclose(g);
        std::printf("steps %lld time %.17g dofs %zu\n", (long long)solver.n_steps_, solver.time_, 3 * p.size());
        // The reference's output path: the final state through DataWriter_t (write_complete_data,
        // into argv[1]_vtk/), then a solve from t = 0 with a writer as its data_writer (output
        // times 0 and t_end / 8, into argv[1]_vtk2/).
        SEM::Helpers::DataWriter_t writer(std::string(argv[1]) + "_vtk/output.pvtu");
        mesh.write_complete_data(solver.time_, NDG, writer);
        SEM::Helpers::DataWriter_t writer2(std::string(argv[1]) + "_vtk2/output.vtu");
        Solvers::Solver2D_t solver2(0.5, {0.0, t_end / 8}, 0.0);
        solver2.solve(NDG, mesh, writer2);
        std::printf("written %.17g\n", solver2.time_);
    } catch (const Error &e) {
        std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
