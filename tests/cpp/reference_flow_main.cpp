// The reference program's call sequence (src/main.cpp:227-232 for the host classes,
// src/main.cu:347-353 for the device classes with a stream), compiled against the C++ mirror
// (ndg_sem.hpp) in place of the reference's headers: NDG_t, the Mesh2D_t constructor with the
// polynomial nodes and worker weights, Solver2D_t, DataWriter_t, initial_conditions(nodes),
// solve(NDG, mesh, data_writer).
//   reference_flow_main <host|device> <mesh.su2> <out.bin> <vtk output file> <N> <t_end> [max_splits adaptivity_interval]
// Writes the final packed p, u, v as raw doubles to out.bin and prints the element count, the
// solver's step count and final time.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "ndg_sem.hpp"

int main(int argc, char **argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: %s host|device mesh.su2 out.bin output.vtu N t_end [max_splits adaptivity_interval]\n",
                     argv[0]);
        return 2;
    }
    const std::string flow = argv[1];
    const std::filesystem::path mesh_file = argv[2];
    const std::filesystem::path output_file = argv[4];
    const int N_initial = std::atoi(argv[5]);
    const int N_max = argc > 8 ? N_initial + 4 : N_initial;
    const size_t n_interpolation_points = 5;
    const int max_splits = argc > 8 ? std::atoi(argv[7]) : 0;
    const size_t adaptivity_interval = argc > 8 ? (size_t)std::atoll(argv[8]) : 0;
    const size_t load_balancing_interval = 0;
    const hostFloat tolerance_min = 1e-6, tolerance_max = 1e-14, load_balancing_threshold = 1.01;
    const hostFloat CFL = 0.5, viscosity = 0.0;
    const std::vector<hostFloat> output_times{std::atof(argv[6])};
    const int global_size = 1;
    std::vector<hostFloat> worker_weights(global_size, 1);
    try {
        std::vector<double> p, u, v;
        size_t n_elements = 0;
        long long steps = 0;
        double time = 0.0;
        if (flow == "host") {
            SEM::Host::Entities::NDG_t<SEM::Host::Polynomials::LegendrePolynomial_t> NDG(N_max, n_interpolation_points);
            SEM::Host::Meshes::Mesh2D_t mesh(mesh_file, N_initial, N_max, n_interpolation_points, max_splits, adaptivity_interval, load_balancing_interval, tolerance_min, tolerance_max, load_balancing_threshold, NDG.nodes_, std::move(worker_weights));
            SEM::Host::Solvers::Solver2D_t solver(CFL, output_times, viscosity);
            SEM::Helpers::DataWriter_t data_writer(output_file);
            mesh.initial_conditions(NDG.nodes_);
            solver.solve(NDG, mesh, data_writer);
            mesh.get_solution(p, u, v);
            n_elements = mesh.n_elements_;
            steps = (long long)solver.n_steps_;
            time = solver.time_;
        } else {
            hipStream_t stream;
            if (hipStreamCreate(&stream) != hipSuccess) return 4;
            {
                SEM::Device::Entities::NDG_t<SEM::Device::Polynomials::LegendrePolynomial_t> NDG(N_max, n_interpolation_points, stream);
                SEM::Device::Meshes::Mesh2D_t mesh(mesh_file, N_initial, N_max, n_interpolation_points, max_splits, adaptivity_interval, load_balancing_interval, tolerance_min, tolerance_max, load_balancing_threshold, NDG.nodes_, std::move(worker_weights), stream);
                SEM::Device::Solvers::Solver2D_t solver(CFL, output_times, viscosity);
                SEM::Helpers::DataWriter_t data_writer(output_file);
                mesh.initial_conditions(NDG.nodes_);
                if (hipStreamSynchronize(stream) != hipSuccess) return 4;
                solver.solve(NDG, mesh, data_writer);
                if (hipStreamSynchronize(stream) != hipSuccess) return 4;
                mesh.get_solution(p, u, v);
                n_elements = mesh.n_elements_;
                steps = (long long)solver.n_steps_;
                time = solver.time_;
            }
            (void)hipStreamDestroy(stream);
        }
        FILE *f = std::fopen(argv[3], "wb");
        if (!f) return 3;
        std::fwrite(p.data(), sizeof(double), p.size(), f);
        std::fwrite(u.data(), sizeof(double), u.size(), f);
        std::fwrite(v.data(), sizeof(double), v.size(), f);
        std::fclose(f);
        std::printf("elements %zu steps %lld time %.17g\n", n_elements, steps, time);
    } catch (const SEM::Device::Error &e) {
        std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
        return 1;
    }
    return 0;
}
