// ndg_sem.hpp -- header-only C++ mirror of the reference host API over the C ABI (ndg_amd.h).
//
// Class and member names, constructor arguments and call order follow guillaumetousignant/NDG_cuda's
// SEM::Device classes:
//   SEM::Device::Entities::NDG_t<Polynomial>(N_max, n_interpolation_points)   include/entities/NDG_t.cuh:11-29
//   SEM::Device::Meshes::Mesh2D_t(mesh, initial_N, maximum_N, n_interpolation_points,
//       max_split_level, adaptivity_interval, load_balancing_interval, tolerance_min,
//       tolerance_max, load_balancing_threshold, NDG, stream)              include/meshes/Mesh2D_t.cuh:23
//   mesh.initial_conditions(...)                                              src/meshes/Mesh2D_t.cpp:1192-1216
//   SEM::Device::Solvers::Solver2D_t(CFL, output_times, viscosity)            include/solvers/Solver2D_t.cuh
//   solver.solve(NDG, mesh, data_writer)                                      include/solvers/Solver2D_t.tcu:7-135
//   SEM::Helpers::DataWriter_t(path)           (ndg_data_writer.hpp)           include/helpers/DataWriter_t.h:14-33
// Differences: the mesh comes from an in-memory description, the square generator or an SU2 file
// (no CGNS), the writer writes VTK XML itself (no VTK library), the mesh's write_data /
// write_complete_data take the NDG_t instead of its device tables, and errors throw
// SEM::Device::Error (carrying the status code) instead of calling exit().
#ifndef NDG_SEM_HPP
#define NDG_SEM_HPP

#include <array>
#include <cctype>
#include <cstddef>
#include <cstdint>
#include <exception>
#include <filesystem>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "ndg_amd.h"
#include "ndg_data_writer.hpp"

namespace SEM {
namespace Device {

class Error : public std::runtime_error {
  public:
    Error(ndg_status code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
    ndg_status code() const { return code_; }

  private:
    ndg_status code_;
};

inline void check(ndg_status status) {
    if (status != NDG_OK) {
        char buf[1024];
        ndg_last_error(buf, sizeof buf);
        throw Error(status, buf);
    }
}

namespace Polynomials {
// Legendre-Gauss nodes (the only polynomial family on the reference's 2D path).
struct LegendrePolynomial_t {};
}  // namespace Polynomials

namespace Entities {

template <typename Polynomial>
class NDG_t {
  public:
    NDG_t(int N_max, std::size_t n_interpolation_points)
        : N_max_(N_max), N_interpolation_points_(n_interpolation_points) {
        check(ndg_tables_create(N_max, (int64_t)n_interpolation_points, &handle_));
        std::vector<std::vector<double>> *tables[12] = {
            &nodes_, &weights_, &barycentric_weights_, &lagrange_interpolant_left_,
            &lagrange_interpolant_right_, &lagrange_interpolant_derivative_left_,
            &lagrange_interpolant_derivative_right_, &derivative_matrices_, &g_hat_derivative_matrices_,
            &derivative_matrices_hat_, &interpolation_matrices_, &polynomials_};
        for (int which = 0; which < 12; ++which) {
            tables[which]->resize(N_max + 1);
            for (int N = 0; N <= N_max; ++N) {
                int64_t count = 0;
                check(ndg_tables_get(handle_, which, N, nullptr, 0, &count));
                (*tables[which])[N].resize((std::size_t)count);
                check(ndg_tables_get(handle_, which, N, (*tables[which])[N].data(), count, nullptr));
            }
        }
    }
    // The device twin's constructor (include/entities/NDG_t.cuh:11-29) takes the rank's stream;
    // the tables live on the host here and each mesh keeps its own device copy, so it is unused.
    template <typename Stream>
    NDG_t(int N_max, std::size_t n_interpolation_points, const Stream &stream)
        : NDG_t(N_max, n_interpolation_points) {
        (void)stream;
    }
    NDG_t(const NDG_t &) = delete;
    NDG_t &operator=(const NDG_t &) = delete;
    ~NDG_t() { ndg_tables_destroy(handle_); }

    const ndg_tables *handle() const { return handle_; }

    int N_max_;
    std::size_t N_interpolation_points_;
    std::vector<std::vector<double>> nodes_;
    std::vector<std::vector<double>> weights_;
    std::vector<std::vector<double>> barycentric_weights_;
    std::vector<std::vector<double>> lagrange_interpolant_left_;
    std::vector<std::vector<double>> lagrange_interpolant_right_;
    std::vector<std::vector<double>> lagrange_interpolant_derivative_left_;
    std::vector<std::vector<double>> lagrange_interpolant_derivative_right_;
    std::vector<std::vector<double>> derivative_matrices_;
    std::vector<std::vector<double>> g_hat_derivative_matrices_;
    std::vector<std::vector<double>> derivative_matrices_hat_;
    std::vector<std::vector<double>> interpolation_matrices_;
    std::vector<std::vector<double>> polynomials_;

  private:
    ndg_tables *handle_ = nullptr;
};

}  // namespace Entities

namespace Meshes {

// Content of the mesh file the reference reads (src/meshes/Mesh2D_t.cpp:422-1048).
struct MeshDescription {
    std::vector<double> xy;         // 2 per node
    std::vector<int64_t> quads;     // 4 per element, counter-clockwise, element order = storage order
    std::vector<int32_t> quad_N;    // empty: initial_N everywhere
    std::vector<int64_t> bars;      // 2 per boundary element
    std::vector<int32_t> bar_type;  // NDG_BC_*
    std::vector<int64_t> bar_donor; // periodic donor bar or -1
};

// extras/mesh_generator/main.cpp:161-382 (square, Hilbert order, optional periodicity and jitter).
inline MeshDescription square_mesh(const ndg_square_mesh_params &params) {
    int64_t nn = 0, nq = 0, nb = 0;
    check(ndg_square_mesh_sizes(&params, &nn, &nq, &nb));
    MeshDescription m;
    m.xy.resize(2 * nn);
    m.quads.resize(4 * nq);
    m.bars.resize(2 * nb);
    m.bar_type.resize(nb);
    m.bar_donor.resize(nb);
    check(ndg_square_mesh_fill(&params, m.xy.data(), m.quads.data(), m.bars.data(), m.bar_type.data(),
                               m.bar_donor.data()));
    return m;
}

// Mesh2D_t::read_su2 (src/meshes/Mesh2D_t.cpp:78-330): nodes, QUAD_4 elements and the
// farfield / wall / inlet / symmetry markers of an SU2 file (bars in the reference's ghost order).
// With `periodic`, the named marker pairs are joined by a translation, as SU2's MARKER_PERIODIC
// does (ndg_su2_fill_periodic): their bars come last, as NDG_BC_PERIODIC with donors.
inline MeshDescription read_su2(const std::filesystem::path &filename,
                                const std::vector<std::pair<std::string, std::string>> &periodic = {}) {
    const std::string path = filename.string();
    std::vector<const char *> tags_a, tags_b;
    for (const auto &pr : periodic) {
        tags_a.push_back(pr.first.c_str());
        tags_b.push_back(pr.second.c_str());
    }
    const ndg_su2_periodic per{(int32_t)periodic.size(), tags_a.data(), tags_b.data()};
    int64_t nn = 0, nq = 0, nb = 0;
    check(ndg_su2_sizes_periodic(path.c_str(), &per, &nn, &nq, &nb));
    MeshDescription m;
    m.xy.resize(2 * nn);
    m.quads.resize(4 * nq);
    m.bars.resize(2 * nb);
    m.bar_type.resize(nb);
    m.bar_donor.resize(nb);
    check(ndg_su2_fill_periodic(path.c_str(), &per, m.xy.data(), m.quads.data(), m.bars.data(), m.bar_type.data(),
                                m.bar_donor.data()));
    return m;
}

// The file branch of the reference's constructor (src/meshes/Mesh2D_t.cpp:56-70): .su2 is read,
// any other extension (.cgns included: no CGNS library here) is refused.
inline MeshDescription read_mesh_file(const std::filesystem::path &filename) {
    std::string ext = filename.extension().string();
    for (char &c : ext) c = (char)std::tolower((unsigned char)c);
    if (ext != ".su2") throw Error(NDG_ERR_UNSUPPORTED, "Mesh2D_t: extension '" + ext + "' not read (only .su2)");
    return read_su2(filename);
}

class Mesh2D_t {
  public:
    template <typename Polynomial>
    Mesh2D_t(const MeshDescription &mesh, int initial_N, int maximum_N, std::size_t n_interpolation_points,
             int max_split_level, std::size_t adaptivity_interval, std::size_t load_balancing_interval,
             double tolerance_min, double tolerance_max, double load_balancing_threshold,
             const Entities::NDG_t<Polynomial> &NDG, void *stream = nullptr)
        : initial_N_(initial_N), maximum_N_(maximum_N), n_interpolation_points_(n_interpolation_points),
          max_split_level_(max_split_level), adaptivity_interval_(adaptivity_interval),
          load_balancing_interval_(load_balancing_interval), tolerance_min_(tolerance_min),
          tolerance_max_(tolerance_max), load_balancing_threshold_(load_balancing_threshold) {
        ndg_mesh_desc d{};
        d.n_nodes = (int64_t)(mesh.xy.size() / 2);
        d.xy = mesh.xy.data();
        d.n_quads = (int64_t)(mesh.quads.size() / 4);
        d.quads = mesh.quads.data();
        d.quad_N = mesh.quad_N.empty() ? nullptr : mesh.quad_N.data();
        d.n_bars = (int64_t)(mesh.bars.size() / 2);
        d.bars = mesh.bars.data();
        d.bar_type = mesh.bar_type.data();
        d.bar_donor = mesh.bar_donor.data();
        d.initial_N = initial_N;
        d.maximum_N = maximum_N;
        d.max_split_level = max_split_level;
        d.tolerance_min = tolerance_min;
        d.tolerance_max = tolerance_max;
        check(ndg_mesh_create(&d, NDG.handle(), &handle_));
        if (stream) check(ndg_mesh_set_stream(handle_, stream));
        if (adaptivity_interval != 0) check(ndg_mesh_set_adaptivity(handle_, (int64_t)adaptivity_interval));
        refresh();
    }

    // Sizes and element orders (they change with adapt and with an adaptive solve).
    void refresh() {
        int64_t ne = 0, faces = 0, dofs = 0;
        check(ndg_mesh_info(handle_, &ne, &faces, &dofs, nullptr, nullptr));
        n_elements_ = (std::size_t)ne;
        n_faces_ = (std::size_t)faces;
        n_dofs_per_field_ = (std::size_t)dofs;
        orders_.resize(n_elements_);
        check(ndg_mesh_orders(handle_, orders_.data()));
        n_trace_values_ = 0;
        for (int32_t N : orders_) n_trace_values_ += 12 * (std::size_t)(N + 1);
    }
    // Mesh2D_t::adapt (N_max: the tables' maximum order): splits the elements that would h-refine
    // (when any would) and raises by 2 those that would p-refine; the element count may change.
    std::int64_t adapt(int N_max) {
        int64_t raised = 0;
        check(ndg_adapt(handle_, N_max, max_split_level_, &raised));
        refresh();
        return raised;
    }
    // Mesh2D_t(path filename, ...) (include/meshes/Mesh2D_t.h:21).
    template <typename Polynomial>
    Mesh2D_t(const std::filesystem::path &filename, int initial_N, int maximum_N, std::size_t n_interpolation_points,
             int max_split_level, std::size_t adaptivity_interval, std::size_t load_balancing_interval,
             double tolerance_min, double tolerance_max, double load_balancing_threshold,
             const Entities::NDG_t<Polynomial> &NDG, void *stream = nullptr)
        : Mesh2D_t(read_mesh_file(filename), initial_N, maximum_N, n_interpolation_points, max_split_level,
                   adaptivity_interval, load_balancing_interval, tolerance_min, tolerance_max,
                   load_balancing_threshold, NDG, stream) {}
    // The reference's own argument list (include/meshes/Mesh2D_t.h:21, device twin
    // include/meshes/Mesh2D_t.cuh:23 with the stream last): the polynomial nodes of every order up
    // to N_max (NDG.nodes_) instead of the NDG_t object, and the MPI workers' weights. The mesh
    // builds its own copy of the tables for those nodes (they must be the Legendre-Gauss nodes
    // NDG_t computes, bit for bit). One GPU: worker_weights must hold exactly one weight.
    Mesh2D_t(const std::filesystem::path &filename, int initial_N, int maximum_N, std::size_t n_interpolation_points,
             int max_split_level, std::size_t adaptivity_interval, std::size_t load_balancing_interval,
             double tolerance_min, double tolerance_max, double load_balancing_threshold,
             const std::vector<std::vector<double>> &polynomial_nodes, std::vector<double> worker_weights)
        : Mesh2D_t(filename, initial_N, maximum_N, n_interpolation_points, max_split_level, adaptivity_interval,
                   load_balancing_interval, tolerance_min, tolerance_max, load_balancing_threshold,
                   Tables(polynomial_nodes, n_interpolation_points, worker_weights).ndg, nullptr) {}
    template <typename Stream>
    Mesh2D_t(const std::filesystem::path &filename, int initial_N, int maximum_N, std::size_t n_interpolation_points,
             int max_split_level, std::size_t adaptivity_interval, std::size_t load_balancing_interval,
             double tolerance_min, double tolerance_max, double load_balancing_threshold,
             const std::vector<std::vector<double>> &polynomial_nodes, std::vector<double> worker_weights,
             const Stream &stream)
        : Mesh2D_t(filename, initial_N, maximum_N, n_interpolation_points, max_split_level, adaptivity_interval,
                   load_balancing_interval, tolerance_min, tolerance_max, load_balancing_threshold,
                   Tables(polynomial_nodes, n_interpolation_points, worker_weights).ndg, (void *)stream) {}
    Mesh2D_t(const Mesh2D_t &) = delete;
    Mesh2D_t &operator=(const Mesh2D_t &) = delete;
    ~Mesh2D_t() { ndg_mesh_destroy(handle_); }

    // Mesh2D_t::initial_conditions(polynomial_nodes) (src/meshes/Mesh2D_t.cpp:1192-1216): the
    // reference plane wave at the mesh's own nodes.
    template <typename Float>
    void initial_conditions(const std::vector<std::vector<Float>> &polynomial_nodes) {
        if ((int)polynomial_nodes.size() <= maximum_N_)
            throw Error(NDG_ERR_INVALID, "Mesh2D_t::initial_conditions: fewer polynomial node sets than maximum_N + 1");
        initial_conditions();
    }

    // NDG_PRECISION_FP64 (the reference arithmetic) or NDG_PRECISION_FP32_COMPENSATED.
    void set_precision(int32_t precision) { check(ndg_mesh_set_precision(handle_, precision)); }

    // The reference plane wave g(x, 0).
    void initial_conditions() {
        ndg_ic_desc ic{NDG_IC_PLANE_WAVE, 0, nullptr, 0.0, 0.0};
        check(ndg_set_ic(handle_, &ic));
    }
    // Radial Gaussian pulses {x0, y0, width, amplitude}; u = v = 0.
    void initial_conditions(const std::vector<std::array<double, 4>> &pulses) {
        ndg_ic_desc ic{NDG_IC_PULSES, (int32_t)pulses.size(), pulses.empty() ? nullptr : pulses[0].data(), 0.0, 0.0};
        check(ndg_set_ic(handle_, &ic));
    }
    // The same pulses periodised with periods (Lx, Ly): each pulse and its 8 nearest images.
    void initial_conditions(const std::vector<std::array<double, 4>> &pulses, double Lx, double Ly) {
        ndg_ic_desc ic{NDG_IC_PERIODIC_PULSES, (int32_t)pulses.size(), pulses.empty() ? nullptr : pulses[0].data(),
                       Lx, Ly};
        check(ndg_set_ic(handle_, &ic));
    }
    // Packed fields in reference node order.
    void get_solution(std::vector<double> &p, std::vector<double> &u, std::vector<double> &v) {
        p.resize(n_dofs_per_field_);
        u.resize(n_dofs_per_field_);
        v.resize(n_dofs_per_field_);
        check(ndg_get_state(handle_, p.data(), u.data(), v.data()));
    }
    void set_solution(const std::vector<double> &p, const std::vector<double> &u, const std::vector<double> &v) {
        if (p.size() != n_dofs_per_field_ || u.size() != n_dofs_per_field_ || v.size() != n_dofs_per_field_)
            throw Error(NDG_ERR_INVALID, "Mesh2D_t::set_solution: wrong field length");
        check(ndg_set_state(handle_, p.data(), u.data(), v.data()));
    }
    // Mesh2D_t::get_solution (src/meshes/Mesh2D_t.cpp:4237-4249): fills x_output_, y_output_,
    // p_output_, u_output_, v_output_ with the values at the n x n uniform output points of every
    // element (n = the tables' n_interpolation_points), element e's point (i, j) at
    // e * n * n + i * n + j.
    template <typename Polynomial>
    void get_solution(const Entities::NDG_t<Polynomial> &NDG) {
        const std::size_t size = n_elements_ * NDG.N_interpolation_points_ * NDG.N_interpolation_points_;
        for (std::vector<double> *a : {&x_output_, &y_output_, &p_output_, &u_output_, &v_output_}) a->resize(size);
        check(ndg_get_solution(handle_, x_output_.data(), y_output_.data(), p_output_.data(), u_output_.data(),
                               v_output_.data()));
    }
    std::vector<double> x_output_, y_output_, p_output_, u_output_, v_output_;
    // Mesh2D_t::get_complete_solution (src/meshes/Mesh2D_t.cu:1861-1950): fills the output
    // members as get_solution does and the other fields (ndg_get_complete_solution: dp_dt etc. are
    // the right-hand side at `time`, the element errors those of the last estimate_error). The
    // tables are the library's; the two table arguments keep the reference's call shape.
    void get_complete_solution(double time, const std::vector<std::vector<double>> &polynomial_nodes,
                               const std::vector<std::vector<double>> &interpolation_matrices, std::vector<int> &N,
                               std::vector<double> &dp_dt, std::vector<double> &du_dt, std::vector<double> &dv_dt,
                               std::vector<double> &p_error, std::vector<double> &u_error,
                               std::vector<double> &v_error, std::vector<double> &p_sigma,
                               std::vector<double> &u_sigma, std::vector<double> &v_sigma, std::vector<int> &refine,
                               std::vector<int> &coarsen, std::vector<int> &split_level,
                               std::vector<double> &p_analytical_error, std::vector<double> &u_analytical_error,
                               std::vector<double> &v_analytical_error, std::vector<int> &status,
                               std::vector<int> &rotation) {
        (void)polynomial_nodes;
        (void)interpolation_matrices;
        const std::size_t points = n_elements_ * n_interpolation_points_ * n_interpolation_points_;
        for (std::vector<double> *a : {&x_output_, &y_output_, &p_output_, &u_output_, &v_output_, &dp_dt, &du_dt,
                                       &dv_dt, &p_analytical_error, &u_analytical_error, &v_analytical_error})
            a->resize(points);
        for (std::vector<double> *a : {&p_error, &u_error, &v_error, &p_sigma, &u_sigma, &v_sigma})
            a->resize(n_elements_);
        for (std::vector<int> *a : {&N, &refine, &coarsen, &split_level, &status, &rotation}) a->resize(n_elements_);
        ndg_complete_solution out{x_output_.data(), y_output_.data(), p_output_.data(), u_output_.data(),
                                  v_output_.data(), dp_dt.data(), du_dt.data(), dv_dt.data(),
                                  p_analytical_error.data(), u_analytical_error.data(), v_analytical_error.data(),
                                  N.data(), p_error.data(), u_error.data(), v_error.data(), p_sigma.data(),
                                  u_sigma.data(), v_sigma.data(), refine.data(), coarsen.data(),
                                  split_level.data(), status.data(), rotation.data()};
        check(ndg_get_complete_solution(handle_, time, &out));
    }
    // Mesh2D_t::write_data (src/meshes/Mesh2D_t.cpp:1601-1609): get_solution, then the writer.
    template <typename Polynomial>
    void write_data(double time, const Entities::NDG_t<Polynomial> &NDG, const Helpers::DataWriter_t &data_writer) {
        get_solution(NDG);
        data_writer.write_data(n_interpolation_points_, n_elements_, time, x_output_, y_output_, p_output_,
                               u_output_, v_output_);
    }
    // Mesh2D_t::write_complete_data (src/meshes/Mesh2D_t.cu:1861-1950): get_complete_solution,
    // then the writer with rank 0.
    template <typename Polynomial>
    void write_complete_data(double time, const Entities::NDG_t<Polynomial> &NDG,
                             const Helpers::DataWriter_t &data_writer) {
        std::vector<int> N, refine, coarsen, split_level, status, rotation;
        std::vector<double> dp_dt, du_dt, dv_dt, p_error, u_error, v_error, p_sigma, u_sigma, v_sigma, pa, ua, va;
        get_complete_solution(time, NDG.nodes_, NDG.interpolation_matrices_, N, dp_dt, du_dt, dv_dt, p_error, u_error,
                              v_error, p_sigma, u_sigma, v_sigma, refine, coarsen, split_level, pa, ua, va, status,
                              rotation);
        data_writer.write_complete_data(n_interpolation_points_, n_elements_, time, 0, x_output_, y_output_, p_output_,
                                        u_output_, v_output_, N, dp_dt, du_dt, dv_dt, p_error, u_error, v_error,
                                        p_sigma, u_sigma, v_sigma, refine, coarsen, split_level, pa, ua, va, status,
                                        rotation);
    }
    // Traces of the current state, [element][side][field][N_e+1] packed element after element.
    void interpolate_to_boundaries(std::vector<double> &traces) {
        traces.resize(n_trace_values_);
        check(ndg_get_traces(handle_, traces.data()));
    }
    // NDG_GRAPH_OFF, NDG_GRAPH_SEQUENTIAL or NDG_GRAPH_PARALLEL (captured RK steps).
    void set_graph(int32_t mode) { check(ndg_mesh_set_graph(handle_, mode)); }
    // Times the launch shapes of each order group's stage kernel and keeps the fastest.
    void autotune(int32_t repeats = 5) { check(ndg_mesh_autotune(handle_, repeats)); }
    void estimate_error(std::vector<double> &error, std::vector<double> &sigma, std::vector<int32_t> &refine,
                        std::vector<int32_t> &coarsen) {
        error.resize(n_elements_);
        sigma.resize(n_elements_);
        refine.resize(n_elements_);
        coarsen.resize(n_elements_);
        check(ndg_estimate_error(handle_, error.data(), sigma.data(), refine.data(), coarsen.data()));
    }

    ndg_mesh *handle() { return handle_; }

    std::vector<int32_t> orders_;

  private:
    // Tables for the reference constructor: NDG_t for N_max = polynomial_nodes.size() - 1, checked
    // against the given nodes; the mesh copies what it needs, so this temporary may go.
    struct Tables {
        Entities::NDG_t<Polynomials::LegendrePolynomial_t> ndg;
        Tables(const std::vector<std::vector<double>> &nodes, std::size_t n_interpolation_points,
               const std::vector<double> &worker_weights)
            : ndg(checked_N_max(nodes, worker_weights), n_interpolation_points) {
            for (std::size_t N = 0; N < nodes.size(); ++N)
                if (nodes[N] != ndg.nodes_[N])
                    throw Error(NDG_ERR_UNSUPPORTED, "Mesh2D_t: polynomial_nodes are not the Legendre-Gauss nodes of NDG_t");
        }
        static int checked_N_max(const std::vector<std::vector<double>> &nodes, const std::vector<double> &weights) {
            if (weights.size() != 1)
                throw Error(NDG_ERR_UNSUPPORTED, "Mesh2D_t: one worker per mesh (one GPU, no MPI ranks): worker_weights "
                                                 "must hold exactly one weight");
            if (nodes.empty()) throw Error(NDG_ERR_INVALID, "Mesh2D_t: no polynomial nodes");
            return (int)nodes.size() - 1;
        }
    };
    std::size_t n_trace_values_ = 0;

  public:

    int initial_N_;
    int maximum_N_;
    std::size_t n_interpolation_points_;
    int max_split_level_;
    std::size_t adaptivity_interval_;
    std::size_t load_balancing_interval_;
    double tolerance_min_;
    double tolerance_max_;
    double load_balancing_threshold_;
    std::size_t n_elements_ = 0;
    std::size_t n_faces_ = 0;
    std::size_t n_dofs_per_field_ = 0;

  private:
    ndg_mesh *handle_ = nullptr;
};

}  // namespace Meshes

namespace Solvers {

class Solver2D_t {
  public:
    Solver2D_t(double CFL, std::vector<double> output_times, double viscosity)
        : CFL_(CFL), output_times_(std::move(output_times)), viscosity_(viscosity) {}

    // Advances to output_times_.back() (include/solvers/Solver2D_t.tpp:37-72); no output is written.
    template <typename Polynomial>
    void solve(const Entities::NDG_t<Polynomial> &NDG, Meshes::Mesh2D_t &mesh) {
        (void)NDG;
        if (output_times_.empty()) throw Error(NDG_ERR_INVALID, "Solver2D_t::solve: no output times");
        check(ndg_solve(mesh.handle(), CFL_, output_times_.back(), &time_, &n_steps_));
        mesh.refresh();
    }

    // The same with the reference's output events (include/solvers/Solver2D_t.tcu:21-29, :74-83,
    // :112-127). A Helpers::DataWriter_t gets what the reference writes there: estimate_error
    // (after the first step) and write_complete_data. Any other data_writer is called as
    // data_writer(time, mesh) and may read the mesh. An exception it throws stops the solve and is
    // rethrown here.
    template <typename Polynomial, typename Writer>
    void solve(const Entities::NDG_t<Polynomial> &NDG, Meshes::Mesh2D_t &mesh, Writer &&data_writer) {
        if (output_times_.empty()) throw Error(NDG_ERR_INVALID, "Solver2D_t::solve: no output times");
        using W = std::remove_reference_t<Writer>;
        struct Context {
            const Entities::NDG_t<Polynomial> *NDG;
            Meshes::Mesh2D_t *mesh;
            W *writer;
            std::exception_ptr error;
        } ctx{&NDG, &mesh, &data_writer, nullptr};
        auto event = [](ndg_mesh *, double time, void *user) -> ndg_status {
            Context &c = *static_cast<Context *>(user);
            try {
                c.mesh->refresh();
                if constexpr (std::is_base_of_v<Helpers::DataWriter_t, std::decay_t<Writer>>) {
                    if (time > 0.0) {
                        std::vector<double> error, sigma;
                        std::vector<int32_t> refine, coarsen;
                        c.mesh->estimate_error(error, sigma, refine, coarsen);
                    }
                    c.mesh->write_complete_data(time, *c.NDG, *c.writer);
                } else {
                    (*c.writer)(time, *c.mesh);
                }
                return NDG_OK;
            } catch (...) {
                c.error = std::current_exception();
                return NDG_ERR_INVALID;
            }
        };
        const ndg_status st = ndg_solve_output(mesh.handle(), CFL_, output_times_.data(), (int64_t)output_times_.size(),
                                               event, &ctx, &time_, &n_steps_);
        if (ctx.error) std::rethrow_exception(ctx.error);
        check(st);
        mesh.refresh();
    }

    // Solver2D_t::pre_condition (include/solvers/Solver2D_t.tpp:138-236): steps with an adapt every
    // pre_condition_interval steps until n_adaptivity_steps adaptations are done, then the mesh's
    // initial condition again on the adapted mesh.
    template <typename Polynomial>
    void pre_condition(const Entities::NDG_t<Polynomial> &NDG, Meshes::Mesh2D_t &mesh, std::size_t n_adaptivity_steps,
                       std::size_t pre_condition_interval) const {
        (void)NDG;
        check(ndg_pre_condition(mesh.handle(), CFL_, (int64_t)n_adaptivity_steps, (int64_t)pre_condition_interval, 0));
        mesh.refresh();
    }
    // Solver2D_t::pre_condition_iterative (:238-336): the initial condition again, and the time set
    // back to 0, after every adaptation.
    template <typename Polynomial>
    void pre_condition_iterative(const Entities::NDG_t<Polynomial> &NDG, Meshes::Mesh2D_t &mesh,
                                 std::size_t n_adaptivity_steps, std::size_t pre_condition_interval) const {
        (void)NDG;
        check(ndg_pre_condition(mesh.handle(), CFL_, (int64_t)n_adaptivity_steps, (int64_t)pre_condition_interval, 1));
        mesh.refresh();
    }

    double get_delta_t(Meshes::Mesh2D_t &mesh) const {
        double dt = 0.0;
        check(ndg_get_delta_t(mesh.handle(), CFL_, &dt));
        return dt;
    }

    double CFL_;
    std::vector<double> output_times_;
    double viscosity_;
    double time_ = 0.0;
    int64_t n_steps_ = 0;
};

}  // namespace Solvers
}  // namespace Device

// The reference's host classes (SEM::Host, src/main.cpp:227-268) are the same classes here: one
// implementation on the GPU serves both spellings.
namespace Host {
namespace Entities = ::SEM::Device::Entities;
namespace Polynomials = ::SEM::Device::Polynomials;
namespace Meshes = ::SEM::Device::Meshes;
namespace Solvers = ::SEM::Device::Solvers;
}  // namespace Host
}  // namespace SEM

// include/helpers/float_types.h
typedef double deviceFloat;
typedef deviceFloat hostFloat;

#endif
