// ref_fixtures.cpp -- fixture generator (test infrastructure, never shipped in the product).
//
// Linked against the reference's own compilable host sources (LegendrePolynomial_t.cpp,
// Element2D_t.cpp, Face2D_t.cpp, Hilbert.cpp, Hilbert_splitting.cpp) and the header-only NDG_t
// tables, it evaluates the reference routines on deterministic inputs and prints them as JSON.
// tests/golden/make_golden.py runs it and stores the result as tests/golden/ref_fixtures.json,
// against which the oracle is pinned bit for bit.
#include "entities/NDG_t.h"
#include "polynomials/LegendrePolynomial_t.h"
#include "entities/Element2D_t.h"
#include "entities/Face2D_t.h"
#include "functions/Hilbert.h"
#include "functions/Hilbert_splitting.h"

#include <cinttypes>
#include <cstdio>
#include <cstdint>
#include <vector>
#include <array>

using SEM::Host::Entities::Element2D_t;
using SEM::Host::Entities::Face2D_t;
using SEM::Host::Entities::NDG_t;
using SEM::Host::Entities::Vec2;
using SEM::Host::Polynomials::LegendrePolynomial_t;

// splitmix64: deterministic inputs independent of the C++ library
static uint64_t sm_state = 0x9E3779B97F4A7C15ull;
static double urand() {
    uint64_t z = (sm_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

static bool first_item = true;
static void key(const char* k) {
    std::printf("%s\"%s\":", first_item ? "" : ",\n", k);
    first_item = false;
}
static void arr(const std::vector<double>& v) {
    std::printf("[");
    for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? "," : "", v[i]);
    std::printf("]");
}

int main() {
    const int N_max = 16;
    const size_t n_interp = 5;
    NDG_t<LegendrePolynomial_t> ndg(N_max, n_interp);

    std::printf("{\n");
    key("N_max"); std::printf("%d", N_max);
    key("n_interp"); std::printf("%zu", n_interp);
    const char* names[12] = {"nodes", "weights", "bary", "lag_left", "lag_right", "dlag_left",
                             "dlag_right", "D", "Ghat", "Dhat", "interp", "poly"};
    const std::vector<std::vector<double>>* tabs[12] = {
        &ndg.nodes_, &ndg.weights_, &ndg.barycentric_weights_, &ndg.lagrange_interpolant_left_,
        &ndg.lagrange_interpolant_right_, &ndg.lagrange_interpolant_derivative_left_,
        &ndg.lagrange_interpolant_derivative_right_, &ndg.derivative_matrices_,
        &ndg.g_hat_derivative_matrices_, &ndg.derivative_matrices_hat_, &ndg.interpolation_matrices_,
        &ndg.polynomials_};
    for (int w = 0; w < 12; ++w) {
        key(names[w]);
        std::printf("[");
        for (int N = 0; N <= N_max; ++N) {
            if (N) std::printf(",");
            arr((*tabs[w])[N]);
        }
        std::printf("]");
    }

    // Element traces, geometry and estimator on deterministic inputs
    const int orders[4] = {4, 6, 8, 12};
    const std::array<std::array<double, 8>, 3> quads = {{
        {{1.0, 0.0, 1.0, 1.0, 0.0, 1.0, 0.0, 0.0}},                   // unit square, generator local order
        {{0.3, -0.2, 0.9, 0.1, 0.6, 0.8, -0.1, 0.5}},                  // general convex quad
        {{2.0, 1.0, 2.0, 1.25, 1.75, 1.25, 1.75, 1.0}},                // small square offset
    }};
    key("elements");
    std::printf("[");
    bool first_el = true;
    for (int N : orders) {
        for (size_t qd = 0; qd < quads.size(); ++qd) {
            const int nf = N + 1;
            Element2D_t e(N, 0, SEM::Host::Hilbert::Status::A, 0, {std::vector<size_t>(1), std::vector<size_t>(1), std::vector<size_t>(1), std::vector<size_t>(1)}, {0, 1, 2, 3});
            std::array<Vec2<double>, 4> pts;
            for (int k = 0; k < 4; ++k) pts[k] = Vec2<double>(quads[qd][2 * k], quads[qd][2 * k + 1]);
            e.compute_geometry(pts, ndg.nodes_[N]);
            std::vector<double> p(nf * nf), u(nf * nf), v(nf * nf);
            for (int k = 0; k < nf * nf; ++k) { p[k] = 2 * urand() - 1; u[k] = 2 * urand() - 1; v[k] = 2 * urand() - 1; }
            // smooth pressure for the estimator: exp of a quadratic in physical coordinates
            std::vector<double> ps(nf * nf);
            for (int i = 0; i <= N; ++i) for (int j = 0; j <= N; ++j) {
                const double xi = ndg.nodes_[N][i], eta = ndg.nodes_[N][j];
                ps[i * nf + j] = std::exp(-((xi - 0.2) * (xi - 0.2) + (eta + 0.1) * (eta + 0.1)) / (0.3 + 0.2 * qd));
            }
            e.p_ = p; e.u_ = u; e.v_ = v;
            e.interpolate_to_boundaries(ndg.lagrange_interpolant_left_[N], ndg.lagrange_interpolant_right_[N]);
            std::vector<double> traces;
            for (int s = 0; s < 4; ++s) {
                traces.insert(traces.end(), e.p_extrapolated_[s].begin(), e.p_extrapolated_[s].end());
                traces.insert(traces.end(), e.u_extrapolated_[s].begin(), e.u_extrapolated_[s].end());
                traces.insert(traces.end(), e.v_extrapolated_[s].begin(), e.v_extrapolated_[s].end());
            }
            std::vector<double> geom;
            for (auto* a : {&e.dxi_dx_, &e.deta_dx_, &e.dxi_dy_, &e.deta_dy_, &e.jacobian_}) geom.insert(geom.end(), a->begin(), a->end());
            for (int s = 0; s < 4; ++s) geom.insert(geom.end(), e.scaling_factor_[s].begin(), e.scaling_factor_[s].end());
            geom.push_back(e.delta_xy_min_); geom.push_back(e.center_.x()); geom.push_back(e.center_.y());

            std::vector<double> est;
            e.p_ = ps;
            e.estimate_p_error(1e-6, 1e-14, ndg.polynomials_[N]);
            est = {e.p_error_, e.p_sigma_, (double)e.refine_, (double)e.coarsen_};
            for (int k = 0; k < std::min(N + 1, 4); ++k) est.push_back(e.spectrum_[k]);

            std::printf("%s{\"N\":%d,\"points\":", first_el ? "" : ",\n", N);
            first_el = false;
            arr(std::vector<double>(quads[qd].begin(), quads[qd].end()));
            std::printf(",\"p\":"); arr(p);
            std::printf(",\"u\":"); arr(u);
            std::printf(",\"v\":"); arr(v);
            std::printf(",\"traces\":"); arr(traces);
            std::printf(",\"geometry\":"); arr(geom);
            std::printf(",\"p_smooth\":"); arr(ps);
            std::printf(",\"estimate\":"); arr(est);
            std::printf("}");
        }
    }
    std::printf("]");

    // Ghost (boundary) element geometry: nodes {a, b, b, a}
    key("boundary_elements");
    std::printf("[");
    const std::array<std::array<double, 4>, 2> bars = {{{{1.0, 0.0, 0.0, 0.0}}, {{0.25, 0.75, 0.1, 0.2}}}};
    for (size_t b = 0; b < bars.size(); ++b) {
        const int N = 8;
        Element2D_t e(N, 0, SEM::Host::Hilbert::Status::A, 0, {std::vector<size_t>(1), std::vector<size_t>(), std::vector<size_t>(), std::vector<size_t>()}, {0, 1, 1, 0});
        e.allocate_boundary_storage();
        const Vec2<double> a(bars[b][0], bars[b][1]), c(bars[b][2], bars[b][3]);
        e.compute_boundary_geometry({a, c, c, a}, ndg.nodes_[N]);
        std::vector<double> out(e.scaling_factor_[0].begin(), e.scaling_factor_[0].end());
        out.push_back(e.delta_xy_min_); out.push_back(e.center_.x()); out.push_back(e.center_.y());
        std::printf("%s{\"N\":%d,\"bar\":", b ? ",\n" : "", N);
        arr(std::vector<double>(bars[b].begin(), bars[b].end()));
        std::printf(",\"out\":"); arr(out);
        std::printf("}");
    }
    std::printf("]");

    // Face geometry: conforming, half face of a longer edge, reversed right element
    key("faces");
    std::printf("[");
    const std::array<std::array<double, 12>, 3> fcases = {{
        {{0.0, 0.0, 1.0, 0.0,  0.0, 0.0, 1.0, 0.0,  1.0, 0.0, 0.0, 0.0}},
        {{0.5, 0.0, 1.0, 0.0,  0.0, 0.0, 1.0, 0.0,  1.0, 0.0, 0.5, 0.0}},
        {{0.1, 0.2, 0.4, 0.9,  0.1, 0.2, 0.4, 0.9,  0.4, 0.9, 0.1, 0.2}},
    }};
    for (size_t fc = 0; fc < fcases.size(); ++fc) {
        Face2D_t f(4, {0, 1}, {0, 1}, {0, 0});
        const auto& c = fcases[fc];
        const std::array<Vec2<double>, 2> fn {Vec2<double>(c[0], c[1]), Vec2<double>(c[2], c[3])};
        const std::array<std::array<Vec2<double>, 2>, 2> en {
            std::array<Vec2<double>, 2>{Vec2<double>(c[4], c[5]), Vec2<double>(c[6], c[7])},
            std::array<Vec2<double>, 2>{Vec2<double>(c[8], c[9]), Vec2<double>(c[10], c[11])}};
        f.compute_geometry({Vec2<double>(0, 0), Vec2<double>(0, 0)}, fn, en);
        std::vector<double> out {f.normal_.x(), f.normal_.y(), f.tangent_.x(), f.tangent_.y(), f.length_,
                                 f.offset_[0], f.offset_[1], f.scale_[0], f.scale_[1]};
        std::printf("%s{\"in\":", fc ? ",\n" : "");
        arr(std::vector<double>(c.begin(), c.end()));
        std::printf(",\"out\":"); arr(out);
        std::printf("}");
    }
    std::printf("]");

    // Hilbert curve order
    key("hilbert");
    std::printf("[");
    const int ns[3] = {4, 16, 512};
    bool first_h = true;
    for (int n : ns) {
        const int count = n <= 16 ? n * n : 64;
        for (int k = 0; k < count; ++k) {
            const int d = n <= 16 ? k : (int)(urand() * n * n);
            const auto xy = SEM::Hilbert::d2xy(n, d);
            std::printf("%s[%d,%d,%d,%d]", first_h ? "" : ",", n, d, xy[0], xy[1]);
            first_h = false;
        }
    }
    std::printf("]");

    // Hilbert splitting tables (Hilbert_splitting.cpp): every status and rotation / side index
    namespace HS = SEM::Host::Hilbert;
    const HS::Status sts[4] = {HS::Status::H, HS::Status::A, HS::Status::R, HS::Status::B};
    key("hilbert_child_order");
    std::printf("[");
    for (int rot = 0; rot < 4; ++rot)
        for (int st = 0; st < 4; ++st) {
            const auto o = HS::child_order(sts[st], rot);
            std::printf("%s[%zu,%zu,%zu,%zu]", rot + st ? "," : "", o[0], o[1], o[2], o[3]);
        }
    std::printf("]");
    key("hilbert_child_status");
    std::printf("[");
    for (int rot = 0; rot < 4; ++rot)
        for (int st = 0; st < 4; ++st) {
            const auto o = HS::child_statuses(sts[st], rot);
            std::printf("%s[%d,%d,%d,%d]", rot + st ? "," : "", (int)o[0], (int)o[1], (int)o[2], (int)o[3]);
        }
    std::printf("]");
    key("hilbert_first_status");
    std::printf("[");
    for (size_t s = 0; s < 4; ++s) std::printf("%s%d", s ? "," : "", (int)HS::deduct_first_element_status(s));
    std::printf("]");
    key("hilbert_last_status");
    std::printf("[");
    for (size_t s = 0; s < 4; ++s) std::printf("%s%d", s ? "," : "", (int)HS::deduct_last_element_status(s));
    std::printf("]");
    key("hilbert_element_status");
    std::printf("[");
    for (size_t a = 0; a < 4; ++a)
        for (size_t b = 0; b < 4; ++b) std::printf("%s%d", a + b ? "," : "", (int)HS::deduct_element_status(a, b));
    std::printf("]");
    std::printf("\n}\n");
    return 0;
}
