// mesh_io.cpp -- SU2 mesh files, as the reference's Mesh2D_t::read_su2 reads them
// (src/meshes/Mesh2D_t.cpp:78-330): "NDIME= 2" first, then in any order "NPOIN= n" followed by n
// lines "x y ...", "NELEM= n" followed by n lines "9 a b c d ..." (QUAD_4 only) and "NMARK= n"
// followed by n markers, each "MARKER_TAG= kind" then "MARKER_ELEMS= m" and m lines "3 a b"
// (BAR_2). Marker kinds: farfield (outflow), wall or airfoil (wall), inlet (inflow), symmetry.
// The boundary bars come out in the reference's ghost order: walls, symmetry, inflow, outflow,
// each in file order. The format has no periodic pairs; as SU2 itself does (MARKER_PERIODIC in
// its configuration file, not in the mesh), the caller may name pairs of markers to be joined by a
// translation: their bars follow the others, pair after pair (all bars of marker a, then all of
// marker b, each in file order), and every bar's donor is the bar of the other marker whose
// midpoint is its own moved by the translation between the two markers' bar-midpoint centroids.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

namespace ndg {
namespace {

struct Su2Mesh {
    std::vector<double> xy;
    std::vector<int64_t> quads;
    std::vector<int64_t> bars[4];  // walls, symmetry, inflow, outflow
    std::vector<std::vector<int64_t>> periodic;  // bars of periodic marker a0, b0, a1, b1, ...
    std::vector<int64_t> donor;                  // per periodic bar, in the same order: its donor
};

constexpr int32_t kSu2Kinds[4] = {NDG_BC_WALL, NDG_BC_SYMMETRY, NDG_BC_INFLOW, NDG_BC_OUTFLOW};

std::string upper(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::toupper(c); });
    return s;
}

std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
    return s;
}

ndg_status bad(const std::string &msg) { return set_error(NDG_ERR_INVALID, ("SU2 mesh: " + msg).c_str()); }

// Reads the next line whose first token is `want` (skipping others, as the reference's marker loop
// does) and returns the token after it.
bool next_keyed(std::ifstream &f, const std::string &want, std::string &value) {
    std::string line, token;
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        if (!(ls >> token)) continue;
        if (upper(token) == want) {
            ls >> value;
            return true;
        }
    }
    return false;
}

// Donors of the periodic markers, pair by pair: the translation is the difference of the two
// markers' bar-midpoint centroids; every bar of either marker must meet exactly one bar of the
// other within 1e-8 of its length. Donor indices count periodic bars from the first one.
ndg_status pair_periodic(Su2Mesh &m) {
    size_t total = 0;
    for (const auto &b : m.periodic) total += b.size() / 2;
    m.donor.assign(total, -1);
    auto mid = [&](const std::vector<int64_t> &bars, size_t j, int c) {
        return (m.xy[2 * (size_t)bars[2 * j] + c] + m.xy[2 * (size_t)bars[2 * j + 1] + c]) / 2;
    };
    auto length = [&](const std::vector<int64_t> &bars, size_t j) {
        const double dx = m.xy[2 * (size_t)bars[2 * j]] - m.xy[2 * (size_t)bars[2 * j + 1]];
        const double dy = m.xy[2 * (size_t)bars[2 * j] + 1] - m.xy[2 * (size_t)bars[2 * j + 1] + 1];
        return std::sqrt(dx * dx + dy * dy);
    };
    size_t base = 0;
    for (size_t p = 0; p + 1 < m.periodic.size(); p += 2) {
        const std::vector<int64_t> &A = m.periodic[p], &B = m.periodic[p + 1];
        const size_t na = A.size() / 2, nb = B.size() / 2;
        if (na != nb || na == 0) return bad("periodic markers of a pair need the same, non-zero number of bars");
        double ca[2] = {0.0, 0.0}, cb[2] = {0.0, 0.0};
        for (size_t j = 0; j < na; ++j)
            for (int c = 0; c < 2; ++c) {
                ca[c] += mid(A, j, c) / (double)na;
                cb[c] += mid(B, j, c) / (double)nb;
            }
        const double tx = cb[0] - ca[0], ty = cb[1] - ca[1];
        for (size_t j = 0; j < na; ++j) {
            const double x = mid(A, j, 0) + tx, y = mid(A, j, 1) + ty, tol = 1e-8 * length(A, j);
            int64_t hit = -1;
            for (size_t k = 0; k < nb; ++k)
                if (std::fabs(mid(B, k, 0) - x) <= tol && std::fabs(mid(B, k, 1) - y) <= tol) {
                    if (hit >= 0) return bad("a periodic bar matches two bars of the other marker");
                    hit = (int64_t)k;
                }
            if (hit < 0) return bad("a periodic bar has no translated partner in the other marker");
            if (m.donor[base + na + (size_t)hit] >= 0) return bad("two periodic bars share a partner");
            m.donor[base + j] = (int64_t)(base + na + (size_t)hit);
            m.donor[base + na + (size_t)hit] = (int64_t)(base + j);
        }
        base += 2 * na;
    }
    return NDG_OK;
}

ndg_status read_su2(const char *path, const ndg_su2_periodic *per, Su2Mesh &m) {
    if (!path) return bad("null path");
    const int32_t n_pairs = per ? per->n_pairs : 0;
    if (n_pairs < 0 || (n_pairs > 0 && (!per->marker_a || !per->marker_b)))
        return bad("bad periodic marker pairs");
    std::vector<std::string> per_tags;
    for (int32_t i = 0; i < n_pairs; ++i) {
        if (!per->marker_a[i] || !per->marker_b[i]) return bad("null periodic marker name");
        per_tags.emplace_back(per->marker_a[i]);
        per_tags.emplace_back(per->marker_b[i]);
    }
    m.periodic.assign(per_tags.size(), {});
    std::vector<char> per_seen(per_tags.size(), 0);
    std::ifstream f(path);
    if (!f.is_open()) return bad(std::string("file '") + path + "' could not be opened");
    std::string line, token;
    do {
        if (!std::getline(f, line)) return bad("empty file");
    } while (line.find_first_not_of(" \t\r") == std::string::npos);
    {
        std::istringstream ls(line);
        int64_t dim = 0;
        ls >> token >> dim;
        if (upper(token) != "NDIME=") return bad("first token should be 'NDIME=', found '" + token + "'");
        if (dim != 2) return bad("only 2 dimensions are supported, found " + std::to_string(dim));
    }
    while (std::getline(f, line)) {
        std::istringstream ls(line);
        if (!(ls >> token)) continue;
        token = upper(token);
        if (token == "NPOIN=") {
            int64_t n = -1;
            ls >> n;
            if (n < 0) return bad("bad NPOIN count");
            m.xy.assign((size_t)(2 * n), 0.0);
            for (int64_t i = 0; i < n; ++i) {
                if (!std::getline(f, line)) return bad("file ends inside NPOIN");
                std::istringstream p(line);
                if (!(p >> m.xy[2 * i] >> m.xy[2 * i + 1])) return bad("bad node line '" + line + "'");
            }
        } else if (token == "NELEM=") {
            int64_t n = -1;
            ls >> n;
            if (n < 0) return bad("bad NELEM count");
            m.quads.assign((size_t)(4 * n), 0);
            for (int64_t i = 0; i < n; ++i) {
                if (!std::getline(f, line)) return bad("file ends inside NELEM");
                std::istringstream p(line);
                std::string type;
                p >> type;
                if (type != "9") return bad("expected element type '9' (QUAD_4), found '" + type + "'");
                if (!(p >> m.quads[4 * i] >> m.quads[4 * i + 1] >> m.quads[4 * i + 2] >> m.quads[4 * i + 3]))
                    return bad("bad element line '" + line + "'");
            }
        } else if (token == "NMARK=") {
            int64_t n = -1;
            ls >> n;
            if (n < 0) return bad("bad NMARK count");
            for (int64_t k = 0; k < n; ++k) {
                std::string tag, count;
                if (!next_keyed(f, "MARKER_TAG=", tag)) return bad("missing MARKER_TAG=");
                int kind = -1;
                for (size_t t = 0; t < per_tags.size(); ++t)
                    if (tag == per_tags[t]) {
                        if (per_seen[t]) return bad("periodic marker '" + tag + "' appears twice");
                        per_seen[t] = 1;
                        kind = 4 + (int)t;
                    }
                if (kind < 0) {
                    tag = lower(tag);
                    if (tag == "wall" || tag == "airfoil") kind = 0;
                    else if (tag == "symmetry") kind = 1;
                    else if (tag == "inlet") kind = 2;
                    else if (tag == "farfield") kind = 3;
                    else return bad("expected marker tag 'farfield', 'wall', 'symmetry' or 'inlet', found '" + tag + "'");
                }
                if (!next_keyed(f, "MARKER_ELEMS=", count)) return bad("missing MARKER_ELEMS=");
                int64_t c = -1;
                std::istringstream(count) >> c;
                if (c < 0) return bad("bad MARKER_ELEMS count");
                for (int64_t j = 0; j < c; ++j) {
                    if (!std::getline(f, line)) return bad("file ends inside a marker");
                    std::istringstream p(line);
                    std::string type;
                    int64_t a = -1, b = -1;
                    p >> type;
                    if (type != "3") return bad("expected boundary element type '3' (BAR_2), found '" + type + "'");
                    if (!(p >> a >> b)) return bad("bad boundary element line '" + line + "'");
                    std::vector<int64_t> &dst = kind < 4 ? m.bars[kind] : m.periodic[(size_t)(kind - 4)];
                    dst.push_back(a);
                    dst.push_back(b);
                }
            }
        } else {
            return bad("expected 'NPOIN=', 'NELEM=' or 'NMARK=', found '" + token + "'");
        }
    }
    const int64_t nn = (int64_t)m.xy.size() / 2;
    for (int64_t v : m.quads)
        if (v < 0 || v >= nn) return bad("element node id out of range");
    for (const auto &b : m.bars)
        for (int64_t v : b)
            if (v < 0 || v >= nn) return bad("boundary node id out of range");
    for (const auto &b : m.periodic)
        for (int64_t v : b)
            if (v < 0 || v >= nn) return bad("boundary node id out of range");
    for (size_t t = 0; t < per_tags.size(); ++t)
        if (!per_seen[t]) return bad("periodic marker '" + per_tags[t] + "' not found");
    return pair_periodic(m);
}

}  // namespace
}  // namespace ndg

using namespace ndg;

namespace {

ndg_status su2_sizes(const char *path, const ndg_su2_periodic *per, int64_t *n_nodes, int64_t *n_quads,
                     int64_t *n_bars) {
    Su2Mesh m;
    const ndg_status st = read_su2(path, per, m);
    if (st != NDG_OK) return st;
    if (n_nodes) *n_nodes = (int64_t)m.xy.size() / 2;
    if (n_quads) *n_quads = (int64_t)m.quads.size() / 4;
    int64_t nb = 0;
    for (const auto &b : m.bars) nb += (int64_t)b.size() / 2;
    for (const auto &b : m.periodic) nb += (int64_t)b.size() / 2;
    if (n_bars) *n_bars = nb;
    return NDG_OK;
}

ndg_status su2_fill(const char *path, const ndg_su2_periodic *per, double *xy, int64_t *quads, int64_t *bars,
                    int32_t *bar_type, int64_t *bar_donor) {
    Su2Mesh m;
    const ndg_status st = read_su2(path, per, m);
    if (st != NDG_OK) return st;
    if (xy) std::copy(m.xy.begin(), m.xy.end(), xy);
    if (quads) std::copy(m.quads.begin(), m.quads.end(), quads);
    int64_t o = 0;
    auto put = [&](const std::vector<int64_t> &src, int32_t kind, int64_t donor_base, bool periodic) {
        for (size_t j = 0; j < src.size() / 2; ++j, ++o) {
            if (bars) {
                bars[2 * o] = src[2 * j];
                bars[2 * o + 1] = src[2 * j + 1];
            }
            if (bar_type) bar_type[o] = kind;
            if (bar_donor) bar_donor[o] = periodic ? donor_base + m.donor[(size_t)(o - donor_base)] : -1;
        }
    };
    for (int k = 0; k < 4; ++k) put(m.bars[k], kSu2Kinds[k], 0, false);
    const int64_t first_periodic = o;
    for (const auto &b : m.periodic) put(b, NDG_BC_PERIODIC, first_periodic, true);
    return NDG_OK;
}

}  // namespace

extern "C" ndg_status ndg_su2_sizes(const char *path, int64_t *n_nodes, int64_t *n_quads, int64_t *n_bars) {
    return su2_sizes(path, nullptr, n_nodes, n_quads, n_bars);
}

extern "C" ndg_status ndg_su2_fill(const char *path, double *xy, int64_t *quads, int64_t *bars, int32_t *bar_type) {
    return su2_fill(path, nullptr, xy, quads, bars, bar_type, nullptr);
}

extern "C" ndg_status ndg_su2_sizes_periodic(const char *path, const ndg_su2_periodic *periodic, int64_t *n_nodes,
                                             int64_t *n_quads, int64_t *n_bars) {
    return su2_sizes(path, periodic, n_nodes, n_quads, n_bars);
}

extern "C" ndg_status ndg_su2_fill_periodic(const char *path, const ndg_su2_periodic *periodic, double *xy,
                                            int64_t *quads, int64_t *bars, int32_t *bar_type, int64_t *bar_donor) {
    return su2_fill(path, periodic, xy, quads, bars, bar_type, bar_donor);
}
