// mesh_io.cpp -- SU2 mesh files, as the reference's Mesh2D_t::read_su2 reads them
// (src/meshes/Mesh2D_t.cpp:78-330): "NDIME= 2" first, then in any order "NPOIN= n" followed by n
// lines "x y ...", "NELEM= n" followed by n lines "9 a b c d ..." (QUAD_4 only) and "NMARK= n"
// followed by n markers, each "MARKER_TAG= kind" then "MARKER_ELEMS= m" and m lines "3 a b"
// (BAR_2). Marker kinds: farfield (outflow), wall or airfoil (wall), inlet (inflow), symmetry.
// The boundary bars come out in the reference's ghost order: walls, symmetry, inflow, outflow,
// each in file order. The format has no periodic pairs.
#include <algorithm>
#include <cctype>
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

ndg_status read_su2(const char *path, Su2Mesh &m) {
    if (!path) return bad("null path");
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
                tag = lower(tag);
                int kind;
                if (tag == "wall" || tag == "airfoil") kind = 0;
                else if (tag == "symmetry") kind = 1;
                else if (tag == "inlet") kind = 2;
                else if (tag == "farfield") kind = 3;
                else return bad("expected marker tag 'farfield', 'wall', 'symmetry' or 'inlet', found '" + tag + "'");
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
                    m.bars[kind].push_back(a);
                    m.bars[kind].push_back(b);
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
    return NDG_OK;
}

}  // namespace
}  // namespace ndg

using namespace ndg;

extern "C" ndg_status ndg_su2_sizes(const char *path, int64_t *n_nodes, int64_t *n_quads, int64_t *n_bars) {
    Su2Mesh m;
    const ndg_status st = read_su2(path, m);
    if (st != NDG_OK) return st;
    if (n_nodes) *n_nodes = (int64_t)m.xy.size() / 2;
    if (n_quads) *n_quads = (int64_t)m.quads.size() / 4;
    int64_t nb = 0;
    for (const auto &b : m.bars) nb += (int64_t)b.size() / 2;
    if (n_bars) *n_bars = nb;
    return NDG_OK;
}

extern "C" ndg_status ndg_su2_fill(const char *path, double *xy, int64_t *quads, int64_t *bars, int32_t *bar_type) {
    Su2Mesh m;
    const ndg_status st = read_su2(path, m);
    if (st != NDG_OK) return st;
    if (xy) std::copy(m.xy.begin(), m.xy.end(), xy);
    if (quads) std::copy(m.quads.begin(), m.quads.end(), quads);
    int64_t o = 0;
    for (int k = 0; k < 4; ++k)
        for (size_t j = 0; j < m.bars[k].size() / 2; ++j, ++o) {
            if (bars) {
                bars[2 * o] = m.bars[k][2 * j];
                bars[2 * o + 1] = m.bars[k][2 * j + 1];
            }
            if (bar_type) bar_type[o] = kSu2Kinds[k];
        }
    return NDG_OK;
}
