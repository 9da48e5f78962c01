// ndg_abi.cpp -- C ABI of libndg_amd: tables, mesh construction (host), device buffers and the
// stage / time-loop drivers. The arithmetic that has to match the reference host path bit for
// bit (face normals and tangents, ghost scaling factors, time accumulation) is done here on the
// host with contraction disabled.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "internal.hpp"

using namespace ndg;

namespace {

thread_local std::string g_last_error;

ndg_status fail(ndg_status code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

ndg_status hip_fail(hipError_t e, const char *where) {
    return fail(NDG_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr, where)                                  \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return hip_fail(_e, where);     \
    } while (0)

// include/solvers/Solver2D_t.tpp:16-18 (Williamson 2N-storage RK3)
constexpr double kAm[3] = {0.0, -5.0 / 9.0, -153.0 / 128.0};
constexpr double kBm[3] = {0.0, 1.0 / 3.0, 0.75};
constexpr double kGm[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

bool ulp2_equal(double a, double b) {
    return std::fabs(a - b) <= DBL_EPSILON * std::fabs(a + b) * 2 || std::fabs(a - b) < DBL_MIN;
}

}  // namespace

ndg_status ndg::set_error(ndg_status code, const char *msg) { return fail(code, msg); }

struct ndg_tables {
    Tables t;
};

struct ndg_mesh {
    DeviceMesh d;
    const ndg_tables *tables = nullptr;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int64_t n_faces = 0;
    double tol_min = 1e-6, tol_max = 1e-14;
    double *dt_dev = nullptr;
    double *est_dev = nullptr;
    double *pulses_dev = nullptr;
    std::vector<double> staging;
};

extern "C" ndg_status ndg_last_error(char *buf, size_t buf_len) {
    if (!buf || buf_len == 0) return NDG_ERR_INVALID;
    std::snprintf(buf, buf_len, "%s", g_last_error.c_str());
    return NDG_OK;
}

extern "C" ndg_status ndg_device_count(int32_t *count) {
    if (!count) return fail(NDG_ERR_INVALID, "ndg_device_count: null argument");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return NDG_OK;
}

// ------------------------------------------------------------------------------------------
// Tables

extern "C" ndg_status ndg_tables_create(int32_t N_max, int64_t n_interp, ndg_tables **out) {
    if (!out) return fail(NDG_ERR_INVALID, "ndg_tables_create: null output");
    if (N_max < 0 || N_max > 64) return fail(NDG_ERR_INVALID, "ndg_tables_create: N_max out of range");
    if (n_interp < 2) return fail(NDG_ERR_INVALID, "ndg_tables_create: n_interpolation_points must be >= 2");
    auto *t = new ndg_tables;
    build_tables(N_max, n_interp, t->t);
    *out = t;
    return NDG_OK;
}

extern "C" ndg_status ndg_tables_destroy(ndg_tables *tables) {
    delete tables;
    return NDG_OK;
}

extern "C" ndg_status ndg_tables_get(const ndg_tables *tables, int32_t which, int32_t N, double *out,
                                     int64_t capacity, int64_t *count) {
    if (!tables) return fail(NDG_ERR_INVALID, "ndg_tables_get: null tables");
    if (which < 0 || which > 11) return fail(NDG_ERR_INVALID, "ndg_tables_get: unknown table");
    if (N < 0 || N > tables->t.N_max) return fail(NDG_ERR_INVALID, "ndg_tables_get: N out of range");
    const auto &v = tables->t.tab[which][N];
    if (count) *count = (int64_t)v.size();
    if (out) {
        if (capacity < (int64_t)v.size()) return fail(NDG_ERR_INVALID, "ndg_tables_get: output too small");
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    }
    return NDG_OK;
}

// ------------------------------------------------------------------------------------------
// Mesh construction

namespace {

struct v2 {
    double x, y;
};

double mag(double x, double y) { return std::sqrt(x * x + y * y); }

// Face geometry with the domain element on the left (src/entities/Face2D_t.cpp:56-66).
void face_frame(v2 f0, v2 f1, v2 l0, v2 l1, BcSide &rec) {
    double tx = f1.x - f0.x, ty = f1.y - f0.y;
    const double len = mag(tx, ty);
    tx /= len;
    ty /= len;
    double nx = ty, ny = -tx;
    const double lx = l1.x - l0.x, ly = l1.y - l0.y;
    const double sign = std::copysign(1.0, tx * lx + ty * ly);
    nx *= sign; ny *= sign; tx *= sign; ty *= sign;
    rec.nx = nx; rec.ny = ny; rec.tx = tx; rec.ty = ty;
}

// Ghost element side-0 scaling factor for nodes {P0, P1, P2, P3}
// (src/entities/Element2D_t.cpp:779-786 with include/functions/quad_metrics.tpp:3-8 at eta = -1).
double ghost_scale(v2 P0, v2 P1, v2 P2, v2 P3) {
    const double eta = -1.0;
    const double mx = ((1 - eta) * (P1.x - P0.x) + (1 + eta) * (P2.x - P3.x)) / 4;
    const double my = ((1 - eta) * (P1.y - P0.y) + (1 + eta) * (P2.y - P3.y)) / 4;
    return std::sqrt(mx * mx + my * my);
}

uint64_t edge_key(int64_t a, int64_t b) { return ((uint64_t)a << 32) | (uint64_t)b; }

template <typename T>
ndg_status device_upload(T **dst, const std::vector<T> &src, const char *what) {
    if (src.empty()) return NDG_OK;
    HIP_TRY(hipMalloc((void **)dst, src.size() * sizeof(T)), what);
    HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), what);
    return NDG_OK;
}

void free_mesh(ndg_mesh *m) {
    if (!m) return;
    DeviceMesh &d = m->d;
    for (void *p : {(void *)d.verts, (void *)d.side_nb, (void *)d.bc, (void *)d.tab, (void *)d.q, (void *)d.g,
                    (void *)d.trace[0], (void *)d.trace[1], (void *)d.rhs, (void *)d.scratch, (void *)m->dt_dev,
                    (void *)m->est_dev, (void *)m->pulses_dev, (void *)d.qh, (void *)d.ql, (void *)d.g32,
                    (void *)d.trace32[0], (void *)d.trace32[1], (void *)d.tab32})
        if (p) (void)hipFree(p);
    if (m->own_stream) (void)hipStreamDestroy(m->own_stream);
    delete m;
}

}  // namespace

extern "C" ndg_status ndg_mesh_create(const ndg_mesh_desc *desc, const ndg_tables *tables, ndg_mesh **out) {
    if (!desc || !tables || !out) return fail(NDG_ERR_INVALID, "ndg_mesh_create: null argument");
    *out = nullptr;
    const int64_t K = desc->n_quads, B = desc->n_bars, NN = desc->n_nodes;
    if (K <= 0 || NN <= 0 || !desc->xy || !desc->quads) return fail(NDG_ERR_INVALID, "ndg_mesh_create: empty mesh");
    if (B > 0 && (!desc->bars || !desc->bar_type)) return fail(NDG_ERR_INVALID, "ndg_mesh_create: bars without types");
    if (NN >= (int64_t(1) << 32)) return fail(NDG_ERR_INVALID, "ndg_mesh_create: too many nodes");
    if (K > (int64_t(1) << 29)) return fail(NDG_ERR_INVALID, "ndg_mesh_create: too many elements");
    int N = desc->initial_N;
    if (desc->quad_N) {
        N = desc->quad_N[0];
        for (int64_t e = 1; e < K; ++e)
            if (desc->quad_N[e] != N)
                return fail(NDG_ERR_UNSUPPORTED, "ndg_mesh_create: elements of different orders need the mortar "
                                                 "(mixed-order) path, which this build does not provide yet");
        if (N != desc->initial_N)
            return fail(NDG_ERR_UNSUPPORTED, "ndg_mesh_create: element orders differ from initial_N (ghost and face "
                                             "orders would differ from the element order)");
    }
    if (!order_supported(N)) return fail(NDG_ERR_UNSUPPORTED, "ndg_mesh_create: order N outside the compiled range 1..16");
    if (N > tables->t.N_max) return fail(NDG_ERR_INVALID, "ndg_mesh_create: N exceeds the tables' N_max");

    for (int64_t e = 0; e < 4 * K; ++e)
        if (desc->quads[e] < 0 || desc->quads[e] >= NN) return fail(NDG_ERR_INVALID, "ndg_mesh_create: quad node id out of range");
    for (int64_t b = 0; b < 2 * B; ++b)
        if (desc->bars[b] < 0 || desc->bars[b] >= NN) return fail(NDG_ERR_INVALID, "ndg_mesh_create: bar node id out of range");

    auto P = [&](int64_t node) { return v2{desc->xy[2 * node], desc->xy[2 * node + 1]}; };

    // Directed edges of the domain elements (build_faces, src/meshes/Mesh2D_t.cpp:1110-1154).
    std::unordered_map<uint64_t, int64_t> edges;
    edges.reserve((size_t)(4 * K) * 2);
    for (int64_t e = 0; e < K; ++e)
        for (int s = 0; s < 4; ++s) {
            const int64_t a = desc->quads[4 * e + s], b = desc->quads[4 * e + (s + 1) % 4];
            if (!edges.emplace(edge_key(a, b), e * 4 + s).second)
                return fail(NDG_ERR_INVALID, "ndg_mesh_create: an edge is used twice with the same orientation");
        }
    std::vector<int32_t> side_nb((size_t)(4 * K), INT32_MIN);
    int64_t n_faces = 0;
    for (int64_t e = 0; e < K; ++e)
        for (int s = 0; s < 4; ++s) {
            const int64_t a = desc->quads[4 * e + s], b = desc->quads[4 * e + (s + 1) % 4];
            auto it = edges.find(edge_key(b, a));
            if (it != edges.end()) {
                side_nb[4 * e + s] = (int32_t)it->second;
                if (it->second > e * 4 + s) ++n_faces;
            }
        }
    // Ghost (bar) elements attach to the domain side with the same two nodes
    // (src/meshes/Mesh2D_t.cpp:1156-1181). Their nodes after attachment are {P(s+1), P(s)}.
    std::vector<int64_t> bar_side((size_t)B, -1);
    for (int64_t b = 0; b < B; ++b) {
        const int64_t a = desc->bars[2 * b], c = desc->bars[2 * b + 1];
        auto it = edges.find(edge_key(c, a));
        if (it == edges.end()) it = edges.find(edge_key(a, c));
        if (it == edges.end()) return fail(NDG_ERR_INVALID, "ndg_mesh_create: boundary bar matches no element edge");
        bar_side[b] = it->second;
        ++n_faces;
    }
    std::vector<BcSide> bcs;
    std::vector<int64_t> side_bar((size_t)(4 * K), -1);
    for (int64_t b = 0; b < B; ++b) {
        const int64_t es = bar_side[b];
        if (side_nb[es] != INT32_MIN || side_bar[es] != -1)
            return fail(NDG_ERR_INVALID, "ndg_mesh_create: a boundary bar lies on an interior or already bounded edge");
        side_bar[es] = b;
    }
    for (int64_t b = 0; b < B; ++b) {
        const int64_t es = bar_side[b], e = es / 4;
        const int s = (int)(es % 4);
        const int64_t u0 = desc->quads[4 * e + s], u1 = desc->quads[4 * e + (s + 1) % 4];
        const int kind = desc->bar_type[b];
        if (kind == NDG_BC_PERIODIC) {
            if (!desc->bar_donor) return fail(NDG_ERR_INVALID, "ndg_mesh_create: periodic bar without donor array");
            const int64_t d = desc->bar_donor[b];
            if (d < 0 || d >= B) return fail(NDG_ERR_INVALID, "ndg_mesh_create: periodic donor out of range");
            // local_interfaces (Mesh2D_t.cpp:4712-4730): the ghost carries the donor side's trace
            // unchanged, so the side couples directly to the donor's element side.
            side_nb[es] = (int32_t)bar_side[d];
            continue;
        }
        if (kind < NDG_BC_NONE || kind > NDG_BC_OUTFLOW) return fail(NDG_ERR_INVALID, "ndg_mesh_create: unknown boundary kind");
        BcSide rec{};
        rec.kind = kind;
        // ghost nodes {a, b, b, a} with a = P(s+1), b = P(s) after face attachment
        const v2 ga = P(u1), gb = P(u0);
        rec.ax = ga.x; rec.ay = ga.y; rec.bx = gb.x; rec.by = gb.y;
        const int64_t bar_a = desc->bars[2 * b];
        v2 G0, G1, G2, G3;
        if (bar_a == u1) {  // bar already reversed with respect to the element edge
            G0 = ga; G1 = gb; G2 = gb; G3 = ga;
        } else {            // "inside out" bar: only its first two nodes are swapped
            G0 = ga; G1 = gb; G2 = ga; G3 = gb;
        }
        rec.sf = ghost_scale(G0, G1, G2, G3);
        face_frame(P(u0), P(u1), P(u0), P(u1), rec);
        side_nb[es] = -(int32_t)(1 + bcs.size());
        bcs.push_back(rec);
    }
    for (int64_t es = 0; es < 4 * K; ++es)
        if (side_nb[es] == INT32_MIN)
            return fail(NDG_ERR_INVALID, "ndg_mesh_create: element side has neither a neighbour nor a boundary bar");

    // The ghost projection of a conforming boundary face (src/meshes/Mesh2D_t.cpp:4383-4418) must
    // reduce to the reversed copy: check the exact-node shortcut and symmetric weights.
    const auto &x = tables->t.tab[NDG_TABLE_NODES][N];
    const auto &w = tables->t.tab[NDG_TABLE_WEIGHTS][N];
    for (int j = 0; j <= N; ++j) {
        const double coord = (x[N - j] + 1) * 1.0 + 2 * 0.0 - 1;
        int hit = -1;
        for (int i = 0; i <= N; ++i)
            if (ulp2_equal(coord, x[i])) { hit = i; break; }
        if (hit != N - j || w[j] / w[N - j] != 1.0)
            return fail(NDG_ERR_UNSUPPORTED, "ndg_mesh_create: boundary ghost projection is not a plain reversed copy "
                                             "for this order");
    }

    int32_t ndev = 0;
    ndg_device_count(&ndev);
    if (ndev <= 0) return fail(NDG_ERR_NO_DEVICE, "ndg_mesh_create: no HIP device visible");

    std::unique_ptr<ndg_mesh, void (*)(ndg_mesh *)> m(new ndg_mesh, free_mesh);
    m->tables = tables;
    m->tol_min = desc->tolerance_min;
    m->tol_max = desc->tolerance_max;
    m->n_faces = n_faces;
    HIP_TRY(hipStreamCreateWithFlags(&m->own_stream, hipStreamNonBlocking), "hipStreamCreate");
    m->stream = m->own_stream;
    DeviceMesh &d = m->d;
    d.N = N;
    d.K = K;
    std::vector<double> verts((size_t)(8 * K));
    for (int64_t e = 0; e < K; ++e)
        for (int k = 0; k < 4; ++k) {
            const v2 p = P(desc->quads[4 * e + k]);
            verts[8 * e + 2 * k] = p.x;
            verts[8 * e + 2 * k + 1] = p.y;
        }
    ndg_status st;
    if ((st = device_upload(&d.verts, verts, "upload vertices")) != NDG_OK) return st;
    if ((st = device_upload(&d.side_nb, side_nb, "upload connectivity")) != NDG_OK) return st;
    if ((st = device_upload(&d.bc, bcs, "upload boundary records")) != NDG_OK) return st;
    d.n_bc = (int64_t)bcs.size();
    if ((st = device_upload(&d.tab, pack_tables(tables->t, N), "upload tables")) != NDG_OK) return st;
    const int64_t nf = N + 1, ne = nf * nf;
    HIP_TRY(hipMalloc((void **)&d.q, sizeof(double) * 3 * ne * K), "hipMalloc q");
    HIP_TRY(hipMalloc((void **)&d.g, sizeof(double) * 3 * ne * K), "hipMalloc g");
    HIP_TRY(hipMalloc((void **)&d.trace[0], sizeof(double) * 12 * nf * K), "hipMalloc traces");
    HIP_TRY(hipMalloc((void **)&d.trace[1], sizeof(double) * 12 * nf * K), "hipMalloc traces");
    d.scratch_len = 1024;
    HIP_TRY(hipMalloc((void **)&d.scratch, sizeof(double) * d.scratch_len), "hipMalloc scratch");
    HIP_TRY(hipMalloc((void **)&m->dt_dev, sizeof(double)), "hipMalloc dt");
    HIP_TRY(hipMemsetAsync(d.q, 0, sizeof(double) * 3 * ne * K, m->stream), "hipMemset q");
    HIP_TRY(hipMemsetAsync(d.g, 0, sizeof(double) * 3 * ne * K, m->stream), "hipMemset g");
    HIP_TRY(launch_traces(d, m->stream), "traces kernel");
    HIP_TRY(hipStreamSynchronize(m->stream), "mesh create sync");
    *out = m.release();
    return NDG_OK;
}

extern "C" ndg_status ndg_mesh_destroy(ndg_mesh *mesh) {
    if (mesh) (void)hipStreamSynchronize(mesh->stream);
    free_mesh(mesh);
    return NDG_OK;
}

extern "C" ndg_status ndg_mesh_set_stream(ndg_mesh *mesh, void *hip_stream) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_mesh_set_stream: null mesh");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_mesh_set_stream sync");
    mesh->stream = hip_stream ? (hipStream_t)hip_stream : mesh->own_stream;
    return NDG_OK;
}

extern "C" ndg_status ndg_mesh_info(const ndg_mesh *mesh, int64_t *n_elements, int64_t *n_faces,
                                    int64_t *n_dofs_per_field, int32_t *N_min, int32_t *N_max) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_mesh_info: null mesh");
    const int64_t ne = (int64_t)(mesh->d.N + 1) * (mesh->d.N + 1);
    if (n_elements) *n_elements = mesh->d.K;
    if (n_faces) *n_faces = mesh->n_faces;
    if (n_dofs_per_field) *n_dofs_per_field = ne * mesh->d.K;
    if (N_min) *N_min = mesh->d.N;
    if (N_max) *N_max = mesh->d.N;
    return NDG_OK;
}

// ------------------------------------------------------------------------------------------
// State

namespace {

// device [K][3][Ne] <-> packed per-field host arrays
ndg_status copy_out3(ndg_mesh *m, const double *dev, double *a, double *b, double *c, const char *what) {
    const int64_t ne = (int64_t)(m->d.N + 1) * (m->d.N + 1), K = m->d.K;
    m->staging.resize((size_t)(3 * ne * K));
    HIP_TRY(hipMemcpyAsync(m->staging.data(), dev, sizeof(double) * 3 * ne * K, hipMemcpyDeviceToHost, m->stream), what);
    HIP_TRY(hipStreamSynchronize(m->stream), what);
    double *outs[3] = {a, b, c};
    for (int64_t e = 0; e < K; ++e)
        for (int f = 0; f < 3; ++f)
            if (outs[f]) std::memcpy(outs[f] + e * ne, m->staging.data() + (e * 3 + f) * ne, sizeof(double) * ne);
    return NDG_OK;
}

// FP32 compensated form: refresh the FP64 state, RK register and traces from (qh, ql, g32).
ndg_status refresh_fp64(ndg_mesh *m) {
    if (m->d.precision != NDG_PRECISION_FP32_COMPENSATED) return NDG_OK;
    HIP_TRY(launch_from_c32(m->d, m->stream), "from_c32 kernel");
    HIP_TRY(launch_traces(m->d, m->stream), "traces kernel");
    return NDG_OK;
}

// After the FP64 state changed: rebuild (qh, ql, g32) and their traces.
ndg_status refresh_c32(ndg_mesh *m) {
    if (m->d.precision != NDG_PRECISION_FP32_COMPENSATED) return NDG_OK;
    HIP_TRY(launch_to_c32(m->d, m->stream), "to_c32 kernel");
    HIP_TRY(launch_traces_c32(m->d, m->stream), "traces_c32 kernel");
    return NDG_OK;
}

}  // namespace

extern "C" ndg_status ndg_mesh_set_precision(ndg_mesh *mesh, int32_t precision) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_mesh_set_precision: null mesh");
    if (precision != NDG_PRECISION_FP64 && precision != NDG_PRECISION_FP32_COMPENSATED)
        return fail(NDG_ERR_INVALID, "ndg_mesh_set_precision: unknown precision");
    DeviceMesh &d = mesh->d;
    if (precision == d.precision) return NDG_OK;
    if (precision == NDG_PRECISION_FP32_COMPENSATED) {
        const int64_t nf = d.N + 1, ne = nf * nf;
        if (!d.qh) {
            HIP_TRY(hipMalloc((void **)&d.qh, sizeof(float) * 3 * ne * d.K), "hipMalloc qh");
            HIP_TRY(hipMalloc((void **)&d.ql, sizeof(float) * 3 * ne * d.K), "hipMalloc ql");
            HIP_TRY(hipMalloc((void **)&d.g32, sizeof(float) * 3 * ne * d.K), "hipMalloc g32");
            HIP_TRY(hipMalloc((void **)&d.trace32[0], sizeof(float) * 12 * nf * d.K), "hipMalloc traces32");
            HIP_TRY(hipMalloc((void **)&d.trace32[1], sizeof(float) * 12 * nf * d.K), "hipMalloc traces32");
            const std::vector<double> t64 = pack_tables(mesh->tables->t, d.N);
            std::vector<float> t32(t64.begin(), t64.end());
            HIP_TRY(hipMalloc((void **)&d.tab32, sizeof(float) * t32.size()), "hipMalloc tab32");
            HIP_TRY(hipMemcpy(d.tab32, t32.data(), sizeof(float) * t32.size(), hipMemcpyHostToDevice), "upload tab32");
        }
        d.precision = NDG_PRECISION_FP32_COMPENSATED;
        ndg_status st = refresh_c32(mesh);
        if (st != NDG_OK) return st;
    } else {
        ndg_status st = refresh_fp64(mesh);
        if (st != NDG_OK) return st;
        d.precision = NDG_PRECISION_FP64;
    }
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_mesh_set_precision sync");
    return NDG_OK;
}

extern "C" ndg_status ndg_set_ic(ndg_mesh *mesh, const ndg_ic_desc *ic) {
    if (!mesh || !ic) return fail(NDG_ERR_INVALID, "ndg_set_ic: null argument");
    if (ic->kind != NDG_IC_PLANE_WAVE && ic->kind != NDG_IC_PULSES) return fail(NDG_ERR_INVALID, "ndg_set_ic: unknown kind");
    if (ic->kind == NDG_IC_PULSES && (ic->n_pulses < 0 || (ic->n_pulses > 0 && !ic->pulses)))
        return fail(NDG_ERR_INVALID, "ndg_set_ic: bad pulse list");
    DeviceMesh &d = mesh->d;
    const int np = ic->kind == NDG_IC_PULSES ? ic->n_pulses : 0;
    if (mesh->pulses_dev) {
        HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_set_ic sync");
        HIP_TRY(hipFree(mesh->pulses_dev), "hipFree");
        mesh->pulses_dev = nullptr;
    }
    if (np > 0) {
        HIP_TRY(hipMalloc((void **)&mesh->pulses_dev, sizeof(double) * 4 * np), "hipMalloc pulses");
        HIP_TRY(hipMemcpy(mesh->pulses_dev, ic->pulses, sizeof(double) * 4 * np, hipMemcpyHostToDevice), "upload pulses");
    }
    const int64_t ne = (int64_t)(d.N + 1) * (d.N + 1);
    HIP_TRY(launch_ic(d, ic->kind, np, mesh->pulses_dev, mesh->stream), "ic kernel");
    HIP_TRY(hipMemsetAsync(d.g, 0, sizeof(double) * 3 * ne * d.K, mesh->stream), "hipMemset g");
    HIP_TRY(launch_traces(d, mesh->stream), "traces kernel");
    ndg_status st = refresh_c32(mesh);
    if (st != NDG_OK) return st;
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_set_ic sync");
    return NDG_OK;
}

extern "C" ndg_status ndg_set_state(ndg_mesh *mesh, const double *p, const double *u, const double *v) {
    if (!mesh || !p || !u || !v) return fail(NDG_ERR_INVALID, "ndg_set_state: null argument");
    DeviceMesh &d = mesh->d;
    const int64_t ne = (int64_t)(d.N + 1) * (d.N + 1), K = d.K;
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_set_state sync");
    mesh->staging.resize((size_t)(3 * ne * K));
    const double *ins[3] = {p, u, v};
    for (int64_t e = 0; e < K; ++e)
        for (int f = 0; f < 3; ++f) std::memcpy(mesh->staging.data() + (e * 3 + f) * ne, ins[f] + e * ne, sizeof(double) * ne);
    HIP_TRY(hipMemcpyAsync(d.q, mesh->staging.data(), sizeof(double) * 3 * ne * K, hipMemcpyHostToDevice, mesh->stream), "upload state");
    HIP_TRY(hipMemsetAsync(d.g, 0, sizeof(double) * 3 * ne * K, mesh->stream), "hipMemset g");
    HIP_TRY(launch_traces(d, mesh->stream), "traces kernel");
    ndg_status st = refresh_c32(mesh);
    if (st != NDG_OK) return st;
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_set_state sync");
    return NDG_OK;
}

extern "C" ndg_status ndg_get_state(ndg_mesh *mesh, double *p, double *u, double *v) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_get_state: null mesh");
    ndg_status st = refresh_fp64(mesh);
    if (st != NDG_OK) return st;
    return copy_out3(mesh, mesh->d.q, p, u, v, "download state");
}

extern "C" ndg_status ndg_get_traces(ndg_mesh *mesh, double *out) {
    if (!mesh || !out) return fail(NDG_ERR_INVALID, "ndg_get_traces: null argument");
    ndg_status st = refresh_fp64(mesh);
    if (st != NDG_OK) return st;
    const DeviceMesh &d = mesh->d;
    const int64_t n = 12 * (int64_t)(d.N + 1) * d.K;
    HIP_TRY(hipMemcpyAsync(out, d.trace[d.cur], sizeof(double) * n, hipMemcpyDeviceToHost, mesh->stream), "download traces");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_get_traces sync");
    return NDG_OK;
}

extern "C" ndg_status ndg_compute_rhs(ndg_mesh *mesh, double t, double *Gp, double *Gu, double *Gv) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_compute_rhs: null mesh");
    ndg_status st = refresh_fp64(mesh);
    if (st != NDG_OK) return st;
    DeviceMesh &d = mesh->d;
    const int64_t ne = (int64_t)(d.N + 1) * (d.N + 1);
    if (!d.rhs) HIP_TRY(hipMalloc((void **)&d.rhs, sizeof(double) * 3 * ne * d.K), "hipMalloc rhs");
    HIP_TRY(launch_rhs(d, t, mesh->stream), "rhs kernel");
    return copy_out3(mesh, d.rhs, Gp, Gu, Gv, "download rhs");
}

extern "C" ndg_status ndg_get_delta_t(ndg_mesh *mesh, double cfl, double *dt) {
    if (!mesh || !dt) return fail(NDG_ERR_INVALID, "ndg_get_delta_t: null argument");
    if (!(cfl > 0.0)) return fail(NDG_ERR_INVALID, "ndg_get_delta_t: CFL must be positive");
    HIP_TRY(launch_delta_t(mesh->d, cfl, mesh->dt_dev, mesh->stream), "delta_t kernels");
    HIP_TRY(hipMemcpyAsync(dt, mesh->dt_dev, sizeof(double), hipMemcpyDeviceToHost, mesh->stream), "download dt");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_get_delta_t sync");
    return NDG_OK;
}

extern "C" ndg_status ndg_step(ndg_mesh *mesh, double t, double dt) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_step: null mesh");
    const bool c32 = mesh->d.precision == NDG_PRECISION_FP32_COMPENSATED;
    for (int s = 0; s < 3; ++s) {
        const double ts = t + kBm[s] * dt;
        if (c32) HIP_TRY(launch_stage_c32(mesh->d, s, ts, kAm[s], kGm[s] * dt, mesh->stream), "stage_c32 kernel");
        else HIP_TRY(launch_stage(mesh->d, s, ts, kAm[s], kGm[s] * dt, mesh->stream), "stage kernel");
    }
    return NDG_OK;
}

extern "C" ndg_status ndg_solve(ndg_mesh *mesh, double cfl, double t_end, double *t_final, int64_t *n_steps) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_solve: null mesh");
    double dt0 = 0.0;
    ndg_status st = ndg_get_delta_t(mesh, cfl, &dt0);
    if (st != NDG_OK) return st;
    double time = 0.0;
    int64_t steps = 0;
    while (time < t_end) {
        ++steps;
        double dt = dt0;
        if (time + dt > t_end) dt = t_end - time;
        if ((st = ndg_step(mesh, time, dt)) != NDG_OK) return st;
        time += dt;
    }
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_solve sync");
    if (t_final) *t_final = time;
    if (n_steps) *n_steps = steps;
    return NDG_OK;
}

extern "C" ndg_status ndg_run_steps(ndg_mesh *mesh, double cfl, int64_t n_steps, double *time_io) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_run_steps: null mesh");
    if (n_steps < 0) return fail(NDG_ERR_INVALID, "ndg_run_steps: negative step count");
    double dt = 0.0;
    ndg_status st = ndg_get_delta_t(mesh, cfl, &dt);
    if (st != NDG_OK) return st;
    double time = time_io ? *time_io : 0.0;
    for (int64_t k = 0; k < n_steps; ++k) {
        if ((st = ndg_step(mesh, time, dt)) != NDG_OK) return st;
        time += dt;
    }
    if (time_io) *time_io = time;
    return NDG_OK;
}

extern "C" ndg_status ndg_synchronize(ndg_mesh *mesh) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_synchronize: null mesh");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_synchronize");
    return NDG_OK;
}

extern "C" ndg_status ndg_estimate_error(ndg_mesh *mesh, double *error, double *sigma, int32_t *refine,
                                         int32_t *coarsen) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_estimate_error: null mesh");
    ndg_status st = refresh_fp64(mesh);
    if (st != NDG_OK) return st;
    DeviceMesh &d = mesh->d;
    if (!mesh->est_dev) HIP_TRY(hipMalloc((void **)&mesh->est_dev, sizeof(double) * 4 * d.K), "hipMalloc estimator");
    HIP_TRY(launch_estimate(d, mesh->tol_min, mesh->tol_max, mesh->est_dev, mesh->stream), "estimator kernel");
    std::vector<double> host((size_t)(4 * d.K));
    HIP_TRY(hipMemcpyAsync(host.data(), mesh->est_dev, sizeof(double) * 4 * d.K, hipMemcpyDeviceToHost, mesh->stream),
            "download estimator");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_estimate_error sync");
    for (int64_t e = 0; e < d.K; ++e) {
        if (error) error[e] = host[4 * e];
        if (sigma) sigma[e] = host[4 * e + 1];
        if (refine) refine[e] = host[4 * e + 2] != 0.0;
        if (coarsen) coarsen[e] = host[4 * e + 3] != 0.0;
    }
    return NDG_OK;
}
