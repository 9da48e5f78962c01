// ndg_abi.cpp -- C ABI of libndg_amd: tables, mesh construction (host), device buffers and the
// stage / time-loop drivers. The arithmetic that has to match the reference host path bit for
// bit (face normals and tangents, ghost scaling factors, time accumulation) is done here on the
// host with contraction disabled.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
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
constexpr double kGm[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

}  // namespace

ndg_status ndg::set_error(ndg_status code, const char *msg) { return fail(code, msg); }

hipError_t ndg::arena_fit(Arena &a, size_t bytes) {
    if (bytes <= a.cap && a.p) return hipSuccess;
    arena_free(a);
    const size_t cap = std::max<size_t>(bytes + bytes / 4, 256);
    hipError_t e = hipMalloc(&a.p, cap);
    if (e != hipSuccess) {
        a.p = nullptr;
        return e;
    }
    a.cap = cap;
    return hipSuccess;
}

hipError_t ndg::arena_fit_exact(Arena &a, size_t bytes) {
    if (bytes <= a.cap && a.p) return hipSuccess;
    arena_free(a);
    hipError_t e = hipMalloc(&a.p, std::max<size_t>(bytes, 256));
    if (e != hipSuccess) {
        a.p = nullptr;
        return e;
    }
    a.cap = std::max<size_t>(bytes, 256);
    return hipSuccess;
}

void ndg::arena_free(Arena &a) {
    if (a.p) (void)hipFree(a.p);
    a.p = nullptr;
    a.cap = 0;
}

void ndg::store_free(DeviceStore &s) {
    for (Arena *a : {&s.verts, &s.side_nb, &s.bc, &s.faces, &s.side_flux, &s.side_flux32, &s.q, &s.g, &s.tr[0],
                     &s.tr[1], &s.rhs, &s.scratch, &s.qh, &s.gl32, &s.tr32[0], &s.tr32[1]})
        arena_free(*a);
}

struct ndg_tables {
    Tables t;
};

// Host description of the mesh (kept for rebuilds after adaptation).
struct HostMesh {
    std::vector<double> xy;        // [n_nodes][2]
    std::vector<int64_t> quads;    // [K][4]
    std::vector<int64_t> bars;     // [B][2]
    std::vector<int32_t> bar_type; // [B]
    std::vector<int64_t> bar_donor;  // [B] (empty: no periodic bars)
    // [B][2] translation from a periodic bar to its partner end (empty: the centre of the donor
    // minus the bar's own centre, which holds while every periodic pair is conforming)
    std::vector<double> bar_shift;
};


struct Topology {
    std::vector<int64_t> link;
    std::vector<char> periodic_link;
    std::vector<char> hang_kind;
    std::vector<int64_t> hang_second;
    std::vector<int8_t> span;
    std::vector<BcSide> bcs;       // boundary (ghost) records
    std::vector<int64_t> bc_bar;   // bar of each record
    std::vector<int64_t> bar_side; // element side of each bar
    int64_t n_faces = 0;
};

struct ndg_mesh {
    DeviceMesh d;                            // views into store[active]
    DeviceStore store[2];
    int active = 0;
    // Device tables shared by every rebuild: packed per-order tables (FP64 and FP32), the 1D
    // tables of all orders, the mortar matrices by key (offsets mortar_key_off, -1 = not built)
    // and the device clock.
    double *tab64[kMaxOrder + 1] = {};
    float *tab32[kMaxOrder + 1] = {};
    double *tab1d = nullptr;
    double *mortar_mats = nullptr;
    std::vector<int64_t> mortar_key_off;
    double *clock = nullptr;
    // The mesh's own copy of the caller's tables (the reference's Mesh2D_t copies the polynomial
    // tables it needs), so the caller's ndg_tables may be destroyed before the mesh.
    ndg_tables tables_own;
    const ndg_tables *tables = &tables_own;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int64_t n_faces = 0;
    double tol_min = 1e-6, tol_max = 1e-14;
    double *dt_dev = nullptr;
    double *est_dev = nullptr;
    int64_t est_len = 0;
    double *pulses_dev = nullptr;
    int ic_kind = -1;                        // last initial condition given (re-applied by pre_condition)
    std::vector<double> ic_pulses;
    double ic_period[2] = {0.0, 0.0};
    std::vector<double> staging;
    // Element e of the caller's order is element dev_of[e] of the device order (grouped by N).
    std::vector<int64_t> dev_of;
    std::vector<int32_t> order;      // N of element e (caller's order)
    std::vector<int64_t> host_off;   // offset of element e in a packed per-field host array
    int64_t n_dofs = 0;              // values per field
    // Captured RK step (ndg_mesh_set_graph): one executable graph per starting trace buffer, for
    // the step size and precision it was captured with.
    int graph_mode = NDG_GRAPH_OFF;
    hipGraphExec_t step_exec[2] = {nullptr, nullptr};
    double graph_dt = 0.0;
    int graph_precision = -1;
    std::vector<hipStream_t> side_streams;   // branches of the parallel capture (default priority)
    std::vector<hipEvent_t> side_events;     // [0] fork, [1 + g] join of group g
    // Host topology, kept for rebuilding the device mesh when orders change (adaptation).
    std::vector<double> host_verts;          // [K][4][2], caller's order
    std::vector<int64_t> link;               // exterior side of each element side (see ndg_mesh_create)
    std::vector<char> periodic_link;
    // Hanging sides: 1 = a small side whose exterior is a piece of a larger element's side (link =
    // that side, span = the piece's mortar class in the larger side's direction, hang_second = the
    // next small side along the larger side or -1), 2 = the larger side (link = its first small
    // side, at its first corner).
    std::vector<char> hang_kind;
    std::vector<int64_t> hang_second;
    std::vector<int8_t> span;
    std::vector<BcSide> bc_base;             // boundary (ghost) records
    std::vector<int64_t> bc_bar;             // bar (ghost) of each record
    std::vector<int64_t> bar_side;           // element side of each bar
    std::vector<char> bar_zero;              // per bar: ghost storage resized to zero by an adaptation
    HostMesh host;                           // the description the topology was built from
    std::vector<int32_t> status, rotation, split_level;  // Hilbert status per element
    // Last estimate per element (caller's order): error, sigma, refine, coarsen. Zero until the
    // first estimate and for elements an adaptation creates, as the reference's element fields.
    std::vector<double> est_error, est_sigma;
    std::vector<int32_t> est_refine, est_coarsen;
    int initial_N = 0, maximum_N = 0, max_split_level = 0;
    int64_t adapt_interval = 0;
    // Device-resident adaptation (adapt.hip): the logical mesh in two stores (the current one and
    // the one an adaptation builds), its scratch, and which copy is current. After a device
    // adaptation the host copies above (orders, statuses, links, ghosts, estimates, dev_of) are
    // refreshed from the device on demand (ensure_host); the node lists and bar flags are not kept
    // by the device path (host_nodes_valid).
    LogicalStore lstore[2];
    LogicalMesh lmesh[2];
    int lactive = 0;
    bool logical_valid = false;
    bool host_stale = false;
    bool host_nodes_valid = true;
    int adapt_mode = NDG_ADAPT_DEVICE;
    AdaptWork work;
    int64_t *mortar_key_dev = nullptr;
    int64_t adjoint_store_budget = 0;        // bytes of stage states kept beyond one segment
    // The adjoint's checkpoints and stage-store blocks, kept for the next gradient call (grown when
    // a call needs more, released with the mesh).
    Arena adj_ck, adj_stage;
    int adjoint_c_mode = NDG_ADJOINT_C_RECOMPUTE;
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
// The translation taking periodic bar b onto its partner end.
v2 bar_translation(const HostMesh &h, int64_t b) {
    if (h.bar_shift.size() == h.bar_type.size() * 2) return v2{h.bar_shift[2 * b], h.bar_shift[2 * b + 1]};
    const int64_t d = h.bar_donor[(size_t)b];
    auto centre = [&](int64_t k) {
        const int64_t u = h.bars[2 * k], w = h.bars[2 * k + 1];
        return v2{(h.xy[2 * u] + h.xy[2 * w]) / 2, (h.xy[2 * u + 1] + h.xy[2 * w + 1]) / 2};
    };
    const v2 cb = centre(b), cd = centre(d);
    return v2{cd.x - cb.x, cd.y - cb.y};
}

double mag(double x, double y) { return std::sqrt(x * x + y * y); }

// Face geometry with the domain element on the left (src/entities/Face2D_t.cpp:56-66).
void face_frame(v2 f0, v2 f1, v2 l0, v2 l1, double &nx, double &ny, double &tx, double &ty) {
    tx = f1.x - f0.x;
    ty = f1.y - f0.y;
    const double len = mag(tx, ty);
    tx /= len;
    ty /= len;
    nx = ty;
    ny = -tx;
    const double lx = l1.x - l0.x, ly = l1.y - l0.y;
    const double sign = std::copysign(1.0, tx * lx + ty * ly);
    nx *= sign; ny *= sign; tx *= sign; ty *= sign;
}

// Face offsets and scales of the two sides (src/entities/Face2D_t.cpp:68-81): en = {L side start,
// L side end, R side start, R side end}, face nodes f0, f1.
void face_offsets(v2 f0, v2 f1, const v2 en[4], double offset[2], double scale[2]) {
    const double d0 = mag(en[1].x - en[0].x, en[1].y - en[0].y);
    const double d1 = mag(en[3].x - en[2].x, en[3].y - en[2].y);
    const double fd = mag(f1.x - f0.x, f1.y - f0.y);
    offset[0] = mag(f0.x - en[0].x, f0.y - en[0].y) / d0;
    offset[1] = mag(f1.x - en[2].x, f1.y - en[2].y) / d1;
    scale[0] = fd / d0;
    scale[1] = fd / d1;
}

// Edge Jacobian of side s of the bilinear quad P (include/functions/quad_metrics.tpp:3-8 on the
// side, src/entities/Element2D_t.cpp:768-772); it is the same at every node of the side.
double side_scale(const v2 P[4], int s) {
    if (s == 0 || s == 2) {
        const double eta = s == 0 ? -1.0 : 1.0;
        const double mx = ((1 - eta) * (P[1].x - P[0].x) + (1 + eta) * (P[2].x - P[3].x)) / 4;
        const double my = ((1 - eta) * (P[1].y - P[0].y) + (1 + eta) * (P[2].y - P[3].y)) / 4;
        return std::sqrt(mx * mx + my * my);
    }
    const double xi = s == 1 ? 1.0 : -1.0;
    const double mx = ((1 - xi) * (P[3].x - P[0].x) + (1 + xi) * (P[2].x - P[1].x)) / 4;
    const double my = ((1 - xi) * (P[3].y - P[0].y) + (1 + xi) * (P[2].y - P[1].y)) / 4;
    return std::sqrt(mx * mx + my * my);
}

uint64_t edge_key(int64_t a, int64_t b) { return ((uint64_t)a << 32) | (uint64_t)b; }

// Element2D_t::almost_equal / Mesh2D_t::almost_equal with 2 ulp (src/meshes/Mesh2D_t.cpp:3785-3792)
bool ulp2_equal(double a, double b) {
    return std::fabs(a - b) <= DBL_EPSILON * std::fabs(a + b) * 2 || std::fabs(a - b) < DBL_MIN;
}

// Interpolation (element side -> face nodes) and projection (face -> element side) matrices of one
// side of a mortar face, with the reference's arithmetic structure: project_to_faces' barycentric
// interpolation at xi = (x_face + 1) scale + 2 offset - 1 with the exact-node shortcut
// (src/meshes/Mesh2D_t.cpp:3820-3890), and project_to_elements' Gauss-quadrature L2 projection
// T_i w_j / w_i with the same shortcut, negated for the face's right side (:3935-4003). The
// right side walks the face nodes backwards. One matrix per key (element order, face order, side,
// class of mortar_class, interpolation or projection), for every order up to N_max: data holds
// them back to back, off[key] their offsets (-1 for the plain copies: equal orders, whole side).
void build_mortar_table(const Tables &t, int N_max, std::vector<double> &data, std::vector<int64_t> &off) {
    data.clear();
    off.assign((size_t)kMortarMatrixKeys, -1);
    for (int Ne = 1; Ne <= N_max; ++Ne)
        for (int Nf = Ne; Nf <= N_max; ++Nf)
            for (int side = 0; side < 2; ++side)
                for (int cls = 0; cls < kMortarClasses; ++cls) {
                    if (Ne == Nf && cls == 0) continue;
                    const double offset = class_offset(cls), scale = class_scale(cls);
                    const auto &xf = t.tab[NDG_TABLE_NODES][Nf], &wf = t.tab[NDG_TABLE_WEIGHTS][Nf];
                    const auto &xe = t.tab[NDG_TABLE_NODES][Ne], &we = t.tab[NDG_TABLE_WEIGHTS][Ne];
                    const auto &le = t.tab[NDG_TABLE_BARYCENTRIC_WEIGHTS][Ne];
                    // interpolation [Nf+1][Ne+1]
                    int64_t at = (int64_t)data.size();
                    off[(size_t)mortar_matrix_key(Ne, Nf, side, cls, 0)] = at;
                    data.resize(data.size() + (size_t)(Nf + 1) * (Ne + 1), 0.0);
                    for (int i = 0; i <= Nf; ++i) {
                        double *row = data.data() + at + (int64_t)i * (Ne + 1);
                        const double coord = (xf[side ? Nf - i : i] + 1) * scale + 2 * offset - 1;
                        int hit = -1;
                        for (int j = 0; j <= Ne; ++j)
                            if (ulp2_equal(coord, xe[j])) {
                                hit = j;
                                break;
                            }
                        if (hit >= 0) {
                            row[hit] = 1.0;
                            continue;
                        }
                        double den = 0.0;
                        for (int j = 0; j <= Ne; ++j) {
                            row[j] = le[j] / (coord - xe[j]);
                            den += row[j];
                        }
                        for (int j = 0; j <= Ne; ++j) row[j] /= den;
                    }
                    // projection [Ne+1][Nf+1]
                    at = (int64_t)data.size();
                    off[(size_t)mortar_matrix_key(Ne, Nf, side, cls, 1)] = at;
                    data.resize(data.size() + (size_t)(Ne + 1) * (Nf + 1), 0.0);
                    double *P = data.data() + at;
                    for (int j = 0; j <= Nf; ++j) {
                        const double coord = (xf[side ? Nf - j : j] + 1) * scale + 2 * offset - 1;
                        int hit = -1;
                        for (int k = 0; k <= Ne; ++k)
                            if (ulp2_equal(coord, xe[k])) {
                                hit = k;
                                break;
                            }
                        if (hit >= 0) {
                            P[(int64_t)hit * (Nf + 1) + j] += (side ? -wf[j] : wf[j]) / we[hit] * scale;
                            continue;
                        }
                        double ssum = 0.0;
                        for (int k = 0; k <= Ne; ++k) ssum += le[k] / (coord - xe[k]);
                        for (int i = 0; i <= Ne; ++i) {
                            const double T = le[i] / ((coord - xe[i]) * ssum);
                            P[(int64_t)i * (Nf + 1) + j] += (side ? -T : T) * wf[j] / we[i] * scale;
                        }
                    }
                }
}

template <typename T>
ndg_status device_upload(T **dst, const std::vector<T> &src, const char *what) {
    if (src.empty()) return NDG_OK;
    HIP_TRY(hipMalloc((void **)dst, src.size() * sizeof(T)), what);
    HIP_TRY(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), what);
    return NDG_OK;
}

// Copies src into the arena (fitted to its size) and points *view at it (null when empty).
template <typename T>
ndg_status upload_to(Arena &a, T **view, const std::vector<T> &src, const char *what) {
    *view = nullptr;
    if (src.empty()) return NDG_OK;
    HIP_TRY(arena_fit(a, src.size() * sizeof(T)), what);
    HIP_TRY(hipMemcpy(a.p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), what);
    *view = (T *)a.p;
    return NDG_OK;
}

// Fits the arena to n values of T and points *view at it.
template <typename T>
ndg_status fit_view(Arena &a, T **view, int64_t n, const char *what) {
    HIP_TRY(arena_fit(a, sizeof(T) * (size_t)std::max<int64_t>(n, 1)), what);
    *view = (T *)a.p;
    return NDG_OK;
}

void logical_free(LogicalStore &s) {
    for (Arena *a : {&s.verts, &s.order, &s.status, &s.rotation, &s.level, &s.link, &s.hang2, &s.hang, &s.periodic,
                     &s.span, &s.est, &s.dev_of, &s.ghost, &s.ghost_side, &s.ghost_zero})
        arena_free(*a);
}

void free_mesh(ndg_mesh *m) {
    if (!m) return;
    for (DeviceStore &st : m->store) store_free(st);
    for (LogicalStore &ls : m->lstore) logical_free(ls);
    for (Arena *a : {&m->work.flags, &m->work.newpos, &m->work.newghost, &m->work.parent, &m->work.corner,
                     &m->work.cnt_rec, &m->work.cnt_slot, &m->work.cnt_face, &m->work.blocks, &m->work.hist,
                     &m->work.counts})
        arena_free(*a);
    if (m->mortar_key_dev) (void)hipFree(m->mortar_key_dev);
    arena_free(m->adj_ck);
    arena_free(m->adj_stage);
    for (int n = 0; n <= kMaxOrder; ++n) {
        if (m->tab64[n]) (void)hipFree(m->tab64[n]);
        if (m->tab32[n]) (void)hipFree(m->tab32[n]);
    }
    for (void *p : {(void *)m->dt_dev, (void *)m->est_dev, (void *)m->pulses_dev, (void *)m->tab1d,
                    (void *)m->mortar_mats, (void *)m->clock})
        if (p) (void)hipFree(p);
    for (hipGraphExec_t &x : m->step_exec)
        if (x) (void)hipGraphExecDestroy(x);
    for (hipEvent_t ev : m->side_events) (void)hipEventDestroy(ev);
    for (hipStream_t st : m->side_streams) (void)hipStreamDestroy(st);
    if (m->own_stream) (void)hipStreamDestroy(m->own_stream);
    delete m;
}

// The ghost projection of a conforming boundary face (src/meshes/Mesh2D_t.cpp:4383-4418) is the
// reversed copy of the element trace up to roundoff: the face coordinate of ghost node j is the
// mirrored element node and the weight ratio w[j] / w[N - j] is 1 (the kernels use the copy). This
// checks that both hold to 1e-13 for the tables in use.
bool ghost_copy_is_close(const Tables &t, int N) {
    const auto &x = t.tab[NDG_TABLE_NODES][N];
    const auto &w = t.tab[NDG_TABLE_WEIGHTS][N];
    for (int j = 0; j <= N; ++j) {
        const double coord = (x[N - j] + 1) * 1.0 + 2 * 0.0 - 1;
        if (std::fabs(coord - x[N - j]) > 1e-13 || std::fabs(w[j] / w[N - j] - 1.0) > 1e-13) return false;
    }
    return true;
}

}  // namespace

namespace {

// The device tables every rebuild shares, uploaded once per mesh: the 1D tables of all orders, the
// mortar matrices of every key up to the tables' N_max, and the packed tables of order N (FP64).
ndg_status ensure_shared_tables(ndg_mesh *m) {
    const Tables &tab = m->tables->t;
    if (!m->tab1d) {
        std::vector<double> t1((size_t)(3 * (kMaxOrder + 1) * kTab1dStride), 0.0);
        for (int n = 0; n <= std::min(kMaxOrder, tab.N_max); ++n)
            for (int i = 0; i <= n; ++i) {
                t1[(3 * n + 0) * kTab1dStride + i] = tab.tab[NDG_TABLE_NODES][n][i];
                t1[(3 * n + 1) * kTab1dStride + i] = tab.tab[NDG_TABLE_WEIGHTS][n][i];
                t1[(3 * n + 2) * kTab1dStride + i] = tab.tab[NDG_TABLE_BARYCENTRIC_WEIGHTS][n][i];
            }
        ndg_status st = device_upload(&m->tab1d, t1, "upload 1D tables");
        if (st != NDG_OK) return st;
    }
    if (!m->mortar_mats) {
        std::vector<double> mats;
        build_mortar_table(tab, std::min(kMaxOrder, tab.N_max), mats, m->mortar_key_off);
        if (mats.empty()) mats.push_back(0.0);
        ndg_status st = device_upload(&m->mortar_mats, mats, "upload mortar matrices");
        if (st != NDG_OK) return st;
        if ((st = device_upload(&m->mortar_key_dev, m->mortar_key_off, "upload mortar matrix keys")) != NDG_OK) return st;
    }
    return NDG_OK;
}

ndg_status ensure_order_tables(ndg_mesh *m, int N) {
    if (m->tab64[N]) return NDG_OK;
    return device_upload(&m->tab64[N], pack_tables(m->tables->t, N), "upload tables");
}

// Builds the device mesh for the given element orders from the host topology kept in m, into the
// buffers of store: order groups, connectivity (group-local neighbour codes, boundary records
// with the ghost scale zeroed where an adaptation resized the ghost, mortar records and faces
// between different orders), vertices, and the state / trace buffers (uninitialised). Sets
// m->order, m->dev_of, m->host_off, m->n_dofs.
ndg_status build_device(ndg_mesh *m, const std::vector<int32_t> &order, DeviceMesh &d, DeviceStore &store) {
    const Tables &tab = m->tables->t;
    const int64_t K = (int64_t)order.size();
    const std::vector<int64_t> &link = m->link;
    const std::vector<char> &periodic_link = m->periodic_link;
    auto corner = [&](int64_t e, int k) {
        const int kk = k & 3;
        return v2{m->host_verts[8 * e + 2 * kk], m->host_verts[8 * e + 2 * kk + 1]};
    };
    d = DeviceMesh();
    d.K = K;
    m->order = order;

    // Order groups: elements sorted by N with a counting sort (stable, so each group keeps the
    // caller's order, e.g. a space-filling curve).
    std::vector<int64_t> perm((size_t)K);
    {
        std::vector<int64_t> at(kMaxOrder + 2, 0);
        for (int64_t e = 0; e < K; ++e) {
            if (order[e] < 1 || order[e] > kMaxOrder)
                return fail(NDG_ERR_UNSUPPORTED, "element order out of the compiled range");
            ++at[(size_t)order[e] + 1];
        }
        for (int n = 0; n <= kMaxOrder; ++n) at[(size_t)n + 1] += at[(size_t)n];
        for (int64_t e = 0; e < K; ++e) perm[(size_t)at[(size_t)order[e]]++] = e;
    }
    m->dev_of.assign((size_t)K, 0);
    for (int64_t k = 0; k < K; ++k) m->dev_of[perm[k]] = k;
    std::vector<int> group_of_N(kMaxOrder + 1, -1);
    for (int64_t k = 0; k < K; ++k) {
        const int N = order[perm[k]];
        if (group_of_N[N] < 0) {
            Group G;
            G.N = N;
            G.e_begin = k;
            if (!d.groups.empty()) {
                const Group &H = d.groups.back();
                G.q_off = H.q_off + 3 * (int64_t)(H.N + 1) * (H.N + 1) * H.K;
                G.tr_off = H.tr_off + 12 * (int64_t)(H.N + 1) * H.K;
            }
            group_of_N[N] = (int)d.groups.size();
            d.groups.push_back(G);
        }
        ++d.groups[group_of_N[N]].K;
    }
    {
        const Group &H = d.groups.back();
        d.n_q = H.q_off + 3 * (int64_t)(H.N + 1) * (H.N + 1) * H.K;
        d.n_tr = H.tr_off + 12 * (int64_t)(H.N + 1) * H.K;
    }
    d.N_min = d.groups.front().N;
    d.N_max = d.groups.back().N;
    {
        std::vector<GroupSpan> spans((size_t)(kMaxOrder + 1), GroupSpan{0, 0, 0, 0});
        for (const Group &G : d.groups) spans[(size_t)G.N] = GroupSpan{G.K, G.e_begin, G.q_off, G.tr_off};
        GroupSpan *view = nullptr;
        const ndg_status sst = upload_to(store.spans, &view, spans, "upload group spans");
        if (sst != NDG_OK) return sst;
    }
    for (const Group &G : d.groups)
        if (!ghost_copy_is_close(tab, G.N))
            return fail(NDG_ERR_UNSUPPORTED, "ndg_mesh_create: boundary ghost projection is not a reversed copy for "
                                             "this order");
    m->host_off.assign((size_t)K, 0);
    int64_t acc = 0;
    for (int64_t e = 0; e < K; ++e) {
        m->host_off[e] = acc;
        acc += (int64_t)(order[e] + 1) * (order[e] + 1);
    }
    m->n_dofs = acc;
    auto group_of = [&](int64_t e) -> const Group & { return d.groups[group_of_N[order[e]]]; };
    auto local_of = [&](int64_t e) { return m->dev_of[e] - group_of(e).e_begin; };
    auto trace_at = [&](int64_t e, int s) {
        const int64_t nf = order[e] + 1;
        return group_of(e).tr_off + local_of(e) * 12 * nf + s * 3 * nf;
    };

    // Boundary records; a ghost whose storage an adaptation resized has a zero scale.
    std::vector<BcSide> bcs = m->bc_base;
    for (size_t b = 0; b < bcs.size(); ++b)
        if (m->bar_zero[(size_t)m->bc_bar[b]]) bcs[b].sf = 0.0;
    for (const BcSide &b : bcs) d.has_inflow = d.has_inflow || b.kind == NDG_BC_INFLOW;

    // Device connectivity. Sides whose exterior state has another order are non-conforming: they
    // get a mortar record, and one mortar face per pair (two-sided for a neighbour, one-sided for
    // a periodic donor) fills their flux.
    std::vector<int32_t> side_nb((size_t)(4 * K));
    std::vector<MortarFace> faces;
    std::vector<int64_t> mortar_rec((size_t)(4 * K), -1);
    int64_t n_slot = 0;
    auto add_mortar_side = [&](int64_t es) {
        BcSide rec{};
        rec.kind = kKindMortar;
        rec.slot = n_slot;
        n_slot += 3 * (int64_t)(order[es / 4] + 1);
        mortar_rec[es] = (int64_t)bcs.size();
        bcs.push_back(rec);
        return rec.slot;
    };
    // A larger hanging side sums the projected fluxes of its faces, one block per small side
    // along it (the chain from link through hang_second).
    auto chain_pos = [&](int64_t small) {
        int64_t pos = 0;
        for (int64_t c = link[(size_t)link[small]]; c != small && c >= 0; c = m->hang_second[(size_t)c]) ++pos;
        return pos;
    };
    for (int64_t es = 0; es < 4 * K; ++es)
        if (m->hang_kind[es] == 2) {
            int n = 0;
            for (int64_t c = link[es]; c >= 0; c = m->hang_second[(size_t)c]) ++n;
            add_mortar_side(es);
            bcs.back().n_faces = n;
            n_slot += 3 * (int64_t)(n - 1) * (order[es / 4] + 1);
        }
    for (int64_t e = 0; e < K; ++e)
        for (int s = 0; s < 4; ++s) {
            const int64_t es = 4 * e + s, l = link[es];
            const int hk = m->hang_kind[es];
            const bool periodic = periodic_link[es] != 0;
            if (hk == 2 && periodic) {
                // a larger periodic side owns one face per piece at the other end, against a ghost
                // holding that piece's trace, each writing only this side's block for the piece
                const v2 f0 = corner(e, s), f1 = corner(e, s + 1);
                v2 PL[4] = {corner(e, 0), corner(e, 1), corner(e, 2), corner(e, 3)};
                const int64_t slot = bcs[(size_t)mortar_rec[es]].slot;
                int64_t h = 0;
                for (int64_t c = l; c >= 0; c = m->hang_second[(size_t)c], ++h) {
                    MortarFace fc{};
                    fc.NL = order[e];
                    fc.NR = order[c / 4];
                    fc.Nf = std::max(fc.NL, fc.NR);
                    fc.trL = trace_at(e, s);
                    fc.trR = trace_at(c / 4, (int)(c % 4));
                    double tx, ty;
                    face_frame(f0, f1, f0, f1, fc.nx, fc.ny, tx, ty);
                    fc.offset[0] = class_offset(m->span[(size_t)c]);
                    fc.scale[0] = class_scale(m->span[(size_t)c]);
                    fc.offset[1] = 0.0;
                    fc.scale[1] = 1.0;
                    fc.sfL = side_scale(PL, s);
                    fc.slotL = slot + h * 3 * (int64_t)(order[e] + 1);
                    fc.slotR = -1;
                    faces.push_back(fc);
                }
                continue;
            }
            if (hk == 2) continue;  // built from its small sides
            if (l < 0 || (hk == 0 && order[l / 4] == order[e])) continue;
            const int64_t r = l / 4;
            const int rs = (int)(l % 4);
            if (hk == 0 && !periodic && r < e) continue;  // the face belongs to its first element
            MortarFace fc{};
            fc.NL = order[e];
            fc.NR = order[r];
            fc.Nf = std::max(fc.NL, fc.NR);
            fc.trL = trace_at(e, s);
            fc.trR = trace_at(r, rs);
            const v2 f0 = corner(e, s), f1 = corner(e, s + 1);
            double tx, ty;
            face_frame(f0, f1, f0, f1, fc.nx, fc.ny, tx, ty);
            v2 en[4] = {f0, f1, f1, f0};
            if (!periodic) {
                en[2] = corner(r, rs);
                en[3] = corner(r, rs + 1);
            }
            face_offsets(f0, f1, en, fc.offset, fc.scale);
            v2 PL[4] = {corner(e, 0), corner(e, 1), corner(e, 2), corner(e, 3)};
            fc.sfL = side_scale(PL, s);
            fc.slotL = add_mortar_side(es);
            fc.slotR = -1;
            if (hk == 1 && periodic) {  // this whole side against its piece of the larger partner
                fc.offset[0] = 0.0;
                fc.scale[0] = 1.0;
                fc.offset[1] = class_offset(m->span[(size_t)es]);
                fc.scale[1] = class_scale(m->span[(size_t)es]);
            } else if (hk == 1) {  // a piece of the larger side l
                v2 PR[4] = {corner(r, 0), corner(r, 1), corner(r, 2), corner(r, 3)};
                fc.sfR = side_scale(PR, rs);
                const BcSide &big = bcs[(size_t)mortar_rec[l]];
                fc.slotR = big.slot + chain_pos(es) * 3 * (int64_t)(order[r] + 1);
            } else if (!periodic) {
                v2 PR[4] = {corner(r, 0), corner(r, 1), corner(r, 2), corner(r, 3)};
                fc.sfR = side_scale(PR, rs);
                fc.slotR = add_mortar_side(l);
            }
            faces.push_back(fc);
        }
    ndg_status st;
    if ((st = ensure_shared_tables(m)) != NDG_OK) return st;
    for (MortarFace &fc : faces)
        for (int side = 0; side < 2; ++side) {
            const int Ne = side ? fc.NR : fc.NL;
            const int cls = mortar_class(fc.offset[side], fc.scale[side]);
            if (cls < 0) return fail(NDG_ERR_UNSUPPORTED, "ndg_mesh_create: a non-conforming face covers neither a whole "
                                                          "element side nor a dyadic piece of one down to 1/16");
            const bool copy = Ne == fc.Nf && cls == 0;
            fc.mat[side] = copy ? -1 : m->mortar_key_off[(size_t)mortar_matrix_key(Ne, fc.Nf, side, cls, 0)];
            fc.mat[2 + side] = copy ? -1 : m->mortar_key_off[(size_t)mortar_matrix_key(Ne, fc.Nf, side, cls, 1)];
        }
    for (int64_t e = 0; e < K; ++e)
        for (int s = 0; s < 4; ++s) {
            const int64_t es = 4 * e + s, l = link[es];
            int32_t code;
            if (mortar_rec[es] >= 0) code = -(int32_t)(1 + mortar_rec[es]);
            else if (l < 0) code = (int32_t)l;
            else if (m->hang_kind[es]) return fail(NDG_ERR_INVALID, "ndg_mesh_create: hanging side without a mortar");
            else code = (int32_t)(local_of(l / 4) * 4 + l % 4);
            side_nb[4 * m->dev_of[e] + s] = code;
        }

    std::vector<double> verts((size_t)(8 * K));
    for (int64_t e = 0; e < K; ++e)
        for (int k = 0; k < 8; ++k) verts[8 * m->dev_of[e] + k] = m->host_verts[8 * e + k];
    if ((st = upload_to(store.verts, &d.verts, verts, "upload vertices")) != NDG_OK) return st;
    if ((st = upload_to(store.side_nb, &d.side_nb, side_nb, "upload connectivity")) != NDG_OK) return st;
    if ((st = upload_to(store.bc, &d.bc, bcs, "upload boundary records")) != NDG_OK) return st;
    d.n_bc = (int64_t)bcs.size();
    if ((st = upload_to(store.faces, &d.faces, faces, "upload mortar faces")) != NDG_OK) return st;
    d.mortar_mats = m->mortar_mats;
    d.n_faces = (int64_t)faces.size();
    d.mortar_nodes = 0;
    for (const MortarFace &fc : faces)
        d.mortar_nodes = std::max(d.mortar_nodes, std::max(fc.Nf, std::max(fc.NL, fc.NR)) + 1);
    d.n_side_flux = n_slot;
    d.side_flux = nullptr;
    if (n_slot > 0 && (st = fit_view(store.side_flux, &d.side_flux, n_slot, "hipMalloc side fluxes")) != NDG_OK)
        return st;
    d.tab1d = m->tab1d;
    for (Group &G : d.groups) {
        if ((st = ensure_order_tables(m, G.N)) != NDG_OK) return st;
        G.tab = m->tab64[G.N];
    }
    if ((st = fit_view(store.q, &d.q, d.n_q, "hipMalloc q")) != NDG_OK) return st;
    if ((st = fit_view(store.g, &d.g, d.n_q, "hipMalloc g")) != NDG_OK) return st;
    if ((st = fit_view(store.tr[0], &d.trace[0], d.n_tr, "hipMalloc traces")) != NDG_OK) return st;
    if ((st = fit_view(store.tr[1], &d.trace[1], d.n_tr, "hipMalloc traces")) != NDG_OK) return st;
    d.scratch_len = 1024;
    if ((st = fit_view(store.scratch, &d.scratch, d.scratch_len, "hipMalloc scratch")) != NDG_OK) return st;
    d.clock = m->clock;
    return NDG_OK;
}

}  // namespace

namespace {

template <typename T>
ndg_status download(std::vector<T> &dst, const void *src, int64_t n, hipStream_t s, const char *what) {
    dst.resize((size_t)n);
    if (n > 0) HIP_TRY(hipMemcpyAsync(dst.data(), src, sizeof(T) * (size_t)n, hipMemcpyDeviceToHost, s), what);
    return NDG_OK;
}

// Fits the arrays of a logical store to K elements and G ghosts and sets the view.
ndg_status fit_logical(LogicalStore &ls, LogicalMesh &L, int64_t K, int64_t G) {
    ndg_status st;
    L.K = K;
    L.n_ghost = G;
    if ((st = fit_view(ls.verts, &L.verts, 8 * K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.order, &L.order, K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.status, &L.status, K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.rotation, &L.rotation, K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.level, &L.level, K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.link, &L.link, 4 * K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.hang2, &L.hang2, 4 * K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.hang, &L.hang, 4 * K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.periodic, &L.periodic, 4 * K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.span, &L.span, 4 * K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.est, &L.est, 4 * K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.dev_of, &L.dev_of, K, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.ghost, &L.ghost, G, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.ghost_side, &L.ghost_side, G, "hipMalloc logical mesh")) != NDG_OK) return st;
    if ((st = fit_view(ls.ghost_zero, &L.ghost_zero, G, "hipMalloc logical mesh")) != NDG_OK) return st;
    return NDG_OK;
}

template <typename T, typename S>
ndg_status upload_into(T *dst, const std::vector<S> &src, const char *what) {
    if (src.empty()) return NDG_OK;
    std::vector<T> tmp(src.begin(), src.end());
    HIP_TRY(hipMemcpy(dst, tmp.data(), sizeof(T) * tmp.size(), hipMemcpyHostToDevice), what);
    return NDG_OK;
}

// The logical mesh of the device adaptation path from the host topology (once, or again after a
// host-path adaptation changed the mesh).
ndg_status ensure_logical(ndg_mesh *m) {
    if (m->logical_valid) return NDG_OK;
    const int64_t K = m->d.K, G = (int64_t)m->bc_base.size();
    LogicalMesh &L = m->lmesh[m->lactive];
    ndg_status st;
    if ((st = fit_logical(m->lstore[m->lactive], L, K, G)) != NDG_OK) return st;
    std::vector<double> est((size_t)(4 * K), 0.0);
    if (m->est_error.size() == (size_t)K)
        for (int64_t e = 0; e < K; ++e) {
            est[4 * e] = m->est_error[e];
            est[4 * e + 1] = m->est_sigma[e];
            est[4 * e + 2] = m->est_refine[e] ? 1.0 : 0.0;
            est[4 * e + 3] = m->est_coarsen[e] ? 1.0 : 0.0;
        }
    std::vector<int64_t> ghost_side((size_t)G);
    std::vector<int8_t> ghost_zero((size_t)G);
    for (int64_t g = 0; g < G; ++g) {
        const int64_t bar = m->bc_bar[(size_t)g];
        ghost_side[g] = m->bar_side[(size_t)bar];
        ghost_zero[g] = m->bar_zero[(size_t)bar] ? 1 : 0;
    }
    if ((st = upload_into(L.verts, m->host_verts, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.order, m->order, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.status, m->status, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.rotation, m->rotation, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.level, m->split_level, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.link, m->link, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.hang2, m->hang_second, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.span, m->span, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.hang, m->hang_kind, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.periodic, m->periodic_link, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.est, est, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.dev_of, m->dev_of, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.ghost, m->bc_base, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.ghost_side, ghost_side, "upload logical mesh")) != NDG_OK) return st;
    if ((st = upload_into(L.ghost_zero, ghost_zero, "upload logical mesh")) != NDG_OK) return st;
    m->logical_valid = true;
    return NDG_OK;
}

// Refreshes the host copies of the mesh description from the logical mesh after device
// adaptations (orders, statuses, rotations, split levels, vertices, links, ghosts, estimates and
// the device permutation).
ndg_status ensure_host(ndg_mesh *m) {
    if (!m->host_stale) return NDG_OK;
    const LogicalMesh &L = m->lmesh[m->lactive];
    const hipStream_t s = m->stream;
    std::vector<int8_t> hang, per;
    std::vector<double> est;
    ndg_status st;
    if ((st = download(m->host_verts, L.verts, 8 * L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(m->order, L.order, L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(m->status, L.status, L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(m->rotation, L.rotation, L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(m->split_level, L.level, L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(m->link, L.link, 4 * L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(m->hang_second, L.hang2, 4 * L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(m->span, L.span, 4 * L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(hang, L.hang, 4 * L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(per, L.periodic, 4 * L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(est, L.est, 4 * L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(m->dev_of, L.dev_of, L.K, s, "download logical mesh")) != NDG_OK) return st;
    if ((st = download(m->bc_base, L.ghost, L.n_ghost, s, "download logical mesh")) != NDG_OK) return st;
    HIP_TRY(hipStreamSynchronize(s), "download logical mesh");
    m->hang_kind.assign(hang.begin(), hang.end());
    m->periodic_link.assign(per.begin(), per.end());
    const size_t K = (size_t)L.K;
    m->est_error.resize(K);
    m->est_sigma.resize(K);
    m->est_refine.resize(K);
    m->est_coarsen.resize(K);
    for (size_t e = 0; e < K; ++e) {
        m->est_error[e] = est[4 * e];
        m->est_sigma[e] = est[4 * e + 1];
        m->est_refine[e] = est[4 * e + 2] != 0.0;
        m->est_coarsen[e] = est[4 * e + 3] != 0.0;
    }
    m->host_off.assign(K, 0);
    int64_t acc = 0;
    for (size_t e = 0; e < K; ++e) {
        m->host_off[e] = acc;
        acc += (int64_t)(m->order[e] + 1) * (m->order[e] + 1);
    }
    m->n_dofs = acc;
    m->host_stale = false;
    return NDG_OK;
}

}  // namespace

namespace {

// Hilbert-curve status of the elements (src/functions/Hilbert_splitting.cpp, statuses H, A, R, B =
// 0..3): tables of the first / last / inner element status by entry and exit side, and of the
// children's order and statuses by parent rotation and status.
constexpr int kFirstStatus[4] = {3, 0, 1, 2};
constexpr int kLastStatus[4] = {1, 2, 3, 0};
constexpr int kInnerStatus[4][4] = {{0, 0, 1, 1}, {3, 3, 2, 2}, {3, 3, 2, 2}, {0, 0, 1, 1}};
constexpr int kChildOrder[4][4][4] = NDG_HILBERT_CHILD_ORDER;
constexpr int kChildStatus[4][4][4] = NDG_HILBERT_CHILD_STATUS;

}  // namespace

extern "C" ndg_status ndg_hilbert_child_order(int32_t status, int32_t rotation, int32_t order[4], int32_t statuses[4]) {
    if (status < 0 || status > 3 || rotation < 0 || rotation > 3)
        return fail(NDG_ERR_INVALID, "ndg_hilbert_child_order: status and rotation must be in 0..3");
    for (int k = 0; k < 4; ++k) {
        if (order) order[k] = kChildOrder[rotation][status][k];
        if (statuses) statuses[k] = kChildStatus[rotation][status][k];
    }
    return NDG_OK;
}

extern "C" ndg_status ndg_hilbert_element_status(int32_t incoming_side, int32_t outgoing_side, int32_t *status) {
    if (!status) return fail(NDG_ERR_INVALID, "ndg_hilbert_element_status: null output");
    if (incoming_side < -1 || incoming_side > 3 || outgoing_side < -1 || outgoing_side > 3 ||
        (incoming_side < 0 && outgoing_side < 0))
        return fail(NDG_ERR_INVALID, "ndg_hilbert_element_status: sides must be in 0..3 (-1 for a curve end, not both)");
    if (incoming_side < 0) *status = kFirstStatus[outgoing_side];
    else if (outgoing_side < 0) *status = kLastStatus[incoming_side];
    else *status = kInnerStatus[incoming_side][outgoing_side];
    return NDG_OK;
}

namespace {

// The side of an element through which the centre offset (dx, dy) points (the element's bottom,
// right, top, left are sides 0..3).
int side_towards(double dx, double dy) {
    if (std::fabs(dy) > std::fabs(dx)) return dy < 0 ? 0 : 2;
    return dx < 0 ? 3 : 1;
}

// compute_element_status (src/meshes/Mesh2D_t.cpp:4055-4178): rotation from the corner angles
// around the centre, status from the sides through which the curve enters (previous element) and
// leaves (next element).
void element_statuses(const HostMesh &h, std::vector<int32_t> &status, std::vector<int32_t> &rotation) {
    const double pi = 3.14159265358979323846;
    const double targets[4] = {-3 * pi / 4, -pi / 4, pi / 4, 3 * pi / 4};
    const int64_t K = (int64_t)h.quads.size() / 4;
    std::vector<double> cx((size_t)K), cy((size_t)K);
    for (int64_t e = 0; e < K; ++e) {
        double x = 0.0, y = 0.0;
        for (int k = 0; k < 4; ++k) {
            x += h.xy[2 * h.quads[4 * e + k]];
            y += h.xy[2 * h.quads[4 * e + k] + 1];
        }
        cx[e] = x / 4;
        cy[e] = y / 4;
    }
    status.assign((size_t)K, 0);
    rotation.assign((size_t)K, 0);
    for (int64_t e = 0; e < K; ++e) {
        long rot = 0;
        for (int k = 0; k < 4; ++k) {
            const int64_t node = h.quads[4 * e + k];
            const double angle = std::atan2(h.xy[2 * node + 1] - cy[e], h.xy[2 * node] - cx[e]);
            const double off = std::atan2(std::sin(angle - targets[k]), std::cos(angle - targets[k]));
            const long turns = std::lround(off * 2 / pi + 4 * (off < 0));
            rot += turns * (turns < 4);
        }
        rotation[e] = (int32_t)(rot / 4);
        if (K == 1) status[e] = 0;
        else if (e == 0) status[e] = kFirstStatus[side_towards(cx[1] - cx[0], cy[1] - cy[0])];
        else if (e == K - 1) status[e] = kLastStatus[side_towards(cx[e - 1] - cx[e], cy[e - 1] - cy[e])];
        else
            status[e] = kInnerStatus[side_towards(cx[e - 1] - cx[e], cy[e - 1] - cy[e])]
                                    [side_towards(cx[e + 1] - cx[e], cy[e + 1] - cy[e])];
    }
}

// Faces of the description: conforming neighbours (build_faces, src/meshes/Mesh2D_t.cpp:1110-1154),
// boundary ghosts, periodic couplings and 2:1 hanging sides, as element-side links.
ndg_status build_topology(const HostMesh &h, Topology &t) {
    const int64_t K = (int64_t)h.quads.size() / 4, B = (int64_t)h.bar_type.size();
    auto P = [&](int64_t node) { return v2{h.xy[2 * node], h.xy[2 * node + 1]}; };

    // Directed edges of the domain elements (build_faces, src/meshes/Mesh2D_t.cpp:1110-1154): the
    // element sides bucketed by start node (CSR, sides in increasing order within a bucket), so the
    // side a -> b is found by scanning the few sides that start at a.
    const int64_t n_nodes = (int64_t)h.xy.size() / 2;
    for (int64_t v : h.quads)
        if (v < 0 || v >= n_nodes) return fail(NDG_ERR_INVALID, "mesh topology: element node id out of range");
    for (int64_t v : h.bars)
        if (v < 0 || v >= n_nodes) return fail(NDG_ERR_INVALID, "mesh topology: boundary node id out of range");
    auto end_node = [&](int64_t es) { return h.quads[4 * (es / 4) + (es % 4 + 1) % 4]; };
    std::vector<int64_t> first((size_t)n_nodes + 1, 0), by_start((size_t)(4 * K));
    for (int64_t es = 0; es < 4 * K; ++es) ++first[(size_t)h.quads[es] + 1];
    for (int64_t v = 0; v < n_nodes; ++v) first[(size_t)v + 1] += first[(size_t)v];
    {
        std::vector<int64_t> fill(first.begin(), first.end() - 1);
        for (int64_t es = 0; es < 4 * K; ++es) by_start[(size_t)fill[(size_t)h.quads[es]]++] = es;
    }
    auto find_side = [&](int64_t a, int64_t b) -> int64_t {  // the element side a -> b, or -1
        for (int64_t i = first[(size_t)a]; i < first[(size_t)a + 1]; ++i)
            if (end_node(by_start[(size_t)i]) == b) return by_start[(size_t)i];
        return -1;
    };
    for (int64_t a = 0; a < n_nodes; ++a)
        for (int64_t i = first[(size_t)a]; i < first[(size_t)a + 1]; ++i)
            for (int64_t j = first[(size_t)a]; j < i; ++j)
                if (end_node(by_start[(size_t)i]) == end_node(by_start[(size_t)j]))
                    return fail(NDG_ERR_INVALID, "mesh topology: an edge is used twice with the same orientation");
    // link[e*4+s]: the element side whose trace is the exterior state of side s of element e (a
    // neighbour, or the donor side of a periodic pair), or -(1 + boundary record) / INT64_MIN.
    std::vector<int64_t> &link = t.link;
    std::vector<char> &periodic_link = t.periodic_link;
    link.assign((size_t)(4 * K), INT64_MIN);
    periodic_link.assign((size_t)(4 * K), 0);
    int64_t &n_faces = t.n_faces;
    n_faces = 0;
    for (int64_t es = 0; es < 4 * K; ++es) {
        const int64_t r = find_side(end_node(es), h.quads[es]);
        if (r >= 0) {
            link[es] = r;
            if (r > es) ++n_faces;
        }
    }
    // Ghost (bar) elements attach to the domain side with the same two nodes
    // (src/meshes/Mesh2D_t.cpp:1156-1181). Their nodes after attachment are {P(s+1), P(s)}.
    std::vector<int64_t> &bar_side = t.bar_side;
    bar_side.assign((size_t)B, -1);
    for (int64_t b = 0; b < B; ++b) {
        const int64_t a = h.bars[2 * b], c = h.bars[2 * b + 1];
        int64_t es = find_side(c, a);
        if (es < 0) es = find_side(a, c);
        if (es < 0) return fail(NDG_ERR_INVALID, "mesh topology: boundary bar matches no element edge");
        bar_side[b] = es;
        ++n_faces;
    }
    std::vector<BcSide> &bcs = t.bcs;
    std::vector<int64_t> &bc_bar = t.bc_bar;
    bcs.clear();
    bc_bar.clear();
    std::vector<int64_t> side_bar((size_t)(4 * K), -1);
    for (int64_t b = 0; b < B; ++b) {
        const int64_t es = bar_side[b];
        if (link[es] != INT64_MIN || side_bar[es] != -1)
            return fail(NDG_ERR_INVALID, "mesh topology: a boundary bar lies on an interior or already bounded edge");
        side_bar[es] = b;
    }
    for (int64_t b = 0; b < B; ++b) {
        const int64_t es = bar_side[b], e = es / 4;
        const int s = (int)(es % 4);
        const int64_t u0 = h.quads[4 * e + s], u1 = h.quads[4 * e + (s + 1) % 4];
        const int kind = h.bar_type[b];
        if (kind == NDG_BC_PERIODIC) {
            if (h.bar_donor.empty()) return fail(NDG_ERR_INVALID, "mesh topology: periodic bar without donor array");
            const int64_t d = h.bar_donor[b];
            if (d < 0 || d >= B) return fail(NDG_ERR_INVALID, "mesh topology: periodic donor out of range");
            // local_interfaces (Mesh2D_t.cpp:4712-4730): the ghost carries the donor side's trace
            // unchanged (and takes the donor element's order), so the side couples to the donor
            // element side directly.
            link[es] = bar_side[d];
            periodic_link[es] = 1;
            continue;
        }
        if (kind < NDG_BC_NONE || kind > NDG_BC_OUTFLOW) return fail(NDG_ERR_INVALID, "mesh topology: unknown boundary kind");
        BcSide rec{};
        rec.kind = kind;
        rec.slot = -1;
        // ghost nodes {a, b, b, a} with a = P(s+1), b = P(s) after face attachment
        const v2 ga = P(u1), gb = P(u0);
        rec.ax = ga.x; rec.ay = ga.y; rec.bx = gb.x; rec.by = gb.y;
        const int64_t bar_a = h.bars[2 * b];
        v2 G[4];
        if (bar_a == u1) {  // bar already reversed with respect to the element edge
            G[0] = ga; G[1] = gb; G[2] = gb; G[3] = ga;
        } else {            // "inside out" bar: only its first two nodes are swapped
            G[0] = ga; G[1] = gb; G[2] = ga; G[3] = gb;
        }
        rec.sf = side_scale(G, 0);
        face_frame(P(u0), P(u1), P(u0), P(u1), rec.nx, rec.ny, rec.tx, rec.ty);
        link[es] = -(int64_t)(1 + bcs.size());
        bcs.push_back(rec);
        bc_bar.push_back(b);
    }
    // Hanging sides: a side a -> b with neither a neighbour nor a bar, covered by a chain of
    // unmatched sides b -> m1 -> ... -> a of other elements whose nodes lie on it at dyadic
    // positions, pieces of 1/2 down to 1/2^kMaxSpanLevel of the side (what h-refinement leaves,
    // src/meshes/Mesh2D_t.cpp:5661-5985). The larger side links to the piece at a, each piece to
    // the larger side and (hang_second) to the next piece towards b; span is the piece's mortar
    // class along a -> b.
    std::vector<char> &hang_kind = t.hang_kind;
    std::vector<int64_t> &hang_second = t.hang_second;
    hang_kind.assign((size_t)(4 * K), 0);
    hang_second.assign((size_t)(4 * K), -1);
    t.span.assign((size_t)(4 * K), 0);
    {
        std::vector<int64_t> chain;
        std::vector<int> cls;
        for (int64_t es = 0; es < 4 * K; ++es) {
            if (link[es] != INT64_MIN || hang_kind[es]) continue;
            const int64_t a = h.quads[es], b = end_node(es);
            const v2 pa = P(a), pb = P(b);
            const double ex = pb.x - pa.x, ey = pb.y - pa.y, len = mag(ex, ey);
            // position of node n along a -> b, or -1 off the segment
            auto along = [&](int64_t n) {
                const v2 p = P(n);
                const double u = ((p.x - pa.x) * ex + (p.y - pa.y) * ey) / (len * len);
                if (std::fabs(pa.x + u * ex - p.x) > 1e-12 * len || std::fabs(pa.y + u * ey - p.y) > 1e-12 * len)
                    return -1.0;
                return u;
            };
            chain.clear();
            cls.clear();
            int64_t cur = b;
            double ucur = 1.0;
            bool ok = true;
            while (cur != a && ok) {
                int64_t pick = -1;
                double unext = -1.0;
                for (int64_t i = first[(size_t)cur]; i < first[(size_t)cur + 1]; ++i) {  // cur -> m along the side
                    const int64_t c = by_start[(size_t)i];
                    if (c == es || hang_kind[c] || link[c] != INT64_MIN) continue;
                    const int64_t mn = end_node(c);
                    const double u = mn == a ? 0.0 : along(mn);
                    if (u < 0.0 || u >= ucur - 1e-12) continue;
                    if (u > unext) pick = c, unext = u;
                }
                // the piece [unext, ucur] must be j / 2^d .. (j + 1) / 2^d with 1 <= d <= kMaxSpanLevel
                int d = 1;
                while (d <= kMaxSpanLevel && std::fabs((ucur - unext) * (1 << d) - 1.0) > 1e-9) ++d;
                const double j = unext * (1 << d);
                if (pick < 0 || d > kMaxSpanLevel || std::fabs(j - std::round(j)) > 1e-9) ok = false;
                else {
                    chain.push_back(pick);
                    cls.push_back(make_class(d, (int)std::lround(j)));
                    cur = end_node(pick);
                    ucur = unext;
                }
            }
            if (!ok || chain.size() < 2) continue;
            // chain runs from b to a; the larger side's first piece is the one at a
            const size_t n = chain.size();
            hang_kind[es] = 2;
            link[es] = chain[n - 1];
            for (size_t k = 0; k < n; ++k) {
                hang_kind[chain[k]] = 1;
                link[chain[k]] = es;
                t.span[chain[k]] = (int8_t)cls[k];
                hang_second[chain[k]] = k == 0 ? -1 : chain[k - 1];
            }
            n_faces += (int64_t)n;
        }
    }
    // Periodic sides of different lengths (splits next to a period, split_interfaces,
    // src/meshes/Mesh2D_t.cpp:6343-6756): a periodic side whose partner end is covered by shorter
    // bars is a larger side, hanging as an interior one would be; its pieces are the shorter bars
    // whose donor it is, placed along it by their translation. Each end keeps its own faces, so the
    // larger side adds one face per piece beyond the one its bar counted.
    if (!h.bar_donor.empty()) {
        auto blen = [&](int64_t k) {
            const v2 u = P(h.bars[2 * k]), w = P(h.bars[2 * k + 1]);
            return mag(w.x - u.x, w.y - u.y);
        };
        std::vector<std::vector<int64_t>> pieces((size_t)B);
        for (int64_t x = 0; x < B; ++x)
            if (h.bar_type[x] == NDG_BC_PERIODIC && blen(x) < blen(h.bar_donor[x]) * (1 - 1e-9))
                pieces[(size_t)h.bar_donor[x]].push_back(x);
        for (int64_t b = 0; b < B; ++b) {
            if (pieces[(size_t)b].empty()) continue;
            const int64_t es = bar_side[b];
            const v2 pa = P(h.quads[es]), pb = P(end_node(es));
            const double ex = pb.x - pa.x, ey = pb.y - pa.y, l2 = ex * ex + ey * ey;
            struct Piece {
                double lo;
                int64_t side;
                int cls;
            };
            std::vector<Piece> ps;
            for (int64_t x : pieces[(size_t)b]) {
                const v2 T = bar_translation(h, x);
                double u[2];
                for (int k = 0; k < 2; ++k) {
                    const v2 q = P(h.bars[2 * x + k]);
                    u[k] = ((q.x + T.x - pa.x) * ex + (q.y + T.y - pa.y) * ey) / l2;
                }
                const double lo = std::min(u[0], u[1]), size = std::fabs(u[1] - u[0]);
                int d = 1;
                while (d <= kMaxSpanLevel && std::fabs(size * (1 << d) - 1.0) > 1e-9) ++d;
                const double j = lo * (1 << d);
                if (d > kMaxSpanLevel || std::fabs(j - std::round(j)) > 1e-9)
                    return fail(NDG_ERR_INVALID, "mesh topology: a periodic piece is not a dyadic part of its partner");
                ps.push_back(Piece{lo, bar_side[x], make_class(d, (int)std::lround(j))});
            }
            std::sort(ps.begin(), ps.end(), [](const Piece &p1, const Piece &p2) { return p1.lo < p2.lo; });
            double covered = 0.0;
            for (const Piece &pc : ps) {
                if (std::fabs(pc.lo - covered) > 1e-9)
                    return fail(NDG_ERR_INVALID, "mesh topology: a periodic side is not covered by its partner's pieces");
                covered = class_offset(pc.cls) + class_scale(pc.cls);
            }
            if (std::fabs(covered - 1.0) > 1e-9)
                return fail(NDG_ERR_INVALID, "mesh topology: a periodic side is not covered by its partner's pieces");
            hang_kind[es] = 2;
            link[es] = ps[0].side;
            for (size_t k = 0; k < ps.size(); ++k) {
                const int64_t c = ps[k].side;
                hang_kind[c] = 1;
                link[c] = es;
                t.span[c] = (int8_t)ps[k].cls;
                hang_second[c] = k + 1 < ps.size() ? ps[k + 1].side : -1;
            }
            n_faces += (int64_t)ps.size() - 1;
        }
    }
    for (int64_t es = 0; es < 4 * K; ++es)
        if (link[es] == INT64_MIN)
            return fail(NDG_ERR_INVALID, "mesh topology: element side has neither a neighbour nor a boundary bar");
    return NDG_OK;
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
    std::vector<int32_t> order((size_t)K, desc->initial_N);
    if (desc->quad_N) std::copy(desc->quad_N, desc->quad_N + K, order.begin());
    for (int64_t e = 0; e < K; ++e) {
        const int N = order[e];
        if (!order_supported(N))
            return fail(NDG_ERR_UNSUPPORTED, "ndg_mesh_create: element order outside the compiled range 1..16");
        if (N > tables->t.N_max) return fail(NDG_ERR_INVALID, "ndg_mesh_create: element order exceeds the tables' N_max");
        if (N < desc->initial_N)
            return fail(NDG_ERR_UNSUPPORTED, "ndg_mesh_create: element order below initial_N (boundary ghosts would "
                                             "carry a higher order than their element)");
    }
    for (int64_t e = 0; e < 4 * K; ++e)
        if (desc->quads[e] < 0 || desc->quads[e] >= NN) return fail(NDG_ERR_INVALID, "ndg_mesh_create: quad node id out of range");
    for (int64_t b = 0; b < 2 * B; ++b)
        if (desc->bars[b] < 0 || desc->bars[b] >= NN) return fail(NDG_ERR_INVALID, "ndg_mesh_create: bar node id out of range");

    HostMesh host;
    host.xy.assign(desc->xy, desc->xy + 2 * NN);
    host.quads.assign(desc->quads, desc->quads + 4 * K);
    if (B > 0) {
        host.bars.assign(desc->bars, desc->bars + 2 * B);
        host.bar_type.assign(desc->bar_type, desc->bar_type + B);
        if (desc->bar_donor) host.bar_donor.assign(desc->bar_donor, desc->bar_donor + B);
    }
    Topology topo;
    ndg_status tst = build_topology(host, topo);
    if (tst != NDG_OK) return tst;
    std::unique_ptr<ndg_mesh, void (*)(ndg_mesh *)> m(new ndg_mesh, free_mesh);
    m->tables_own = *tables;
    m->tol_min = desc->tolerance_min;
    m->tol_max = desc->tolerance_max;
    m->n_faces = topo.n_faces;
    m->initial_N = desc->initial_N;
    m->maximum_N = desc->maximum_N;
    m->max_split_level = desc->max_split_level;
    m->link = std::move(topo.link);
    m->periodic_link = std::move(topo.periodic_link);
    m->hang_kind = std::move(topo.hang_kind);
    m->hang_second = std::move(topo.hang_second);
    m->span = std::move(topo.span);
    m->bc_base = std::move(topo.bcs);
    m->bc_bar = std::move(topo.bc_bar);
    m->bar_side = std::move(topo.bar_side);
    m->bar_zero.assign(host.bar_type.size(), 0);
    element_statuses(host, m->status, m->rotation);
    m->split_level.assign((size_t)K, 0);
    m->host_verts.resize((size_t)(8 * K));
    for (int64_t e = 0; e < K; ++e)
        for (int k = 0; k < 4; ++k) {
            const int64_t node = host.quads[4 * e + k];
            m->host_verts[8 * e + 2 * k] = host.xy[2 * node];
            m->host_verts[8 * e + 2 * k + 1] = host.xy[2 * node + 1];
        }
    m->host = std::move(host);

    // Everything above is host work; the device is needed from here on.
    int32_t ndev = 0;
    ndg_device_count(&ndev);
    if (ndev <= 0) return fail(NDG_ERR_NO_DEVICE, "ndg_mesh_create: no HIP device visible");
    HIP_TRY(hipStreamCreateWithFlags(&m->own_stream, hipStreamNonBlocking), "hipStreamCreate");
    m->stream = m->own_stream;
    HIP_TRY(hipMalloc((void **)&m->dt_dev, sizeof(double)), "hipMalloc dt");
    HIP_TRY(hipMalloc((void **)&m->clock, sizeof(double) * 4), "hipMalloc clock");
    HIP_TRY(hipMemsetAsync(m->clock, 0, sizeof(double) * 4, m->stream), "hipMemset clock");
    ndg_status st = build_device(m.get(), order, m->d, m->store[0]);
    if (st != NDG_OK) return st;
    DeviceMesh &d = m->d;
    HIP_TRY(hipMemsetAsync(d.q, 0, sizeof(double) * d.n_q, m->stream), "hipMemset q");
    HIP_TRY(hipMemsetAsync(d.g, 0, sizeof(double) * d.n_q, m->stream), "hipMemset g");
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
    if (n_elements) *n_elements = mesh->d.K;
    if (n_faces) *n_faces = mesh->n_faces;
    if (n_dofs_per_field) *n_dofs_per_field = mesh->n_dofs;
    if (N_min) *N_min = mesh->d.N_min;
    if (N_max) *N_max = mesh->d.N_max;
    return NDG_OK;
}

// ------------------------------------------------------------------------------------------
// State

namespace {

// Device layout [element (device order)][3][(N+1)^2] <-> packed per-field host arrays in the
// caller's element order. download: device -> host, else host -> device.
ndg_status copy_state(ndg_mesh *m, double *dev, double *const fields[3], bool download, const char *what) {
    const DeviceMesh &d = m->d;
    m->staging.resize((size_t)d.n_q);
    if (download) {
        HIP_TRY(hipMemcpyAsync(m->staging.data(), dev, sizeof(double) * d.n_q, hipMemcpyDeviceToHost, m->stream), what);
        HIP_TRY(hipStreamSynchronize(m->stream), what);
    }
    for (int64_t e = 0; e < d.K; ++e) {
        const int N = m->order[e];
        const int64_t ne = (int64_t)(N + 1) * (N + 1);
        const Group *G = nullptr;
        for (const Group &H : d.groups)
            if (H.N == N) { G = &H; break; }
        double *blk = m->staging.data() + G->q_off + (m->dev_of[e] - G->e_begin) * 3 * ne;
        for (int f = 0; f < 3; ++f) {
            if (!fields[f]) continue;
            if (download) std::memcpy(fields[f] + m->host_off[e], blk + f * ne, sizeof(double) * ne);
            else std::memcpy(blk + f * ne, fields[f] + m->host_off[e], sizeof(double) * ne);
        }
    }
    if (!download) {
        HIP_TRY(hipMemcpyAsync(dev, m->staging.data(), sizeof(double) * d.n_q, hipMemcpyHostToDevice, m->stream), what);
        HIP_TRY(hipStreamSynchronize(m->stream), what);
    }
    return NDG_OK;
}

// FP32 compensated form: refresh the FP64 state, RK register and traces from (qh, gl32).
ndg_status refresh_fp64(ndg_mesh *m) {
    if (m->d.precision != NDG_PRECISION_FP32_COMPENSATED) return NDG_OK;
    HIP_TRY(launch_from_c32(m->d, m->stream), "from_c32 kernel");
    HIP_TRY(launch_traces(m->d, m->stream), "traces kernel");
    return NDG_OK;
}

// After the FP64 state changed: rebuild (qh, gl32) and their traces.
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
        if (!d.qh) {
            DeviceStore &sto = mesh->store[mesh->active];
            ndg_status st;
            if ((st = fit_view(sto.qh, &d.qh, d.n_q, "hipMalloc qh")) != NDG_OK) return st;
            if ((st = fit_view(sto.gl32, &d.gl32, d.n_q, "hipMalloc gl32")) != NDG_OK) return st;
            if ((st = fit_view(sto.tr32[0], &d.trace32[0], d.n_tr, "hipMalloc traces32")) != NDG_OK) return st;
            if ((st = fit_view(sto.tr32[1], &d.trace32[1], d.n_tr, "hipMalloc traces32")) != NDG_OK) return st;
            d.side_flux32 = nullptr;
            if (d.n_side_flux > 0 &&
                (st = fit_view(sto.side_flux32, &d.side_flux32, d.n_side_flux, "hipMalloc side fluxes")) != NDG_OK)
                return st;
            for (Group &G : d.groups) {
                if (!mesh->tab32[G.N]) {
                    const std::vector<double> t64 = pack_tables(mesh->tables->t, G.N);
                    const std::vector<float> t32(t64.begin(), t64.end());
                    if ((st = device_upload(&mesh->tab32[G.N], t32, "upload tab32")) != NDG_OK) return st;
                }
                G.tab32 = mesh->tab32[G.N];
            }
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
    if (ic->kind != NDG_IC_PLANE_WAVE && ic->kind != NDG_IC_PULSES && ic->kind != NDG_IC_PERIODIC_PULSES)
        return fail(NDG_ERR_INVALID, "ndg_set_ic: unknown kind");
    const bool with_pulses = ic->kind != NDG_IC_PLANE_WAVE;
    if (with_pulses && (ic->n_pulses < 0 || (ic->n_pulses > 0 && !ic->pulses)))
        return fail(NDG_ERR_INVALID, "ndg_set_ic: bad pulse list");
    if (ic->kind == NDG_IC_PERIODIC_PULSES && !(ic->period_x > 0.0 && ic->period_y > 0.0))
        return fail(NDG_ERR_INVALID, "ndg_set_ic: periodic pulses need positive periods");
    DeviceMesh &d = mesh->d;
    const int np = with_pulses ? ic->n_pulses : 0;
    if (mesh->pulses_dev) {
        HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_set_ic sync");
        HIP_TRY(hipFree(mesh->pulses_dev), "hipFree");
        mesh->pulses_dev = nullptr;
    }
    if (np > 0) {
        HIP_TRY(hipMalloc((void **)&mesh->pulses_dev, sizeof(double) * 4 * np), "hipMalloc pulses");
        HIP_TRY(hipMemcpy(mesh->pulses_dev, ic->pulses, sizeof(double) * 4 * np, hipMemcpyHostToDevice), "upload pulses");
    }
    HIP_TRY(launch_ic(d, ic->kind, np, mesh->pulses_dev, ic->period_x, ic->period_y, mesh->stream), "ic kernel");
    mesh->ic_kind = ic->kind;
    mesh->ic_pulses.assign(ic->pulses, ic->pulses + 4 * np);
    mesh->ic_period[0] = ic->period_x;
    mesh->ic_period[1] = ic->period_y;
    HIP_TRY(hipMemsetAsync(d.g, 0, sizeof(double) * d.n_q, mesh->stream), "hipMemset g");
    HIP_TRY(launch_traces(d, mesh->stream), "traces kernel");
    ndg_status st = refresh_c32(mesh);
    if (st != NDG_OK) return st;
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_set_ic sync");
    return NDG_OK;
}

extern "C" ndg_status ndg_set_state(ndg_mesh *mesh, const double *p, const double *u, const double *v) {
    if (!mesh || !p || !u || !v) return fail(NDG_ERR_INVALID, "ndg_set_state: null argument");
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    DeviceMesh &d = mesh->d;
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_set_state sync");
    double *const ins[3] = {const_cast<double *>(p), const_cast<double *>(u), const_cast<double *>(v)};
    ndg_status st = copy_state(mesh, d.q, ins, false, "upload state");
    if (st != NDG_OK) return st;
    HIP_TRY(hipMemsetAsync(d.g, 0, sizeof(double) * d.n_q, mesh->stream), "hipMemset g");
    HIP_TRY(launch_traces(d, mesh->stream), "traces kernel");
    if ((st = refresh_c32(mesh)) != NDG_OK) return st;
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_set_state sync");
    return NDG_OK;
}

extern "C" ndg_status ndg_get_state(ndg_mesh *mesh, double *p, double *u, double *v) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_get_state: null mesh");
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    ndg_status st = refresh_fp64(mesh);
    if (st != NDG_OK) return st;
    double *const outs[3] = {p, u, v};
    return copy_state(mesh, mesh->d.q, outs, true, "download state");
}

namespace {

// The right-hand side at time t without disturbing a running solve: launch_rhs sets the device
// clock (which the next step of ndg_solve continues from) to t, so its 4 values are saved before
// and put back after.
ndg_status rhs_keeping_clock(ndg_mesh *mesh, double t) {
    DeviceMesh &d = mesh->d;
    double saved[4];
    HIP_TRY(hipMemcpyAsync(saved, d.clock, sizeof saved, hipMemcpyDeviceToHost, mesh->stream), "save clock");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "save clock sync");
    HIP_TRY(launch_rhs(d, t, mesh->stream), "rhs kernel");
    HIP_TRY(hipMemcpyAsync(d.clock, saved, sizeof saved, hipMemcpyHostToDevice, mesh->stream), "restore clock");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "restore clock sync");
    return NDG_OK;
}

// The uniform-grid output of get_solution (5 planes: x, y, p, u, v) or, with `complete`, of
// get_complete_solution (11 planes: also dq/dt of the current state at `time` and the plane-wave
// errors), each plane [K][n][n] in the caller's element order; a null plane pointer is skipped.
ndg_status interpolated_output(ndg_mesh *mesh, const char *who, bool complete, double time,
                               double *const *planes) {
    const Tables &tab = mesh->tables->t;
    const int64_t n = tab.n_interp;
    if (n < 2) return fail(NDG_ERR_INVALID, (std::string(who) + ": the tables have fewer than 2 output points").c_str());
    if (n > 4096) return fail(NDG_ERR_INVALID, (std::string(who) + ": too many output points").c_str());
    ndg_status st = refresh_fp64(mesh);
    if (st != NDG_OK) return st;
    DeviceMesh &d = mesh->d;
    if (complete) {
        if (!d.rhs) {
            const ndg_status rs = fit_view(mesh->store[mesh->active].rhs, &d.rhs, d.n_q, "hipMalloc rhs");
            if (rs != NDG_OK) return rs;
        }
        if ((st = rhs_keeping_clock(mesh, time)) != NDG_OK) return st;
    }
    const int64_t K = d.K, nn = n * n;
    const int n_planes = complete ? 11 : 5;
    std::vector<double> interp;
    std::vector<int64_t> i_off(kMaxOrder + 1, -1);
    for (const Group &G : d.groups) {
        i_off[G.N] = (int64_t)interp.size();
        const auto &I = tab.tab[NDG_TABLE_INTERPOLATION][G.N];
        interp.insert(interp.end(), I.begin(), I.end());
    }
    std::vector<OutputRec> recs((size_t)K);
    for (int64_t e = 0; e < K; ++e) {
        const int N = mesh->order[e];
        const Group *G = nullptr;
        for (const Group &H : d.groups)
            if (H.N == N) G = &H;
        OutputRec &R = recs[(size_t)e];
        R.q_off = G->q_off + (mesh->dev_of[e] - G->e_begin) * 3 * (int64_t)(N + 1) * (N + 1);
        R.dev = mesh->dev_of[e];
        R.i_off = i_off[N];
        R.N = N;
        R.pad = 0;
    }
    OutputRec *recs_dev = nullptr;
    double *interp_dev = nullptr, *out_dev = nullptr;
    st = device_upload(&recs_dev, recs, "upload output records");
    if (st == NDG_OK) st = device_upload(&interp_dev, interp, "upload interpolation matrices");
    if (st == NDG_OK) {
        const hipError_t e = hipMalloc((void **)&out_dev, sizeof(double) * n_planes * K * nn);
        if (e != hipSuccess) st = hip_fail(e, "hipMalloc output");
    }
    if (st == NDG_OK) {
        hipError_t e = launch_output(d.q, complete ? d.rhs : nullptr, d.verts, recs_dev, interp_dev, d.tab1d, (int)n,
                                     K, time, d.wave_speed, out_dev, mesh->stream);
        for (int f = 0; f < n_planes && e == hipSuccess; ++f)
            if (planes[f])
                e = hipMemcpyAsync(planes[f], out_dev + f * K * nn, sizeof(double) * K * nn, hipMemcpyDeviceToHost,
                                   mesh->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(mesh->stream);
        if (e != hipSuccess) st = hip_fail(e, "output kernel");
    }
    for (void *ptr : {(void *)recs_dev, (void *)interp_dev, (void *)out_dev})
        if (ptr) (void)hipFree(ptr);
    return st;
}

}  // namespace

extern "C" ndg_status ndg_get_solution(ndg_mesh *mesh, double *x, double *y, double *p, double *u, double *v) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_get_solution: null mesh");
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    double *const planes[5] = {x, y, p, u, v};
    return interpolated_output(mesh, "ndg_get_solution", false, 0.0, planes);
}

extern "C" ndg_status ndg_get_complete_solution(ndg_mesh *mesh, double time, const ndg_complete_solution *out) {
    if (!mesh || !out) return fail(NDG_ERR_INVALID, "ndg_get_complete_solution: null argument");
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    if (!std::isfinite(time)) return fail(NDG_ERR_INVALID, "ndg_get_complete_solution: time is not finite");
    double *const planes[11] = {out->x,     out->y,     out->p,     out->u,
                                out->v,     out->dp_dt, out->du_dt, out->dv_dt,
                                out->p_analytical_error, out->u_analytical_error, out->v_analytical_error};
    ndg_status st = interpolated_output(mesh, "ndg_get_complete_solution", true, time, planes);
    if (st != NDG_OK) return st;
    const size_t K = (size_t)mesh->d.K;
    const bool have_est = mesh->est_error.size() == K;
    for (size_t e = 0; e < K; ++e) {
        if (out->N) out->N[e] = mesh->order[e];
        if (out->p_error) out->p_error[e] = have_est ? mesh->est_error[e] : 0.0;
        if (out->p_sigma) out->p_sigma[e] = have_est ? mesh->est_sigma[e] : 0.0;
        if (out->refine) out->refine[e] = have_est ? mesh->est_refine[e] : 0;
        if (out->coarsen) out->coarsen[e] = have_est ? mesh->est_coarsen[e] : 0;
        // the reference estimates the pressure only (Mesh2D_t.cu:5158-5171 calls estimate_p_error)
        if (out->u_error) out->u_error[e] = 0.0;
        if (out->v_error) out->v_error[e] = 0.0;
        if (out->u_sigma) out->u_sigma[e] = 0.0;
        if (out->v_sigma) out->v_sigma[e] = 0.0;
        if (out->split_level) out->split_level[e] = mesh->split_level[e];
        if (out->status) out->status[e] = mesh->status[e];
        if (out->rotation) out->rotation[e] = mesh->rotation[e];
    }
    return NDG_OK;
}

extern "C" ndg_status ndg_get_traces(ndg_mesh *mesh, double *out) {
    if (!mesh || !out) return fail(NDG_ERR_INVALID, "ndg_get_traces: null argument");
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    ndg_status st = refresh_fp64(mesh);
    if (st != NDG_OK) return st;
    const DeviceMesh &d = mesh->d;
    mesh->staging.resize((size_t)d.n_tr);
    HIP_TRY(hipMemcpyAsync(mesh->staging.data(), d.trace[d.cur], sizeof(double) * d.n_tr, hipMemcpyDeviceToHost,
                           mesh->stream), "download traces");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_get_traces sync");
    int64_t o = 0;
    for (int64_t e = 0; e < d.K; ++e) {
        const int N = mesh->order[e];
        const int64_t len = 12 * (int64_t)(N + 1);
        for (const Group &G : d.groups)
            if (G.N == N) {
                std::memcpy(out + o, mesh->staging.data() + G.tr_off + (mesh->dev_of[e] - G.e_begin) * len,
                            sizeof(double) * len);
                break;
            }
        o += len;
    }
    return NDG_OK;
}

extern "C" ndg_status ndg_compute_rhs(ndg_mesh *mesh, double t, double *Gp, double *Gu, double *Gv) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_compute_rhs: null mesh");
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    ndg_status st = refresh_fp64(mesh);
    if (st != NDG_OK) return st;
    DeviceMesh &d = mesh->d;
    if (!d.rhs) {
        const ndg_status rs = fit_view(mesh->store[mesh->active].rhs, &d.rhs, d.n_q, "hipMalloc rhs");
        if (rs != NDG_OK) return rs;
    }
    if ((st = rhs_keeping_clock(mesh, t)) != NDG_OK) return st;
    double *const outs[3] = {Gp, Gu, Gv};
    return copy_state(mesh, d.rhs, outs, true, "download rhs");
}

extern "C" ndg_status ndg_get_delta_t(ndg_mesh *mesh, double cfl, double *dt) {
    if (!mesh || !dt) return fail(NDG_ERR_INVALID, "ndg_get_delta_t: null argument");
    if (!(cfl > 0.0)) return fail(NDG_ERR_INVALID, "ndg_get_delta_t: CFL must be positive");
    HIP_TRY(launch_delta_t(mesh->d, cfl, mesh->dt_dev, mesh->stream), "delta_t kernels");
    HIP_TRY(hipMemcpyAsync(dt, mesh->dt_dev, sizeof(double), hipMemcpyDeviceToHost, mesh->stream), "download dt");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_get_delta_t sync");
    return NDG_OK;
}

namespace {

void drop_graphs(ndg_mesh *m) {
    for (hipGraphExec_t &x : m->step_exec)
        if (x) {
            (void)hipGraphExecDestroy(x);
            x = nullptr;
        }
    m->graph_precision = -1;
}

// One RK step from the device clock (clock_begin with set = 0), launched on s; with `parallel`
// the order groups of each stage are forked onto side streams and joined (graph branches).
ndg_status launch_step_body(ndg_mesh *m, double dt, bool parallel, hipStream_t s) {
    DeviceMesh &d = m->d;
    HIP_TRY(launch_clock_begin(d, 0.0, 0, dt, s), "clock kernel");
    const int ng = (int)d.groups.size();
    for (int st = 0; st < 3; ++st) {
        HIP_TRY(launch_mortar(d, s), "mortar kernel");
        if (parallel && ng > 1) {
            // one branch per group, the groups with the most bytes forked first (their launches
            // are the longest)
            std::vector<int> gorder(ng);
            for (int g = 0; g < ng; ++g) gorder[g] = g;
            auto bytes = [&](int g) {
                const double nf = d.groups[g].N + 1;
                return (double)d.groups[g].K * (88 * nf * nf + 192 * nf + 80);
            };
            std::stable_sort(gorder.begin(), gorder.end(), [&](int x, int y) { return bytes(x) > bytes(y); });
            HIP_TRY(hipEventRecord(m->side_events[0], s), "fork");
            for (int g : gorder) {
                HIP_TRY(hipStreamWaitEvent(m->side_streams[g], m->side_events[0], 0), "fork wait");
                HIP_TRY(launch_group_stage(d, g, st, kAm[st], kGm[st] * dt, m->side_streams[g]), "stage kernel");
                HIP_TRY(hipEventRecord(m->side_events[1 + g], m->side_streams[g]), "join");
                HIP_TRY(hipStreamWaitEvent(s, m->side_events[1 + g], 0), "join wait");
            }
        } else {
            for (int g = 0; g < ng; ++g)
                HIP_TRY(launch_group_stage(d, g, st, kAm[st], kGm[st] * dt, s), "stage kernel");
        }
        d.cur = 1 - d.cur;
    }
    HIP_TRY(launch_clock_end(d, dt, s), "clock kernel");
    return NDG_OK;
}

ndg_status ensure_side_streams(ndg_mesh *m) {
    const size_t ng = m->d.groups.size();
    while (m->side_streams.size() < ng) {
        hipStream_t st;
        HIP_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "hipStreamCreate");
        m->side_streams.push_back(st);
    }
    while (m->side_events.size() < ng + 1) {
        hipEvent_t ev;
        HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate");
        m->side_events.push_back(ev);
    }
    return NDG_OK;
}

// Captures the step graphs for dt (both starting trace buffers); leaves `cur` as it was.
ndg_status capture_steps(ndg_mesh *m, double dt) {
    drop_graphs(m);
    const bool parallel = m->graph_mode == NDG_GRAPH_PARALLEL;
    if (parallel) {
        ndg_status st = ensure_side_streams(m);
        if (st != NDG_OK) return st;
    }
    DeviceMesh &d = m->d;
    const int cur0 = d.cur;
    for (int k = 0; k < 2; ++k) {
        d.cur = k;
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        ndg_status st = launch_step_body(m, dt, parallel, m->stream);
        hipGraph_t captured = nullptr;
        const hipError_t e = hipStreamEndCapture(m->stream, &captured);
        graph = captured;
        if (st != NDG_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            d.cur = cur0;
            return st;
        }
        if (e != hipSuccess) {
            d.cur = cur0;
            return hip_fail(e, "hipStreamEndCapture");
        }
        const hipError_t ei = hipGraphInstantiate(&m->step_exec[k], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            d.cur = cur0;
            return hip_fail(ei, "hipGraphInstantiate");
        }
    }
    d.cur = cur0;
    m->graph_dt = dt;
    m->graph_precision = d.precision;
    return NDG_OK;
}

// One step starting from the device clock: a graph replay when graphs are on, else direct
// launches.
ndg_status step_from_clock(ndg_mesh *m, double dt) {
    DeviceMesh &d = m->d;
    if (m->graph_mode == NDG_GRAPH_OFF) return launch_step_body(m, dt, false, m->stream);
    if (!m->step_exec[0] || m->graph_dt != dt || m->graph_precision != d.precision) {
        ndg_status st = capture_steps(m, dt);
        if (st != NDG_OK) return st;
    }
    HIP_TRY(hipGraphLaunch(m->step_exec[d.cur], m->stream), "hipGraphLaunch");
    d.cur = 1 - d.cur;  // three stages flip the trace buffer three times
    return NDG_OK;
}

}  // namespace

extern "C" ndg_status ndg_step(ndg_mesh *mesh, double t, double dt) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_step: null mesh");
    HIP_TRY(launch_clock_begin(mesh->d, t, 1, dt, mesh->stream), "clock kernel");
    return step_from_clock(mesh, dt);
}

extern "C" ndg_status ndg_mesh_set_graph(ndg_mesh *mesh, int32_t mode) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_mesh_set_graph: null mesh");
    if (mode != NDG_GRAPH_OFF && mode != NDG_GRAPH_SEQUENTIAL && mode != NDG_GRAPH_PARALLEL)
        return fail(NDG_ERR_INVALID, "ndg_mesh_set_graph: unknown mode");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_mesh_set_graph sync");
    drop_graphs(mesh);
    mesh->graph_mode = mode;
    return NDG_OK;
}

extern "C" ndg_status ndg_mesh_autotune(ndg_mesh *mesh, int32_t repeats) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_mesh_autotune: null mesh");
    if (repeats < 1) return fail(NDG_ERR_INVALID, "ndg_mesh_autotune: repeats must be >= 1");
    DeviceMesh &d = mesh->d;
    const int prec = d.precision;
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_mesh_autotune sync");
    drop_graphs(mesh);
    // the probes read the mortar side fluxes of the current traces
    HIP_TRY(launch_mortar(d, mesh->stream), "mortar kernel");
    hipEvent_t ev[2];
    HIP_TRY(hipEventCreate(&ev[0]), "hipEventCreate");
    HIP_TRY(hipEventCreate(&ev[1]), "hipEventCreate");
    ndg_status status = NDG_OK;
    for (int gi = 0; gi < (int)d.groups.size() && status == NDG_OK; ++gi) {
        Group &G = d.groups[gi];
        const int emax = stage_max_epb(G.N);
        std::vector<int> epbs;
        for (int e : {emax, (emax + 1) / 2, (emax + 3) / 4, 1})
            if (std::find(epbs.begin(), epbs.end(), e) == epbs.end()) epbs.push_back(e);
        // the FP32 form keeps its low state words in the order of the full tiles: block order only
        if (prec == NDG_PRECISION_FP32_COMPENSATED) epbs.assign(1, emax);
        float best = INFINITY;
        Group::Tune best_t = G.tune[prec];
        for (int e : epbs)
            for (int remap : {1, 0}) {
                G.tune[prec].epb = e;
                G.tune[prec].remap = remap;
                hipError_t he = launch_group_probe(d, gi, mesh->stream);
                if (he == hipSuccess) he = hipEventRecord(ev[0], mesh->stream);
                for (int r = 0; r < repeats && he == hipSuccess; ++r) he = launch_group_probe(d, gi, mesh->stream);
                if (he == hipSuccess) he = hipEventRecord(ev[1], mesh->stream);
                if (he == hipSuccess) he = hipEventSynchronize(ev[1]);
                float ms = 0.0f;
                if (he == hipSuccess) he = hipEventElapsedTime(&ms, ev[0], ev[1]);
                if (he != hipSuccess) {
                    status = hip_fail(he, "autotune probe");
                    break;
                }
                if (ms < best) {
                    best = ms;
                    best_t = G.tune[prec];
                }
            }
        G.tune[prec] = best_t;
    }
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    return status;
}

extern "C" ndg_status ndg_mesh_tuning(const ndg_mesh *mesh, int32_t *orders, int32_t *epb, int32_t *remap,
                                      int32_t capacity, int32_t *n_groups) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_mesh_tuning: null mesh");
    const DeviceMesh &d = mesh->d;
    if (n_groups) *n_groups = (int32_t)d.groups.size();
    if (capacity < (int32_t)d.groups.size() && (orders || epb || remap))
        return fail(NDG_ERR_INVALID, "ndg_mesh_tuning: output too small");
    for (size_t g = 0; g < d.groups.size(); ++g) {
        const Group &G = d.groups[g];
        const int e = G.tune[d.precision].epb;
        if (orders) orders[g] = G.N;
        if (epb) epb[g] = e >= 1 ? e : stage_max_epb(G.N);
        if (remap) remap[g] = G.tune[d.precision].remap;
    }
    return NDG_OK;
}

namespace {

// Launch shapes chosen by ndg_mesh_autotune stay with their order when the mesh is rebuilt.
void carry_tuning(const DeviceMesh &from, DeviceMesh &to) {
    for (Group &G : to.groups)
        for (const Group &H : from.groups)
            if (H.N == G.N) std::copy(H.tune, H.tune + 2, G.tune);
}

// The host-side description, topology and ordering of a mesh, swapped out while the splitting
// branch of ndg_adapt builds its replacement.
struct MeshHostState {
    HostMesh host;
    std::vector<int64_t> link, hang_second, bc_bar, bar_side, dev_of, host_off;
    std::vector<char> periodic_link, hang_kind, bar_zero;
    std::vector<int8_t> span;
    std::vector<BcSide> bc_base;
    std::vector<int32_t> status, rotation, split_level, order;
    std::vector<double> host_verts;
    int64_t n_faces = 0, n_dofs = 0;
};

void swap_host_state(ndg_mesh *m, MeshHostState &s) {
    std::swap(m->host, s.host);
    std::swap(m->link, s.link);
    std::swap(m->hang_second, s.hang_second);
    std::swap(m->span, s.span);
    std::swap(m->bc_bar, s.bc_bar);
    std::swap(m->bar_side, s.bar_side);
    std::swap(m->dev_of, s.dev_of);
    std::swap(m->host_off, s.host_off);
    std::swap(m->periodic_link, s.periodic_link);
    std::swap(m->hang_kind, s.hang_kind);
    std::swap(m->bar_zero, s.bar_zero);
    std::swap(m->bc_base, s.bc_base);
    std::swap(m->status, s.status);
    std::swap(m->rotation, s.rotation);
    std::swap(m->split_level, s.split_level);
    std::swap(m->order, s.order);
    std::swap(m->host_verts, s.host_verts);
    std::swap(m->n_faces, s.n_faces);
    std::swap(m->n_dofs, s.n_dofs);
}

// quad_map (include/functions/quad_map.tpp:3-8)
v2 quad_point(double xi, double eta, const v2 P[4]) {
    return v2{(P[0].x * (1 - xi) * (1 - eta) + P[1].x * (xi + 1) * (1 - eta) + P[2].x * (xi + 1) * (eta + 1) +
               P[3].x * (1 - xi) * (eta + 1)) / 4,
              (P[0].y * (1 - xi) * (1 - eta) + P[1].y * (xi + 1) * (1 - eta) + P[2].y * (xi + 1) * (eta + 1) +
               P[3].y * (1 - xi) * (eta + 1)) / 4};
}

double cross(v2 a, v2 b) { return a.x * b.y - a.y * b.x; }

// inverse_quad_map (include/functions/inverse_quad_map.tpp:46-88): the bilinear map's inverse by
// its quadratic, the root inside the unit square.
v2 reference_point(v2 g, const v2 P[4]) {
    const v2 e{P[1].x - P[0].x, P[1].y - P[0].y}, f{P[3].x - P[0].x, P[3].y - P[0].y};
    const v2 gg{P[0].x - P[1].x + P[2].x - P[3].x, P[0].y - P[1].y + P[2].y - P[3].y};
    const v2 h{g.x - P[0].x, g.y - P[0].y};
    const double k2 = cross(gg, f), k1 = cross(e, f) + cross(h, gg), k0 = cross(h, e);
    if (std::fabs(k2) < DBL_MIN) {
        const double ux = e.x * k1 - gg.x * k0, uy = e.y * k1 - gg.y * k0;
        const double u = std::fabs(ux) > std::fabs(uy) ? (h.x * k1 + f.x * k0) / ux : (h.y * k1 + f.y * k0) / uy;
        return v2{u * 2 - 1, -k0 / k1 * 2 - 1};
    }
    double w = k1 * k1 - 4 * k0 * k2;
    if (w < 0.0) return v2{-1, -1};
    w = std::sqrt(w);
    const double ik2 = 1 / (2 * k2);
    double v = (-k1 - w) * ik2;
    double ux = e.x + gg.x * v, uy = e.y + gg.y * v;
    double u = std::fabs(ux) > std::fabs(uy) ? (h.x - f.x * v) / ux : (h.y - f.y * v) / uy;
    if (u < 0.0 || u > 1.0 || v < 0.0 || v > 1.0) {
        v = (-k1 + w) * ik2;
        ux = e.x + gg.x * v;
        uy = e.y + gg.y * v;
        u = std::fabs(ux) > std::fabs(uy) ? (h.x - f.x * v) / ux : (h.y - f.y * v) / uy;
    }
    return v2{u * 2 - 1, v * 2 - 1};
}

// The splitting branch of Mesh2D_t::adapt (src/meshes/Mesh2D_t.cpp:1993-2268; hp_adapt
// :5186-5659; split_boundaries :6232-6343): every element with would_h_refine becomes four children
// (corner s: {corner s, midpoint of side s, centre, midpoint of side s - 1} at local positions s,
// s + 1, s + 2, s + 3) placed in the Hilbert child order, with the parent's order, split level + 1,
// the child status and the parent's rotation; would_p_refine elements are raised by 2 orders; a
// boundary ghost on a split side becomes two ghosts (fresh geometry), a ghost whose element is
// raised gets fresh geometry. The faces are rebuilt from the new elements and ghosts. Elements
// with a periodic side (split across the period by the device path only), or whose split would
// cover a side of an unsplit neighbour with more than two faces, are refused (NDG_ERR_UNSUPPORTED,
// nothing changed).
// The last estimate of the elements an adaptation keeps (new element k is old element from[k],
// unchanged when keep[k]); created and raised elements start from zero, as the reference's new
// Element2D_t objects do (src/meshes/Mesh2D_t.cu:6206-6232).
void carry_estimates(ndg_mesh *m, size_t K_old, const std::vector<int64_t> &from, const std::vector<char> &keep) {
    const size_t Kn = from.size();
    const bool have = m->est_error.size() == K_old;
    std::vector<double> er(Kn, 0.0), sg(Kn, 0.0);
    std::vector<int32_t> rf(Kn, 0), cs(Kn, 0);
    for (size_t k = 0; have && k < Kn; ++k)
        if (keep[k]) {
            er[k] = m->est_error[(size_t)from[k]];
            sg[k] = m->est_sigma[(size_t)from[k]];
            rf[k] = m->est_refine[(size_t)from[k]];
            cs[k] = m->est_coarsen[(size_t)from[k]];
        }
    m->est_error = std::move(er);
    m->est_sigma = std::move(sg);
    m->est_refine = std::move(rf);
    m->est_coarsen = std::move(cs);
}

ndg_status adapt_split(ndg_mesh *mesh, const std::vector<int32_t> &ref, const std::vector<double> &sig,
                       int32_t N_max, int32_t max_split_level, int64_t *n_changed) {
    const HostMesh &h = mesh->host;
    const int64_t K = (int64_t)h.quads.size() / 4, B = (int64_t)h.bar_type.size();
    std::vector<char> split((size_t)K, 0), raise((size_t)K, 0);
    int64_t n_split = 0, n_raise = 0;
    for (int64_t e = 0; e < K; ++e) {
        if (ref[e] && sig[e] < 1.0 && mesh->split_level[e] < max_split_level) split[e] = 1, ++n_split;
        else if (ref[e] && sig[e] >= 1.0 && mesh->order[e] + 2 <= N_max) raise[e] = 1, ++n_raise;
    }
    for (int64_t e = 0; e < K; ++e) {
        if (raise[e] && (mesh->order[e] + 2 > kMaxOrder || mesh->order[e] + 2 > mesh->tables->t.N_max))
            return fail(NDG_ERR_UNSUPPORTED, "ndg_adapt: a raised order exceeds the compiled range or the tables");
        if (!split[e]) continue;
        for (int s = 0; s < 4; ++s) {
            const int64_t es = 4 * e + s;
            // a piece of 1/2^kMaxSpanLevel of a larger side that splits alone would leave faces of
            // half that size
            if (mesh->hang_kind[es] == 1 && class_level(mesh->span[es]) >= kMaxSpanLevel &&
                !split[mesh->link[es] / 4])
                return fail(NDG_ERR_UNSUPPORTED, "ndg_adapt: a split would leave a face smaller than 1/16 of a side");
        }
    }
    auto node_xy = [&](const std::vector<double> &xy, int64_t n) { return v2{xy[2 * n], xy[2 * n + 1]}; };
    auto end_node = [&](int64_t es) { return h.quads[4 * (es / 4) + (es % 4 + 1) % 4]; };
    HostMesh nh;
    nh.xy = h.xy;
    std::unordered_map<uint64_t, int64_t> mids;
    auto mid_key = [](int64_t a, int64_t b) { return edge_key(std::min(a, b), std::max(a, b)); };
    // a larger side's midpoint is the node where its pieces cross 1/2: the start of the piece that
    // ends there (pieces run against the larger side)
    // (a larger periodic side's pieces lie at the other end of the period: it splits at its own
    // midpoint)
    for (int64_t es = 0; es < 4 * K; ++es)
        if (mesh->hang_kind[es] == 2 && !mesh->periodic_link[es])
            for (int64_t c = mesh->link[es]; c >= 0; c = mesh->hang_second[(size_t)c])
                if (class_offset(mesh->span[c]) + class_scale(mesh->span[c]) == 0.5)
                    mids[mid_key(h.quads[es], end_node(es))] = h.quads[c];
    auto midpoint = [&](int64_t a, int64_t b) {
        auto it = mids.find(mid_key(a, b));
        if (it != mids.end()) return it->second;
        const int64_t md = (int64_t)nh.xy.size() / 2;
        nh.xy.push_back((nh.xy[2 * a] + nh.xy[2 * b]) / 2);
        nh.xy.push_back((nh.xy[2 * a + 1] + nh.xy[2 * b + 1]) / 2);
        mids[mid_key(a, b)] = md;
        return md;
    };
    const int64_t Knew = K + 3 * n_split;
    nh.quads.assign((size_t)(4 * Knew), -1);
    std::vector<int32_t> order((size_t)Knew), status((size_t)Knew), rotation((size_t)Knew), level((size_t)Knew);
    std::vector<int64_t> parent((size_t)Knew);
    std::vector<char> child((size_t)Knew, 0);
    int64_t pos = 0;
    for (int64_t e = 0; e < K; ++e) {
        const int64_t *q = &h.quads[4 * e];
        if (!split[e]) {
            for (int k = 0; k < 4; ++k) nh.quads[4 * pos + k] = q[k];
            order[pos] = mesh->order[e] + (raise[e] ? 2 : 0);
            status[pos] = mesh->status[e];
            rotation[pos] = mesh->rotation[e];
            level[pos] = mesh->split_level[e];
            parent[pos] = e;
            ++pos;
            continue;
        }
        int64_t md[4];
        for (int s = 0; s < 4; ++s) md[s] = midpoint(q[s], q[(s + 1) % 4]);
        double cx = 0.0, cy = 0.0;
        for (int k = 0; k < 4; ++k) {
            cx += nh.xy[2 * q[k]];
            cy += nh.xy[2 * q[k] + 1];
        }
        const int64_t c = (int64_t)nh.xy.size() / 2;
        nh.xy.push_back(cx / 4);
        nh.xy.push_back(cy / 4);
        const int *co = kChildOrder[mesh->rotation[e]][mesh->status[e]];
        const int *cs = kChildStatus[mesh->rotation[e]][mesh->status[e]];
        for (int s = 0; s < 4; ++s) {
            const int64_t at = pos + co[s];
            nh.quads[4 * at + s] = q[s];
            nh.quads[4 * at + (s + 1) % 4] = md[s];
            nh.quads[4 * at + (s + 2) % 4] = c;
            nh.quads[4 * at + (s + 3) % 4] = md[(s + 3) % 4];
            order[at] = mesh->order[e];
            status[at] = cs[s];
            rotation[at] = mesh->rotation[e];
            level[at] = mesh->split_level[e] + 1;
            parent[at] = e;
            child[at] = 1;
        }
        pos += 4;
    }
    std::vector<char> bar_zero;
    std::vector<int64_t> bar_new((size_t)B, -1), bar_count((size_t)B, 1), bar_parent;
    const bool periodic_bars = !h.bar_donor.empty();
    for (int64_t b = 0; b < B; ++b) {
        const int64_t es = mesh->bar_side[b], e = es / 4;
        bar_new[b] = (int64_t)nh.bar_type.size();
        const v2 T = periodic_bars && h.bar_type[b] == NDG_BC_PERIODIC ? bar_translation(h, b) : v2{0.0, 0.0};
        auto push = [&](int64_t n0, int64_t n1, char zero) {
            nh.bars.push_back(n0);
            nh.bars.push_back(n1);
            nh.bar_type.push_back(h.bar_type[b]);
            if (periodic_bars) {
                nh.bar_donor.push_back(-1);
                nh.bar_shift.push_back(T.x);
                nh.bar_shift.push_back(T.y);
            }
            bar_parent.push_back(b);
            bar_zero.push_back(zero);
        };
        if (split[e]) {  // two ghosts: (P(s+1), m) and (m, P(s))
            const int64_t u0 = h.quads[es], u1 = end_node(es), md = midpoint(u0, u1);
            push(u1, md, 0);
            push(md, u0, 0);
            bar_count[b] = 2;
        } else {
            push(h.bars[2 * b], h.bars[2 * b + 1], mesh->bar_zero[b] && !raise[e]);
        }
    }
    // A periodic bar's donor is the bar at the partner end holding its translated centre (its
    // match, or the whole side it is a piece of), or, for a side whose partner is in pieces, the
    // first piece that touches it. The candidates are the new bars of the parent's partners: its
    // donor, the bars whose donor it is, and the other pieces of its donor.
    std::vector<std::vector<int64_t>> donor_of((size_t)(periodic_bars ? B : 0));
    for (int64_t b = 0; b < (periodic_bars ? B : 0); ++b)
        if (h.bar_type[b] == NDG_BC_PERIODIC) donor_of[(size_t)h.bar_donor[b]].push_back(b);
    for (size_t x = 0; x < nh.bar_donor.size(); ++x) {
        if (nh.bar_type[x] != NDG_BC_PERIODIC) continue;
        const int64_t b = bar_parent[x], d = h.bar_donor[(size_t)b];
        std::vector<int64_t> partners{d};
        for (int64_t k : donor_of[(size_t)b]) partners.push_back(k);
        for (int64_t k : donor_of[(size_t)d]) partners.push_back(k);
        const int64_t u = nh.bars[2 * x], w = nh.bars[2 * x + 1];
        const v2 c{(nh.xy[2 * u] + nh.xy[2 * w]) / 2 + nh.bar_shift[2 * x],
                   (nh.xy[2 * u + 1] + nh.xy[2 * w + 1]) / 2 + nh.bar_shift[2 * x + 1]};
        int64_t pick = -1;
        for (int pass = 0; pass < 2 && pick < 0; ++pass)
            for (int64_t pb : partners) {
                if (pb == b) continue;
                for (int64_t y = bar_new[(size_t)pb]; y < bar_new[(size_t)pb] + bar_count[(size_t)pb] && pick < 0; ++y) {
                    const v2 p0 = node_xy(nh.xy, nh.bars[2 * y]), p1 = node_xy(nh.xy, nh.bars[2 * y + 1]);
                    const double vx = p1.x - p0.x, vy = p1.y - p0.y, l2 = vx * vx + vy * vy;
                    const double t = ((c.x - p0.x) * vx + (c.y - p0.y) * vy) / l2;
                    const double off = std::fabs((c.x - p0.x) * vy - (c.y - p0.y) * vx) / l2;  // over the length
                    const double m = pass == 0 ? 1e-9 : -1e-9;
                    if (off < 1e-9 && t > m && t < 1 - m) pick = y;
                }
            }
        if (pick < 0) return fail(NDG_ERR_INVALID, "ndg_adapt: a periodic side has no partner after the split");
        nh.bar_donor[x] = pick;
    }
    Topology topo;
    ndg_status st = build_topology(nh, topo);
    if (st != NDG_OK) return st;

    // Child reference coordinates in the parent (quad_map, inverse_quad_map) for the transfer.
    const std::vector<int32_t> old_order = mesh->order;
    const std::vector<int64_t> old_dev = mesh->dev_of;
    std::vector<double> coords;
    std::vector<int64_t> coord_at((size_t)Knew, -1);
    for (int64_t k = 0; k < Knew; ++k) {
        if (!child[k]) continue;
        v2 Pc[4], Pp[4];
        for (int q = 0; q < 4; ++q) {
            Pc[q] = node_xy(nh.xy, nh.quads[4 * k + q]);
            Pp[q] = node_xy(h.xy, h.quads[4 * parent[k] + q]);
        }
        const auto &x = mesh->tables->t.tab[NDG_TABLE_NODES][order[k]];
        coord_at[k] = (int64_t)coords.size();
        for (int i = 0; i <= order[k]; ++i)
            for (int j = 0; j <= order[k]; ++j) {
                const v2 r = reference_point(quad_point(x[i], x[j], Pc), Pp);
                coords.push_back(r.x);
                coords.push_back(r.y);
            }
    }
    const int precision = mesh->d.precision;
    if (precision == NDG_PRECISION_FP32_COMPENSATED) {
        if ((st = ndg_mesh_set_precision(mesh, NDG_PRECISION_FP64)) != NDG_OK) return st;
    }
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_adapt sync");
    drop_graphs(mesh);
    // commit the new host topology and build the device mesh (the previous host state is kept
    // until the new device mesh holds the transferred state, and restored on a failure before)
    DeviceMesh &d = mesh->d;
    MeshHostState previous;
    swap_host_state(mesh, previous);
    mesh->host = std::move(nh);
    mesh->link = std::move(topo.link);
    mesh->periodic_link = std::move(topo.periodic_link);
    mesh->hang_kind = std::move(topo.hang_kind);
    mesh->hang_second = std::move(topo.hang_second);
    mesh->span = std::move(topo.span);
    mesh->bc_base = std::move(topo.bcs);
    mesh->bc_bar = std::move(topo.bc_bar);
    mesh->bar_side = std::move(topo.bar_side);
    mesh->bar_zero = std::move(bar_zero);
    mesh->n_faces = topo.n_faces;
    mesh->status = std::move(status);
    mesh->rotation = std::move(rotation);
    mesh->split_level = std::move(level);
    mesh->host_verts.assign((size_t)(8 * Knew), 0.0);
    for (int64_t k = 0; k < Knew; ++k)
        for (int q = 0; q < 4; ++q) {
            const v2 p = node_xy(mesh->host.xy, mesh->host.quads[4 * k + q]);
            mesh->host_verts[8 * k + 2 * q] = p.x;
            mesh->host_verts[8 * k + 2 * q + 1] = p.y;
        }
    auto restore = [&](ndg_status code) {
        swap_host_state(mesh, previous);
        if (precision == NDG_PRECISION_FP32_COMPENSATED) (void)ndg_mesh_set_precision(mesh, precision);
        return code;
    };
    DeviceMesh nd;
    if ((st = build_device(mesh, order, nd, mesh->store[1 - mesh->active])) != NDG_OK) return restore(st);
    auto offset = [](const DeviceMesh &dm, int N, int64_t dev) -> int64_t {
        for (const Group &G : dm.groups)
            if (G.N == N) return G.q_off + (dev - G.e_begin) * 3 * (int64_t)(N + 1) * (N + 1);
        return -1;
    };
    std::vector<TransferRec> recs((size_t)Knew);
    for (int64_t k = 0; k < Knew; ++k) {
        TransferRec &R = recs[(size_t)mesh->dev_of[k]];
        const int64_t e = parent[k];
        R.old_off = offset(d, old_order[e], old_dev[e]);
        R.new_off = offset(nd, order[k], mesh->dev_of[k]);
        R.N_old = old_order[e];
        R.N_new = order[k];
        R.coord_off = coord_at[k];
    }
    TransferRec *recs_dev = nullptr;
    double *coords_dev = nullptr;
    st = device_upload(&recs_dev, recs, "upload transfer records");
    if (st == NDG_OK && !coords.empty()) st = device_upload(&coords_dev, coords, "upload child coordinates");
    if (st == NDG_OK) {
        const hipError_t e1 = launch_transfer(d.q, nd.q, recs_dev, coords_dev, Knew, nd.tab1d, mesh->stream);
        const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(mesh->stream) : e1;
        if (e2 != hipSuccess) st = hip_fail(e2, "transfer kernel");
    }
    if (recs_dev) (void)hipFree(recs_dev);
    if (coords_dev) (void)hipFree(coords_dev);
    if (st != NDG_OK) return restore(st);
    nd.wave_speed = d.wave_speed;
    carry_tuning(d, nd);
    d = nd;
    mesh->active ^= 1;
    {
        std::vector<char> keep((size_t)Knew);
        for (int64_t k = 0; k < Knew; ++k) keep[(size_t)k] = !child[(size_t)k] && !raise[(size_t)parent[(size_t)k]];
        carry_estimates(mesh, (size_t)K, parent, keep);
    }
    HIP_TRY(hipMemsetAsync(d.g, 0, sizeof(double) * d.n_q, mesh->stream), "hipMemset g");
    HIP_TRY(launch_traces(d, mesh->stream), "traces kernel");
    if (precision == NDG_PRECISION_FP32_COMPENSATED) {
        if ((st = ndg_mesh_set_precision(mesh, precision)) != NDG_OK) return st;
    }
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_adapt sync");
    if (n_changed) *n_changed = n_split + n_raise;
    return NDG_OK;
}

}  // namespace

namespace {

// ndg_adapt on the device (adapt.hip): estimator, decisions, new logical mesh, layout and transfer
// without the host topology; the host reads back two small count blocks. Same outcomes as the
// host path (the splitting branch of Mesh2D_t::adapt when any element would h-refine, else the
// p-refinement branch); refused cases change nothing.
ndg_status adapt_on_device(ndg_mesh *mesh, int32_t N_max, int32_t max_split_level, int64_t *n_changed) {
    ndg_status st;
    if ((st = ensure_logical(mesh)) != NDG_OK) return st;
    if ((st = refresh_fp64(mesh)) != NDG_OK) return st;
    DeviceMesh &d = mesh->d;
    const hipStream_t s = mesh->stream;
    AdaptWork &w = mesh->work;
    const LogicalMesh &L = mesh->lmesh[mesh->lactive];
    if (mesh->est_len < d.K) {
        if (mesh->est_dev) HIP_TRY(hipFree(mesh->est_dev), "hipFree");
        mesh->est_dev = nullptr;
        HIP_TRY(hipMalloc((void **)&mesh->est_dev, sizeof(double) * 4 * d.K), "hipMalloc estimator");
        mesh->est_len = d.K;
    }
    HIP_TRY(launch_estimate(d, mesh->tol_min, mesh->tol_max, mesh->est_dev, s), "estimator kernel");
    HIP_TRY(adapt_decide(L, mesh->est_dev, N_max, mesh->tables->t.N_max, max_split_level, w, s), "adaptation decisions");
    AdaptCounts c;
    HIP_TRY(hipMemcpyAsync(&c, w.counts.p, sizeof c, hipMemcpyDeviceToHost, s), "download adaptation counts");
    HIP_TRY(hipStreamSynchronize(s), "ndg_adapt sync");
    mesh->host_stale = true;  // the logical copy of the estimate changed
    if (c.refused & kRefuseOrder)
        return fail(NDG_ERR_UNSUPPORTED, "ndg_adapt: a raised order exceeds the compiled range or the tables");
    if (c.n_split > 0 && (c.refused & kRefuseSpan))
        return fail(NDG_ERR_UNSUPPORTED, "ndg_adapt: a split would leave a face smaller than 1/16 of its element side");
    if (n_changed) *n_changed = (int64_t)(c.n_split + c.n_raise);
    if (c.n_split + c.n_raise == 0) return NDG_OK;
    const bool p_only = c.n_split == 0;
    const int precision = d.precision;
    if (precision == NDG_PRECISION_FP32_COMPENSATED && (st = ndg_mesh_set_precision(mesh, NDG_PRECISION_FP64)) != NDG_OK)
        return st;
    auto restore = [&](ndg_status code) {
        if (precision == NDG_PRECISION_FP32_COMPENSATED) (void)ndg_mesh_set_precision(mesh, precision);
        return code;
    };
    drop_graphs(mesh);
    const int nl = 1 - mesh->lactive;
    LogicalMesh &Ln = mesh->lmesh[nl];
    const int64_t Kn = L.K + 3 * (int64_t)c.n_split, Gn = L.n_ghost + (int64_t)c.ghost_split;
    if ((st = fit_logical(mesh->lstore[nl], Ln, Kn, Gn)) != NDG_OK) return restore(st);
    auto hip_restore = [&](hipError_t e, const char *where) { return restore(hip_fail(e, where)); };
    hipError_t he;
    if ((he = adapt_rebuild(L, Ln, p_only, w, s)) != hipSuccess) return hip_restore(he, "adaptation rebuild");
    DeviceStore &ns = mesh->store[1 - mesh->active];
    GroupSpan *spans_new = nullptr;
    if ((st = fit_view(ns.spans, &spans_new, kMaxOrder + 1, "hipMalloc group spans")) != NDG_OK) return restore(st);
    if ((he = adapt_layout_count(Ln, spans_new, w, s)) != hipSuccess) return hip_restore(he, "adaptation layout");
    if ((he = hipMemcpyAsync(&c, w.counts.p, sizeof c, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (he = hipStreamSynchronize(s)) != hipSuccess)
        return hip_restore(he, "download adaptation counts");
    // the device mesh of the new layout, in the spare store
    DeviceMesh nd;
    nd.K = Kn;
    int64_t e_begin = 0, q_off = 0, tr_off = 0;
    for (int n = 0; n <= kMaxOrder; ++n) {
        const int64_t Kg = (int64_t)c.order_count[n];
        if (Kg == 0) continue;
        if ((st = ensure_order_tables(mesh, n)) != NDG_OK) return restore(st);
        Group G;
        G.N = n;
        G.K = Kg;
        G.e_begin = e_begin;
        G.q_off = q_off;
        G.tr_off = tr_off;
        G.tab = mesh->tab64[n];
        nd.groups.push_back(G);
        e_begin += Kg;
        q_off += 3 * (int64_t)(n + 1) * (n + 1) * Kg;
        tr_off += 12 * (int64_t)(n + 1) * Kg;
    }
    nd.n_q = q_off;
    nd.n_tr = tr_off;
    nd.N_min = nd.groups.front().N;
    nd.N_max = nd.groups.back().N;
    nd.n_bc = Gn + (int64_t)c.n_rec;
    nd.n_faces = (int64_t)c.n_face;
    nd.n_side_flux = (int64_t)c.n_slot;
    nd.mortar_nodes = (int)c.mortar_nodes;
    if ((st = fit_view(ns.verts, &nd.verts, 8 * Kn, "hipMalloc vertices")) != NDG_OK ||
        (st = fit_view(ns.side_nb, &nd.side_nb, 4 * Kn, "hipMalloc connectivity")) != NDG_OK ||
        (st = fit_view(ns.bc, &nd.bc, nd.n_bc, "hipMalloc boundary records")) != NDG_OK ||
        (st = fit_view(ns.faces, &nd.faces, nd.n_faces, "hipMalloc mortar faces")) != NDG_OK ||
        (st = fit_view(ns.q, &nd.q, nd.n_q, "hipMalloc q")) != NDG_OK ||
        (st = fit_view(ns.g, &nd.g, nd.n_q, "hipMalloc g")) != NDG_OK ||
        (st = fit_view(ns.tr[0], &nd.trace[0], nd.n_tr, "hipMalloc traces")) != NDG_OK ||
        (st = fit_view(ns.tr[1], &nd.trace[1], nd.n_tr, "hipMalloc traces")) != NDG_OK)
        return restore(st);
    nd.scratch_len = 1024;
    if ((st = fit_view(ns.scratch, &nd.scratch, nd.scratch_len, "hipMalloc scratch")) != NDG_OK) return restore(st);
    if (nd.n_side_flux > 0 &&
        (st = fit_view(ns.side_flux, &nd.side_flux, nd.n_side_flux, "hipMalloc side fluxes")) != NDG_OK)
        return restore(st);
    if (nd.n_faces == 0) nd.faces = nullptr;
    nd.tab1d = mesh->tab1d;
    nd.mortar_mats = mesh->mortar_mats;
    nd.clock = mesh->clock;
    nd.has_inflow = d.has_inflow;
    nd.wave_speed = d.wave_speed;
    nd.cur = 0;
    carry_tuning(d, nd);
    if ((he = adapt_layout_fill(Ln, spans_new, mesh->mortar_key_dev, nd, w, s)) != hipSuccess)
        return hip_restore(he, "adaptation layout");
    if ((he = adapt_transfer(L, (const GroupSpan *)mesh->store[mesh->active].spans.p, d.q, Ln, spans_new, nd.q,
                             mesh->tab1d, w, s)) != hipSuccess)
        return hip_restore(he, "transfer kernel");
    if ((he = hipMemsetAsync(nd.g, 0, sizeof(double) * nd.n_q, s)) != hipSuccess) return hip_restore(he, "hipMemset g");
    if ((he = launch_traces(nd, s)) != hipSuccess) return hip_restore(he, "traces kernel");
    AdaptCounts c2;
    if ((he = hipMemcpyAsync(&c2, w.counts.p, sizeof c2, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (he = hipStreamSynchronize(s)) != hipSuccess)
        return hip_restore(he, "ndg_adapt sync");
    if (c2.bad_face)
        return restore(fail(NDG_ERR_UNSUPPORTED, "ndg_adapt: a non-conforming face covers neither a whole element "
                                                 "side nor half of one"));
    // commit
    d = nd;
    mesh->active ^= 1;
    mesh->lactive = nl;
    mesh->n_faces = (int64_t)c.conforming_pairs + Gn + (int64_t)c.periodic_faces + (int64_t)c.hanging_faces;
    mesh->n_dofs = (int64_t)c.dofs;
    mesh->host_stale = true;
    mesh->host_nodes_valid = false;
    if (precision == NDG_PRECISION_FP32_COMPENSATED && (st = ndg_mesh_set_precision(mesh, precision)) != NDG_OK)
        return st;
    return NDG_OK;
}

}  // namespace

extern "C" ndg_status ndg_inverse_quad_map(const double *vertices, int64_t n, const double *xy, double *xi_eta) {
    if (!vertices || n < 0 || (n > 0 && (!xy || !xi_eta))) return fail(NDG_ERR_INVALID, "ndg_inverse_quad_map: bad argument");
    const v2 P[4] = {{vertices[0], vertices[1]}, {vertices[2], vertices[3]}, {vertices[4], vertices[5]},
                     {vertices[6], vertices[7]}};
    for (int64_t i = 0; i < n; ++i) {
        const v2 r = reference_point(v2{xy[2 * i], xy[2 * i + 1]}, P);
        xi_eta[2 * i] = r.x;
        xi_eta[2 * i + 1] = r.y;
    }
    return NDG_OK;
}

extern "C" ndg_status ndg_adapt(ndg_mesh *mesh, int32_t N_max, int32_t max_split_level, int64_t *n_raised) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_adapt: null mesh");
    if (mesh->adapt_mode == NDG_ADAPT_DEVICE) return adapt_on_device(mesh, N_max, max_split_level, n_raised);
    if (!mesh->host_nodes_valid)
        return fail(NDG_ERR_UNSUPPORTED, "ndg_adapt: the host path needs the mesh's node lists, which a device "
                                         "adaptation does not keep");
    mesh->logical_valid = false;  // the host path changes the mesh behind the logical copy
    DeviceMesh &d = mesh->d;
    const int64_t K = d.K;
    std::vector<double> err((size_t)K), sig((size_t)K);
    std::vector<int32_t> ref((size_t)K), coa((size_t)K);
    ndg_status st = ndg_estimate_error(mesh, err.data(), sig.data(), ref.data(), coa.data());
    if (st != NDG_OK) return st;
    // Element2D_t.cpp:802-812: any would_h_refine takes the splitting branch (hp_adapt)
    for (int64_t e = 0; e < K; ++e)
        if (ref[e] && sig[e] < 1.0 && mesh->split_level[e] < max_split_level)
            return adapt_split(mesh, ref, sig, N_max, max_split_level, n_raised);
    std::vector<int32_t> order = mesh->order;
    int64_t raised = 0;
    for (int64_t e = 0; e < K; ++e)
        if (ref[e] && sig[e] >= 1.0 && order[e] + 2 <= N_max) {
            order[e] += 2;
            ++raised;
        }
    if (n_raised) *n_raised = raised;
    if (raised == 0) return NDG_OK;
    for (int64_t e = 0; e < K; ++e)
        if (order[e] > kMaxOrder || order[e] > mesh->tables->t.N_max)
            return fail(NDG_ERR_UNSUPPORTED, "ndg_adapt: a raised order exceeds the compiled range or the tables");
    const int precision = d.precision;
    if (precision == NDG_PRECISION_FP32_COMPENSATED) {
        if ((st = ndg_mesh_set_precision(mesh, NDG_PRECISION_FP64)) != NDG_OK) return st;
    }
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_adapt sync");
    drop_graphs(mesh);
    // Everything the rebuild changes on the host (orders, device permutation, packed offsets, the
    // resized-ghost flags) is restored, with the precision, if the rebuild or the transfer fails.
    const std::vector<int32_t> old_order = mesh->order;
    const std::vector<int64_t> old_dev = mesh->dev_of, old_host_off = mesh->host_off;
    const std::vector<char> old_bar_zero = mesh->bar_zero;
    const int64_t old_dofs = mesh->n_dofs;
    auto restore = [&](ndg_status code) {
        mesh->order = old_order;
        mesh->dev_of = old_dev;
        mesh->host_off = old_host_off;
        mesh->bar_zero = old_bar_zero;
        mesh->n_dofs = old_dofs;
        if (precision == NDG_PRECISION_FP32_COMPENSATED) (void)ndg_mesh_set_precision(mesh, precision);
        return code;
    };
    // adjust_boundaries (Mesh2D_t.cpp:7482-7504): the ghost of a raised element is resized, which
    // zeroes its edge scaling factor (Element2D_t.cpp:735-741).
    for (int64_t es = 0; es < 4 * K; ++es) {
        const int64_t l = mesh->link[es];
        if (l < 0 && l != INT64_MIN && order[es / 4] != mesh->order[es / 4])
            mesh->bar_zero[(size_t)mesh->bc_bar[(size_t)(-l - 1)]] = 1;
    }
    DeviceMesh nd;
    st = build_device(mesh, order, nd, mesh->store[1 - mesh->active]);
    if (st != NDG_OK) return restore(st);
    auto offset = [](const DeviceMesh &dm, int N, int64_t dev) -> int64_t {
        for (const Group &G : dm.groups)
            if (G.N == N) return G.q_off + (dev - G.e_begin) * 3 * (int64_t)(N + 1) * (N + 1);
        return -1;
    };
    std::vector<TransferRec> recs((size_t)K);
    for (int64_t e = 0; e < K; ++e) {
        TransferRec &R = recs[(size_t)mesh->dev_of[e]];
        R.old_off = offset(d, old_order[e], old_dev[e]);
        R.new_off = offset(nd, order[e], mesh->dev_of[e]);
        R.N_old = old_order[e];
        R.N_new = order[e];
        R.coord_off = -1;
    }
    TransferRec *recs_dev = nullptr;
    st = device_upload(&recs_dev, recs, "upload transfer records");
    if (st == NDG_OK) {
        const hipError_t e1 = launch_transfer(d.q, nd.q, recs_dev, nullptr, K, nd.tab1d, mesh->stream);
        const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(mesh->stream) : e1;
        if (e2 != hipSuccess) st = hip_fail(e2, "transfer kernel");
    }
    if (recs_dev) (void)hipFree(recs_dev);
    if (st != NDG_OK) return restore(st);
    nd.wave_speed = d.wave_speed;
    nd.cur = 0;
    carry_tuning(d, nd);
    d = nd;
    mesh->active ^= 1;
    {
        std::vector<int64_t> same((size_t)K);
        std::vector<char> keep((size_t)K);
        for (int64_t e = 0; e < K; ++e) {
            same[(size_t)e] = e;
            keep[(size_t)e] = order[e] == old_order[e];
        }
        carry_estimates(mesh, (size_t)K, same, keep);
    }
    HIP_TRY(hipMemsetAsync(d.g, 0, sizeof(double) * d.n_q, mesh->stream), "hipMemset g");
    HIP_TRY(launch_traces(d, mesh->stream), "traces kernel");
    if (precision == NDG_PRECISION_FP32_COMPENSATED) {
        if ((st = ndg_mesh_set_precision(mesh, precision)) != NDG_OK) return st;
    }
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_adapt sync");
    return NDG_OK;
}

extern "C" ndg_status ndg_mesh_orders(const ndg_mesh *mesh, int32_t *orders) {
    if (!mesh || !orders) return fail(NDG_ERR_INVALID, "ndg_mesh_orders: null argument");
    if (mesh->host_stale) {  // after a device adaptation: the orders alone, from the logical mesh
        const LogicalMesh &L = mesh->lmesh[mesh->lactive];
        HIP_TRY(hipMemcpyAsync(orders, L.order, sizeof(int32_t) * L.K, hipMemcpyDeviceToHost, mesh->stream),
                "download orders");
        HIP_TRY(hipStreamSynchronize(mesh->stream), "download orders");
        return NDG_OK;
    }
    std::copy(mesh->order.begin(), mesh->order.end(), orders);
    return NDG_OK;
}

extern "C" ndg_status ndg_mesh_set_adapt_mode(ndg_mesh *mesh, int32_t mode) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_mesh_set_adapt_mode: null mesh");
    if (mode != NDG_ADAPT_DEVICE && mode != NDG_ADAPT_HOST) return fail(NDG_ERR_INVALID, "ndg_mesh_set_adapt_mode: unknown mode");
    if (mode == NDG_ADAPT_HOST && !mesh->host_nodes_valid)
        return fail(NDG_ERR_UNSUPPORTED, "ndg_mesh_set_adapt_mode: the host path needs the mesh's node lists, which a "
                                         "device adaptation does not keep");
    mesh->adapt_mode = mode;
    return NDG_OK;
}

extern "C" ndg_status ndg_mesh_sides(ndg_mesh *mesh, int64_t *link, int8_t *hang, int64_t *next_small,
                                     int8_t *periodic, int8_t *span) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_mesh_sides: null mesh");
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    const size_t n = mesh->link.size();
    for (size_t i = 0; i < n; ++i) {
        if (link) link[i] = mesh->link[i];
        if (hang) hang[i] = (int8_t)mesh->hang_kind[i];
        if (next_small) next_small[i] = mesh->hang_second[i];
        if (periodic) periodic[i] = (int8_t)mesh->periodic_link[i];
        if (span) span[i] = mesh->span[i];
    }
    return NDG_OK;
}

extern "C" ndg_status ndg_mesh_set_adaptivity(ndg_mesh *mesh, int64_t adaptivity_interval) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_mesh_set_adaptivity: null mesh");
    if (adaptivity_interval < 0) return fail(NDG_ERR_INVALID, "ndg_mesh_set_adaptivity: negative interval");
    mesh->adapt_interval = adaptivity_interval;
    return NDG_OK;
}

namespace {

// Solver2D_t::solve (include/solvers/Solver2D_t.tpp:7-135): steps of the CFL size up to the last
// output time (the last step clipped to it), the output events of the reference (before the first
// step, every output time t_o in [0, dt); after a step, the first t_o with time in [t_o, t_o + dt);
// at the end, once more if the last step wrote nothing), and adaptation every adapt_interval steps
// with the tables' N_max, after which dt follows the new orders.
ndg_status solve_impl(ndg_mesh *mesh, double cfl, const double *times, int64_t n_times, ndg_output_fn fn,
                      void *user, double *t_final, int64_t *n_steps) {
    const double t_end = times[n_times - 1];
    double dt0 = 0.0;
    ndg_status st = ndg_get_delta_t(mesh, cfl, &dt0);
    if (st != NDG_OK) return st;
    double time = 0.0, dt = dt0;
    int64_t steps = 0;
    auto output = [&]() -> ndg_status {
        const ndg_status r = fn(mesh, time, user);
        return r == NDG_OK ? NDG_OK : fail(r, "ndg_solve_output: the output callback returned an error");
    };
    if (fn)
        for (int64_t k = 0; k < n_times; ++k)
            if (time >= times[k] && time < times[k] + dt && (st = output()) != NDG_OK) return st;
    HIP_TRY(launch_clock_begin(mesh->d, time, 1, dt0, mesh->stream), "clock kernel");
    while (time < t_end) {
        ++steps;
        dt = dt0;
        if (time + dt > t_end) {
            dt = t_end - time;
            if ((st = ndg_step(mesh, time, dt)) != NDG_OK) return st;
        } else if ((st = step_from_clock(mesh, dt)) != NDG_OK) {
            return st;
        }
        time += dt;
        if (fn)
            for (int64_t k = 0; k < n_times; ++k)
                if (time >= times[k] && time < times[k] + dt) {
                    if ((st = output()) != NDG_OK) return st;
                    break;
                }
        // include/solvers/Solver2D_t.tpp:88-96: estimate_error + adapt every adaptivity_interval
        // steps, with the tables' N_max; Δt then follows the new orders.
        if (mesh->adapt_interval > 0 && steps % mesh->adapt_interval == 0) {
            int64_t raised = 0;
            if ((st = ndg_adapt(mesh, mesh->tables->t.N_max, mesh->max_split_level, &raised)) != NDG_OK) return st;
            if (raised > 0) {
                if ((st = ndg_get_delta_t(mesh, cfl, &dt0)) != NDG_OK) return st;
                HIP_TRY(launch_clock_begin(mesh->d, time, 1, dt0, mesh->stream), "clock kernel");
            }
        }
    }
    if (fn) {
        bool wrote = false;
        for (int64_t k = 0; k < n_times && !wrote; ++k) wrote = time >= times[k] && time < times[k] + dt;
        if (!wrote && (st = output()) != NDG_OK) return st;
    }
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_solve sync");
    if (t_final) *t_final = time;
    if (n_steps) *n_steps = steps;
    return NDG_OK;
}

}  // namespace

extern "C" ndg_status ndg_solve(ndg_mesh *mesh, double cfl, double t_end, double *t_final, int64_t *n_steps) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_solve: null mesh");
    return solve_impl(mesh, cfl, &t_end, 1, nullptr, nullptr, t_final, n_steps);
}

extern "C" ndg_status ndg_solve_output(ndg_mesh *mesh, double cfl, const double *output_times, int64_t n_times,
                                       ndg_output_fn fn, void *user, double *t_final, int64_t *n_steps) {
    if (!mesh || !output_times || n_times < 1) return fail(NDG_ERR_INVALID, "ndg_solve_output: null mesh or no output times");
    return solve_impl(mesh, cfl, output_times, n_times, fn, user, t_final, n_steps);
}

// Solver2D_t::pre_condition (include/solvers/Solver2D_t.tpp:138-236) and pre_condition_iterative
// (:238-336): steps of the CFL size, estimate + adapt every pre_condition_interval steps until
// n_adaptivity_steps adaptations are done, and the initial condition applied again to the adapted
// mesh: once at the end, or (iterative) after every adaptation, with the time set back to 0.
extern "C" ndg_status ndg_pre_condition(ndg_mesh *mesh, double cfl, int64_t n_adaptivity_steps,
                                        int64_t pre_condition_interval, int32_t iterative) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_pre_condition: null mesh");
    if (pre_condition_interval <= 0)
        return fail(NDG_ERR_INVALID, "ndg_pre_condition: the pre-condition interval must be positive");
    if (n_adaptivity_steps < 0) return fail(NDG_ERR_INVALID, "ndg_pre_condition: negative adaptation count");
    if (mesh->ic_kind < 0) return fail(NDG_ERR_INVALID, "ndg_pre_condition: no initial condition was set");
    const std::vector<double> pulses = mesh->ic_pulses;  // ndg_set_ic below rewrites the mesh's copy
    const ndg_ic_desc ic_keep{mesh->ic_kind, (int32_t)(pulses.size() / 4), pulses.empty() ? nullptr : pulses.data(),
                              mesh->ic_period[0], mesh->ic_period[1]};
    double time = 0.0, dt = 0.0;
    int64_t timestep = 0, n_adaptivity = 0;
    ndg_status st;
    while (n_adaptivity < n_adaptivity_steps) {
        ++timestep;
        if ((st = ndg_get_delta_t(mesh, cfl, &dt)) != NDG_OK) return st;
        if ((st = ndg_step(mesh, time, dt)) != NDG_OK) return st;
        time += dt;
        if (timestep % pre_condition_interval == 0) {
            int64_t changed = 0;
            if ((st = ndg_adapt(mesh, mesh->tables->t.N_max, mesh->max_split_level, &changed)) != NDG_OK) return st;
            ++n_adaptivity;
            if (iterative) {
                if ((st = ndg_set_ic(mesh, &ic_keep)) != NDG_OK) return st;
                time = 0.0;
            }
        }
    }
    if (!iterative && (st = ndg_set_ic(mesh, &ic_keep)) != NDG_OK) return st;
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_pre_condition sync");
    return NDG_OK;
}

extern "C" ndg_status ndg_run_steps(ndg_mesh *mesh, double cfl, int64_t n_steps, double *time_io) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_run_steps: null mesh");
    if (n_steps < 0) return fail(NDG_ERR_INVALID, "ndg_run_steps: negative step count");
    double dt = 0.0;
    ndg_status st = ndg_get_delta_t(mesh, cfl, &dt);
    if (st != NDG_OK) return st;
    double time = time_io ? *time_io : 0.0;
    HIP_TRY(launch_clock_begin(mesh->d, time, 1, dt, mesh->stream), "clock kernel");
    for (int64_t k = 0; k < n_steps; ++k) {
        if ((st = step_from_clock(mesh, dt)) != NDG_OK) return st;
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
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    ndg_status st = refresh_fp64(mesh);
    if (st != NDG_OK) return st;
    DeviceMesh &d = mesh->d;
    if (mesh->est_len < d.K) {
        if (mesh->est_dev) HIP_TRY(hipFree(mesh->est_dev), "hipFree");
        mesh->est_dev = nullptr;
        HIP_TRY(hipMalloc((void **)&mesh->est_dev, sizeof(double) * 4 * d.K), "hipMalloc estimator");
        mesh->est_len = d.K;
    }
    HIP_TRY(launch_estimate(d, mesh->tol_min, mesh->tol_max, mesh->est_dev, mesh->stream), "estimator kernel");
    if (mesh->logical_valid)
        HIP_TRY(adapt_copy_estimate(mesh->lmesh[mesh->lactive], mesh->est_dev, mesh->stream), "estimate copy kernel");
    std::vector<double> host((size_t)(4 * d.K));
    HIP_TRY(hipMemcpyAsync(host.data(), mesh->est_dev, sizeof(double) * 4 * d.K, hipMemcpyDeviceToHost, mesh->stream),
            "download estimator");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_estimate_error sync");
    const size_t K = (size_t)d.K;
    mesh->est_error.resize(K);
    mesh->est_sigma.resize(K);
    mesh->est_refine.resize(K);
    mesh->est_coarsen.resize(K);
    for (int64_t e = 0; e < d.K; ++e) {
        const double *h = host.data() + 4 * mesh->dev_of[e];
        mesh->est_error[e] = h[0];
        mesh->est_sigma[e] = h[1];
        mesh->est_refine[e] = h[2] != 0.0;
        mesh->est_coarsen[e] = h[3] != 0.0;
        if (error) error[e] = mesh->est_error[e];
        if (sigma) sigma[e] = mesh->est_sigma[e];
        if (refine) refine[e] = mesh->est_refine[e];
        if (coarsen) coarsen[e] = mesh->est_coarsen[e];
    }
    return NDG_OK;
}

// ------------------------------------------------------------------------------------------
// Wave speed and the adjoint form (no reference counterpart)

extern "C" ndg_status ndg_mesh_set_wave_speed(ndg_mesh *mesh, double c) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_mesh_set_wave_speed: null mesh");
    if (!(c > 0.0) || !std::isfinite(c)) return fail(NDG_ERR_INVALID, "ndg_mesh_set_wave_speed: c must be positive");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_mesh_set_wave_speed sync");
    drop_graphs(mesh);
    mesh->d.wave_speed = c;
    return NDG_OK;
}

namespace {

// Device buffer freed on scope exit (adjoint temporaries).
struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T *as() const { return (T *)p; }
};

ndg_status dev_alloc(DevBuf &b, size_t bytes, const char *what) {
    HIP_TRY(hipMalloc(&b.p, bytes), what);
    return NDG_OK;
}

// Lagrange interpolating polynomials of the order-N Gauss nodes at xi (barycentric form with the
// exact-node case, include/entities/NDG_t.tpp:209-226).
void lagrange_at(const Tables &t, int N, double xi, std::vector<double> &out) {
    const auto &x = t.tab[NDG_TABLE_NODES][N];
    const auto &lam = t.tab[NDG_TABLE_BARYCENTRIC_WEIGHTS][N];
    out.assign((size_t)N + 1, 0.0);
    for (int i = 0; i <= N; ++i)
        if (xi == x[i]) {
            out[i] = 1.0;
            return;
        }
    double den = 0.0;
    for (int i = 0; i <= N; ++i) {
        out[i] = lam[i] / (xi - x[i]);
        den += out[i];
    }
    for (int i = 0; i <= N; ++i) out[i] /= den;
}

// The adjoint and the receivers need one order group, no boundary records and no mortar faces.
ndg_status check_adjoint_mesh(const ndg_mesh *m, const char *who) {
    const DeviceMesh &d = m->d;
    if (d.groups.size() != 1 || d.n_bc != 0 || d.n_faces != 0)
        return fail(NDG_ERR_UNSUPPORTED, std::string(who) + ": needs a single-order mesh whose sides are all interior "
                                                            "or periodic");
    if (d.precision != NDG_PRECISION_FP64) return fail(NDG_ERR_UNSUPPORTED, std::string(who) + ": FP64 precision only");
    return NDG_OK;
}

struct DeviceReceivers {
    DevBuf elem, w, w_src;  // w_src: the weights divided by J at their node (adjoint source)
    DevBuf w_t;             // the weights in the RK register's transposed node order
    int n = 0;
};

// Device element index and the (N+1)^2 interpolation weights of every receiver.
ndg_status receiver_weights(const ndg_mesh *m, const ndg_receivers *rec, std::vector<int64_t> &elem,
                            std::vector<double> &w) {
    if (!rec || rec->n < 1 || !rec->element || !rec->xi_eta) return fail(NDG_ERR_INVALID, "receivers: empty list");
    const int N = m->d.groups[0].N, ne = (N + 1) * (N + 1);
    elem.assign((size_t)rec->n, 0);
    w.assign((size_t)rec->n * ne, 0.0);
    std::vector<double> lx, ly;
    for (int r = 0; r < rec->n; ++r) {
        const int64_t e = rec->element[r];
        const double xi = rec->xi_eta[2 * r], eta = rec->xi_eta[2 * r + 1];
        if (e < 0 || e >= m->d.K) return fail(NDG_ERR_INVALID, "receivers: element out of range");
        if (!(xi >= -1.0 && xi <= 1.0 && eta >= -1.0 && eta <= 1.0))
            return fail(NDG_ERR_INVALID, "receivers: reference coordinates outside [-1, 1]");
        elem[r] = m->dev_of[e];
        lagrange_at(m->tables->t, N, xi, lx);
        lagrange_at(m->tables->t, N, eta, ly);
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N; ++j) w[(size_t)r * ne + i * (N + 1) + j] = lx[i] * ly[j];
    }
    return NDG_OK;
}

ndg_status upload_receivers(const ndg_mesh *m, const ndg_receivers *rec, DeviceReceivers &out) {
    const int N = m->d.groups[0].N, ne = (N + 1) * (N + 1);
    std::vector<int64_t> elem;
    std::vector<double> w;
    ndg_status st = receiver_weights(m, rec, elem, w);
    if (st != NDG_OK) return st;
    if ((st = dev_alloc(out.elem, elem.size() * sizeof(int64_t), "hipMalloc receivers")) != NDG_OK) return st;
    if ((st = dev_alloc(out.w, w.size() * sizeof(double), "hipMalloc receivers")) != NDG_OK) return st;
    HIP_TRY(hipMemcpy(out.elem.p, elem.data(), elem.size() * sizeof(int64_t), hipMemcpyHostToDevice), "upload receivers");
    HIP_TRY(hipMemcpy(out.w.p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice), "upload receivers");
    out.n = rec->n;
    std::vector<double> wt(w.size());
    for (int r = 0; r < rec->n; ++r)
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N; ++j)
                wt[(size_t)r * ne + j * (N + 1) + i] = w[(size_t)r * ne + i * (N + 1) + j];
    if ((st = dev_alloc(out.w_t, wt.size() * sizeof(double), "hipMalloc receivers")) != NDG_OK) return st;
    HIP_TRY(hipMemcpy(out.w_t.p, wt.data(), wt.size() * sizeof(double), hipMemcpyHostToDevice), "upload receivers");
    if ((st = dev_alloc(out.w_src, w.size() * sizeof(double), "hipMalloc receivers")) != NDG_OK) return st;
    HIP_TRY(hipMemcpy(out.w_src.p, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice), "upload receivers");
    HIP_TRY(launch_jacobian_scale(m->d, out.w_src.as<double>(), out.elem.as<int64_t>(), out.n, -1, m->stream),
            "jacobian scale kernel");
    return NDG_OK;
}

}  // namespace

extern "C" ndg_status ndg_record_receivers(ndg_mesh *mesh, const ndg_receivers *rec, double dt, int64_t n_steps,
                                           double *p_out) {
    if (!mesh || !p_out) return fail(NDG_ERR_INVALID, "ndg_record_receivers: null argument");
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    if (n_steps < 0 || !(dt > 0.0)) return fail(NDG_ERR_INVALID, "ndg_record_receivers: bad step count or dt");
    ndg_status st = check_adjoint_mesh(mesh, "ndg_record_receivers");
    if (st != NDG_OK) return st;
    DeviceReceivers R;
    if ((st = upload_receivers(mesh, rec, R)) != NDG_OK) return st;
    DevBuf out;
    if (n_steps == 0) return NDG_OK;
    if ((st = dev_alloc(out, sizeof(double) * R.n * n_steps, "hipMalloc records")) != NDG_OK) return st;
    DeviceMesh &d = mesh->d;
    HIP_TRY(launch_clock_begin(d, 0.0, 1, dt, mesh->stream), "clock kernel");
    for (int64_t n = 0; n < n_steps; ++n) {
        if ((st = step_from_clock(mesh, dt)) != NDG_OK) return st;
        HIP_TRY(launch_record(d, R.elem.as<int64_t>(), R.w.as<double>(), R.n, out.as<double>() + n * R.n,
                              mesh->stream), "record kernel");
    }
    HIP_TRY(hipMemcpyAsync(p_out, out.p, sizeof(double) * R.n * n_steps, hipMemcpyDeviceToHost, mesh->stream),
            "download records");
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_record_receivers sync");
    return NDG_OK;
}

extern "C" ndg_status ndg_record_receivers_noalloc(ndg_mesh *mesh, const ndg_receivers *rec, double dt,
                                                   int64_t n_steps, double *p_out) {
    if (!mesh || !p_out) return fail(NDG_ERR_INVALID, "ndg_record_receivers_noalloc: null argument");
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    if (n_steps < 0 || !(dt > 0.0)) return fail(NDG_ERR_INVALID, "ndg_record_receivers_noalloc: bad step count or dt");
    ndg_status st = check_adjoint_mesh(mesh, "ndg_record_receivers_noalloc");
    if (st != NDG_OK) return st;
    std::vector<int64_t> elem;
    std::vector<double> w;
    if ((st = receiver_weights(mesh, rec, elem, w)) != NDG_OK) return st;
    // Scratch layout: receiver elements (int64), weights, then the records of every step. Every
    // value is written before it is read, so nothing is cleared.
    const int64_t n = rec->n, ne = (int64_t)(w.size() / elem.size());
    const int64_t need = 8 * (n + n * ne + n * n_steps);
    if (need > (int64_t)kNoAllocScratchBytes)
        return fail(NDG_ERR_INVALID, "ndg_record_receivers_noalloc: " + std::to_string(need) +
                                         " bytes of receivers and records exceed its fixed scratch of " +
                                         std::to_string(kNoAllocScratchBytes) + " bytes");
    if (n_steps == 0) return NDG_OK;
    double *scr = nullptr;
    HIP_TRY(noalloc_scratch(&scr), "ndg_record_receivers_noalloc: scratch address");
    int64_t *elem_dev = reinterpret_cast<int64_t *>(scr);
    double *w_dev = scr + n, *rec_dev = scr + n + n * ne;
    DeviceMesh &d = mesh->d;
    hipStream_t s = mesh->stream;
    HIP_TRY(hipMemcpyAsync(elem_dev, elem.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, s), "upload receivers");
    HIP_TRY(hipMemcpyAsync(w_dev, w.data(), sizeof(double) * n * ne, hipMemcpyHostToDevice, s), "upload receivers");
    // direct launches (a captured step graph would be instantiated on first use)
    HIP_TRY(launch_clock_begin(d, 0.0, 1, dt, s), "clock kernel");
    for (int64_t k = 0; k < n_steps; ++k) {
        if ((st = launch_step_body(mesh, dt, false, s)) != NDG_OK) return st;
        HIP_TRY(launch_record(d, elem_dev, w_dev, (int)n, rec_dev + k * n, s), "record kernel");
    }
    HIP_TRY(hipMemcpyAsync(p_out, rec_dev, sizeof(double) * n * n_steps, hipMemcpyDeviceToHost, s), "download records");
    HIP_TRY(hipStreamSynchronize(s), "ndg_record_receivers_noalloc sync");
    return NDG_OK;
}

extern "C" ndg_status ndg_apply_rhs_transpose(ndg_mesh *mesh, const double *mp, const double *mu, const double *mv,
                                              double *lp, double *lu, double *lv) {
    if (!mesh || !mp || !mu || !mv) return fail(NDG_ERR_INVALID, "ndg_apply_rhs_transpose: null argument");
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    ndg_status st = check_adjoint_mesh(mesh, "ndg_apply_rhs_transpose");
    if (st != NDG_OK) return st;
    DeviceMesh &d = mesh->d;
    DevBuf lam, muv, chi;
    if ((st = dev_alloc(lam, sizeof(double) * d.n_q, "hipMalloc adjoint")) != NDG_OK) return st;
    if ((st = dev_alloc(muv, sizeof(double) * d.n_q, "hipMalloc adjoint")) != NDG_OK) return st;
    if ((st = dev_alloc(chi, sizeof(double) * d.n_tr, "hipMalloc adjoint")) != NDG_OK) return st;
    HIP_TRY(hipStreamSynchronize(mesh->stream), "ndg_apply_rhs_transpose sync");
    double *const ins[3] = {const_cast<double *>(mp), const_cast<double *>(mu), const_cast<double *>(mv)};
    if ((st = copy_state(mesh, muv.as<double>(), ins, false, "upload adjoint")) != NDG_OK) return st;
    // the kernel works on values divided by J at their node
    HIP_TRY(launch_jacobian_scale(d, muv.as<double>(), nullptr, 0, -1, mesh->stream), "jacobian scale kernel");
    HIP_TRY(hipMemsetAsync(lam.p, 0, sizeof(double) * d.n_q, mesh->stream), "hipMemset adjoint");
    AdjointArgs A;
    A.lam = lam.as<double>();
    A.mu = muv.as<double>();
    A.chi_out = chi.as<double>();
    A.flags = 1 | 4;  // mu input, export only
    HIP_TRY(launch_adjoint_stage(d, A, mesh->stream), "adjoint export kernel");
    A.chi_in = chi.as<double>();
    A.chi_out = nullptr;
    A.flags = 1 | 8;  // mu input, apply
    HIP_TRY(launch_adjoint_stage(d, A, mesh->stream), "adjoint kernel");
    HIP_TRY(launch_jacobian_scale(d, lam.as<double>(), nullptr, 0, 1, mesh->stream), "jacobian scale kernel");
    double *const outs[3] = {lp, lu, lv};
    return copy_state(mesh, lam.as<double>(), outs, true, "download adjoint");
}

extern "C" ndg_status ndg_adjoint_set_c_derivative(ndg_mesh *mesh, int32_t mode) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_adjoint_set_c_derivative: null mesh");
    if (mode != NDG_ADJOINT_C_RECOMPUTE && mode != NDG_ADJOINT_C_STAGE_RECORDS)
        return fail(NDG_ERR_INVALID, "ndg_adjoint_set_c_derivative: unknown mode");
    mesh->adjoint_c_mode = mode;
    return NDG_OK;
}

extern "C" ndg_status ndg_adjoint_set_store_budget(ndg_mesh *mesh, int64_t bytes) {
    if (!mesh || bytes < 0) return fail(NDG_ERR_INVALID, "ndg_adjoint_set_store_budget: null mesh or negative budget");
    mesh->adjoint_store_budget = bytes;
    return NDG_OK;
}

extern "C" ndg_status ndg_adjoint_gradient(ndg_mesh *mesh, const ndg_receivers *rec, double dt, int64_t n_steps,
                                           int64_t checkpoint_interval, const double *data, double *J,
                                           double *dJ_dp0, double *dJ_du0, double *dJ_dv0, double *dJ_dc) {
    if (!mesh) return fail(NDG_ERR_INVALID, "ndg_adjoint_gradient: null mesh");
    if (const ndg_status hs = ensure_host(mesh); hs != NDG_OK) return hs;
    if (n_steps < 1 || !(dt > 0.0) || checkpoint_interval < 1)
        return fail(NDG_ERR_INVALID, "ndg_adjoint_gradient: bad step count, dt or checkpoint interval");
    ndg_status st = check_adjoint_mesh(mesh, "ndg_adjoint_gradient");
    if (st != NDG_OK) return st;
    DeviceMesh &d = mesh->d;
    hipStream_t s = mesh->stream;
    DeviceReceivers R;
    if ((st = upload_receivers(mesh, rec, R)) != NDG_OK) return st;
    // With the c-derivative from stage records no forward state is needed in the reverse sweep:
    // one segment, no stage store, the initial state as the only checkpoint.
    const bool records = mesh->adjoint_c_mode == NDG_ADJOINT_C_STAGE_RECORDS;
    const int64_t interval = records ? n_steps : checkpoint_interval;
    const int64_t n_ck = (n_steps + interval - 1) / interval;
    const int64_t seg = std::min(interval, n_steps);
    const size_t qb = sizeof(double) * d.n_q;
    // Segments k >= first_kept keep their stage states from the forward pass (one block each);
    // the others are recomputed into block 0 once the kept ones have been consumed.
    const int64_t seg_bytes = (int64_t)qb * 3 * seg;
    const int64_t n_kept =
        records ? 0 : std::max<int64_t>(1, std::min<int64_t>(n_ck, 1 + mesh->adjoint_store_budget / seg_bytes));
    const int64_t first_kept = n_ck - n_kept;
    // Recomputed segments alternate between two stage-store blocks, so that the recompute of
    // segment k - 1 (on a second stream) runs beside the reverse sweep of segment k: the kept
    // blocks of segments first_kept and first_kept + 1, or one extra block when only one is kept.
    const bool overlap = !records && first_kept > 0;
    DevBuf lam, mu, chi0, chi1, cpart, recs, res, recs_g;
    // checkpoints and stage-store blocks in the mesh's arenas (reused by the next call)
    const int64_t n_blocks = n_kept + (overlap && n_kept < 2 ? 1 : 0);
    HIP_TRY(arena_fit_exact(mesh->adj_ck, qb * (size_t)n_ck), "hipMalloc checkpoints");
    if (n_blocks > 0) HIP_TRY(arena_fit_exact(mesh->adj_stage, (size_t)seg_bytes * (size_t)n_blocks), "hipMalloc stage store");
    double *const ck_base = (double *)mesh->adj_ck.p;
    auto block = [&](int64_t b) { return (double *)mesh->adj_stage.p + b * (seg_bytes / (int64_t)sizeof(double)); };
    double *const rbuf[2] = {n_kept >= 1 ? block(0) : nullptr, n_kept >= 2 ? block(1) : (n_blocks > 1 ? block(1) : nullptr)};
    if ((st = dev_alloc(lam, qb, "hipMalloc adjoint")) != NDG_OK) return st;
    if ((st = dev_alloc(mu, qb, "hipMalloc adjoint")) != NDG_OK) return st;
    if ((st = dev_alloc(chi0, sizeof(double) * d.n_tr, "hipMalloc adjoint")) != NDG_OK) return st;
    if ((st = dev_alloc(chi1, sizeof(double) * d.n_tr, "hipMalloc adjoint")) != NDG_OK) return st;
    const int64_t nblk = adjoint_blocks(d);
    if ((st = dev_alloc(cpart, sizeof(double) * nblk, "hipMalloc adjoint")) != NDG_OK) return st;
    if ((st = dev_alloc(recs, sizeof(double) * R.n * n_steps, "hipMalloc records")) != NDG_OK) return st;
    if ((st = dev_alloc(res, sizeof(double) * R.n * n_steps, "hipMalloc residuals")) != NDG_OK) return st;
    if (records && (st = dev_alloc(recs_g, sizeof(double) * R.n * 3 * n_steps, "hipMalloc records")) != NDG_OK)
        return st;
    auto ckp = [&](int64_t k) { return ck_base + k * d.n_q; };
    // stage state (step j of segment k, relative to the segment start)
    auto stp = [&](int64_t k, int64_t j, int stage) {
        double *b = k >= first_kept ? block(k - first_kept) : rbuf[(first_kept - k) % 2];
        return b + (j * 3 + stage) * d.n_q;
    };
    double *const q_main = d.q;
    // On every exit path, errors included, the mesh's state pointers are its own buffers again:
    // run_stored points them into the checkpoint and stage-store blocks.
    struct StatePointers {
        DeviceMesh &dm;
        double *q;
        ~StatePointers() {
            dm.q = q;
            dm.q_out = nullptr;
        }
    } state_pointers{d, q_main};

    // Steps n0..n1-1 from the state at q_start, keeping every stage input: the stage kernels write
    // their output state into the next store slot (q_start itself stands for slot (n0, 0)); the
    // output of the last step goes to the mesh's own state buffer. Records the receivers after
    // every step when asked.
    // dm: the mesh (forward pass) or a copy of it whose state pointers the recompute moves (on
    // stream st, beside the reverse sweep, which reads none of the buffers the stage kernels write).
    auto run_stored = [&](DeviceMesh &dm, hipStream_t st_, int64_t k, const double *q_start, int64_t n0, int64_t n1,
                          bool record) -> ndg_status {
        dm.q = const_cast<double *>(q_start);
        HIP_TRY(launch_traces(dm, st_), "traces kernel");
        for (int64_t j = n0; j < n1; ++j)
            for (int stage = 0; stage < 3; ++stage) {
                double *in = (j == n0 && stage == 0) ? const_cast<double *>(q_start) : stp(k, j - n0, stage);
                double *out = stage < 2 ? stp(k, j - n0, stage + 1) : (j + 1 < n1 ? stp(k, j + 1 - n0, 0) : q_main);
                dm.q = in;
                dm.q_out = out;
                HIP_TRY(launch_stage(dm, stage, kAm[stage], kGm[stage] * dt, st_), "stage kernel");
                if (stage == 2 && record) {
                    dm.q = out;
                    HIP_TRY(launch_record(dm, R.elem.as<int64_t>(), R.w.as<double>(), R.n,
                                          recs.as<double>() + j * R.n, st_), "record kernel");
                }
            }
        dm.q = q_main;
        dm.q_out = nullptr;
        return NDG_OK;
    };
    auto stage_input = [&](int64_t k, int64_t j, int stage) -> const double * {
        return (j == k * interval && stage == 0) ? ckp(k) : stp(k, j - k * interval, stage);
    };

    // Forward pass: checkpoints; the kept segments keep their stage inputs right away. With stage
    // records: the RK register's pressure at the receivers after every stage as well.
    HIP_TRY(launch_clock_begin(d, 0.0, 1, dt, s), "clock kernel");
    if (records) {
        HIP_TRY(hipMemcpyAsync(ckp(0), d.q, qb, hipMemcpyDeviceToDevice, s), "checkpoint");
        // the RK register is recorded after every stage, the last included
        struct KeepLastRk {
            DeviceMesh &d;
            KeepLastRk(DeviceMesh &dm) : d(dm) { d.keep_last_rk = true; }
            ~KeepLastRk() { d.keep_last_rk = false; }
        } keep_last_rk{d};
        for (int64_t n = 0; n < n_steps; ++n) {
            for (int stage = 0; stage < 3; ++stage) {
                HIP_TRY(launch_stage(d, stage, kAm[stage], kGm[stage] * dt, s), "stage kernel");
                HIP_TRY(launch_record_rk(d, R.elem.as<int64_t>(), R.w_t.as<double>(), R.n,
                                         recs_g.as<double>() + (n * 3 + stage) * R.n, s),
                        "record kernel");
            }
            HIP_TRY(launch_record(d, R.elem.as<int64_t>(), R.w.as<double>(), R.n, recs.as<double>() + n * R.n, s),
                    "record kernel");
        }
    } else {
        for (int64_t n = 0; n < first_kept * interval; ++n) {
            if (n % interval == 0)
                HIP_TRY(hipMemcpyAsync(ckp(n / interval), d.q, qb, hipMemcpyDeviceToDevice, s), "checkpoint");
            if ((st = step_from_clock(mesh, dt)) != NDG_OK) return st;
            HIP_TRY(launch_record(d, R.elem.as<int64_t>(), R.w.as<double>(), R.n, recs.as<double>() + n * R.n, s),
                    "record kernel");
        }
        for (int64_t k = first_kept; k < n_ck; ++k) {
            HIP_TRY(hipMemcpyAsync(ckp(k), d.q, qb, hipMemcpyDeviceToDevice, s), "checkpoint");
            const int64_t n0 = k * interval, n1 = std::min(n0 + interval, n_steps);
            if ((st = run_stored(d, s, k, ckp(k), n0, n1, true)) != NDG_OK) return st;
        }
    }
    std::vector<double> hres((size_t)(R.n * n_steps));
    HIP_TRY(hipMemcpyAsync(hres.data(), recs.p, sizeof(double) * hres.size(), hipMemcpyDeviceToHost, s), "download records");
    HIP_TRY(hipStreamSynchronize(s), "adjoint sync");
    double Jv = 0.0;
    for (size_t k = 0; k < hres.size(); ++k) {
        hres[k] -= data ? data[k] : 0.0;
        Jv += hres[k] * hres[k];
    }
    Jv /= 2;
    HIP_TRY(hipMemcpyAsync(res.p, hres.data(), sizeof(double) * hres.size(), hipMemcpyHostToDevice, s), "upload residuals");
    // dJ/dc from the stage records. With z = dt G = c dt G1 one step is q_n = P(z) q_{n-1},
    // P(z) = 1 + z + z^2/2 + z^3/6 (the RK3 stability polynomial), so dq_n/dc = (n / c) P'(z) z
    // q_{n-1} = (n / c) (z + z^2 + z^3 / 2) q_{n-1}. The stage right-hand sides k_s = dt G q_s of step
    // n, k_s = dt (g_s - a_s g_{s-1}), satisfy k0 = z q, k1 = z q + gm0 z^2 q,
    // k2 = z q + beta z^2 q + gm0 gm1 z^3 q with beta = gm0 + gm1 (a1 + 1).
    double dc_records = 0.0;
    if (records) {
        std::vector<double> hg((size_t)(R.n * 3 * n_steps));
        HIP_TRY(hipMemcpy(hg.data(), recs_g.p, sizeof(double) * hg.size(), hipMemcpyDeviceToHost), "download records");
        const double beta = kGm[0] + kGm[1] * (kAm[1] + 1.0);
        for (int64_t n = 0; n < n_steps; ++n)
            for (int r = 0; r < R.n; ++r) {
                const double g0 = hg[(size_t)((n * 3 + 0) * R.n + r)], g1 = hg[(size_t)((n * 3 + 1) * R.n + r)],
                             g2 = hg[(size_t)((n * 3 + 2) * R.n + r)];
                const double k0 = dt * g0, k1 = dt * (g1 - kAm[1] * g0), k2 = dt * (g2 - kAm[2] * g1);
                const double z2 = (k1 - k0) / kGm[0];
                const double z3 = (k2 - k0 - beta * z2) / (kGm[0] * kGm[1]);
                dc_records += hres[(size_t)(n * R.n + r)] * ((double)(n + 1) / d.wave_speed) * (k0 + z2 + z3 / 2);
            }
    }

    // Reverse sweep: per segment (the last one is already stored) recompute from its checkpoint
    // keeping the stage inputs, then the transposed stages backwards. The misfit source of each
    // step end is added by the stage-0 kernel of the following step (in reverse order), which
    // also exports for the next stage 2.
    HIP_TRY(hipMemsetAsync(lam.p, 0, qb, s), "hipMemset adjoint");
    HIP_TRY(hipMemsetAsync(mu.p, 0, qb, s), "hipMemset adjoint");
    HIP_TRY(hipMemsetAsync(cpart.p, 0, sizeof(double) * nblk, s), "hipMemset adjoint");
    double *chi[2] = {chi0.as<double>(), chi1.as<double>()};
    int cc = 0;
    HIP_TRY(launch_source(lam.as<double>(), d, R.elem.as<int64_t>(), R.w_src.as<double>(),
                          res.as<double>() + (n_steps - 1) * R.n, R.n, s), "source kernel");
    {
        AdjointArgs A;
        A.lam = lam.as<double>();
        A.mu = mu.as<double>();
        A.chi_out = chi[1];
        A.flags = 4;  // export only, mu = 0 at a step end
        A.gnext = kGm[2] * dt;
        HIP_TRY(launch_adjoint_stage(d, A, s), "adjoint export kernel");
        cc = 1;
    }
    // The recompute stream (default priority) and its events: ev_rec[k] after segment k is
    // recomputed, ev_rev[k] after its reverse sweep (its block is then free for segment k - 2).
    struct Overlap {
        hipStream_t rs = nullptr;
        std::vector<hipEvent_t> ev_rec, ev_rev;
        ~Overlap() {
            if (rs) {
                (void)hipStreamSynchronize(rs);
                (void)hipStreamDestroy(rs);
            }
            for (hipEvent_t e : ev_rec) (void)hipEventDestroy(e);
            for (hipEvent_t e : ev_rev) (void)hipEventDestroy(e);
        }
    } ov;
    if (overlap) {
        HIP_TRY(hipStreamCreateWithFlags(&ov.rs, hipStreamNonBlocking), "hipStreamCreate");
        for (std::vector<hipEvent_t> *v : {&ov.ev_rec, &ov.ev_rev})
            for (int64_t k = 0; k < n_ck; ++k) {
                hipEvent_t e;
                HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
                v->push_back(e);
            }
    }
    auto recompute = [&](int64_t k) -> ndg_status {
        const int64_t n0 = k * interval, n1 = std::min(n0 + interval, n_steps);
        if (k + 2 < n_ck) HIP_TRY(hipStreamWaitEvent(ov.rs, ov.ev_rev[(size_t)(k + 2)], 0), "hipStreamWaitEvent");
        DeviceMesh dr = d;
        ndg_status rst = run_stored(dr, ov.rs, k, ckp(k), n0, n1, false);
        if (rst != NDG_OK) return rst;
        HIP_TRY(hipEventRecord(ov.ev_rec[(size_t)k], ov.rs), "hipEventRecord");
        return NDG_OK;
    };
    for (int64_t k = n_ck - 1; k >= 0; --k) {
        const int64_t n0 = k * interval, n1 = std::min(n0 + interval, n_steps);
        if (!records && k < first_kept) {
            if (overlap) HIP_TRY(hipStreamWaitEvent(s, ov.ev_rec[(size_t)k], 0), "hipStreamWaitEvent");
            else if ((st = run_stored(d, s, k, ckp(k), n0, n1, false)) != NDG_OK) return st;
        }
        if (overlap && k >= 1 && k - 1 < first_kept && (st = recompute(k - 1)) != NDG_OK) return st;
        for (int64_t j = n1 - 1; j >= n0; --j)
            for (int stage = 2; stage >= 0; --stage) {
                AdjointArgs A;
                A.lam = lam.as<double>();
                A.mu = mu.as<double>();
                A.chi_in = chi[cc];
                A.chi_out = chi[1 - cc];
                A.qs = records ? nullptr : stage_input(k, j, stage);
                A.cpart = records ? nullptr : cpart.as<double>();
                const bool more = stage != 0 || j >= 1;  // a reverse stage follows
                A.flags = 8 | (stage != 2 ? 1 : 0) | (stage != 0 ? 2 : 0) | (more ? 4 : 0) |
                          (stage == 0 && j >= 1 ? 16 : 0);
                A.gs = kGm[stage] * dt;
                A.as = kAm[stage];
                A.gnext = stage > 0 ? kGm[stage - 1] * dt : kGm[2] * dt;
                A.rec_elem = R.elem.as<int64_t>();
                A.rec_w = R.w_src.as<double>();
                A.rec_res = j >= 1 ? res.as<double>() + (j - 1) * R.n : nullptr;
                A.n_rec = R.n;
                HIP_TRY(launch_adjoint_stage(d, A, s), "adjoint kernel");
                if (more) cc = 1 - cc;
            }
        if (overlap) HIP_TRY(hipEventRecord(ov.ev_rev[(size_t)k], s), "hipEventRecord");
    }

    // Outputs (lambda back from its J-scaled form); the mesh state goes back to the initial fields.
    HIP_TRY(launch_jacobian_scale(d, lam.as<double>(), nullptr, 0, 1, s), "jacobian scale kernel");
    std::vector<double> hc((size_t)nblk);
    HIP_TRY(hipMemcpyAsync(hc.data(), cpart.p, sizeof(double) * nblk, hipMemcpyDeviceToHost, s), "download c-gradient");
    double *const outs[3] = {dJ_dp0, dJ_du0, dJ_dv0};
    if ((st = copy_state(mesh, lam.as<double>(), outs, true, "download gradient")) != NDG_OK) return st;
    double dc = 0.0;
    for (double v : hc) dc += v;
    HIP_TRY(hipMemcpyAsync(d.q, ckp(0), qb, hipMemcpyDeviceToDevice, s), "restore state");
    HIP_TRY(hipMemsetAsync(d.g, 0, qb, s), "hipMemset g");
    HIP_TRY(launch_traces(d, s), "traces kernel");
    HIP_TRY(hipStreamSynchronize(s), "ndg_adjoint_gradient sync");
    if (J) *J = Jv;
    if (dJ_dc) *dJ_dc = records ? dc_records : dc / d.wave_speed;
    return NDG_OK;
}
