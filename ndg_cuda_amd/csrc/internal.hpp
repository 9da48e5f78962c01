// internal.hpp -- structures shared by the host side of libndg_amd (mesh build, C ABI) and the
// HIP kernels. Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "ndg_amd.h"

namespace ndg {

constexpr double kWaveSpeed = 1.0;  // include/helpers/constants.h:15 (the flux form; see wave_speed)
constexpr int kMaxOrder = 16;       // largest order with a compiled kernel instance

// Boundary (ghost) side record. One per element side whose neighbour is a wall, symmetry,
// inflow or outflow ghost element, or whose face is non-conforming (kind kKindMortar). n/t are the
// face normal and tangent of the reference (Face2D_t::compute_geometry with the domain element on
// the left); a/b the ghost element's first two nodes after face building; sf the ghost's edge
// scaling factor (half the edge length, src/entities/Element2D_t.cpp:779-790). For mortar sides
// `slot` is the offset (in values) of the side's projected flux [3][N+1] in the side-flux buffer;
// a larger side covered by n_faces faces sums n_faces such blocks, one per face in the side's
// direction, from `slot` on.
struct BcSide {
    double nx, ny, tx, ty;
    double ax, ay, bx, by;
    double sf;
    int32_t kind;
    int32_t n_faces = 1;
    int64_t slot;
};

// Stage kernel flags: the step's first stage (the RK register is not read) and its last stage.
constexpr int kRkFirst = 1, kRkLast = 2;

// Element side whose face is non-conforming: its element flux trace is produced by the mortar
// kernel (interpolation to the face, flux, L2 projection back) before the stage kernels run.
constexpr int32_t kKindMortar = 100;

// One non-conforming face (src/meshes/Mesh2D_t.cpp:3803-4006). L is the element
// that created the face in the reference's face building, R the other element (or, for a periodic
// side, the element whose trace the ghost copies, in which case only L is projected: slotR = -1).
struct MortarFace {
    int64_t trL, trR;        // offsets (values) of the L / R side traces [3][N+1] in the trace buffer
    int64_t slotL, slotR;    // offsets of the L / R projected side fluxes, slotR = -1 if one-sided
    double nx, ny;           // face normal (outward from L)
    double offset[2], scale[2];
    double sfL, sfR;         // edge Jacobians of the L / R element sides
    int32_t NL, NR, Nf, pad;
    // Offsets into the mortar matrix buffer of the L and R interpolation matrices ([Nf+1][N+1],
    // element trace -> face nodes, R reversed) and projection matrices ([N+1][Nf+1], face flux ->
    // element side, R negated and reversed, face scale included); -1 where the side is a plain copy.
    int64_t mat[4];
};

// Mortar side classes. Splitting halves sides, so a non-conforming face covers a dyadic piece of an
// element side: [j / 2^d, (j + 1) / 2^d] (offset j / 2^d, scale 1 / 2^d), class 2^d - 1 + j, with d
// up to kMaxSpanLevel (class 0 the whole side, 1 and 2 its halves, 3..6 its quarters, ...). The
// offsets and scales computed from the vertices (Face2D_t.cpp:68-81) can miss these by an ulp; the
// class snaps them, so one interpolation/projection matrix per (element order, face order, side,
// class) serves every face. Returns -1 for a side that is none of these.
constexpr int kMaxSpanLevel = 4;
constexpr int kMortarClasses = (2 << kMaxSpanLevel) - 1;
__host__ __device__ inline int mortar_class(double offset, double scale) {
    const double tol = 1e-9;
    for (int d = 0; d <= kMaxSpanLevel; ++d) {
        const double w = 1.0 / (double)(1 << d);
        if (fabs(scale - w) > tol) continue;
        const double jj = offset / w;
        const int j = (int)floor(jj + 0.5);
        if (j < 0 || j >= (1 << d) || fabs(offset - j * w) > tol) return -1;
        return (1 << d) - 1 + j;
    }
    return -1;
}
__host__ __device__ inline int class_level(int cls) {
    int d = 0;
    while ((2 << d) - 1 <= cls) ++d;
    return d;
}
__host__ __device__ inline int class_index(int cls) { return cls + 1 - (1 << class_level(cls)); }
__host__ __device__ inline int make_class(int d, int j) { return (1 << d) - 1 + j; }
__host__ __device__ inline double class_offset(int cls) {
    return (double)class_index(cls) / (double)(1 << class_level(cls));
}
__host__ __device__ inline double class_scale(int cls) { return 1.0 / (double)(1 << class_level(cls)); }
// Index of a matrix in the table of mortar matrices (which = 0 interpolation, 1 projection).
__host__ __device__ inline int64_t mortar_matrix_key(int Ne, int Nf, int side, int cls, int which) {
    return ((((int64_t)Ne * (kMaxOrder + 1) + Nf) * 2 + side) * kMortarClasses + cls) * 2 + which;
}
constexpr int64_t kMortarMatrixKeys = (int64_t)(kMaxOrder + 1) * (kMaxOrder + 1) * 4 * kMortarClasses;

// Per-order constant tables in device memory, packed as Dhat[Nf*Nf], l(-1)[Nf], l(+1)[Nf], w[Nf],
// x[Nf], 1/w[Nf], then the modal matrix poly[Nf*Nf] used by the estimator and the transpose
// Dhat^T[Nf*Nf] used by the adjoint (its columns as contiguous rows).
inline int64_t table_pack_len(int N) {
    const int nf = N + 1;
    return (int64_t)nf * nf * 3 + 5 * nf;
}
inline int64_t table_dhat_t_offset(int N) {
    const int nf = N + 1;
    return (int64_t)nf * nf * 2 + 5 * nf;
}
inline int64_t table_poly_offset(int N) {
    const int nf = N + 1;
    return (int64_t)nf * nf + 5 * nf;
}

// 1D tables of every order 0..kMaxOrder for the mortar kernel: nodes, weights and barycentric
// weights of order n at tab1d[(3 * n + which) * kTab1dStride].
constexpr int kTab1dStride = kMaxOrder + 1;

struct Tables {
    int N_max = 0;
    int64_t n_interp = 0;
    std::vector<std::vector<double>> tab[12];
};

void build_tables(int N_max, int64_t n_interp, Tables &t);
// Packs the device table block for order N (layout above).
std::vector<double> pack_tables(const Tables &t, int N);

// A device allocation that keeps its capacity: arena_fit reallocates only when more bytes are
// asked for than it holds (with 25% headroom), so meshes rebuilt by adaptation reuse their
// buffers. Contents are not kept across a reallocation.
struct Arena {
    void *p = nullptr;
    size_t cap = 0;
};
hipError_t arena_fit(Arena &a, size_t bytes);
// The same without headroom (for very large blocks sized by the caller).
hipError_t arena_fit_exact(Arena &a, size_t bytes);
void arena_free(Arena &a);

// The buffers of one device mesh. A mesh keeps two: the current one, and the one an adaptation
// builds the next mesh into (the state is transferred from the first to the second).
struct DeviceStore {
    Arena verts, side_nb, bc, faces, side_flux, side_flux32, q, g, tr[2], rhs, scratch, qh, gl32, tr32[2];
    Arena spans;  // GroupSpan[kMaxOrder + 1] of the layout
};
void store_free(DeviceStore &s);

// Hilbert child tables (src/functions/Hilbert_splitting.cpp:23-81), statuses H, A, R, B = 0..3:
// the place of child s (the child at corner s) among the 4 children, and its status, by parent
// rotation and status. One initializer for the host and the device copies.
#define NDG_HILBERT_CHILD_ORDER                                                    \
    {{{0, 3, 2, 1}, {0, 1, 2, 3}, {2, 1, 0, 3}, {2, 3, 0, 1}},                     \
     {{3, 2, 1, 0}, {1, 2, 3, 0}, {1, 0, 3, 2}, {3, 0, 1, 2}},                     \
     {{2, 1, 0, 3}, {2, 3, 0, 1}, {0, 3, 2, 1}, {0, 1, 2, 3}},                     \
     {{1, 0, 3, 2}, {3, 0, 1, 2}, {3, 2, 1, 0}, {1, 2, 3, 0}}}
#define NDG_HILBERT_CHILD_STATUS                                                   \
    {{{1, 3, 0, 0}, {0, 1, 1, 2}, {2, 2, 3, 1}, {3, 0, 2, 3}},                     \
     {{3, 0, 0, 1}, {1, 1, 2, 0}, {2, 3, 1, 2}, {0, 2, 3, 3}},                     \
     {{0, 0, 1, 3}, {1, 2, 0, 1}, {3, 1, 2, 2}, {2, 3, 3, 0}},                     \
     {{0, 1, 3, 0}, {2, 0, 1, 1}, {1, 2, 2, 3}, {3, 3, 0, 2}}}

// Elements of one order, contiguous in device order (Hilbert order kept within the group).
struct Group {
    int N = 0;
    int64_t K = 0;
    int64_t e_begin = 0;     // first element (device order)
    int64_t q_off = 0;       // offset (values) into q / g: 3 (N+1)^2 per element
    int64_t tr_off = 0;      // offset (values) into the trace buffers: 12 (N+1) per element
    double *tab = nullptr;   // packed tables for N
    float *tab32 = nullptr;
    // Launch shape of the group's stage kernel per precision (index NDG_PRECISION_*): elements
    // per workgroup (0 = the shape's maximum) and the XCD-contiguous block order on/off.
    struct Tune {
        int epb = 0;
        int remap = 1;
    } tune[2];
};

// Device-side description of a mesh: elements grouped by order, each group stored as the
// uniform-order layout [K_g][3][(N+1)^2] (state, RK register) and [K_g][4][3][N+1] (traces).
struct DeviceMesh {
    std::vector<Group> groups;
    int64_t K = 0;
    int N_min = 0, N_max = 0;
    int64_t n_q = 0, n_tr = 0;   // values per state buffer / trace buffer
    double *verts = nullptr;     // [K][4][2], device order
    int32_t *side_nb = nullptr;  // [K][4]: group-local nb*4+side (>=0) or -(1+bc index)
    BcSide *bc = nullptr;        // [n_bc]
    int64_t n_bc = 0;
    MortarFace *faces = nullptr; // [n_faces]
    int64_t n_faces = 0;
    int mortar_nodes = 0;        // largest N + 1 over the faces' NL, NR and Nf (sizes the mortar LDS)
    double *side_flux = nullptr; // projected mortar side fluxes
    float *side_flux32 = nullptr;
    int64_t n_side_flux = 0;
    double *tab1d = nullptr;     // 1D tables of all orders (solution transfer)
    double *mortar_mats = nullptr;  // interpolation / projection matrices of the mortar faces
    double *q = nullptr;
    double *g = nullptr;         // RK register (alive within a step; the stage kernel's tile order)
    double *trace[2] = {nullptr, nullptr};
    double *rhs = nullptr;       // compute_rhs only (allocated on demand)
    double *scratch = nullptr;   // reductions
    int64_t scratch_len = 0;
    int cur = 0;                 // trace buffer holding the traces of the current state
    double *clock = nullptr;     // [4]: step start time, then the 3 stage times
    double *q_out = nullptr;     // when set, FP64 stage kernels write the new state here (not in q)
    bool has_inflow = false;     // the stage time is read only by inflow ghosts
    // Wave speed c of the symmetric system p_t + c div(u, v) = 0, (u, v)_t + c grad p = 0: the
    // right-hand side is c times the reference operator (whose constant c is 1), so c = 1 is the
    // reference arithmetic bit for bit.
    double wave_speed = 1.0;
    // FP32 form with compensated state (allocated by ndg_mesh_set_precision): the state is the
    // unevaluated sum qh + ql of two floats per value, the RK register and traces are floats.
    int precision = NDG_PRECISION_FP64;
    // Store the RK register at a step's last stage too (the adjoint's stage records read it after
    // every stage); otherwise that store is skipped, no later stage reading it.
    bool keep_last_rk = false;
    float *qh = nullptr;
    // RK register (x) and low state word (y) of each value, interleaved so that a line owner
    // reads and writes both with one 8-byte access; transposed within each (element, field) block
    float2 *gl32 = nullptr;
    float *trace32[2] = {nullptr, nullptr};
};

// Position of the order-N group in the device layout (one per order 0..kMaxOrder; K = 0 when the
// mesh has no element of that order), in device memory for the adaptation kernels.
struct GroupSpan {
    int64_t K, e_begin, q_off, tr_off;
};

// The mesh's element-side topology in the caller's element order (the curve order), kept in device
// memory so that an adaptation can rebuild the mesh there (adapt.hip). link[4e+s] is the element
// side whose trace is the exterior state of side s of element e (a neighbour, or the donor side of
// a periodic pair), or -(1 + ghost); hang: 0 conforming, 1 a small side whose exterior is a piece
// of a larger element's side (link = that side, span = the piece's mortar class in the larger
// side's direction, hang2 = the next small side along the larger side, -1 after the last), 2 the
// larger side (link = its first small side, the one at the side's first corner; hang2 = -1).
struct LogicalMesh {
    int64_t K = 0, n_ghost = 0;
    double *verts = nullptr;                      // [K][4][2]
    int32_t *order = nullptr, *status = nullptr, *rotation = nullptr, *level = nullptr;
    int64_t *link = nullptr, *hang2 = nullptr;    // [4K]
    int8_t *hang = nullptr, *periodic = nullptr;  // [4K]
    int8_t *span = nullptr;                       // [4K]
    double *est = nullptr;                        // [K][4] last estimate: error, sigma, refine, coarsen
    int64_t *dev_of = nullptr;                    // [K] index in the order-grouped device layout
    BcSide *ghost = nullptr;                      // [n_ghost] ghost records, in boundary-bar order
    int64_t *ghost_side = nullptr;                // [n_ghost] element side of each ghost
    int8_t *ghost_zero = nullptr;                 // [n_ghost] storage resized to zero (the scale is 0)
};
struct LogicalStore {
    Arena verts, order, status, rotation, level, link, hang2, hang, periodic, span, est, dev_of, ghost, ghost_side,
        ghost_zero;
};

// Counters one adaptation gathers on the device and the host reads back (two round trips).
struct AdaptCounts {
    unsigned long long n_split, n_raise, refused, ghost_split;
    unsigned long long dofs, nodes1d;            // sum (N+1)^2 and sum (N+1) of the new mesh
    unsigned long long order_count[kMaxOrder + 1];
    unsigned long long n_rec, n_slot, n_face;    // mortar records, side-flux values, mortar faces
    unsigned long long conforming_pairs, mortar_nodes, bad_face;
    unsigned long long periodic_faces;           // faces of periodic sides (one per face on the side)
    unsigned long long hanging_faces;            // faces of interior larger sides
};
// refused bits: a raised order beyond the tables, a face smaller than 1 / 2^kMaxSpanLevel of its side
constexpr unsigned kRefuseOrder = 1, kRefuseSpan = 4;

// Scratch of the adaptation kernels (capacities kept across adaptations).
struct AdaptWork {
    Arena flags, newpos, newghost, parent, corner, cnt_rec, cnt_slot, cnt_face, blocks, hist, counts;
};

// Device adaptation steps (adapt.hip). Each launches on stream s and returns the launch error.
// 1. Decisions from the estimate est (device order): flags[e] = 0 keep, 1 raise by 2, 2 split
//    (Element2D_t.cpp:802-812 with the split level bound), the logical copy of the estimate, and
//    the counts (splits, raises, refusals, the new mesh's sizes).
hipError_t adapt_decide(const LogicalMesh &L, const double *est, int N_max, int tables_N_max, int max_split_level,
                        AdaptWork &w, hipStream_t s);
// 2. The new logical mesh (children in Hilbert order, raised orders, links, ghosts), into Ln
//    (buffers sized for the counts of step 1). p_only: the no-split branch, whose raised elements'
//    ghosts keep a zero scale.
hipError_t adapt_rebuild(const LogicalMesh &L, LogicalMesh &Ln, bool p_only, AdaptWork &w, hipStream_t s);
// 3. The order-grouped layout of Ln: dev_of, the group spans (spans[kMaxOrder + 1]) and the
//    counts of mortar records, slots and faces.
hipError_t adapt_layout_count(LogicalMesh &Ln, GroupSpan *spans, AdaptWork &w, hipStream_t s);
// 4. The device mesh arrays of Ln (vertices, neighbour codes, boundary and mortar records, mortar
//    faces) into d (buffers sized for the counts of step 3); mortar_key_off maps a matrix key to
//    its offset in the mortar matrix table.
hipError_t adapt_layout_fill(const LogicalMesh &Ln, const GroupSpan *spans, const int64_t *mortar_key_off,
                             DeviceMesh &d, AdaptWork &w, hipStream_t s);
// 5. The state of every new element from its parent (copy, new order, or a child's nodes through
//    the inverse bilinear map), from the layout (L, spans_old, q_old) to (Ln, spans_new, q_new).
hipError_t adapt_transfer(const LogicalMesh &L, const GroupSpan *spans_old, const double *q_old,
                          const LogicalMesh &Ln, const GroupSpan *spans_new, double *q_new, const double *tab1d,
                          AdaptWork &w, hipStream_t s);

// The logical copy of an estimate est (device order), L.est (caller's order).
hipError_t adapt_copy_estimate(const LogicalMesh &L, const double *est, hipStream_t s);

// Kernel launchers (kernels.hip). All return hipSuccess or the first launch error.
hipError_t launch_traces(const DeviceMesh &m, hipStream_t s);
// Stage s of the RK step (mortar faces, then every order group, in the mesh's precision); the
// stage time is read from clock[1 + s] (see launch_clock_begin). Flips `cur`.
hipError_t launch_stage(DeviceMesh &m, int stage, double a, double gdt, hipStream_t s);
hipError_t launch_rhs(DeviceMesh &m, double t, hipStream_t s);
// FP32 compensated form: conversions between the FP64 state (q, g) and (qh, ql, g32), traces of
// qh, and the stage.
hipError_t launch_to_c32(DeviceMesh &m, hipStream_t s);
hipError_t launch_from_c32(DeviceMesh &m, hipStream_t s);
hipError_t launch_traces_c32(const DeviceMesh &m, hipStream_t s);
// One order group's stage kernel (of the mesh's precision); does not flip `cur`.
hipError_t launch_group_stage(DeviceMesh &m, int group, int stage, double a, double gdt, hipStream_t s);
// Device clock: begin writes the 3 stage times of a step starting at t (set != 0) or at clock[0];
// end advances clock[0] by dt. Both are single-thread kernels, so a captured step replays with
// the time advancing on the device. Meshes without inflow ghosts never read the time, and the
// clock kernels are skipped for them.
hipError_t launch_clock_begin(DeviceMesh &m, double t, int set, double dt, hipStream_t s);
hipError_t launch_clock_end(DeviceMesh &m, double dt, hipStream_t s);
// Autotuning: the largest elements-per-workgroup of the stage kernel for order N, and one launch
// of group g's stage kernel that leaves the state and the current traces unchanged.
int stage_max_epb(int N);
hipError_t launch_group_probe(DeviceMesh &m, int group, hipStream_t s);
// Initial condition of kind NDG_IC_*; Lx, Ly are the periods of NDG_IC_PERIODIC_PULSES.
hipError_t launch_ic(const DeviceMesh &m, int kind, int n_pulses, const double *pulses_dev, double Lx, double Ly,
                     hipStream_t s);
hipError_t launch_delta_t(const DeviceMesh &m, double cfl, double *dt_dev, hipStream_t s);
// Mortar faces: interpolation to the face, upwind flux, projection to the element sides.
hipError_t launch_mortar(DeviceMesh &m, hipStream_t s);
hipError_t launch_estimate(const DeviceMesh &m, double tol_min, double tol_max, double *out_dev,
                           hipStream_t s);
bool order_supported(int N);

// Adjoint (transpose) stage on a single-order mesh with interior and periodic sides only
// (adjoint.hip). flags: 1 mu input non-zero, 2 store mu <- as mu~, 4 export chi for the next
// reverse stage (from mu' + gnext lambda'), 8 apply the transpose (else export only), 16 add the
// receivers' misfit source to lambda' (after the transpose, before the export).
struct AdjointArgs {
    double *lam = nullptr, *mu = nullptr;
    const double *chi_in = nullptr;
    double *chi_out = nullptr;
    const double *qs = nullptr;  // stage input state for the c-derivative (may be null)
    double *cpart = nullptr;     // per-wave accumulators of <d lambda, q_s> (may be null)
    int flags = 0;
    double gs = 0.0, as = 0.0, gnext = 0.0;
    // flag 16: receivers (element in device order, weights, residuals) added to lambda'
    const int64_t *rec_elem = nullptr;
    const double *rec_w = nullptr, *rec_res = nullptr;
    int n_rec = 0;
};
hipError_t launch_adjoint_stage(DeviceMesh &m, const AdjointArgs &A, hipStream_t s);
int64_t adjoint_blocks(const DeviceMesh &m);  // accumulator slots of cpart (3 per workgroup)
// x *= J^power per node (power 1 or -1), single-order meshes: the packed state layout, or with
// rec_elem the receiver weight arrays (n_rec of (N+1)^2, element rec_elem[r] in device order).
hipError_t launch_jacobian_scale(const DeviceMesh &m, double *x, const int64_t *rec_elem, int64_t n_rec, int power,
                                 hipStream_t s);
// Receivers (element in device order, (N+1)^2 interpolation weights each): p at the receivers of
// the current state, and lambda_p += res_r w_r.
// Receiver values sum_k w[r][k] q[elem r][field 0][k] of the state q.
hipError_t launch_record(const DeviceMesh &m, const int64_t *elem, const double *w, int n, double *out,
                         hipStream_t s);
// The same of the RK register g (single-order meshes), which the FP64 stage kernel keeps in its
// tile order; w_t holds the weights in transposed node order (k = j (N+1) + i for node (i, j)).
hipError_t launch_record_rk(const DeviceMesh &m, const int64_t *elem, const double *w_t, int n, double *out,
                            hipStream_t s);
// Elements per workgroup tile of the FP64 stage kernel for group G (the RK register's tile order).
int stage_tile_epb(const Group &G);
hipError_t launch_source(double *lam, const DeviceMesh &m, const int64_t *elem, const double *w,
                         const double *res, int n, hipStream_t s);

// Solution transfer between two layouts after adaptation (one record per new element, offsets in
// values of q): copy, interpolation to a new order, or interpolation at given reference
// coordinates of the old element (a child of a split element).
// One element of the uniform-grid output (Mesh2D_t::get_solution), in the caller's element order.
struct OutputRec {
    int64_t q_off;   // the element's state block in DeviceMesh::q
    int64_t dev;     // its device index (vertices)
    int64_t i_off;   // its order's [n][N+1] interpolation matrix in the concatenated table
    int32_t N, pad;
};
// rhs == nullptr: 5 output planes (x, y, p, u, v); else 11 (plus the right-hand side and the
// plane-wave errors at `time`).
hipError_t launch_output(const double *q, const double *rhs, const double *verts, const OutputRec *recs,
                         const double *interp, const double *tab1d, int n, int64_t K, double time, double cw,
                         double *out, hipStream_t s);

struct TransferRec {
    int64_t old_off, new_off;
    int32_t N_old, N_new;
    int64_t coord_off;  // >= 0: the new nodes' reference coordinates in the old element, [(N+1)^2][2]
};
hipError_t launch_transfer(const double *q_old, double *q_new, const TransferRec *recs, const double *coords,
                           int64_t K, const double *tab1d, hipStream_t s);

// Allocation-free entry points: size of the fixed per-device scratch array (noalloc.hip), and its
// device address on the current device (looked up once per device, then cached).
constexpr size_t kNoAllocScratchBytes = size_t(4) << 20;
hipError_t noalloc_scratch(double **ptr);

// Records msg as the calling thread's last error and returns code.
ndg_status set_error(ndg_status code, const char *msg);

}  // namespace ndg
