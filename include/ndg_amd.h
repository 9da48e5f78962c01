/*
 * ndg_amd.h -- C ABI of the MI355X DG-SEM wave-stage library (libndg_amd.so).
 *
 * One RK3 time step of the nodal discontinuous-Galerkin spectral-element solver for the
 * first-order 2D wave system (p, u, v) on Legendre-Gauss nodes, run as HIP kernels on one AMD
 * Instinct MI355X (gfx950). Each entry point below replaces one piece of the reference host API
 * of guillaumetousignant/NDG_cuda (cited as file:line into that repository).
 *
 * Conventions
 *  - Every function returns an ndg_status; 0 is success. The reference exits the process on
 *    errors (e.g. src/meshes/Mesh2D_t.cpp:68-69); this ABI returns a code instead and keeps a
 *    message readable with ndg_last_error (thread-local).
 *  - Host pointers are borrowed for the duration of the call only. Device memory is owned by
 *    the library. All device work of a mesh is ordered on one HIP stream (the mesh's own, or
 *    one given with ndg_mesh_set_stream). Handles are not thread-safe.
 *  - Packed field layout ("reference order"): element by element in mesh order, each element's
 *    (N+1)^2 values in the reference node order q[i*(N+1)+j], i the xi index and j the eta index
 *    (include/entities/Element2D_t.h:37-39).
 */
#ifndef NDG_AMD_H
#define NDG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int ndg_status;
#define NDG_OK 0
#define NDG_ERR_INVALID 1     /* bad argument or inconsistent mesh description */
#define NDG_ERR_HIP 2         /* a HIP runtime call or kernel launch failed */
#define NDG_ERR_UNSUPPORTED 3 /* valid input the library does not handle (message says which) */
#define NDG_ERR_NO_DEVICE 4   /* no HIP device visible */

/* Boundary kinds of boundary (BAR_2) edges. Values follow the CGNS BC types the reference reads
 * (src/meshes/Mesh2D_t.cpp:689-791) plus the periodic self-interface (:820-875). */
#define NDG_BC_NONE 0
#define NDG_BC_WALL 1
#define NDG_BC_SYMMETRY 2
#define NDG_BC_INFLOW 3
#define NDG_BC_OUTFLOW 4
#define NDG_BC_PERIODIC 5

/* Table selectors (the public tables of NDG_t, include/entities/NDG_t.h:11-26). */
#define NDG_TABLE_NODES 0
#define NDG_TABLE_WEIGHTS 1
#define NDG_TABLE_BARYCENTRIC_WEIGHTS 2
#define NDG_TABLE_LAGRANGE_LEFT 3
#define NDG_TABLE_LAGRANGE_RIGHT 4
#define NDG_TABLE_LAGRANGE_DERIVATIVE_LEFT 5
#define NDG_TABLE_LAGRANGE_DERIVATIVE_RIGHT 6
#define NDG_TABLE_DERIVATIVE 7
#define NDG_TABLE_G_HAT_DERIVATIVE 8
#define NDG_TABLE_DERIVATIVE_HAT 9
#define NDG_TABLE_INTERPOLATION 10
#define NDG_TABLE_POLYNOMIALS 11

/* Arithmetic forms of the stage (ndg_mesh_set_precision). */
#define NDG_PRECISION_FP64 0             /* the reference's FP64 arithmetic (default) */
#define NDG_PRECISION_FP32_COMPENSATED 1 /* FP32 right-hand side and RK register; state kept as a
                                            two-sum-compensated pair of floats (no reference
                                            counterpart) */

/* Launch modes of the RK step (ndg_mesh_set_graph). */
#define NDG_GRAPH_OFF 0        /* direct launches on the mesh stream (default) */
#define NDG_GRAPH_SEQUENTIAL 1 /* one captured hipGraph per step; order groups one after another */
#define NDG_GRAPH_PARALLEL 2   /* one captured hipGraph per step; the order groups of a stage are
                                  independent graph branches that may run concurrently */

/* Initial-condition kinds. */
#define NDG_IC_PLANE_WAVE 0 /* the reference g(x, 0): include/functions/analytical_solution.tpp:5-11 */
#define NDG_IC_PULSES 1     /* sum of radial Gaussian pulses, u = v = 0 (no reference counterpart) */
#define NDG_IC_PERIODIC_PULSES 2 /* the same pulses periodised with periods (period_x, period_y): each
                                    pulse plus its 8 nearest images, so the field is periodic to
                                    within a exp(-ln2 min(period)^2 / w^2) (no reference counterpart) */

typedef struct ndg_tables ndg_tables;
typedef struct ndg_mesh ndg_mesh;

/* In-memory form of the mesh file the reference reads (CGNS unstructured zone,
 * src/meshes/Mesh2D_t.cpp:422-1048): nodes, QUAD_4 domain elements, BAR_2 boundary elements with
 * a boundary kind, and for periodic bars the donor bar (the CGNS Abutting1to1 point list donor). */
typedef struct ndg_mesh_desc {
    int64_t n_nodes;
    const double *xy;         /* 2*n_nodes: x0, y0, x1, y1, ... */
    int64_t n_quads;
    const int64_t *quads;     /* 4*n_quads node ids, counter-clockwise; storage order = element order */
    const int32_t *quad_N;    /* per-element order, or NULL for initial_N everywhere */
    int64_t n_bars;
    const int64_t *bars;      /* 2*n_bars node ids */
    const int32_t *bar_type;  /* NDG_BC_* per bar */
    const int64_t *bar_donor; /* per bar: donor bar index for NDG_BC_PERIODIC, else ignored (may be NULL) */
    int32_t initial_N;
    int32_t maximum_N;
    int32_t max_split_level;
    double tolerance_min;     /* estimator thresholds (reference defaults 1e-6 / 1e-14) */
    double tolerance_max;
} ndg_mesh_desc;

/* Parameters of the square-mesh generator, the equivalent of extras/mesh_generator/main.cpp
 * (--resolution, --x_min ... --y_max, boundary types, --x_periodic / --y_periodic) plus an
 * optional seeded vertex jitter (no reference counterpart). bc = {bottom, right, top, left}. */
typedef struct ndg_square_mesh_params {
    int32_t resolution;       /* power of two */
    double x_min, x_max, y_min, y_max;
    int32_t bc[4];
    int32_t x_periodic, y_periodic;
    double jitter;            /* interior vertex jitter as a fraction of the spacing (0 = none) */
    uint64_t seed;
} ndg_square_mesh_params;

typedef struct ndg_ic_desc {
    int32_t kind;             /* NDG_IC_* */
    int32_t n_pulses;
    const double *pulses;     /* n_pulses x {x0, y0, width, amplitude}: p = a exp(-ln2 r^2 / w^2) */
    double period_x, period_y; /* NDG_IC_PERIODIC_PULSES only */
} ndg_ic_desc;

/* ---- errors / device ---- */
/* Copies the last error message of this thread (always NUL-terminated). */
ndg_status ndg_last_error(char *buf, size_t buf_len);
/* Number of visible HIP devices (0 without a GPU; never fails on a CPU-only host). */
ndg_status ndg_device_count(int32_t *count);

/* ---- NDG_t: include/entities/NDG_t.h:11-26, include/entities/NDG_t.tpp:9-59 ---- */
ndg_status ndg_tables_create(int32_t N_max, int64_t n_interpolation_points, ndg_tables **out);
ndg_status ndg_tables_destroy(ndg_tables *tables);
/* Copies one table for order N into out (capacity doubles). *count receives its length. */
ndg_status ndg_tables_get(const ndg_tables *tables, int32_t which, int32_t N, double *out,
                          int64_t capacity, int64_t *count);

/* ---- mesh generator: extras/mesh_generator/main.cpp:161-382 ---- */
ndg_status ndg_square_mesh_sizes(const ndg_square_mesh_params *params, int64_t *n_nodes,
                                 int64_t *n_quads, int64_t *n_bars);
/* Fills caller-allocated arrays of the sizes given by ndg_square_mesh_sizes. Elements follow the
 * Hilbert curve (src/functions/Hilbert.cpp:16-28); local nodes are BR, TR, TL, BL. */
ndg_status ndg_square_mesh_fill(const ndg_square_mesh_params *params, double *xy, int64_t *quads,
                                int64_t *bars, int32_t *bar_type, int64_t *bar_donor);

/* ---- Hilbert-curve statuses: src/functions/Hilbert_splitting.cpp (statuses H, A, R, B = 0..3) ----
 * Hilbert::child_order / child_statuses (:23-81): the curve order of the four children of an
 * element with the given status and rotation (children numbered from the element's first corner,
 * counter-clockwise) and their statuses. Either output may be NULL. */
ndg_status ndg_hilbert_child_order(int32_t status, int32_t rotation, int32_t order[4], int32_t statuses[4]);
/* Hilbert::deduct_element_status (:13-21) for the sides through which the curve enters and leaves
 * an element; incoming_side = -1 is deduct_first_element_status (:3-6), outgoing_side = -1
 * deduct_last_element_status (:8-11). */
ndg_status ndg_hilbert_element_status(int32_t incoming_side, int32_t outgoing_side, int32_t *status);

/* ---- mesh files: Mesh2D_t::read_su2 (src/meshes/Mesh2D_t.cpp:78-330), the .su2 branch of the
 * Mesh2D_t constructor (:56-70). NDIME= 2, NPOIN= nodes, NELEM= QUAD_4 (type 9) elements, NMARK=
 * BAR_2 (type 3) markers: farfield -> NDG_BC_OUTFLOW, wall or airfoil -> NDG_BC_WALL, inlet ->
 * NDG_BC_INFLOW, symmetry -> NDG_BC_SYMMETRY. Errors (the reference's exits 69-75) return
 * NDG_ERR_INVALID with the reason in ndg_last_error. CGNS files are not read (no CGNS library). ---- */
ndg_status ndg_su2_sizes(const char *path, int64_t *n_nodes, int64_t *n_quads, int64_t *n_bars);
/* Fills caller-allocated arrays of the sizes given by ndg_su2_sizes; the bars come in the
 * reference's ghost order (walls, symmetry, inflow, outflow, each in file order). Any pointer may
 * be NULL. */
ndg_status ndg_su2_fill(const char *path, double *xy, int64_t *quads, int64_t *bars, int32_t *bar_type);
/* Periodic boundaries for SU2 meshes. The reference has none (its SU2 reader has no
 * self-interfaces, src/meshes/Mesh2D_t.cpp:398-400; periodic meshes come from CGNS). As SU2 itself
 * does in its configuration (MARKER_PERIODIC), pairs of marker tags name the boundaries joined by
 * a translation. Those tags (case-sensitive) are read as NDG_BC_PERIODIC bars, after the others,
 * pair after pair (all bars of marker_a[i], then all of marker_b[i], each in file order). Each
 * bar's donor (bar_donor, as ndg_mesh_desc takes it) is the bar of the other marker whose midpoint
 * is its own moved by the difference of the two markers' bar-midpoint centroids; a bar without
 * exactly one such partner (within 1e-8 of its length) is an error. n_pairs = 0 or a NULL pointer
 * reads as ndg_su2_sizes / ndg_su2_fill do. */
typedef struct ndg_su2_periodic {
    int32_t n_pairs;
    const char *const *marker_a;
    const char *const *marker_b;
} ndg_su2_periodic;
ndg_status ndg_su2_sizes_periodic(const char *path, const ndg_su2_periodic *periodic, int64_t *n_nodes,
                                  int64_t *n_quads, int64_t *n_bars);
ndg_status ndg_su2_fill_periodic(const char *path, const ndg_su2_periodic *periodic, double *xy, int64_t *quads,
                                 int64_t *bars, int32_t *bar_type, int64_t *bar_donor);

/* ---- Mesh2D_t: include/meshes/Mesh2D_t.h:21 (constructor), :105-128 (phase methods) ---- */
ndg_status ndg_mesh_create(const ndg_mesh_desc *desc, const ndg_tables *tables, ndg_mesh **out);
ndg_status ndg_mesh_destroy(ndg_mesh *mesh);
/* Orders all later work of this mesh on the given hipStream_t (NULL: the mesh's own stream). */
ndg_status ndg_mesh_set_stream(ndg_mesh *mesh, void *hip_stream);
/* n_dofs_per_field = sum over elements of (N+1)^2. */
ndg_status ndg_mesh_info(const ndg_mesh *mesh, int64_t *n_elements, int64_t *n_faces,
                         int64_t *n_dofs_per_field, int32_t *N_min, int32_t *N_max);
/* Selects the arithmetic form of later steps (NDG_PRECISION_*); the current state carries over.
 * No reference counterpart: the reference is FP64 only (include/helpers/float_types.h:4-5). */
ndg_status ndg_mesh_set_precision(ndg_mesh *mesh, int32_t precision);
/* Mesh2D_t::initial_conditions (src/meshes/Mesh2D_t.cpp:1192-1216); zeroes G and the RK register. */
ndg_status ndg_set_ic(ndg_mesh *mesh, const ndg_ic_desc *ic);
/* Mesh2D_t::get_solution's data without interpolation: packed p, u, v (host arrays). */
ndg_status ndg_set_state(ndg_mesh *mesh, const double *p, const double *u, const double *v);
ndg_status ndg_get_state(ndg_mesh *mesh, double *p, double *u, double *v);
/* Mesh2D_t::get_solution (src/meshes/Mesh2D_t.cpp:4237-4249, Element2D_t::interpolate_solution
 * src/entities/Element2D_t.cpp:581-604): the state interpolated to the tables' n x n uniform output
 * points of every element, with their coordinates. Host arrays of n_elements * n * n values,
 * element e's point (i, j) at e * n * n + i * n + j (caller's element order); a null array is
 * skipped. Requires n >= 2. */
ndg_status ndg_get_solution(ndg_mesh *mesh, double *x, double *y, double *p, double *u, double *v);
/* Output of Mesh2D_t::get_complete_solution (src/meshes/Mesh2D_t.cu:1861-1950, :5096-5155,
 * Element2D_t::interpolate_complete_solution src/entities/Element2D_t.cu:623-665). Point fields
 * hold n_elements * n * n values laid out as in ndg_get_solution; element fields hold n_elements
 * values in the caller's element order. Any pointer may be null (skipped). */
typedef struct ndg_complete_solution {
    double *x, *y, *p, *u, *v;
    double *dp_dt, *du_dt, *dv_dt;
    double *p_analytical_error, *u_analytical_error, *v_analytical_error;
    int32_t *N;
    double *p_error, *u_error, *v_error;
    double *p_sigma, *u_sigma, *v_sigma;
    int32_t *refine, *coarsen, *split_level, *status, *rotation;
} ndg_complete_solution;
/* Point fields: x, y, p, u, v as ndg_get_solution; dp_dt, du_dt, dv_dt the right-hand side of the
 * current state at `time` (the reference writes its last RK stage's G there, the right-hand side
 * of that stage's state); *_analytical_error the node-wise |g(x, time) - q| of the reference plane
 * wave (include/functions/analytical_solution.tpp:5-11, moving at the mesh's wave speed),
 * interpolated like the state. Element fields: N; the last ndg_estimate_error result (p_error,
 * p_sigma, refine, coarsen; zero before the first estimate and for elements created or raised by
 * an adaptation since); u/v error and sigma are 0 as in the reference, whose device estimator is
 * the pressure one (src/meshes/Mesh2D_t.cu:5158-5171); split level, Hilbert status and rotation. */
ndg_status ndg_get_complete_solution(ndg_mesh *mesh, double time, const ndg_complete_solution *out);
/* Element traces (Mesh2D_t::interpolate_to_boundaries, src/meshes/Mesh2D_t.cpp:3794-3801) of the
 * current state, packed [element][side][field][N+1]. */
ndg_status ndg_get_traces(ndg_mesh *mesh, double *out);
/* One stage's right-hand side G at time t (phases interpolate_to_boundaries .. compute_dg_wave_
 * derivative, include/solvers/Solver2D_t.tpp:46-51), packed; the state is not advanced, and the
 * stage clock a running ndg_solve / ndg_solve_output continues from is left as it was. */
ndg_status ndg_compute_rhs(ndg_mesh *mesh, double t, double *Gp, double *Gu, double *Gv);
/* Solver2D_t::get_delta_t (src/solvers/Solver2D_t.cpp:19-29): a device reduction. */
ndg_status ndg_get_delta_t(ndg_mesh *mesh, double cfl, double *dt);
/* One RK3 step (3 stages, include/solvers/Solver2D_t.tpp:44-70) from time t with step dt.
 * Stream-ordered; returns without waiting for the GPU. */
ndg_status ndg_step(ndg_mesh *mesh, double t, double dt);
/* The reference time loop Solver2D_t::solve (include/solvers/Solver2D_t.tpp:37-100) without
 * output: dt from get_delta_t, the last step clipped to t_end, time accumulated on the host exactly
 * as the reference does, and the mesh's adaptivity (ndg_mesh_set_adaptivity). Waits for
 * completion. */
ndg_status ndg_solve(ndg_mesh *mesh, double cfl, double t_end, double *t_final, int64_t *n_steps);
/* Solver2D_t::solve with its output events (include/solvers/Solver2D_t.tpp:21-29, :74-83,
 * :112-127): runs to output_times[n_times - 1] as ndg_solve does and calls fn(mesh, time, user)
 * where the reference calls write_complete_data: before the first step for every output time t_o
 * with 0 in [t_o, t_o + dt); after a step for the first t_o with time in [t_o, t_o + dt); at the
 * end if the last step wrote nothing. Output times do not shorten steps. The callback may read
 * the mesh (ndg_get_state, ndg_get_solution, ndg_estimate_error, ...); a status other than NDG_OK
 * from it stops the solve with that status. No reference counterpart for the callback itself: the
 * reference writes VTK files there (DataWriter_t). */
typedef ndg_status (*ndg_output_fn)(ndg_mesh *mesh, double time, void *user);
ndg_status ndg_solve_output(ndg_mesh *mesh, double cfl, const double *output_times, int64_t n_times,
                            ndg_output_fn fn, void *user, double *t_final, int64_t *n_steps);
/* n_steps steps of the same loop without clipping, from *time_io (updated). Stream-ordered. */
ndg_status ndg_run_steps(ndg_mesh *mesh, double cfl, int64_t n_steps, double *time_io);
/* Solver2D_t::pre_condition (iterative = 0, include/solvers/Solver2D_t.tpp:138-236) and
 * pre_condition_iterative (iterative = 1, :238-336): from time 0, steps of the CFL size with an
 * estimate + adapt (the tables' N_max, the mesh's max_split_level) every pre_condition_interval
 * steps, until n_adaptivity_steps adaptations are done; the initial condition last given to
 * ndg_set_ic is then applied to the adapted mesh (iterative: after every adaptation, with the
 * time set back to 0). The reference exits where this returns NDG_ERR_INVALID (interval 0; here
 * also no initial condition set). Waits for completion. */
ndg_status ndg_pre_condition(ndg_mesh *mesh, double cfl, int64_t n_adaptivity_steps, int64_t pre_condition_interval,
                             int32_t iterative);
/* Mesh2D_t::adapt (src/meshes/Mesh2D_t.cpp:1644-2268): estimates the error, then
 * - when some element would h-refine (refine && sigma < 1 && split level < max_split_level,
 *   Element2D_t.cpp:802-812), the hp_adapt branch (:5186-5659): each such element becomes four
 *   children in Hilbert order (child order and statuses, Hilbert_splitting.cpp:23-81) carrying the
 *   parent polynomial (inverse bilinear map + interpolate_from), ghosts on split sides become two
 *   (split_boundaries, :6232-6343), the would-p-refine elements are raised by 2, and 2:1 hanging
 *   faces appear, also across periodic sides (split_interfaces, :6343-6756: each end of the
 *   period gets its own faces against the other end's trace), with any number of faces on a side
 *   down to pieces of 1/16 of it; a split that would leave a smaller piece returns
 *   NDG_ERR_UNSUPPORTED and changes nothing (in either NDG_ADAPT_* mode);
 * - otherwise the p_adapt branch (:5026-5049): every element with refine && sigma >= 1 &&
 *   N + 2 <= N_max is raised by 2 (Element2D_t.cpp:464-579); ghosts and faces follow
 *   (adjust_boundaries / adjust_interfaces / adjust_faces, :7482-7538), including the reference's
 *   zero edge scaling factor of a resized ghost.
 * The element count can change (ndg_mesh_info). *n_raised (may be NULL) gets the number of split
 * plus raised elements. */
ndg_status ndg_adapt(ndg_mesh *mesh, int32_t N_max, int32_t max_split_level, int64_t *n_raised);
/* SEM::inverse_quad_map (include/functions/inverse_quad_map.tpp:46-88): the reference coordinates
 * (xi, eta) of n physical points xy (x, y pairs) in the bilinear quad of the 4 vertices (x, y pairs,
 * the element's local node order), from the closed-form root of the map's quadratic. The
 * adaptation's solution transfer maps every child node into its parent with this function. */
ndg_status ndg_inverse_quad_map(const double *vertices, int64_t n, const double *xy, double *xi_eta);
/* The order of every element (caller's order); it changes with ndg_adapt. */
ndg_status ndg_mesh_orders(const ndg_mesh *mesh, int32_t *orders);
/* Where ndg_adapt runs. NDG_ADAPT_DEVICE (the default) keeps the element-side topology in device
 * memory and rebuilds it there with the decisions, the children, the links, the layout and the
 * transfer, reading back only two small blocks of counts per adaptation, as the reference's device
 * adaptation does (src/meshes/Mesh2D_t.cu:6206-7232, :7498, :8832-8902). NDG_ADAPT_HOST rebuilds
 * the topology on the host from the node lists (the reference's host adapt,
 * src/meshes/Mesh2D_t.cpp:1644-2268); it is refused once a device adaptation has run. Both give the
 * same meshes and states bit for bit. */
#define NDG_ADAPT_DEVICE 0
#define NDG_ADAPT_HOST 1
ndg_status ndg_mesh_set_adapt_mode(ndg_mesh *mesh, int32_t mode);
/* The element-side topology (Element2D_t::faces_ and Face2D_t::elements_,
 * include/entities/Element2D_t.h:12-127, include/entities/Face2D_t.h:18-83, as links between
 * element sides), 4 values per element in the caller's order; any output may be NULL.
 * link[4e+s]: the element side 4r+rs whose trace is the exterior state of side s of element e (a
 * neighbour, or the donor side of a periodic pair), or -(1 + ghost). hang[4e+s]: 0 conforming, 1 a
 * small side covering a piece of the larger side link, 2 a larger side covered by several small
 * sides, the first of them (at its first corner) being link. For a small side, span[4e+s] is the
 * piece [j / 2^d, (j + 1) / 2^d] of the larger side it covers, in the larger side's direction, as
 * 2^d - 1 + j (1, 2: the halves; 3..6: the quarters; d <= 4), and next_small[4e+s] the next small
 * side along the larger side (-1 after the last, and on other sides). periodic[4e+s]: 1 for a
 * periodic pair. */
ndg_status ndg_mesh_sides(ndg_mesh *mesh, int64_t *link, int8_t *hang, int64_t *next_small, int8_t *periodic,
                          int8_t *span);
/* Mesh2D_t's adaptivity_interval: ndg_solve then adapts every that many steps with the tables'
 * N_max and the description's max_split_level (include/solvers/Solver2D_t.tpp:88-96). 0 = off. */
ndg_status ndg_mesh_set_adaptivity(ndg_mesh *mesh, int64_t adaptivity_interval);
/* Waits for all work queued on the mesh's stream. */
ndg_status ndg_synchronize(ndg_mesh *mesh);
/* Replaces Solver2D_t::solve's per-phase launches (include/solvers/Solver2D_t.tpp:44-70, one
 * launch per phase and per element kernel) by a captured hipGraph per RK step (mortar faces, then
 * each order group's fused stage kernel, three times). The stage times are read from a device
 * clock that the graph advances, so one graph serves every step of size dt; it is recaptured when
 * dt or the precision changes. No reference counterpart for the mode switch itself. */
ndg_status ndg_mesh_set_graph(ndg_mesh *mesh, int32_t mode);
/* Times every candidate launch shape of each order group's stage kernel (elements per workgroup
 * from the shape's maximum down to 1, XCD-contiguous block order on or off) on this mesh, `repeats`
 * launches each, and keeps the fastest for the current precision. The state is left unchanged.
 * No reference counterpart (the reference launches every element kernel with 32 threads per
 * block, include/solvers/Solver2D_t.tcu:45-52). */
ndg_status ndg_mesh_autotune(ndg_mesh *mesh, int32_t repeats);
/* The launch shape in use per order group (ascending N): orders, elements per workgroup, block
 * order flag. Any output may be NULL; *n_groups receives the group count. */
ndg_status ndg_mesh_tuning(const ndg_mesh *mesh, int32_t *orders, int32_t *epb, int32_t *remap,
                           int32_t capacity, int32_t *n_groups);
/* Mesh2D_t::estimate_error -> Element2D_t::estimate_p_error (src/entities/Element2D_t.cpp:274-343),
 * pressure estimator per element. Any output pointer may be NULL. */
ndg_status ndg_estimate_error(ndg_mesh *mesh, double *error, double *sigma, int32_t *refine,
                              int32_t *coarsen);

/* ---------------------------------------------------------------------------------------------
 * Wave speed and the discrete adjoint (no reference counterpart; the reference's c is the
 * compile-time constant 1, include/helpers/constants.h:15).
 *
 * With wave speed c the system is p_t + c div(u, v) = 0, (u, v)_t + c grad p = 0: the right-hand
 * side is c times the reference operator, the CFL step scales by 1/c, and the plane wave moves
 * with speed c. c = 1 is the reference arithmetic bit for bit. */
ndg_status ndg_mesh_set_wave_speed(ndg_mesh *mesh, double c);

/* Pressure receivers: element (caller's order) and reference coordinates (xi, eta) in [-1, 1]^2
 * of each point; p there is the element's Lagrange interpolant. */
typedef struct ndg_receivers {
    int32_t n;
    const int64_t *element;
    const double *xi_eta; /* [n][2] */
} ndg_receivers;

/* Advances the mesh n_steps RK3 steps of size dt and records p at the receivers after every step:
 * p_out[step][receiver]. Single-order meshes with interior/periodic sides, FP64. */
ndg_status ndg_record_receivers(ndg_mesh *mesh, const ndg_receivers *rec, double dt, int64_t n_steps,
                                double *p_out);

/* ndg_record_receivers without any device allocation (no reference counterpart): for callers
 * that cannot allocate device memory (serving loops that must not call hipMalloc, harnesses that
 * forbid it). The receivers' elements and weights and the records of every step live in a fixed
 * per-GPU scratch array of 4 MiB held by the library until the process ends: 8 * (n + n (N+1)^2 +
 * n n_steps) bytes must fit, else NDG_ERR_INVALID and nothing runs. Steps are launched directly
 * (no captured graph). Same results as ndg_record_receivers, bit for bit. Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (other threads, other meshes' streams)
 * race and must not be made. */
ndg_status ndg_record_receivers_noalloc(ndg_mesh *mesh, const ndg_receivers *rec, double dt, int64_t n_steps,
                                        double *p_out);

/* Misfit J = 1/2 sum_steps sum_receivers (p_r(t_n) - data[n][r])^2 over n_steps steps of size dt
 * (frozen) from the mesh's current state, and its gradient with respect to the initial fields
 * (packed like ndg_get_state) and to c, by the discrete adjoint of the RK3 scheme: a forward pass
 * keeping a checkpoint every checkpoint_interval steps, then per segment a recompute from its
 * checkpoint that keeps the stage states (on a second stream, beside the reverse sweep of the
 * following segment), and the transposed stages backwards. The mesh state is the initial one
 * again on return. data may be NULL (zero); any output may be NULL. The checkpoints and
 * stage-store blocks stay allocated with the mesh for the next call (grown when a call needs
 * more, released by ndg_mesh_destroy). */
ndg_status ndg_adjoint_gradient(ndg_mesh *mesh, const ndg_receivers *rec, double dt, int64_t n_steps,
                                int64_t checkpoint_interval, const double *data, double *J, double *dJ_dp0,
                                double *dJ_du0, double *dJ_dv0, double *dJ_dc);

/* How ndg_adjoint_gradient obtains dJ/dc (no reference counterpart):
 * NDG_ADJOINT_C_RECOMPUTE (default): checkpoints, per-segment recompute of the stage states, and
 *   <G^T mu~, q_s> / c gathered by the transposed stages;
 * NDG_ADJOINT_C_STAGE_RECORDS: from the forward pass alone. The operator is linear and
 *   time-invariant, so one step is q_n = P(z) q_{n-1} with z = dt G = c dt G1 and P the RK3
 *   stability polynomial, dq_n/dc = (n / c) (z + z^2 + z^3 / 2) q_{n-1}, and the three stage
 *   right-hand sides of step n (recorded at the receivers) give z q, z^2 q and z^3 q. The
 *   transposed stages then need no forward state: no checkpoints, no recompute, no stage store
 *   (checkpoint_interval and the store budget are not used). dJ/dp0, du0, dv0 are the same bit for
 *   bit; dJ/dc agrees to rounding. */
#define NDG_ADJOINT_C_RECOMPUTE 0
#define NDG_ADJOINT_C_STAGE_RECORDS 1
ndg_status ndg_adjoint_set_c_derivative(ndg_mesh *mesh, int32_t mode);

/* Device memory (bytes) ndg_adjoint_gradient may spend on stage states beyond one segment (a
 * segment is checkpoint_interval steps, 3 states per step): the forward pass keeps the stage states
 * of the last 1 + floor(bytes / segment bytes) segments as it goes, and the reverse sweep recomputes
 * only the segments before them. 0 (the default) keeps the last segment only: every other segment
 * is recomputed from its checkpoint. The result is the same bit for bit. No reference counterpart. */
ndg_status ndg_adjoint_set_store_budget(ndg_mesh *mesh, int64_t bytes);

/* lambda = G^T mu for the stage right-hand side G (ndg_compute_rhs without boundary terms), packed
 * fields; for dot-product tests of the adjoint. */
ndg_status ndg_apply_rhs_transpose(ndg_mesh *mesh, const double *mp, const double *mu, const double *mv,
                                   double *lp, double *lu, double *lv);

#ifdef __cplusplus
}
#endif
#endif
