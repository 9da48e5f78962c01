/*
 * ndg_oracle.h -- CPU restatement of the reference DG-SEM wave solver (test infrastructure).
 *
 * THIS IS THE ORACLE. It is test infrastructure only: the tests, the smoke check and the
 * CPU-baseline leg of the benchmark load it as the checker. The product path (the HIP library
 * under ndg_cuda_amd/) never links, loads or calls it.
 *
 * It restates, in plain C, the host (CPU) algorithm of guillaumetousignant/NDG_cuda for one
 * low-storage RK3 time step of the first-order 2D wave system (p, u, v) on a quad mesh with
 * ghost boundary elements and periodic self-interfaces. Each function cites the reference
 * file:line it follows. The data model mirrors the reference's (elements with per-side face
 * lists, faces with left/right states, ghost elements holding one trace), so that its
 * arithmetic follows the reference's operation order.
 *
 * Parity pinning of each function is described where the function is defined (ndg_oracle.c).
 */
#ifndef NDG_ORACLE_H
#define NDG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Table selectors for o_tables_get (same tables as the reference NDG_t, include/entities/NDG_t.h:11-26). */
enum {
    O_NODES = 0, O_WEIGHTS = 1, O_BARY = 2, O_LAG_LEFT = 3, O_LAG_RIGHT = 4,
    O_DLAG_LEFT = 5, O_DLAG_RIGHT = 6, O_D = 7, O_GHAT = 8, O_DHAT = 9,
    O_INTERP = 10, O_POLY = 11
};

/* Boundary (ghost) element kinds, as the CGNS BC types the reference reads (src/meshes/Mesh2D_t.cpp:689-791). */
enum { O_BC_NONE = 0, O_BC_WALL = 1, O_BC_SYMMETRY = 2, O_BC_INFLOW = 3, O_BC_OUTFLOW = 4, O_BC_PERIODIC = 5 };

typedef struct o_tables o_tables;
typedef struct o_mesh o_mesh;

o_tables *o_tables_create(int N_max, int n_interp);
void o_tables_destroy(o_tables *t);
/* Copies table `which` for order N into out (Nf, Nf*Nf or n_interp*Nf doubles). Returns count or -1. */
int o_tables_get(const o_tables *t, int which, int N, double *out);

/* Builds the reference mesh model from an in-memory description equivalent to the CGNS file the
 * reference reads: node coordinates, QUAD_4 domain elements (4 node ids, counter-clockwise,
 * element order = storage order), BAR_2 boundary elements (2 node ids) with a BC kind each, and
 * for periodic bars the index of the donor bar. quad_N may be NULL (uniform initial_N). */
o_mesh *o_mesh_create(const o_tables *t, int initial_N,
                      int64_t n_nodes, const double *xy,
                      int64_t n_quads, const int64_t *quads, const int32_t *quad_N,
                      int64_t n_bars, const int64_t *bars, const int32_t *bar_type,
                      const int64_t *bar_donor, double tol_min, double tol_max);
void o_mesh_destroy(o_mesh *m);
int64_t o_mesh_n_elements(const o_mesh *m);
int64_t o_mesh_n_faces(const o_mesh *m);
/* faces per element side [K][4] */
void o_mesh_side_faces(const o_mesh *m, int32_t *nf);
int64_t o_mesh_n_ghosts(const o_mesh *m);
/* Sum over domain elements of (N+1)^2: the length of one packed field. */
int64_t o_mesh_n_nodes_total(const o_mesh *m);
/* Order of each domain element. */
void o_mesh_get_N(const o_mesh *m, int32_t *N_out);

/* Initial conditions. Plane wave: the reference g(x, 0) (src/meshes/Mesh2D_t.cpp:1192-1216).
 * Pulses: p = sum_k a_k exp(-ln2 ((x-x_k)^2 + (y-y_k)^2) / w_k^2), u = v = 0 (no reference
 * counterpart; pulses = n x {x_k, y_k, w_k, a_k}). Both zero G and the RK register. */
void o_initial_conditions_plane(o_mesh *m);
void o_initial_conditions_pulses(o_mesh *m, int n, const double *pulses);
void o_initial_conditions_periodic_pulses(o_mesh *m, int n, const double *pulses, double Lx, double Ly);

/* Packed per-field state: element by element in mesh order, each element's (N+1)^2 values in
 * the reference node order q[i*(N+1)+j]. */
void o_set_state(o_mesh *m, const double *p, const double *u, const double *v);
void o_get_state(const o_mesh *m, double *p, double *u, double *v);
void o_get_rhs(const o_mesh *m, double *Gp, double *Gu, double *Gv);
/* Element traces p,u,v on the 4 sides, packed [element][side][field][Nf]. */
/* Mesh2D_t::get_solution: x, y, p, u, v at the tables' n x n uniform output points per element,
 * element k's point (i, j) at k*n*n + i*n + j. */
void o_get_solution(const o_mesh *m, double *x, double *y, double *p, double *u, double *v);
void o_get_complete_solution(const o_mesh *m, double time, double *out);
void o_get_traces(const o_mesh *m, double *out);
/* Element flux traces, packed [element][side][field][Nf]. */
void o_get_flux_traces(const o_mesh *m, double *out);

double o_get_delta_t(const o_mesh *m, double cfl);
/* Phases 1-6 of one RK stage at time t: traces, boundary conditions, face projection, fluxes,
 * element projection, volume + lift -> G. */
void o_compute_rhs(o_mesh *m, double t);
/* Wave speed c of the symmetric system (p_t + c div u = 0, u_t + c grad p = 0): scales the
 * right-hand side by c and the CFL step by 1/c; the reference is c = 1 (no reference counterpart
 * for other values). */
void o_mesh_set_wave_speed(o_mesh *m, double c);
/* One full RK3 step (3 stages) from time t with step dt. */
void o_step(o_mesh *m, double t, double dt);
/* The reference time loop (include/solvers/Solver2D_t.tpp:37-72) without output or adaptation.
 * Returns the number of steps taken; *t_final receives the accumulated time. */
int64_t o_solve(o_mesh *m, double cfl, double t_end, double *t_final);
/* n_steps steps of the same loop with no end-time clipping. */
void o_run_steps(o_mesh *m, double cfl, int64_t n_steps, double *time_io);

/* Pressure estimator per domain element (src/entities/Element2D_t.cpp:274-343). */
void o_estimate_error(o_mesh *m, double *error, double *sigma, int32_t *refine, int32_t *coarsen);
/* One adaptation (src/meshes/Mesh2D_t.cpp:1644-2268) from the last estimate. Without a split
 * request (p_adapt, :5026-5049): elements with refine && sigma >= 1 && N + 2 <= N_max are raised by
 * 2 orders (solution interpolated, geometry recomputed); ghosts and faces follow (a resized ghost's
 * storage, including its edge scaling factor, becomes zero). With one (hp_adapt): elements with
 * refine && sigma < 1 && split_level < max_split_level split into 4 children in Hilbert order, the
 * others p-refine as above. Returns the number of changed elements, or -2 (nothing changed) for a
 * split this restatement does not cover (a periodic side, or a side that would get more than two
 * faces). */
int64_t o_adapt(o_mesh *m, int N_max, int max_split_level);
/* Element nodes (4 per element), Hilbert status and rotation (may be NULL), node coordinates. */
void o_mesh_get_elements(const o_mesh *m, int64_t *nodes, int32_t *status, int32_t *rotation, double *xy);
int64_t o_mesh_n_nodes(const o_mesh *m);
/* Sets the estimator's refine flags and decay rates directly (tests of the adaptation rules). */
void o_mesh_set_flags(o_mesh *m, const int32_t *refine, const double *sigma);
/* The reference time loop with an adaptation every adapt_interval steps; -1 if one is not covered. */
int64_t o_solve_adaptive(o_mesh *m, double cfl, double t_end, int64_t adapt_interval, int N_max,
                         int max_split_level, double *t_final);

/* ---- element / face level entry points (for known-answer and fixture checks) ---- */
/* traces of one element: q = p,u,v each Nf*Nf; out = [4][3][Nf] */
void o_element_traces(const o_tables *t, int N, const double *p, const double *u,
                      const double *v, double *out);
/* geometry of one element from its 4 vertices (x0,y0,...,x3,y3): out = dxi_dx, deta_dx, dxi_dy,
 * deta_dy, jacobian (5 x Nf*Nf), scaling factors (4 x Nf), delta_xy_min, cx, cy. */
void o_element_geometry(const o_tables *t, int N, const double *pts, double *out);
/* geometry of a ghost element from its 4 vertices: out = sf0 (Nf), delta_xy_min, cx, cy */
void o_boundary_geometry(const o_tables *t, int N, const double *pts, double *out);
/* face geometry: face_nodes (4), element_nodes (8: L side start/end, R side start/end);
 * out = nx, ny, tx, ty, length, offset0, offset1, scale0, scale1 */
void o_face_geometry(const double *face_nodes, const double *element_nodes, double *out);
/* one element's estimator: out = error, sigma, refine, coarsen, spectrum[0..n-1] */
void o_element_estimate(const o_tables *t, int N, const double *p, double tol_min,
                        double tol_max, double *out);
/* upwind flux at one point: in = pL,uL,vL,pR,uR,vR,nx,ny, out = Fp,Fu,Fv */
void o_wave_flux(const double *in, double *out);
/* Threads of the stage loops (OpenMP, bitwise the same result for any count); 0 = default. */
int o_set_threads(int n);
void o_hilbert_d2xy(int n, int d, int *xy);
/* Hilbert_splitting.cpp tables: child_order / child_status [rotation][status][4] (64 each),
 * first / last status [side] (4 each), inner status [incoming][outgoing] (16). */
void o_hilbert_tables(int *child_order, int *child_status, int *first, int *last, int *inner);
int o_hilbert_xy2d(int n, int x, int y);


/* Known-answer-test entry points: one project_to_faces side (trace of order Ne -> Nf face nodes),
 * one project_to_elements face (flux of order Nf accumulated into an order-Ne side, edge scale sf),
 * Element2D_t::interpolate_from (same geometry when Po or Pn is NULL, else a child element Pn
 * inside the parent Po, vertices as 4 x (x, y)), and inverse_quad_map. */
void o_kat_face_values(const o_tables *t, int Ne, int Nf, int side, double offset, double scale,
                       const double *trace, double *out);
void o_kat_side_flux(const o_tables *t, int Ne, int Nf, int fwd, double offset, double scale, double sf,
                     const double *flux, double *out);
void o_kat_interpolate(const o_tables *t, int No, const double *Po, int N, const double *Pn, const double *q_old,
                       double *q_new);
void o_kat_inverse_quad_map(const double *P, double x, double y, double *xi_eta);

#ifdef __cplusplus
}
#endif
#endif
