// kernels.hip -- CDNA4 (gfx950) kernels of one low-storage RK3 stage of the DG-SEM wave solver.
//
// Data layout in HBM (structure of arrays, element-major, elements in Hilbert-curve order):
//   q      [K][3][(N+1)^2]   state (p, u, v), reference node order i*(N+1)+j
//   g      [K][3][(N+1)^2]   RK register, alive within a step only; the stage kernel's own order
//                            within each workgroup tile (see stage_kernel)
//   trace  [K][4][3][N+1]    element traces of the current state (two buffers, ping-pong)
//   verts  [K][4][2]         element vertices (metrics are recomputed from them per node)
//   side_nb[K][4]            neighbour element*4+side, or -(1+index) into the BcSide records
//
// One fused stage kernel replaces the reference's per-phase kernels of one stage
// (include/solvers/Solver2D_t.tcu:45-52): a workgroup stages the (N+1)^2 tiles of EPB elements in
// LDS, computes the upwind flux on its 4 sides from its own traces (recomputed from the LDS tile)
// and the neighbours' traces (one HBM read), applies volume + lift, the RK update, and writes the
// traces of the updated state for the next stage. Per element and stage HBM traffic is the state
// and RK register (read + write), one read of the neighbours' traces and one write of its own.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <cmath>

#include "device.hpp"
#include "internal.hpp"

namespace ndg {

template <int N>
struct Shape {
    static constexpr int NF = N + 1;
    static constexpr int NE = NF * NF;
    static constexpr int EPB = NE >= 256 ? 1 : 256 / NE;  // elements per workgroup
    static constexpr int THREADS = ((EPB * NE + 63) / 64) * 64;
};

// The reference plane wave g(x, t) (include/functions/analytical_solution.tpp:5-11).
__device__ __forceinline__ void plane_wave(double x, double y, double t, double cw, double &p, double &u,
                                           double &v) {
    const double kx = 0.707106781186548, ky = 0.707106781186548, x0 = -0.2, y0 = -0.2;
    const double d = 0.120112240878645;
    const double a = kx * (x - x0) + ky * (y - y0) - cw * t;
    p = exp(-a * a / (d * d));
    u = p * kx;  // the reference's p k / c at c = 1 (and the plane wave of the speed-c system)
    v = p * ky;
}

// Trace k of side s of one element tile (src/entities/Element2D_t.cpp:86-125): sides 2 and 3
// run counter-clockwise, i.e. reversed with respect to the xi / eta index.
template <int N>
__device__ __forceinline__ double side_trace(const double *tile, const double *lm, const double *lp,
                                             int s, int k) {
    constexpr int NF = N + 1;
    double acc = 0.0;
    if (s == 0) {
#pragma unroll
        for (int j = 0; j < NF; ++j) acc += lm[j] * tile[k * NF + j];
    } else if (s == 1) {
#pragma unroll
        for (int j = 0; j < NF; ++j) acc += lp[j] * tile[j * NF + k];
    } else if (s == 2) {
#pragma unroll
        for (int j = 0; j < NF; ++j) acc += lp[j] * tile[(N - k) * NF + j];
    } else {
#pragma unroll
        for (int j = 0; j < NF; ++j) acc += lm[j] * tile[j * NF + (N - k)];
    }
    return acc;
}

// ---------------------------------------------------------------------------------------------
// Traces of the current state (used after the state is set; the stage kernel writes them
// afterwards).
template <int N>
__global__ void __launch_bounds__(Shape<N>::THREADS)
traces_kernel(const double *__restrict__ tab, const double *__restrict__ q,
              double *__restrict__ tr, int64_t K) {
    using S = Shape<N>;
    constexpr int NF = S::NF, NE = S::NE, EPB = S::EPB, T = S::THREADS;
    __shared__ double s_lm[NF], s_lp[NF];
    __shared__ double s_q[EPB * 3 * NE];
    const int tid = threadIdx.x;
    const int64_t e0 = (int64_t)blockIdx.x * EPB;
    const int nel = (int)min<int64_t>(EPB, K - e0);
    for (int r = tid; r < NF; r += T) {
        s_lm[r] = tab[NF * NF + r];
        s_lp[r] = tab[NF * NF + NF + r];
    }
    const double *qg = q + e0 * 3 * NE;
    for (int r = tid; r < nel * 3 * NE; r += T) s_q[r] = qg[r];
    __syncthreads();
    for (int r = tid; r < nel * 12 * NF; r += T) {
        const int el = r / (12 * NF), rs = r % (12 * NF);
        const int s = rs / (3 * NF), f = (rs / NF) % 3, k = rs % NF;
        tr[(e0 + el) * 12 * NF + rs] = side_trace<N>(s_q + el * 3 * NE + f * NE, s_lm, s_lp, s, k);
    }
}

// ---------------------------------------------------------------------------------------------
// One fused RK stage. A workgroup is three waves, one per field (p, u, v); lane (element e, line l)
// of the wave for field f handles one (N+1)-point line of that field for EPB = floor(64/(N+1))
// elements. Every field-dependent branch is therefore wave-uniform (a scalar branch, no masked
// lanes), the D^ contractions run in registers with the D^ rows as wave-uniform scalar loads, and
// LDS carries the element tiles:
//   P0  every HBM read of the tile is issued before any is waited on: state tile -> LDS, the
//       thread's RK-register line and the neighbour traces of its flux points -> registers
//   P1  the element's own traces (row l and column l of field f); wave 0 the metric lines along
//       xi, wave 1 along eta, wave 2 the side normals and edge Jacobians
//   P2  upwind flux at the 4(N+1) side points, times the edge Jacobian
//       (boundary_conditions / project_to_faces / calculate_wave_fluxes / project_to_elements,
//       Mesh2D_t.cpp:3731-3935, Solver2D_t.cpp:48-63)
//   P3  xi lines: D^ (y_eta f - x_eta g) + lift (Solver2D_t.cpp:71-100)
//   P4  eta lines: D^ (-y_xi f + x_xi g) + lift, G = (-d_xi - d_eta)/J (:104-133), RK update
//       (Solver2D_t.cpp:138-172)
//   P5  coalesced stores of q and g, traces of the updated state for the next stage.
// RHS_ONLY writes G (the compute_dg_wave_derivative output) and leaves the state alone. rk_flags:
// kRkFirst, the step's first stage (a = 0: the RK register is not read); kRkLast, the step's last
// stage, whose RK register no one reads (the next step's first stage overwrites it), so it is not
// stored.
// The new state goes to q_out, which is q for the in-place step (q and q_out are therefore not
// restrict; every value of q a workgroup reads precedes, through barriers, its writes).
// A workgroup takes epb <= EPB consecutive elements (fewer gives more workgroups for a small
// order group); remap != 0 applies the XCD-contiguous block order. t_ptr points at the stage time
// in device memory (the inflow ghost reads it), so that a captured step can be replayed.


template <int N, bool RHS_ONLY>
__global__ void __launch_bounds__(192)
stage_kernel(const double *__restrict__ verts, const int32_t *__restrict__ side_nb,
             const BcSide *__restrict__ bc, const double *__restrict__ tab,
             const double *q, double *q_out, double *__restrict__ g, const double *__restrict__ tr_in,
             double *__restrict__ tr_out, double *__restrict__ rhs, const double *__restrict__ side_flux,
             int64_t K, int rk_flags, double a, double gdt, const double *__restrict__ t_ptr, int epb,
             int remap, double cw) {
    const bool first = rk_flags & kRkFirst, store_g = !(rk_flags & kRkLast);
    using S = LineShape<N>;
    constexpr int NF = S::NF, NE = S::NE, EPB = S::EPB, T = S::T, FPT = S::FPT;
    constexpr double c = kWaveSpeed;
    __shared__ double s_q[EPB * 3 * NE];
    __shared__ double s_d[EPB * 3 * NE];
    __shared__ double s_tr[EPB * 12 * NF];
    __shared__ double s_met[EPB * 4 * NF];
    __shared__ double s_side[EPB * 4 * 3];  // outward normal and edge Jacobian per side
    __shared__ double s_v[EPB * 8];
    const double *Dh = tab;
    const double *lm = tab + NF * NF;
    const double *lp = lm + NF;
    const double *x = lp + 2 * NF;
    const double *winv = x + NF;

    const int tid = threadIdx.x;
    const int f = __builtin_amdgcn_readfirstlane(tid >> 6);  // field of this wave
    const int lane = tid & 63;
    const int el = lane / NF, l = lane % NF;
    const int64_t blk = remap ? xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const int64_t e0 = blk * epb;
    const int nel = (int)min<int64_t>(epb, K - e0);
    const bool active = el < nel;

    __builtin_amdgcn_s_setprio(3);  // issue this workgroup's loads ahead of other waves' math
    // P0
    int32_t code[FPT];
#pragma unroll
    for (int it = 0; it < FPT; ++it) {
        const int item = tid + it * T;
        code[it] = item < nel * 4 * NF ? side_nb[(e0 + item / NF / 4) * 4 + (item / NF) % 4] : INT32_MIN;
    }
    double qreg[S::QPT];
    const double *qg = q + e0 * 3 * NE;
#pragma unroll
    for (int it = 0; it < S::QPT; ++it) {
        const int r = tid + it * T;
        qreg[it] = r < nel * 3 * NE ? qg[r] : 0.0;
    }
    double vreg[S::VPT];
#pragma unroll
    for (int it = 0; it < S::VPT; ++it) {
        const int r = tid + it * T;
        vreg[it] = r < nel * 8 ? verts[e0 * 8 + r] : 0.0;
    }
    // The RK register lives only within a step (the first stage writes it without reading it), so
    // its layout is the kernel's own: within the workgroup's tile of nel elements it is stored as
    // [f][j][el][i], so that the row owners of all the tile's elements read and write one line
    // index j with consecutive lanes on consecutive addresses (one contiguous run per wave
    // instruction instead of one short run per element).
    double g_old[NF];
    if (!RHS_ONLY && !first && active) {
        const double *gg = g + e0 * 3 * NE + ((int64_t)f * NF * nel + el) * NF + l;
#pragma unroll
        for (int j = 0; j < NF; ++j) g_old[j] = gg[j * nel * NF];
    } else {
#pragma unroll
        for (int j = 0; j < NF; ++j) g_old[j] = 0.0;
    }
    double nbv[FPT][3];
#pragma unroll
    for (int it = 0; it < FPT; ++it) {
        nbv[it][0] = nbv[it][1] = nbv[it][2] = 0.0;
        const int32_t cd = code[it];
        if (cd >= 0) {
            const int k = (tid + it * T) % NF;
            const double *nb = tr_in + (int64_t)(cd >> 2) * 12 * NF + (cd & 3) * 3 * NF + (N - k);
            nbv[it][0] = nb[0];
            nbv[it][1] = nb[NF];
            nbv[it][2] = nb[2 * NF];
        }
    }
    __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int it = 0; it < S::VPT; ++it) {
        const int r = tid + it * T;
        if (r < nel * 8) s_v[r] = vreg[it];
    }
#pragma unroll
    for (int it = 0; it < S::QPT; ++it) {
        const int r = tid + it * T;
        if (r < nel * 3 * NE) s_q[r] = qreg[it];
    }
    __syncthreads();

    // P1
    if (active) {
        const double *P = s_v + el * 8;
        double *met = s_met + el * 4 * NF;
        if (f == 0) {  // x_eta, y_eta depend on xi only (include/functions/quad_metrics.tpp:3-8)
            const double xi = x[l];
            met[l] = ((1 - xi) * (P[6] - P[0]) + (1 + xi) * (P[4] - P[2])) / 4;
            met[NF + l] = ((1 - xi) * (P[7] - P[1]) + (1 + xi) * (P[5] - P[3])) / 4;
        } else if (f == 1) {  // x_xi, y_xi depend on eta only
            const double eta = x[l];
            met[2 * NF + l] = ((1 - eta) * (P[2] - P[0]) + (1 + eta) * (P[4] - P[6])) / 4;
            met[3 * NF + l] = ((1 - eta) * (P[3] - P[1]) + (1 + eta) * (P[5] - P[7])) / 4;
        } else {  // sides l, l + N + 1, ...: outward normal of the edge P_s -> P_s+1 and its Jacobian
            for (int s = l; s < 4; s += NF) {
                const int s1 = (s + 1) & 3;
                const double ex = P[2 * s1] - P[2 * s], ey = P[2 * s1 + 1] - P[2 * s + 1];
                const double len = sqrt(ex * ex + ey * ey);
                double *sd = s_side + (el * 4 + s) * 3;
                sd[0] = ey / len;
                sd[1] = -(ex / len);
                sd[2] = side_scale(P, s, x[0]);
            }
        }
        const double *tq = s_q + el * 3 * NE + f * NE;
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            const double row = tq[l * NF + j], col = tq[j * NF + l];
            t0 += lm[j] * row;
            t2 += lp[j] * row;
            t1 += lp[j] * col;
            t3 += lm[j] * col;
        }
        double *tr = s_tr + el * 12 * NF + f * NF;
        tr[l] = t0;
        tr[6 * NF + (N - l)] = t2;
        tr[3 * NF + l] = t1;
        tr[9 * NF + (N - l)] = t3;
    }
    __syncthreads();

    // P2
#pragma unroll
    for (int it = 0; it < FPT; ++it) {
        const int item = tid + it * T;
        if (item < nel * 4 * NF) {
            const int fe = item / (4 * NF), s = (item / NF) % 4, k = item % NF;
            double *tr = s_tr + fe * 12 * NF + s * 3 * NF + k;
            const double pL = tr[0], uL = tr[NF], vL = tr[2 * NF];
            const double *sd = s_side + (fe * 4 + s) * 3;
            double nx = sd[0], ny = sd[1];
            const double sf = sd[2];
            double pR, uR, vR;
            const int32_t cd = code[it];
            if (cd >= 0) {
                pR = nbv[it][0];
                uR = nbv[it][1];
                vR = nbv[it][2];
            } else {
                const BcSide rec = bc[-1 - cd];
                if (rec.kind == kKindMortar) {  // projected by the mortar kernel (already x sf)
                    const double *sfl = side_flux + rec.slot + k;
                    double f0 = sfl[0], f1 = sfl[NF], f2 = sfl[2 * NF];
                    for (int m = 1; m < rec.n_faces; ++m) {  // a side covered by several faces
                        const double *s2 = sfl + m * 3 * NF;
                        f0 += s2[0];
                        f1 += s2[NF];
                        f2 += s2[2 * NF];
                    }
                    tr[0] = f0;
                    tr[NF] = f1;
                    tr[2 * NF] = f2;
                    continue;
                }
                nx = rec.nx;
                ny = rec.ny;
                if (rec.kind == NDG_BC_WALL || rec.kind == NDG_BC_SYMMETRY) {
                    // ghost = sf * own trace, normal velocity reversed (Mesh2D_t.cpp:4377-4433)
                    pR = pL * rec.sf;
                    const double ux = uL * rec.sf, uy = vL * rec.sf;
                    const double l0 = -(ux * rec.nx + uy * rec.ny);
                    const double l1 = ux * rec.tx + uy * rec.ty;
                    uR = rec.nx * l0 + rec.tx * l1;
                    vR = rec.ny * l0 + rec.ty * l1;
                } else if (rec.kind == NDG_BC_OUTFLOW) {
                    pR = pL * rec.sf;
                    uR = uL * rec.sf;
                    vR = vL * rec.sf;
                } else if (rec.kind == NDG_BC_INFLOW) {
                    // ghost trace = g(x, t) at the ghost's nodes, seen reversed (Mesh2D_t.cpp:4575-4599)
                    const double interp = (x[N - k] + 1) / 2;
                    const double X = rec.bx * interp + rec.ax * (1 - interp);
                    const double Y = rec.by * interp + rec.ay * (1 - interp);
                    plane_wave(X, Y, *t_ptr, cw, pR, uR, vR);
                } else {
                    pR = 0.0;
                    uR = 0.0;
                    vR = 0.0;
                }
            }
            const double wL = pL + c * (uL * nx + vL * ny);
            const double wR = pR - c * (uR * nx + vR * ny);
            tr[0] = (c * (wL - wR) / 2) * sf;
            tr[NF] = (nx * (wL + wR) / 2) * sf;
            tr[2 * NF] = (ny * (wL + wR) / 2) * sf;
        }
    }
    __syncthreads();

    const double *tq = s_q + el * 3 * NE;
    const double *met = s_met + el * 4 * NF;
    const double *fs = s_tr + el * 12 * NF + f * NF;
    // P3: thread owns xi-line j = l of field f
    if (active) {
        const int j = l;
        double F[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const int node = k * NF + j;
            const double xeta = met[k], yeta = met[NF + k];
            if (f == 0) F[k] = yeta * (c * tq[NE + node]) - xeta * (c * tq[2 * NE + node]);
            else if (f == 1) F[k] = yeta * tq[node];
            else F[k] = -(xeta * tq[node]);
        }
        const double Fr = fs[3 * NF + j], Fl = fs[9 * NF + (N - j)];
        double *d = s_d + el * 3 * NE + f * NE + j;
#pragma unroll
        for (int i = 0; i < NF; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < NF; ++k) acc += Dh[i * NF + k] * F[k];
            acc += (Fr * lp[i] + Fl * lm[i]) * winv[i];
            d[i * NF] = acc;
        }
    }
    __syncthreads();

    // P4: thread owns eta-line i = l of field f. The new RK register goes straight to HBM (tile
    // layout, coalesced); the new state waits for the barrier.
    double qn[NF];
    if (active) {
        const int i = l;
        double Gt[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const int node = i * NF + k;
            const double xxi = met[2 * NF + k], yxi = met[3 * NF + k];
            if (f == 0) Gt[k] = -yxi * (c * tq[NE + node]) + xxi * (c * tq[2 * NE + node]);
            else if (f == 1) Gt[k] = -yxi * tq[node];
            else Gt[k] = xxi * tq[node];
        }
        const double Ft = fs[6 * NF + (N - i)], Fb = fs[i];
        const double xeta_i = met[i], yeta_i = met[NF + i];
        double *d = s_d + el * 3 * NE + f * NE + i * NF;
        const double *qrow = tq + f * NE + i * NF;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < NF; ++k) acc += Dh[j * NF + k] * Gt[k];
            acc += (Ft * lp[j] + Fb * lm[j]) * winv[j];
            const double jac = met[2 * NF + j] * yeta_i - xeta_i * met[3 * NF + j];
            const double Gv = cw * ((-d[j] - acc) * recip(jac));
            if (RHS_ONLY) {
                d[j] = Gv;
            } else {
                const double gn = first ? Gv : a * g_old[j] + Gv;
                if (store_g) g[e0 * 3 * NE + ((f * NF + j) * nel + el) * NF + i] = gn;
                qn[j] = qrow[j] + gdt * gn;
            }
        }
    }
    __syncthreads();
    if (!RHS_ONLY && active) {
        double *qw = s_q + el * 3 * NE + f * NE + l * NF;
#pragma unroll
        for (int j = 0; j < NF; ++j) qw[j] = qn[j];
    }
    __syncthreads();

    // P5
    if (RHS_ONLY) {
        double *rg = rhs + e0 * 3 * NE;
#pragma unroll
        for (int it = 0; it < S::QPT; ++it) {
            const int r = tid + it * T;
            if (r < nel * 3 * NE) rg[r] = s_d[r];
        }
        return;
    }
    {
        double *qw = q_out + e0 * 3 * NE;
#pragma unroll
        for (int it = 0; it < S::QPT; ++it) {
            const int r = tid + it * T;
            if (r < nel * 3 * NE) qw[r] = s_q[r];
        }
    }
    if (active) {
        const double *tqf = s_q + el * 3 * NE + f * NE;
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            const double row = tqf[l * NF + j], col = tqf[j * NF + l];
            t0 += lm[j] * row;
            t2 += lp[j] * row;
            t1 += lp[j] * col;
            t3 += lm[j] * col;
        }
        double *to = tr_out + (e0 + el) * 12 * NF + f * NF;
        to[l] = t0;
        to[6 * NF + (N - l)] = t2;
        to[3 * NF + l] = t1;
        to[9 * NF + (N - l)] = t3;
    }
}

// ---------------------------------------------------------------------------------------------
// FP32 form with a compensated state. The right-hand side (traces, flux, volume, lift) and the RK
// register are FP32; the state is the unevaluated sum qh + ql of two floats, updated with an
// error-free two-sum so that the accumulation over many steps keeps ~48 bits. Geometry (metric
// lines, normals, edge Jacobians) is evaluated in FP64 from the FP64 vertices and rounded once.
// Same thread layout and phases as stage_kernel.

// qh + ql += d exactly up to the final rounding of ql (Knuth two-sum, then fast two-sum
// renormalisation); contraction is off so that the error terms are exact.
__device__ __forceinline__ void two_sum_add(float &hi, float &lo, float d) {
#pragma clang fp contract(off)
    const float s = hi + d;
    const float bb = s - hi;
    const float err = (hi - (s - bb)) + (d - bb);
    const float l = lo + err;
    const float h = s + l;
    lo = l - (h - s);
    hi = h;
}

// Position in the FP32 form's gl32 (tile order [f][j][el][i] of tiles of epb elements, the last
// tile of K - tile start elements) of value (element e, field f, node (i, j)); nf = N + 1.
__device__ __forceinline__ int64_t c32_tile_index(int64_t e, int f, int i, int j, int nf, int epb, int64_t K) {
    const int64_t e0 = e - e % epb;
    const int nel = (int)min<int64_t>(epb, K - e0), el = (int)(e - e0);
    return e0 * 3 * nf * nf + ((int64_t)(f * nf + j) * nel + el) * nf + i;
}

// q and qh are in reference node order; gl32 (RK register in x, low state word in y) is in the
// FP32 stage kernel's tile order. Thread r writes one float2. The RK register carries no value
// from one step to the next, so g is copied word for word whatever its order.
__global__ void to_c32_kernel(const double *__restrict__ q, const double *__restrict__ g,
                              float *__restrict__ qh, float2 *__restrict__ gl32, int64_t n, int nf, int epb) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int ne = nf * nf, node = (int)(r % ne), i = node / nf, j = node % nf;
    const int64_t e = r / (3 * ne);
    const int f = (int)((r / ne) % 3);
    const double v = q[r];
    const float h = (float)v;
    qh[r] = h;
    gl32[c32_tile_index(e, f, i, j, nf, epb, n / (3 * ne))] = make_float2((float)g[r], (float)(v - (double)h));
}

__global__ void from_c32_kernel(const float *__restrict__ qh, const float2 *__restrict__ gl32,
                                double *__restrict__ q, double *__restrict__ g, int64_t n, int nf, int epb) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int ne = nf * nf, node = (int)(r % ne), i = node / nf, j = node % nf;
    const int64_t e = r / (3 * ne);
    const int f = (int)((r / ne) % 3);
    const float2 w = gl32[c32_tile_index(e, f, i, j, nf, epb, n / (3 * ne))];
    q[r] = (double)qh[r] + (double)w.y;
    g[r] = (double)w.x;
}

template <int N>
__global__ void __launch_bounds__(192)
traces_c32_kernel(const float *__restrict__ tab, const float *__restrict__ qh, float *__restrict__ tr,
                  int64_t K) {
    using S = LineShape<N>;
    constexpr int NF = S::NF, NE = S::NE, EPB = S::EPB, T = S::T;
    __shared__ float s_q[EPB * 3 * NE];
    const float *lm = tab + NF * NF;
    const float *lp = lm + NF;
    const int tid = threadIdx.x;
    const int f = tid >> 6, lane = tid & 63, el = lane / NF, l = lane % NF;
    const int64_t e0 = (int64_t)blockIdx.x * EPB;
    const int nel = (int)min<int64_t>(EPB, K - e0);
    for (int r = tid; r < nel * 3 * NE; r += T) s_q[r] = qh[e0 * 3 * NE + r];
    __syncthreads();
    if (el < nel) {
        const float *tq = s_q + el * 3 * NE + f * NE;
        float t0 = 0.f, t1 = 0.f, t2 = 0.f, t3 = 0.f;
        for (int j = 0; j < NF; ++j) {
            const float row = tq[l * NF + j], col = tq[j * NF + l];
            t0 += lm[j] * row;
            t2 += lp[j] * row;
            t1 += lp[j] * col;
            t3 += lm[j] * col;
        }
        float *to = tr + (e0 + el) * 12 * NF + f * NF;
        to[l] = t0;
        to[6 * NF + (N - l)] = t2;
        to[3 * NF + l] = t1;
        to[9 * NF + (N - l)] = t3;
    }
}

template <int N>
__global__ void __launch_bounds__(192)
stage_c32_kernel(const double *__restrict__ verts, const int32_t *__restrict__ side_nb,
                 const BcSide *__restrict__ bc, const double *__restrict__ tab64,
                 const float *__restrict__ tab, float *__restrict__ qh, float2 *__restrict__ gl,
                 const float *__restrict__ tr_in, float *__restrict__ tr_out,
                 const float *__restrict__ side_flux, int64_t K, int first, float a, float gdt,
                 const double *__restrict__ t_ptr, int epb, int remap, double cw) {
    using S = LineShape<N>;
    constexpr int NF = S::NF, NE = S::NE, EPB = S::EPB, T = S::T, FPT = S::FPT;
    __shared__ float s_q[EPB * 3 * NE];
    __shared__ float s_d[EPB * 3 * NE];
    __shared__ float s_tr[EPB * 12 * NF];
    __shared__ float s_met[EPB * 4 * NF];
    __shared__ float s_side[EPB * 4 * 3];
    __shared__ double s_v[EPB * 8];
    const float *Dh = tab;
    const float *lm = tab + NF * NF;
    const float *lp = lm + NF;
    const float *winv = lp + 3 * NF;
    const double *x64 = tab64 + NF * NF + 3 * NF;

    const int tid = threadIdx.x;
    const int f = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int el = lane / NF, l = lane % NF;
    const int64_t blk = remap ? xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const int64_t e0 = blk * epb;
    const int nel = (int)min<int64_t>(epb, K - e0);
    const bool active = el < nel;

    __builtin_amdgcn_s_setprio(3);  // issue this workgroup's loads ahead of other waves' math
    // P0
    int32_t code[FPT];
#pragma unroll
    for (int it = 0; it < FPT; ++it) {
        const int item = tid + it * T;
        code[it] = item < nel * 4 * NF ? side_nb[(e0 + item / NF / 4) * 4 + (item / NF) % 4] : INT32_MIN;
    }
    float qreg[S::QPT];
    const float *qg = qh + e0 * 3 * NE;
#pragma unroll
    for (int it = 0; it < S::QPT; ++it) {
        const int r = tid + it * T;
        qreg[it] = r < nel * 3 * NE ? qg[r] : 0.f;
    }
    double vreg[S::VPT];
#pragma unroll
    for (int it = 0; it < S::VPT; ++it) {
        const int r = tid + it * T;
        vreg[it] = r < nel * 8 ? verts[e0 * 8 + r] : 0.0;
    }
    // RK register and low words, interleaved, in the tile order [f][j][el][i] of the shape's
    // full tiles (the launch always uses them, so the low words keep their places from one step to
    // the next): the row owners of the tile's elements read and write one line index j with
    // consecutive lanes on consecutive addresses, both words with one 8-byte access.
    float g_old[NF], lo_row[NF];
    if (active) {
        const int64_t off = e0 * 3 * NE + ((int64_t)f * NF * nel + el) * NF + l;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            const float2 w = gl[off + j * nel * NF];
            g_old[j] = first ? 0.f : w.x;
            lo_row[j] = w.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < NF; ++j) g_old[j] = lo_row[j] = 0.f;
    }
    float nbv[FPT][3];
#pragma unroll
    for (int it = 0; it < FPT; ++it) {
        nbv[it][0] = nbv[it][1] = nbv[it][2] = 0.f;
        const int32_t cd = code[it];
        if (cd >= 0) {
            const int k = (tid + it * T) % NF;
            const float *nb = tr_in + (int64_t)(cd >> 2) * 12 * NF + (cd & 3) * 3 * NF + (N - k);
            nbv[it][0] = nb[0];
            nbv[it][1] = nb[NF];
            nbv[it][2] = nb[2 * NF];
        }
    }
    __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int it = 0; it < S::VPT; ++it) {
        const int r = tid + it * T;
        if (r < nel * 8) s_v[r] = vreg[it];
    }
#pragma unroll
    for (int it = 0; it < S::QPT; ++it) {
        const int r = tid + it * T;
        if (r < nel * 3 * NE) s_q[r] = qreg[it];
    }
    __syncthreads();

    // P1 (geometry in FP64, rounded once)
    if (active) {
        const double *P = s_v + el * 8;
        float *met = s_met + el * 4 * NF;
        if (f == 0) {
            const double xi = x64[l];
            met[l] = (float)(((1 - xi) * (P[6] - P[0]) + (1 + xi) * (P[4] - P[2])) / 4);
            met[NF + l] = (float)(((1 - xi) * (P[7] - P[1]) + (1 + xi) * (P[5] - P[3])) / 4);
        } else if (f == 1) {
            const double eta = x64[l];
            met[2 * NF + l] = (float)(((1 - eta) * (P[2] - P[0]) + (1 + eta) * (P[4] - P[6])) / 4);
            met[3 * NF + l] = (float)(((1 - eta) * (P[3] - P[1]) + (1 + eta) * (P[5] - P[7])) / 4);
        } else {  // sides l, l + N + 1, ... (N + 1 lanes cover the 4 sides)
            for (int s = l; s < 4; s += NF) {
                const int s1 = (s + 1) & 3;
                const double ex = P[2 * s1] - P[2 * s], ey = P[2 * s1 + 1] - P[2 * s + 1];
                const double len = sqrt(ex * ex + ey * ey);
                float *sd = s_side + (el * 4 + s) * 3;
                sd[0] = (float)(ey / len);
                sd[1] = (float)(-(ex / len));
                sd[2] = (float)side_scale(P, s, x64[0]);
            }
        }
        const float *tq = s_q + el * 3 * NE + f * NE;
        float t0 = 0.f, t1 = 0.f, t2 = 0.f, t3 = 0.f;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            const float row = tq[l * NF + j], col = tq[j * NF + l];
            t0 += lm[j] * row;
            t2 += lp[j] * row;
            t1 += lp[j] * col;
            t3 += lm[j] * col;
        }
        float *tr = s_tr + el * 12 * NF + f * NF;
        tr[l] = t0;
        tr[6 * NF + (N - l)] = t2;
        tr[3 * NF + l] = t1;
        tr[9 * NF + (N - l)] = t3;
    }
    __syncthreads();

    // P2
#pragma unroll
    for (int it = 0; it < FPT; ++it) {
        const int item = tid + it * T;
        if (item < nel * 4 * NF) {
            const int fe = item / (4 * NF), s = (item / NF) % 4, k = item % NF;
            float *tr = s_tr + fe * 12 * NF + s * 3 * NF + k;
            const float pL = tr[0], uL = tr[NF], vL = tr[2 * NF];
            const float *sd = s_side + (fe * 4 + s) * 3;
            float nx = sd[0], ny = sd[1];
            const float sf = sd[2];
            float pR, uR, vR;
            const int32_t cd = code[it];
            if (cd >= 0) {
                pR = nbv[it][0];
                uR = nbv[it][1];
                vR = nbv[it][2];
            } else {
                const BcSide rec = bc[-1 - cd];
                if (rec.kind == kKindMortar) {
                    const float *sfl = side_flux + rec.slot + k;
                    float f0 = sfl[0], f1 = sfl[NF], f2 = sfl[2 * NF];
                    for (int m = 1; m < rec.n_faces; ++m) {
                        const float *s2 = sfl + m * 3 * NF;
                        f0 += s2[0];
                        f1 += s2[NF];
                        f2 += s2[2 * NF];
                    }
                    tr[0] = f0;
                    tr[NF] = f1;
                    tr[2 * NF] = f2;
                    continue;
                }
                nx = (float)rec.nx;
                ny = (float)rec.ny;
                const float tx = (float)rec.tx, ty = (float)rec.ty, gs = (float)rec.sf;
                if (rec.kind == NDG_BC_WALL || rec.kind == NDG_BC_SYMMETRY) {
                    pR = pL * gs;
                    const float ux = uL * gs, uy = vL * gs;
                    const float l0 = -(ux * nx + uy * ny);
                    const float l1 = ux * tx + uy * ty;
                    uR = nx * l0 + tx * l1;
                    vR = ny * l0 + ty * l1;
                } else if (rec.kind == NDG_BC_OUTFLOW) {
                    pR = pL * gs;
                    uR = uL * gs;
                    vR = vL * gs;
                } else if (rec.kind == NDG_BC_INFLOW) {
                    const double interp = (x64[N - k] + 1) / 2;
                    const double X = rec.bx * interp + rec.ax * (1 - interp);
                    const double Y = rec.by * interp + rec.ay * (1 - interp);
                    double p64, u64, v64;
                    plane_wave(X, Y, *t_ptr, cw, p64, u64, v64);
                    pR = (float)p64;
                    uR = (float)u64;
                    vR = (float)v64;
                } else {
                    pR = uR = vR = 0.f;
                }
            }
            const float wL = pL + (uL * nx + vL * ny);
            const float wR = pR - (uR * nx + vR * ny);
            tr[0] = ((wL - wR) * 0.5f) * sf;
            tr[NF] = (nx * (wL + wR) * 0.5f) * sf;
            tr[2 * NF] = (ny * (wL + wR) * 0.5f) * sf;
        }
    }
    __syncthreads();

    const float *tq = s_q + el * 3 * NE;
    const float *met = s_met + el * 4 * NF;
    const float *fs = s_tr + el * 12 * NF + f * NF;
    // P3
    if (active) {
        const int j = l;
        float F[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const int node = k * NF + j;
            const float xeta = met[k], yeta = met[NF + k];
            if (f == 0) F[k] = yeta * tq[NE + node] - xeta * tq[2 * NE + node];
            else if (f == 1) F[k] = yeta * tq[node];
            else F[k] = -(xeta * tq[node]);
        }
        const float Fr = fs[3 * NF + j], Fl = fs[9 * NF + (N - j)];
        float *d = s_d + el * 3 * NE + f * NE + j;
#pragma unroll
        for (int i = 0; i < NF; ++i) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < NF; ++k) acc += Dh[i * NF + k] * F[k];
            acc += (Fr * lp[i] + Fl * lm[i]) * winv[i];
            d[i * NF] = acc;
        }
    }
    __syncthreads();

    // P4
    float hn[NF];
    if (active) {
        const int i = l;
        float Gt[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const int node = i * NF + k;
            const float xxi = met[2 * NF + k], yxi = met[3 * NF + k];
            if (f == 0) Gt[k] = -yxi * tq[NE + node] + xxi * tq[2 * NE + node];
            else if (f == 1) Gt[k] = -yxi * tq[node];
            else Gt[k] = xxi * tq[node];
        }
        const float Ft = fs[6 * NF + (N - i)], Fb = fs[i];
        const float xeta_i = met[i], yeta_i = met[NF + i];
        float *d = s_d + el * 3 * NE + f * NE + i * NF;
        const float *qrow = tq + f * NE + i * NF;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            float acc = 0.f;
#pragma unroll
            for (int k = 0; k < NF; ++k) acc += Dh[j * NF + k] * Gt[k];
            acc += (Ft * lp[j] + Fb * lm[j]) * winv[j];
            const float jac = met[2 * NF + j] * yeta_i - xeta_i * met[3 * NF + j];
            const float Gv = (float)cw * ((-d[j] - acc) / jac);
            const float gn = first ? Gv : a * g_old[j] + Gv;
            float h = qrow[j], lo = lo_row[j];
            two_sum_add(h, lo, gdt * gn);
            hn[j] = h;
            gl[e0 * 3 * NE + ((f * NF + j) * nel + el) * NF + i] = make_float2(gn, lo);
        }
    }
    __syncthreads();
    if (active) {
        float *qw = s_q + el * 3 * NE + f * NE + l * NF;
#pragma unroll
        for (int j = 0; j < NF; ++j) qw[j] = hn[j];
    }
    __syncthreads();

    // P5
    {
        float *qw = qh + e0 * 3 * NE;
#pragma unroll
        for (int it = 0; it < S::QPT; ++it) {
            const int r = tid + it * T;
            if (r < nel * 3 * NE) qw[r] = s_q[r];
        }
    }
    if (active) {
        const float *tqf = s_q + el * 3 * NE + f * NE;
        float t0 = 0.f, t1 = 0.f, t2 = 0.f, t3 = 0.f;
#pragma unroll
        for (int j = 0; j < NF; ++j) {
            const float row = tqf[l * NF + j], col = tqf[j * NF + l];
            t0 += lm[j] * row;
            t2 += lp[j] * row;
            t1 += lp[j] * col;
            t3 += lm[j] * col;
        }
        float *to = tr_out + (e0 + el) * 12 * NF + f * NF;
        to[l] = t0;
        to[6 * NF + (N - l)] = t2;
        to[3 * NF + l] = t1;
        to[9 * NF + (N - l)] = t3;
    }
}

// ---------------------------------------------------------------------------------------------
// Initial conditions (src/meshes/Mesh2D_t.cpp:1192-1216; pulses have no reference counterpart).
template <int N>
__global__ void ic_kernel(const double *__restrict__ verts, const double *__restrict__ tab,
                          double *__restrict__ q, int64_t K, int kind, int n_pulses,
                          const double *__restrict__ pulses, double Lx, double Ly) {
    constexpr int NF = N + 1, NE = NF * NF;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= K * NE) return;
    const int64_t e = gid / NE;
    const int node = (int)(gid % NE), i = node / NF, j = node % NF;
    const double *x = tab + NF * NF + 3 * NF;
    const double *P = verts + e * 8;
    const double xi = x[i], eta = x[j];
    const double X = (P[0] * (1 - xi) * (1 - eta) + P[2] * (xi + 1) * (1 - eta) +
                      P[4] * (xi + 1) * (eta + 1) + P[6] * (1 - xi) * (eta + 1)) / 4;
    const double Y = (P[1] * (1 - xi) * (1 - eta) + P[3] * (xi + 1) * (1 - eta) +
                      P[5] * (xi + 1) * (eta + 1) + P[7] * (1 - xi) * (eta + 1)) / 4;
    double p = 0.0, u = 0.0, v = 0.0;
    if (kind == NDG_IC_PLANE_WAVE) {
        plane_wave(X, Y, 0.0, 1.0, p, u, v);
    } else if (kind == NDG_IC_PERIODIC_PULSES) {
        // the 3 x 3 periodic images of each pulse, (mx, my) in {-1, 0, 1}^2
        for (int k = 0; k < n_pulses; ++k) {
            const double w = pulses[4 * k + 2];
            for (int mx = -1; mx <= 1; ++mx)
                for (int my = -1; my <= 1; ++my) {
                    const double dx = X - (pulses[4 * k] + mx * Lx), dy = Y - (pulses[4 * k + 1] + my * Ly);
                    p += pulses[4 * k + 3] * exp(-M_LN2 * (dx * dx + dy * dy) / (w * w));
                }
        }
    } else {
        for (int k = 0; k < n_pulses; ++k) {
            const double dx = X - pulses[4 * k], dy = Y - pulses[4 * k + 1], w = pulses[4 * k + 2];
            p += pulses[4 * k + 3] * exp(-M_LN2 * (dx * dx + dy * dy) / (w * w));
        }
    }
    double *qe = q + e * 3 * NE + node;
    qe[0] = p;
    qe[NE] = u;
    qe[2 * NE] = v;
}

// ---------------------------------------------------------------------------------------------
// CFL time step: min over elements of CFL * delta_min / (N^2 c) (src/solvers/Solver2D_t.cpp:19-29,
// Element2D_t.cpp:772-775), reduced in two passes on the device; no host round trip per block.
__device__ __forceinline__ double edge_len(const double *P, int a, int b) {
#pragma clang fp contract(off)
    const double dx = P[2 * a] - P[2 * b], dy = P[2 * a + 1] - P[2 * b + 1];
    return sqrt(dx * dx + dy * dy);
}

__device__ __forceinline__ double wave_min(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
    return v;
}

__global__ void __launch_bounds__(256)
dt_partial_kernel(const double *__restrict__ verts, int64_t K, int N, double cfl, double cw,
                  double *__restrict__ partial) {
#pragma clang fp contract(off)
    __shared__ double s_min[4];
    double v = INFINITY;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < K; e += (int64_t)gridDim.x * 256) {
        const double *P = verts + e * 8;
        const double d = fmin(fmin(fmin(edge_len(P, 1, 0), edge_len(P, 2, 3)),
                                   fmin(edge_len(P, 1, 2), edge_len(P, 0, 3))),
                              fmin(edge_len(P, 1, 3), edge_len(P, 2, 0)));
        v = fmin(cfl * d / (N * N * cw), v);
    }
    v = wave_min(v);
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = fmin(fmin(s_min[0], s_min[1]), fmin(s_min[2], s_min[3]));
}

__global__ void __launch_bounds__(256)
dt_final_kernel(const double *__restrict__ partial, int n, double *__restrict__ out) {
    __shared__ double s_min[4];
    double v = INFINITY;
    for (int k = threadIdx.x; k < n; k += 256) v = fmin(v, partial[k]);
    v = wave_min(v);
    if ((threadIdx.x & 63) == 0) s_min[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = fmin(fmin(s_min[0], s_min[1]), fmin(s_min[2], s_min[3]));
}

// ---------------------------------------------------------------------------------------------
// Pressure estimator (src/entities/Element2D_t.cpp:274-343): spectrum of the last
// min(N+1, 4) modal indices, least-squares fit of its log-decay, error and flags. One thread per
// element, FP64 throughout. out[e] = {error, sigma, refine, coarsen}.
template <int N>
__global__ void __launch_bounds__(128)
estimate_kernel(const double *__restrict__ tab, const double *__restrict__ q, int64_t K,
                double tol_min, double tol_max, double *__restrict__ out) {
    constexpr int NF = N + 1, NE = NF * NF;
    constexpr int n = NF < 4 ? NF : 4;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= K) return;
    const double *M = tab + NF * NF + 5 * NF;
    const double *p = q + e * 3 * NE;
    double spec[n];
    for (int ni = 0; ni < n; ++ni) {
        spec[ni] = 0.0;
        const int pm = N + ni + 1 - n;
        for (int i = 0; i < pm; ++i) {
            double lx = 0.0, ly = 0.0;
            for (int k = 0; k <= N; ++k)
                for (int l = 0; l <= N; ++l) {
                    lx += p[k * NF + l] * M[i * NF + k] * M[pm * NF + l];
                    ly += p[k * NF + l] * M[pm * NF + k] * M[i * NF + l];
                }
            spec[ni] += fabs(lx) + fabs(ly);
        }
        double lx = 0.0;
        for (int k = 0; k <= N; ++k)
            for (int l = 0; l <= N; ++l) lx += p[k * NF + l] * M[pm * NF + k] * M[pm * NF + l];
        spec[ni] += fabs(lx);
    }
    double xa = 0.0, ya = 0.0;
    for (int i = 0; i < n; ++i) {
        xa += N + i + 1 - n;
        ya += log(spec[i]);
    }
    xa /= n;
    ya /= n;
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n; ++i) {
        const int pp = N + i + 1 - n;
        num += (pp - xa) * (log(spec[i]) - ya);
        den += (pp - xa) * (pp - xa);
    }
    const double sig = num / den;
    const double C = exp(ya - sig * xa);
    const double sigma = fabs(sig);
    const double err = sqrt(spec[n - 1] * spec[n - 1] + C * C * 0.5 / sigma * exp(-2 * sigma * (N + 1)));
    out[4 * e + 0] = err;
    out[4 * e + 1] = sigma;
    out[4 * e + 2] = err > tol_min ? 1.0 : 0.0;
    out[4 * e + 3] = err < tol_max ? 1.0 : 0.0;
}

// ---------------------------------------------------------------------------------------------
// Device clock of a step: clock[0] is the time at the start of the step, clock[1 + s] the time of
// stage s, t + b_s dt (include/solvers/Solver2D_t.tpp:16, :45-63), rounded as the host loop does
// (no contraction). set != 0 starts from t, else from clock[0]; clock_end advances clock[0] by dt.
__global__ void clock_begin_kernel(double *clock, double t, int set, double dt) {
#pragma clang fp contract(off)
    if (threadIdx.x != 0) return;
    const double t0 = set ? t : clock[0];
    clock[0] = t0;
    clock[1] = t0 + 0.0 * dt;
    clock[2] = t0 + (1.0 / 3.0) * dt;
    clock[3] = t0 + 0.75 * dt;
}

__global__ void clock_end_kernel(double *clock, double dt) {
#pragma clang fp contract(off)
    if (threadIdx.x == 0) clock[0] = clock[0] + dt;
}

// ---------------------------------------------------------------------------------------------
// Non-conforming (mortar) faces, FP64 arithmetic (the traces and the output are of type T: double,
// or float in the FP32 form). Restates, for one face:
//   project_to_faces (src/meshes/Mesh2D_t.cpp:3803-3893): each side's trace is copied when the
//     element order equals the face order, else interpolated barycentrically at
//     xi = (x_face + 1) scale + 2 offset - 1 with the exact-node shortcut (right side reversed);
//   calculate_wave_fluxes (src/solvers/Solver2D_t.cpp:48-63);
//   project_to_elements (src/meshes/Mesh2D_t.cpp:3895-4003): conforming copy (left forward, right
//     backward and negated) or the Gauss-quadrature L2 projection, times the edge Jacobian and the
//     face scale, accumulated over face nodes in order.

__device__ __forceinline__ bool ulp2_equal(double a, double b) {
    return fabs(a - b) <= DBL_EPSILON * fabs(a + b) * 2 || fabs(a - b) < DBL_MIN;
}

// Loads n <= 32 * kMaxItems values src[0..n) into dst through registers, every load in flight
// before the first LDS write (lane i of a 32-lane half takes items i, i + 32, ...).
template <typename T, int kMaxItems>
__device__ __forceinline__ void stage_to_lds(const T *__restrict__ src, double *dst, int n, int i) {
    double v[kMaxItems];
#pragma unroll
    for (int it = 0; it < kMaxItems; ++it) {
        const int idx = i + 32 * it;
        v[it] = idx < n ? (double)src[idx] : 0.0;
    }
#pragma unroll
    for (int it = 0; it < kMaxItems; ++it) {
        const int idx = i + 32 * it;
        if (idx < n) dst[idx] = v[it];
    }
}

// One wave per face: lanes 0..31 handle the left side, lanes 32..63 the right side. Each half first
// stages its element trace and its interpolation and projection matrices (built on the host with
// the reference's exact-node shortcuts, see MortarFace) into LDS with all loads in flight, then
// lane i computes face node i (interpolation) and element node i (projection); face states and
// flux go through LDS. M bounds N + 1 of every face and side (checked on the host when the kernel
// is chosen): a smaller M gives a smaller LDS footprint and more faces resident per CU.
template <typename T, int M>
__global__ void __launch_bounds__(64)
mortar_kernel(const MortarFace *__restrict__ faces, const double *__restrict__ mats,
              const T *__restrict__ tr, T *__restrict__ side_flux) {
    constexpr int kMat = M * M, kMatItems = (kMat + 31) / 32, kTrItems = (3 * M + 31) / 32;
    __shared__ double s_tr[2][3 * M];   // element traces, left and right
    __shared__ double s_mi[2][kMat];    // interpolation matrices [face node][element node]
    __shared__ double s_mp[2][kMat];    // projection matrices [element node][face node]
    __shared__ double s_st[2][3][M];    // face states, left and right
    __shared__ double s_f[3][M];        // face flux
    const MortarFace fc = faces[blockIdx.x];
    const int lane = threadIdx.x;
    const int side = lane >> 5, i = lane & 31;
    const int Nf = fc.Nf;
    const int Ne = side ? fc.NR : fc.NL;
    // selects, not fc.mat[side]: a lane-dependent index would put the whole record in scratch
    const int64_t mi = side ? fc.mat[1] : fc.mat[0], mp = side ? fc.mat[3] : fc.mat[2];
    stage_to_lds<T, kTrItems>(tr + (side ? fc.trR : fc.trL), s_tr[side], 3 * (Ne + 1), i);
    if (mi >= 0) stage_to_lds<double, kMatItems>(mats + mi, s_mi[side], (Nf + 1) * (Ne + 1), i);
    if (mp >= 0) stage_to_lds<double, kMatItems>(mats + mp, s_mp[side], (Ne + 1) * (Nf + 1), i);
    __syncthreads();
    // project_to_faces (src/meshes/Mesh2D_t.cpp:3803-3893)
    if (i <= Nf) {
        const double *te = s_tr[side];
        double v0 = 0.0, v1 = 0.0, v2 = 0.0;
        if (mi < 0) {
            const int k = side ? Nf - i : i;
            v0 = te[k];
            v1 = te[(Ne + 1) + k];
            v2 = te[2 * (Ne + 1) + k];
        } else {
            // unrolled to M with every LDS read issued unconditionally (all in bounds: i <= Nf < M,
            // j < M), only the terms j <= Ne accumulated, in the same order
            const double *row = s_mi[side] + i * (Ne + 1);
#pragma unroll
            for (int j = 0; j < M; ++j) {
                const double w = row[j], a0 = te[j], a1 = te[(Ne + 1) + j], a2 = te[2 * (Ne + 1) + j];
                if (j <= Ne) {
                    v0 += w * a0;
                    v1 += w * a1;
                    v2 += w * a2;
                }
            }
        }
        s_st[side][0][i] = v0;
        s_st[side][1][i] = v1;
        s_st[side][2][i] = v2;
    }
    __syncthreads();
    // calculate_wave_fluxes (src/solvers/Solver2D_t.cpp:48-63)
    if (lane <= Nf) {
        const double c = kWaveSpeed;
        const double wl = s_st[0][0][lane] + c * (s_st[0][1][lane] * fc.nx + s_st[0][2][lane] * fc.ny);
        const double wr = s_st[1][0][lane] - c * (s_st[1][1][lane] * fc.nx + s_st[1][2][lane] * fc.ny);
        s_f[0][lane] = c * (wl - wr) / 2;
        s_f[1][lane] = fc.nx * (wl + wr) / 2;
        s_f[2][lane] = fc.ny * (wl + wr) / 2;
    }
    __syncthreads();
    // project_to_elements (src/meshes/Mesh2D_t.cpp:3895-4006), times the edge Jacobian
    const int64_t slot = side ? fc.slotR : fc.slotL;
    if (slot < 0 || i > Ne) return;
    const double sf = side ? fc.sfR : fc.sfL;
    T *out = side_flux + slot;
    if (mp < 0) {
        for (int f = 0; f < 3; ++f)
            out[f * (Ne + 1) + i] = side ? (T)(-s_f[f][Nf - i] * sf) : (T)(s_f[f][i] * sf);
        return;
    }
    const double *row = s_mp[side] + i * (Nf + 1);
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int j = 0; j < M; ++j) {  // as the interpolation: i <= Ne < M, j < M
        const double w = row[j], f0 = s_f[0][j], f1 = s_f[1][j], f2 = s_f[2][j];
        if (j <= Nf) {
            a0 += w * f0;
            a1 += w * f1;
            a2 += w * f2;
        }
    }
    out[i] = (T)(a0 * sf);
    out[(Ne + 1) + i] = (T)(a1 * sf);
    out[2 * (Ne + 1) + i] = (T)(a2 * sf);
}

// ---------------------------------------------------------------------------------------------
// Launchers: dispatch on the (uniform) order N.
// ---------------------------------------------------------------------------------------------
// Solution transfer after adaptation, one workgroup per new element: a copy when neither order nor
// geometry changes, else Element2D_t::interpolate_from (src/entities/Element2D_t.cpp:345-579) at
// the new nodes (same geometry) or at the given reference coordinates in the old element (a child):
// the old element's node when both coordinates match a node (almost_equal, the last matching
// index wins), a 1D barycentric interpolation along the other coordinate when one matches, else
// the tensor barycentric interpolation.
__global__ void __launch_bounds__(256)
transfer_kernel(const double *__restrict__ q_old, double *__restrict__ q_new,
                const TransferRec *__restrict__ recs, const double *__restrict__ coords,
                const double *__restrict__ tab1d) {
    const TransferRec R = recs[blockIdx.x];
    const int No = R.N_old, Nn = R.N_new, nfo = No + 1, nf = Nn + 1, neo = nfo * nfo, ne = nf * nf;
    const double *src = q_old + R.old_off;
    double *dst = q_new + R.new_off;
    if (No == Nn && R.coord_off < 0) {
        for (int r = threadIdx.x; r < 3 * ne; r += 256) dst[r] = src[r];
        return;
    }
    const double *xo = tab1d + (3 * No + 0) * kTab1dStride, *lo = tab1d + (3 * No + 2) * kTab1dStride;
    const double *xn = tab1d + (3 * Nn + 0) * kTab1dStride;
    for (int r = threadIdx.x; r < 3 * ne; r += 256) {
        const int f = r / ne, node = r % ne, i = node / nf, j = node % nf;
        const double X = R.coord_off >= 0 ? coords[R.coord_off + 2 * node] : xn[i];
        const double Y = R.coord_off >= 0 ? coords[R.coord_off + 2 * node + 1] : xn[j];
        dst[r] = interpolate_at(src + f * neo, No, X, Y, xo, lo);
    }
}

hipError_t launch_transfer(const double *q_old, double *q_new, const TransferRec *recs, const double *coords,
                           int64_t K, const double *tab1d, hipStream_t s) {
    if (K <= 0) return hipSuccess;
    hipLaunchKernelGGL(transfer_kernel, dim3((unsigned)K), dim3(256), 0, s, q_old, q_new, recs, coords, tab1d);
    return hipGetLastError();
}

// Mesh2D_t::get_solution (src/meshes/Mesh2D_t.cpp:4237-4249) with
// Element2D_t::interpolate_solution (src/entities/Element2D_t.cpp:581-604): for each element (one
// workgroup, caller's order) and output point (i, j) of the n x n uniform grid, x and y by the
// bilinear map at ((i / (n - 1)) * 2 - 1, (j / (n - 1)) * 2 - 1), and each field as
// sum_m sum_k q[m][k] I[i][m] I[j][k] with the order's uniform-grid interpolation matrix I.
// With rhs given, also Mesh2D_t::get_complete_solution (src/meshes/Mesh2D_t.cu:5096-5155,
// src/entities/Element2D_t.cu:623-665): the right-hand side (laid out as q) and
// |g(x_mk, time) - q[m][k]| of the plane wave at the solution nodes, interpolated
// with the same weights. The element's state, right-hand side and node errors are staged in LDS.
// Output layout [x, y, p, u, v (, dp, du, dv, |dp|, |du|, |dv|)][element][i * n + j].
constexpr int kOutMaxNe = (kMaxOrder + 1) * (kMaxOrder + 1);

__global__ __launch_bounds__(64) void output_kernel(const double *__restrict__ q, const double *__restrict__ rhs,
                                                    const double *__restrict__ verts,
                                                    const OutputRec *__restrict__ recs,
                                                    const double *__restrict__ interp,
                                                    const double *__restrict__ tab1d, int n, int64_t K,
                                                    double time, double cw, double *__restrict__ out) {
    __shared__ double s_q[3 * kOutMaxNe], s_r[3 * kOutMaxNe], s_a[3 * kOutMaxNe];
    const int64_t e = blockIdx.x;
    const OutputRec R = recs[e];
    const int nf = R.N + 1, ne = nf * nf, nn = n * n;
    const double *I = interp + R.i_off;
    const double *P = verts + 8 * R.dev;
    const double *src = q + R.q_off;
    const bool complete = rhs != nullptr;
    const double *nodes = tab1d + 3 * R.N * kTab1dStride;
    for (int r = threadIdx.x; r < 3 * ne; r += 64) {
        const double v = src[r];
        s_q[r] = v;
        if (complete) {
            const int c = r / ne, mk = r - c * ne, m = mk / nf, k = mk - m * nf;
            s_r[r] = rhs[R.q_off + r];
            const double xi = nodes[m], eta = nodes[k];
            const double w0 = (1 - xi) * (1 - eta), w1 = (xi + 1) * (1 - eta), w2 = (xi + 1) * (eta + 1),
                         w3 = (1 - xi) * (eta + 1);
            const double X = (P[0] * w0 + P[2] * w1 + P[4] * w2 + P[6] * w3) / 4;
            const double Y = (P[1] * w0 + P[3] * w1 + P[5] * w2 + P[7] * w3) / 4;
            double g[3];
            plane_wave(X, Y, time, cw, g[0], g[1], g[2]);
            s_a[r] = fabs(g[c] - v);
        }
    }
    __syncthreads();
    const int64_t plane = K * nn;
    for (int r = threadIdx.x; r < nn; r += 64) {
        const int i = r / n, j = r % n;
        const double xi = (double)i / (double)(n - 1) * 2 - 1, eta = (double)j / (double)(n - 1) * 2 - 1;
        const double w0 = (1 - xi) * (1 - eta), w1 = (xi + 1) * (1 - eta), w2 = (xi + 1) * (eta + 1),
                     w3 = (1 - xi) * (eta + 1);
        const double x = (P[0] * w0 + P[2] * w1 + P[4] * w2 + P[6] * w3) / 4;
        const double y = (P[1] * w0 + P[3] * w1 + P[5] * w2 + P[7] * w3) / 4;
        double f[9] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (int m = 0; m < nf; ++m) {
            const double a = I[i * nf + m];
            for (int k = 0; k < nf; ++k) {
                const double b = I[j * nf + k];
                const int mk = m * nf + k;
                for (int c = 0; c < 3; ++c) f[c] += s_q[c * ne + mk] * a * b;
                if (complete)
                    for (int c = 0; c < 3; ++c) {
                        f[3 + c] += s_r[c * ne + mk] * a * b;
                        f[6 + c] += s_a[c * ne + mk] * a * b;
                    }
            }
        }
        const int64_t at = e * nn + r;
        out[at] = x;
        out[plane + at] = y;
        const int nfield = complete ? 9 : 3;
        for (int c = 0; c < nfield; ++c) out[(2 + c) * plane + at] = f[c];
    }
}

hipError_t launch_output(const double *q, const double *rhs, const double *verts, const OutputRec *recs,
                         const double *interp, const double *tab1d, int n, int64_t K, double time, double cw,
                         double *out, hipStream_t s) {
    if (K <= 0) return hipSuccess;
    hipLaunchKernelGGL(output_kernel, dim3((unsigned)K), dim3(64), 0, s, q, rhs, verts, recs, interp, tab1d, n, K,
                       time, cw, out);
    return hipGetLastError();
}

#define NDG_FOR_ORDERS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16)
bool order_supported(int N) { return N >= 1 && N <= kMaxOrder; }

#define CASES_1(F, ...) \
    case 1: return F<1>(__VA_ARGS__); case 2: return F<2>(__VA_ARGS__); case 3: return F<3>(__VA_ARGS__); \
    case 4: return F<4>(__VA_ARGS__); case 5: return F<5>(__VA_ARGS__); case 6: return F<6>(__VA_ARGS__); \
    case 7: return F<7>(__VA_ARGS__); case 8: return F<8>(__VA_ARGS__); case 9: return F<9>(__VA_ARGS__); \
    case 10: return F<10>(__VA_ARGS__); case 11: return F<11>(__VA_ARGS__); case 12: return F<12>(__VA_ARGS__); \
    case 13: return F<13>(__VA_ARGS__); case 14: return F<14>(__VA_ARGS__); case 15: return F<15>(__VA_ARGS__); \
    case 16: return F<16>(__VA_ARGS__);

// Launchers: every per-element kernel runs once per order group, on the group's slices of the
// device buffers.

template <int N>
static hipError_t traces_n(const DeviceMesh &m, const Group &G, hipStream_t s) {
    using S = Shape<N>;
    const int64_t blocks = (G.K + S::EPB - 1) / S::EPB;
    hipLaunchKernelGGL(traces_kernel<N>, dim3((unsigned)blocks), dim3(S::THREADS), 0, s, G.tab, m.q + G.q_off,
                       m.trace[m.cur] + G.tr_off, G.K);
    return hipGetLastError();
}

// Elements per workgroup actually used for group G (a tuning choice, at most the shape's EPB).
template <int N>
static int group_epb(const Group &G, int precision) {
    const int e = G.tune[precision].epb;
    return e >= 1 && e <= LineShape<N>::EPB ? e : LineShape<N>::EPB;
}

template <int N>
static int stage_epb_n(const Group &G) {
    return group_epb<N>(G, NDG_PRECISION_FP64);
}

int stage_tile_epb(const Group &G) {
    switch (G.N) { CASES_1(stage_epb_n, G) }
    return 1;
}

template <int N>
static hipError_t stage_n(DeviceMesh &m, const Group &G, int stage, double a, double gdt, hipStream_t s) {
    using S = LineShape<N>;
    const int epb = group_epb<N>(G, NDG_PRECISION_FP64);
    const int64_t blocks = (G.K + epb - 1) / epb;
    hipLaunchKernelGGL((stage_kernel<N, false>), dim3((unsigned)blocks), dim3(S::T), 0, s, m.verts + G.e_begin * 8,
                       m.side_nb + G.e_begin * 4, m.bc, G.tab, m.q + G.q_off, (m.q_out ? m.q_out : m.q) + G.q_off,
                       m.g + G.q_off, m.trace[m.cur] + G.tr_off, m.trace[1 - m.cur] + G.tr_off, nullptr, m.side_flux,
                       G.K, (stage == 0 ? kRkFirst : 0) | (stage == 2 && !m.keep_last_rk ? kRkLast : 0), a, gdt,
                       m.clock + 1 + stage, epb, G.tune[NDG_PRECISION_FP64].remap,
                       m.wave_speed);
    return hipGetLastError();
}

template <int N>
static hipError_t rhs_n(DeviceMesh &m, const Group &G, hipStream_t s) {
    using S = LineShape<N>;
    const int64_t blocks = (G.K + S::EPB - 1) / S::EPB;
    hipLaunchKernelGGL((stage_kernel<N, true>), dim3((unsigned)blocks), dim3(S::T), 0, s, m.verts + G.e_begin * 8,
                       m.side_nb + G.e_begin * 4, m.bc, G.tab, m.q + G.q_off, m.q + G.q_off, m.g + G.q_off,
                       m.trace[m.cur] + G.tr_off, nullptr, m.rhs + G.q_off, m.side_flux, G.K, 1, 0.0, 0.0,
                       m.clock + 1, S::EPB, 1, m.wave_speed);
    return hipGetLastError();
}

template <int N>
static hipError_t ic_n(const DeviceMesh &m, const Group &G, int kind, int n_pulses, const double *pulses,
                       double Lx, double Ly, hipStream_t s) {
    const int64_t total = G.K * (N + 1) * (N + 1);
    hipLaunchKernelGGL(ic_kernel<N>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, m.verts + G.e_begin * 8,
                       G.tab, m.q + G.q_off, G.K, kind, n_pulses, pulses, Lx, Ly);
    return hipGetLastError();
}

template <int N>
static hipError_t estimate_n(const DeviceMesh &m, const Group &G, double tol_min, double tol_max, double *out,
                             hipStream_t s) {
    hipLaunchKernelGGL(estimate_kernel<N>, dim3((unsigned)((G.K + 127) / 128)), dim3(128), 0, s, G.tab,
                       m.q + G.q_off, G.K, tol_min, tol_max, out + 4 * G.e_begin);
    return hipGetLastError();
}

template <int N>
static hipError_t traces_c32_n(const DeviceMesh &m, const Group &G, hipStream_t s) {
    using S = LineShape<N>;
    const int64_t blocks = (G.K + S::EPB - 1) / S::EPB;
    hipLaunchKernelGGL(traces_c32_kernel<N>, dim3((unsigned)blocks), dim3(S::T), 0, s, G.tab32, m.qh + G.q_off,
                       m.trace32[m.cur] + G.tr_off, G.K);
    return hipGetLastError();
}

template <int N>
static hipError_t stage_c32_n(DeviceMesh &m, const Group &G, int stage, double a, double gdt, hipStream_t s) {
    using S = LineShape<N>;
    const int epb = S::EPB;  // the low state words are kept in the full tiles' order
    const int64_t blocks = (G.K + epb - 1) / epb;
    hipLaunchKernelGGL(stage_c32_kernel<N>, dim3((unsigned)blocks), dim3(S::T), 0, s, m.verts + G.e_begin * 8,
                       m.side_nb + G.e_begin * 4, m.bc, G.tab, G.tab32, m.qh + G.q_off, m.gl32 + G.q_off,
                       m.trace32[m.cur] + G.tr_off, m.trace32[1 - m.cur] + G.tr_off,
                       m.side_flux32, G.K, stage == 0 ? 1 : 0, (float)a, (float)gdt, m.clock + 1 + stage, epb,
                       G.tune[NDG_PRECISION_FP32_COMPENSATED].remap, m.wave_speed);
    return hipGetLastError();
}

template <int N>
static int max_epb_n() { return LineShape<N>::EPB; }

int stage_max_epb(int N) {
    switch (N) { CASES_1(max_epb_n) }
    return 0;
}

hipError_t launch_group_stage(DeviceMesh &m, int group, int stage, double a, double gdt, hipStream_t s) {
    const Group &G = m.groups[group];
    if (m.precision == NDG_PRECISION_FP32_COMPENSATED) {
        switch (G.N) { CASES_1(stage_c32_n, m, G, stage, a, gdt, s) }
    } else {
        switch (G.N) { CASES_1(stage_n, m, G, stage, a, gdt, s) }
    }
    return hipErrorInvalidValue;
}

// Autotuning probe: stage 0 form with a = 0 and gdt = 0, so that the state and the current traces
// are left as they are.
hipError_t launch_group_probe(DeviceMesh &m, int group, hipStream_t s) {
    return launch_group_stage(m, group, 0, 0.0, 0.0, s);
}



template <typename Fn>
static hipError_t for_groups(const DeviceMesh &m, Fn &&fn) {
    for (const Group &G : m.groups) {
        const hipError_t e = fn(G);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_traces(const DeviceMesh &m, hipStream_t s) {
    return for_groups(m, [&](const Group &G) -> hipError_t {
        switch (G.N) { CASES_1(traces_n, m, G, s) }
        return hipErrorInvalidValue;
    });
}

template <int M>
static hipError_t mortar_m(DeviceMesh &m, hipStream_t s) {
    const unsigned blocks = (unsigned)m.n_faces;
    if (m.precision == NDG_PRECISION_FP32_COMPENSATED)
        hipLaunchKernelGGL((mortar_kernel<float, M>), dim3(blocks), dim3(64), 0, s, m.faces, m.mortar_mats,
                           m.trace32[m.cur], m.side_flux32);
    else
        hipLaunchKernelGGL((mortar_kernel<double, M>), dim3(blocks), dim3(64), 0, s, m.faces, m.mortar_mats,
                           m.trace[m.cur], m.side_flux);
    return hipGetLastError();
}

hipError_t launch_mortar(DeviceMesh &m, hipStream_t s) {
    if (m.n_faces == 0) return hipSuccess;
    static_assert(kMaxOrder + 1 <= 32, "a mortar side is handled by one 32-lane half wave");
    if (m.mortar_nodes < 1 || m.mortar_nodes > kMaxOrder + 1) return hipErrorInvalidValue;
    if (m.mortar_nodes <= 9) return mortar_m<9>(m, s);
    if (m.mortar_nodes <= 13) return mortar_m<13>(m, s);
    return mortar_m<kMaxOrder + 1>(m, s);
}

hipError_t launch_stage(DeviceMesh &m, int stage, double a, double gdt, hipStream_t s) {
    hipError_t e = launch_mortar(m, s);
    if (e != hipSuccess) return e;
    for (int g = 0; g < (int)m.groups.size(); ++g)
        if ((e = launch_group_stage(m, g, stage, a, gdt, s)) != hipSuccess) return e;
    m.cur = 1 - m.cur;
    return hipSuccess;
}

hipError_t launch_rhs(DeviceMesh &m, double t, hipStream_t s) {
    hipError_t e = launch_clock_begin(m, t, 1, 0.0, s);
    if (e != hipSuccess) return e;
    if ((e = launch_mortar(m, s)) != hipSuccess) return e;
    return for_groups(m, [&](const Group &G) -> hipError_t {
        switch (G.N) { CASES_1(rhs_n, m, G, s) }
        return hipErrorInvalidValue;
    });
}

hipError_t launch_ic(const DeviceMesh &m, int kind, int n_pulses, const double *pulses, double Lx, double Ly,
                     hipStream_t s) {
    return for_groups(m, [&](const Group &G) -> hipError_t {
        switch (G.N) { CASES_1(ic_n, m, G, kind, n_pulses, pulses, Lx, Ly, s) }
        return hipErrorInvalidValue;
    });
}

hipError_t launch_estimate(const DeviceMesh &m, double tol_min, double tol_max, double *out, hipStream_t s) {
    return for_groups(m, [&](const Group &G) -> hipError_t {
        switch (G.N) { CASES_1(estimate_n, m, G, tol_min, tol_max, out, s) }
        return hipErrorInvalidValue;
    });
}

hipError_t launch_traces_c32(const DeviceMesh &m, hipStream_t s) {
    return for_groups(m, [&](const Group &G) -> hipError_t {
        switch (G.N) { CASES_1(traces_c32_n, m, G, s) }
        return hipErrorInvalidValue;
    });
}


// q and qh are in reference node order; gl32 in the FP32 stage's tile order (tiles of the
// shape's EPB elements). Run per group (the order depends on N).
hipError_t launch_to_c32(DeviceMesh &m, hipStream_t s) {
    return for_groups(m, [&](const Group &G) -> hipError_t {
        const int64_t n = G.K * 3 * (int64_t)(G.N + 1) * (G.N + 1);
        hipLaunchKernelGGL(to_c32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m.q + G.q_off,
                           m.g + G.q_off, m.qh + G.q_off, m.gl32 + G.q_off, n, G.N + 1, stage_max_epb(G.N));
        return hipGetLastError();
    });
}

hipError_t launch_from_c32(DeviceMesh &m, hipStream_t s) {
    return for_groups(m, [&](const Group &G) -> hipError_t {
        const int64_t n = G.K * 3 * (int64_t)(G.N + 1) * (G.N + 1);
        hipLaunchKernelGGL(from_c32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, m.qh + G.q_off,
                           m.gl32 + G.q_off, m.q + G.q_off, m.g + G.q_off, n, G.N + 1, stage_max_epb(G.N));
        return hipGetLastError();
    });
}

// One partial minimum per workgroup of every group (at most 1024 in total), then one final pass.
hipError_t launch_delta_t(const DeviceMesh &m, double cfl, double *dt_dev, hipStream_t s) {
    int used = 0;
    for (const Group &G : m.groups) {
        const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((G.K + 255) / 256,
                                                                      1024 / (int64_t)m.groups.size()));
        if (used + blocks > m.scratch_len) return hipErrorInvalidValue;
        hipLaunchKernelGGL(dt_partial_kernel, dim3(blocks), dim3(256), 0, s, m.verts + G.e_begin * 8, G.K, G.N, cfl,
                           m.wave_speed, m.scratch + used);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        used += blocks;
    }
    hipLaunchKernelGGL(dt_final_kernel, dim3(1), dim3(256), 0, s, m.scratch, used, dt_dev);
    return hipGetLastError();
}

hipError_t launch_clock_begin(DeviceMesh &m, double t, int set, double dt, hipStream_t s) {
    if (!m.has_inflow) return hipSuccess;  // nothing reads the stage time
    hipLaunchKernelGGL(clock_begin_kernel, dim3(1), dim3(64), 0, s, m.clock, t, set, dt);
    return hipGetLastError();
}

hipError_t launch_clock_end(DeviceMesh &m, double dt, hipStream_t s) {
    if (!m.has_inflow) return hipSuccess;
    hipLaunchKernelGGL(clock_end_kernel, dim3(1), dim3(64), 0, s, m.clock, dt);
    return hipGetLastError();
}

}  // namespace ndg
