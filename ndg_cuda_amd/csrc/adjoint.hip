// adjoint.hip -- discrete adjoint (transpose) of the fused RK3 stage, for gradients of a receiver
// misfit with respect to the initial fields and the wave speed c. No reference counterpart.
//
// The forward stage (kernels.hip) is, per element, with nu = mu-adjoint scaled below:
//   G_f = c (-d_xi - d_eta) / J,
//   d_xi[i][j]  = sum_k Dh[i][k] F_f[k][j]  + (F*_1[j] lp[i] + F*_3[N-j] lm[i]) / w[i],
//   d_eta[i][j] = sum_k Dh[j][k] Gt_f[i][k] + (F*_2[N-i] lp[j] + F*_0[i] lm[j]) / w[j],
//   F_p = y_eta u - x_eta v, F_u = y_eta p, F_v = -x_eta p (x_eta, y_eta functions of xi_i),
//   Gt_p = -y_xi u + x_xi v, Gt_u = -y_xi p, Gt_v = x_xi p (x_xi, y_xi functions of eta_j),
//   F*_s[k] = sf_s flux(own trace (s, k), neighbour trace (s', N - k)), the upwind flux with
//   w_L = p_L + u_L.n, w_R = p_R - u_R.n: F*p = (w_L - w_R)/2, F*u = n_x (w_L + w_R)/2, F*v = n_y ...,
//   traces tr0[i] = sum_j lm[j] q[i][j], tr2[N-i] = sum_j lp[j] q[i][j],
//          tr1[j] = sum_i lp[i] q[i][j], tr3[N-j] = sum_i lm[i] q[i][j]
// (kernels.hip stage_kernel; src/solvers/Solver2D_t.cpp:48-136, Element2D_t.cpp:86-125).
//
// Transpose applied to an adjoint mu (per node and field): nu = -c mu / J, then
//   Fbar_f[k][j] = sum_i Dh[i][k] nu_f[i][j],  Gbar_f[i][k] = sum_j Dh[j][k] nu_f[i][j],
//   rho_1[j] = sum_i lp[i] nu[i][j] / w[i],  rho_3[N-j] = sum_i lm[i] nu[i][j] / w[i],
//   rho_0[i] = sum_j lm[j] nu[i][j] / w[j],  rho_2[N-i] = sum_j lp[j] nu[i][j] / w[j]
// (rho_s is the adjoint of the side flux F*_s); per side point omega_L = sf (rho_p + n.rho_uv)/2,
// omega_R = sf (-rho_p + n.rho_uv)/2; the own trace adjoint is (omega_L, n omega_L) plus what the
// neighbour exports for this point, chi = (omega_R, -n omega_R) computed on its side; finally the
// traces' and fluxes' transposes scatter into lambda:
//   dl_p += y_eta Fbar_u - x_eta Fbar_v - y_xi Gbar_u + x_xi Gbar_v,
//   dl_u += y_eta Fbar_p - y_xi Gbar_p,  dl_v += -x_eta Fbar_p + x_xi Gbar_p,
//   dl_f[i][j] += lm[j] tau0_f[i] + lp[j] tau2_f[N-i] + lp[i] tau1_f[j] + lm[i] tau3_f[N-j].
//
// Reverse RK stage s (forward: g' = a_s g + G(q), q' = q + gs g' with gs = gamma_s dt), given the
// adjoints (lambda, mu) of (q', g'):  mu~ = mu + gs lambda;  lambda <- lambda + G^T mu~;
// mu <- a_s mu~.  The c-derivative of the misfit gathers <mu~, dG/dc q_s> = <G^T mu~, q_s> / c
// over every stage (G is linear in c), from the stage input state q_s.
//
// The kernel keeps lambda and mu divided by the Jacobian J at their node (lambda / J, mu / J): the
// RK combinations are the same in these variables, nu = -c mu~ / J is then -c times the stored
// value, and only the new contribution G^T mu~ is divided by J, once per node and field. The
// driver scales the receivers' misfit source and the final gradient accordingly
// (jacobian_scale_kernel).
#include <hip/hip_runtime.h>

#include <cmath>

#include "device.hpp"
#include "internal.hpp"

namespace ndg {

// Flags of adjoint_stage_kernel.
constexpr int kAdjMuIn = 1;     // mu input is non-zero (stages after the first reverse stage of a step)
constexpr int kAdjMuOut = 2;    // store mu <- a_s mu~ (a_s != 0)
constexpr int kAdjExport = 4;   // export chi for the next reverse stage
constexpr int kAdjApply = 8;    // do the transpose (off: export only)
constexpr int kAdjSource = 16;  // add the receivers' misfit source to lambda' before the export

template <int N>
__global__ void __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(N <= 10 ? 3 : 2)))
adjoint_stage_kernel(const double *__restrict__ verts, const int32_t *__restrict__ side_nb,
                     const double *__restrict__ tab, double *__restrict__ lam, double *__restrict__ mu,
                     const double *__restrict__ chi_in, double *__restrict__ chi_out,
                     const double *__restrict__ qs, double *__restrict__ cpart, int64_t K, int flags,
                     double gs, double as, double gnext, double cw, const int64_t *__restrict__ rec_elem,
                     const double *__restrict__ rec_w, const double *__restrict__ rec_res, int n_rec) {
    using S = LineShape<N>;
    constexpr int NF = S::NF, NE = S::NE, EPB = S::EPB, T = S::T, FPT = S::FPT, QPT = S::QPT;
    __shared__ double s_n[EPB * 3 * NE];   // nu = -c mu~ / J
    __shared__ double s_a[EPB * 3 * NE];   // d lambda, then lambda' = lambda + d lambda
    __shared__ double s_rho[EPB * 12 * NF];
    __shared__ double s_met[EPB * 4 * NF];
    __shared__ double s_side[EPB * 4 * 3];
    __shared__ double s_v[EPB * 8];
    const double *lm = tab + NF * NF;
    const double *lp = lm + NF;
    const double *x = lp + 2 * NF;
    const double *winv = x + NF;
    const double *DhT = tab + 2 * NF * NF + 5 * NF;  // DhT[a][i] = Dh[i][a]

    const int tid = threadIdx.x;
    const int f = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int el = lane / NF, l = lane % NF;
    const int64_t e0 = (int64_t)blockIdx.x * EPB;
    const int nel = (int)min<int64_t>(EPB, K - e0);
    const bool active = el < nel;
    const bool apply = flags & kAdjApply;

    // P0: connectivity and neighbour exports of the thread's flux points, lambda and mu tiles,
    // vertices.
    double nbv[FPT][3];
#pragma unroll
    for (int it = 0; it < FPT; ++it) {
        nbv[it][0] = nbv[it][1] = nbv[it][2] = 0.0;
        const int item = tid + it * T;
        if (apply && item < nel * 4 * NF) {
            const int32_t cd = side_nb[(e0 + item / NF / 4) * 4 + (item / NF) % 4];
            const int k = item % NF;
            if (cd >= 0) {  // the host admits interior and periodic sides only
                const double *nb = chi_in + (int64_t)(cd >> 2) * 12 * NF + (cd & 3) * 3 * NF + (N - k);
                nbv[it][0] = nb[0];
                nbv[it][1] = nb[NF];
                nbv[it][2] = nb[2 * NF];
            }
        }
    }
    double mreg[QPT], lreg[QPT], qreg[QPT];
    const double *lg = lam + e0 * 3 * NE;
    const double *mg = mu + e0 * 3 * NE;
    // the stage input state for the c-derivative, loaded with the other tiles so that its latency
    // is hidden behind P1-P5
    const double *qg = (qs && cpart && apply) ? qs + e0 * 3 * NE : nullptr;
#pragma unroll
    for (int it = 0; it < QPT; ++it) {
        const int r = tid + it * T;
        const bool in = r < nel * 3 * NE;
        const int rr = in ? r : 0;  // in-range address for every lane, value selected below
        const double lv = lg[rr];
        const double mv = (flags & kAdjMuIn) ? mg[rr] : 0.0;
        lreg[it] = in ? lv : 0.0;
        mreg[it] = in ? mv : 0.0;
    }
#pragma unroll
    for (int it = 0; it < QPT; ++it) {
        const int r = tid + it * T;
        const bool in = r < nel * 3 * NE;
        const double qv = qg ? qg[in ? r : 0] : 0.0;
        qreg[it] = in ? qv : 0.0;
    }
    for (int r = tid; r < nel * 8; r += T) s_v[r] = verts[e0 * 8 + r];

    // P2 (needs only this thread's registers, so it shares P0's barrier): mu~ = mu + gs lambda,
    // nu = -c mu~ (J-scaled variables) to s_n; mu' = a_s mu~ (stored). Lanes past the last element
    // hold zeros and store nothing.
    if (apply) {
        double *mo = mu + e0 * 3 * NE;
#pragma unroll
        for (int it = 0; it < QPT; ++it) {
            const int r = tid + it * T;
            const double mt = mreg[it] + gs * lreg[it];
            mreg[it] = as * mt;
            if (r < nel * 3 * NE) {
                if (flags & kAdjMuOut) mo[r] = mreg[it];
                s_n[r] = -cw * mt;
            }
        }
    }
    __syncthreads();

    // P1: metric lines, side normals and edge Jacobians (as the forward stage).
    if (active) {
        const double *P = s_v + el * 8;
        double *met = s_met + el * 4 * NF;
        if (f == 0) {
            const double xi = x[l];
            met[l] = ((1 - xi) * (P[6] - P[0]) + (1 + xi) * (P[4] - P[2])) / 4;
            met[NF + l] = ((1 - xi) * (P[7] - P[1]) + (1 + xi) * (P[5] - P[3])) / 4;
        } else if (f == 1) {
            const double eta = x[l];
            met[2 * NF + l] = ((1 - eta) * (P[2] - P[0]) + (1 + eta) * (P[4] - P[6])) / 4;
            met[3 * NF + l] = ((1 - eta) * (P[3] - P[1]) + (1 + eta) * (P[5] - P[7])) / 4;
        } else {  // sides l, l + N + 1, ... (N + 1 lanes cover the 4 sides)
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
    }

    if (apply) {
        // P3 (same phase as P1, which writes only the metric and side arrays): per field f, the lift transposes rho of the four sides and the volume transpose
        // Fbar_f[a][l] = sum_i Dh[i][a] nu_f[i][l] (column l, into s_a); row l of nu_f stays in
        // registers for P4. One contraction per field and direction: the same work in each wave.
        double row[NF];
        if (active) {
            const double *nb = s_n + el * 3 * NE + f * NE;
            double col[NF];
#pragma unroll
            for (int m = 0; m < NF; ++m) {
                col[m] = nb[m * NF + l];
                row[m] = nb[l * NF + m];
            }
            double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
#pragma unroll
            for (int m = 0; m < NF; ++m) {
                r1 += lp[m] * winv[m] * col[m];
                r3 += lm[m] * winv[m] * col[m];
                r0 += lm[m] * winv[m] * row[m];
                r2 += lp[m] * winv[m] * row[m];
            }
            double *rho = s_rho + el * 12 * NF + f * NF;
            rho[l] = r0;
            rho[3 * NF + l] = r1;
            rho[6 * NF + (N - l)] = r2;
            rho[9 * NF + (N - l)] = r3;
            double *fb = s_a + el * 3 * NE + f * NE + l;
#pragma unroll 1
            for (int a = 0; a < NF; ++a) {
                double F = 0.0;
#pragma unroll
                for (int i = 0; i < NF; ++i) F += DhT[a * NF + i] * col[i];
                fb[a * NF] = F;
            }
        }
        __syncthreads();

        // P4: Gbar_f[l][b] = sum_j Dh[j][b] nu_f[l][j] in place of row l of nu_f (only this
        // thread reads that row from here on); flux transpose at the side points: own trace
        // adjoint tau = (omega_L, n omega_L) plus the neighbour's export for this point.
        if (active) {
            double *gs_ = s_n + el * 3 * NE + f * NE + l * NF;
#pragma unroll 1
            for (int b = 0; b < NF; ++b) {
                double G = 0.0;
#pragma unroll
                for (int j = 0; j < NF; ++j) G += DhT[b * NF + j] * row[j];
                gs_[b] = G;
            }
        }
#pragma unroll
        for (int it = 0; it < FPT; ++it) {
            const int item = tid + it * T;
            if (item < nel * 4 * NF) {
                const int fe = item / (4 * NF), s = (item / NF) % 4, k = item % NF;
                double *rho = s_rho + fe * 12 * NF + s * 3 * NF + k;
                const double *sd = s_side + (fe * 4 + s) * 3;
                const double nx = sd[0], ny = sd[1], sf = sd[2];
                const double wl = sf * ((rho[0] + nx * rho[NF] + ny * rho[2 * NF]) / 2);
                rho[0] = wl + nbv[it][0];
                rho[NF] = nx * wl + nbv[it][1];
                rho[2 * NF] = ny * wl + nbv[it][2];
            }
        }
        __syncthreads();

        // P5 (thread owns row i = l of field f): d lambda_f[i][b] from the volume transposes of
        // the fields that f's fluxes depend on, plus the trace transposes:
        //   p: y_eta(i) Fbar_u - x_eta(i) Fbar_v - y_xi(b) Gbar_u + x_xi(b) Gbar_v,
        //   u: y_eta(i) Fbar_p - y_xi(b) Gbar_p,   v: -x_eta(i) Fbar_p + x_xi(b) Gbar_p.
        // After the barrier d lambda goes to s_n (for the c-derivative term) and d lambda / J to
        // s_a (J-scaled variables).
        double da[NF];
        if (active) {
            const int i = l;
            const double *Fb = s_a + el * 3 * NE + i * NF;
            const double *Gb = s_n + el * 3 * NE + i * NF;
            const double *met = s_met + el * 4 * NF;
            const double *tau = s_rho + el * 12 * NF + f * NF;
            const double t0 = tau[i], t2 = tau[6 * NF + (N - i)];
            const double ye = met[NF + i], xe = met[i], lpi = lp[i], lmi = lm[i];
#pragma unroll
            for (int b = 0; b < NF; ++b) {
                const double yx = met[3 * NF + b], xx = met[2 * NF + b];
                double v;
                if (f == 0)
                    v = ye * Fb[NE + b] - xe * Fb[2 * NE + b] - yx * Gb[NE + b] + xx * Gb[2 * NE + b];
                else if (f == 1)
                    v = ye * Fb[b] - yx * Gb[b];
                else
                    v = -(xe * Fb[b]) + xx * Gb[b];
                da[b] = v + lm[b] * t0 + lp[b] * t2 + lpi * tau[3 * NF + b] + lmi * tau[9 * NF + (N - b)];
            }
        }
        __syncthreads();
        if (active) {
            const double *met = s_met + el * 4 * NF;
            const double ye = met[NF + l], xe = met[l];
            double *dst = s_a + el * 3 * NE + f * NE + l * NF, *raw = s_n + el * 3 * NE + f * NE + l * NF;
#pragma unroll
            for (int b = 0; b < NF; ++b) {
                raw[b] = da[b];
                dst[b] = da[b] * recip(met[2 * NF + b] * ye - xe * met[3 * NF + b]);
            }
        }
        __syncthreads();

        // P6 (coalesced): lambda' = lambda + d lambda / J, and the c-derivative term
        // <d lambda, q_s>. Every thread reads only the LDS values at its own positions r from here
        // on, which P7 then overwrites: no barrier until P7's.
        double dot = 0.0;
#pragma unroll
        for (int it = 0; it < QPT; ++it) {
            const int r = tid + it * T;
            if (r < nel * 3 * NE) {
                if (qg) dot += s_n[r] * qreg[it];
                lreg[it] += s_a[r];
            }
        }
        // one slot per wave, one adder per slot and launch: the no-return atomic accumulates the
        // same sum over the launches without waiting for a read of the slot
        if (cpart) {
            for (int off = 32; off > 0; off >>= 1) dot += __shfl_down(dot, off, 64);
            if (lane == 0) unsafeAtomicAdd(cpart + (int64_t)blockIdx.x * 3 + f, dot);
        }
        // misfit source of the step end this stage reaches: lambda_p += res_r w_r / J (the weights
        // come divided by J), added by the thread that holds the node, receivers in order
        if (flags & kAdjSource) {
            for (int rc = 0; rc < n_rec; ++rc) {
                const int64_t le = rec_elem[rc] - e0;
                if (le < 0 || le >= nel) continue;
                const double *w = rec_w + (int64_t)rc * NE;
                const double c = rec_res[rc];
                const int base = (int)le * 3 * NE;  // field p of the element
#pragma unroll
                for (int it = 0; it < QPT; ++it) {
                    const int k = tid + it * T - base;
                    if (k >= 0 && k < NE) lreg[it] += c * w[k];
                }
            }
        }

        double *lo = lam + e0 * 3 * NE;
#pragma unroll
        for (int it = 0; it < QPT; ++it) {
            const int r = tid + it * T;
            if (r < nel * 3 * NE) lo[r] = lreg[it];
        }
    }
    if (!(flags & kAdjExport)) return;

    // P7: nu of the next reverse stage, -c (mu' + gnext lambda') (J-scaled variables).
#pragma unroll
    for (int it = 0; it < QPT; ++it) {
        const int r = tid + it * T;
        const double v = -cw * (mreg[it] + gnext * lreg[it]);
        if (r < nel * 3 * NE) s_n[r] = v;
    }
    __syncthreads();
    // P8: lift transposes of all four sides (thread owns column l and row l).
    if (active) {
        const double *nf_ = s_n + el * 3 * NE + f * NE;
        double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
#pragma unroll
        for (int m = 0; m < NF; ++m) {
            const double col = nf_[m * NF + l], row = nf_[l * NF + m];
            r1 += lp[m] * winv[m] * col;
            r3 += lm[m] * winv[m] * col;
            r0 += lm[m] * winv[m] * row;
            r2 += lp[m] * winv[m] * row;
        }
        double *rho = s_rho + el * 12 * NF + f * NF;
        rho[l] = r0;
        rho[3 * NF + l] = r1;
        rho[6 * NF + (N - l)] = r2;
        rho[9 * NF + (N - l)] = r3;
    }
    __syncthreads();
    // P9: the export chi = (omega_R, -n omega_R) of every side point.
#pragma unroll
    for (int it = 0; it < FPT; ++it) {
        const int item = tid + it * T;
        if (item < nel * 4 * NF) {
            const int fe = item / (4 * NF), s = (item / NF) % 4, k = item % NF;
            const double *rho = s_rho + fe * 12 * NF + s * 3 * NF + k;
            const double *sd = s_side + (fe * 4 + s) * 3;
            const double nx = sd[0], ny = sd[1], sf = sd[2];
            const double wr = sf * ((-rho[0] + nx * rho[NF] + ny * rho[2 * NF]) / 2);
            double *co = chi_out + (e0 + fe) * 12 * NF + s * 3 * NF + k;
            co[0] = wr;
            co[NF] = -(nx * wr);
            co[2 * NF] = -(ny * wr);
        }
    }
}

// Receivers: p at point r = sum over the element's nodes of w_r[node] p[node] (one wave per
// receiver), and the misfit source lambda_p += res_r w_r (one workgroup, receivers in order so
// that receivers sharing an element add without races; w_r divided by J for the J-scaled
// adjoint).
__global__ void __launch_bounds__(64)
record_kernel(const double *__restrict__ q, const int64_t *__restrict__ elem, const double *__restrict__ w,
              int ne, double *__restrict__ out) {
    const int r = blockIdx.x;
    const double *pe = q + elem[r] * 3 * ne;
    const double *wr = w + (int64_t)r * ne;
    double acc = 0.0;
    for (int k = threadIdx.x; k < ne; k += 64) acc += wr[k] * pe[k];
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (threadIdx.x == 0) out[r] = acc;
}

// record_kernel on the RK register in the stage kernel's tile order: element e is element
// el = e mod epb of its tile of nel elements, node (i, j) of field 0 at ((j nel + el)(N+1) + i).
__global__ void __launch_bounds__(64)
record_rk_kernel(const double *__restrict__ g, const int64_t *__restrict__ elem, const double *__restrict__ w_t,
                 int nf, int64_t K, int epb, double *__restrict__ out) {
    const int r = blockIdx.x, ne = nf * nf;
    const int64_t e = elem[r], e0 = e - e % epb;
    const int el = (int)(e - e0), nel = (int)min<int64_t>(epb, K - e0);
    const double *tile = g + e0 * 3 * ne;
    const double *wr = w_t + (int64_t)r * ne;
    double acc = 0.0;
    for (int k = threadIdx.x; k < ne; k += 64) {
        const int j = k / nf, i = k % nf;
        acc += wr[k] * tile[(j * nel + el) * nf + i];
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (threadIdx.x == 0) out[r] = acc;
}

__global__ void __launch_bounds__(256)
source_kernel(double *__restrict__ lam, const int64_t *__restrict__ elem, const double *__restrict__ w,
              const double *__restrict__ res, int n, int ne) {
    for (int r = 0; r < n; ++r) {
        double *le = lam + elem[r] * 3 * ne;
        const double *wr = w + (int64_t)r * ne;
        const double c = res[r];
        for (int k = threadIdx.x; k < ne; k += 256) le[k] += c * wr[k];
        __syncthreads();
    }
}

// x[e][f][i*(N+1)+j] *= J(i, j)^power (power = 1 or -1) for every element of the single order
// group, J from the metric lines exactly as in the stage kernels; with rec (n_rec receivers, their
// device elements), scales the receiver weight arrays [n_rec][(N+1)^2] instead (one field).
template <int N>
__global__ void __launch_bounds__(256)
jacobian_scale_kernel(const double *__restrict__ verts, const double *__restrict__ tab, double *__restrict__ x,
                      int64_t n_blocks, const int64_t *__restrict__ rec_elem, int power) {
    constexpr int NF = N + 1, NE = NF * NF;
    const double *xn = tab + NF * NF + 3 * NF;  // nodes (table layout: Dhat, l(-1), l(+1), w, x, ...)
    const int64_t blk = blockIdx.x;
    if (blk >= n_blocks) return;
    const int64_t e = rec_elem ? rec_elem[blk] : blk;
    const int nf = rec_elem ? 1 : 3;
    const double *P = verts + 8 * e;
    double *xe = x + blk * nf * NE;
    for (int r = threadIdx.x; r < nf * NE; r += 256) {
        const int node = r % NE, i = node / NF, j = node % NF;
        const double xi = xn[i], eta = xn[j];
        const double xeta = ((1 - xi) * (P[6] - P[0]) + (1 + xi) * (P[4] - P[2])) / 4;
        const double yeta = ((1 - xi) * (P[7] - P[1]) + (1 + xi) * (P[5] - P[3])) / 4;
        const double xxi = ((1 - eta) * (P[2] - P[0]) + (1 + eta) * (P[4] - P[6])) / 4;
        const double yxi = ((1 - eta) * (P[3] - P[1]) + (1 + eta) * (P[5] - P[7])) / 4;
        const double jac = xxi * yeta - xeta * yxi;
        xe[r] = power > 0 ? xe[r] * jac : xe[r] * recip(jac);
    }
}

template <int N>
static hipError_t jacobian_scale_n(const DeviceMesh &m, double *x, int64_t n, const int64_t *rec_elem, int power,
                                   hipStream_t s) {
    hipLaunchKernelGGL(jacobian_scale_kernel<N>, dim3((unsigned)n), dim3(256), 0, s, m.verts, m.groups[0].tab, x, n,
                       rec_elem, power);
    return hipGetLastError();
}

template <int N>
static hipError_t adjoint_n(DeviceMesh &m, const AdjointArgs &A, hipStream_t s) {
    using S = LineShape<N>;
    const Group &G = m.groups[0];
    const int64_t blocks = (G.K + S::EPB - 1) / S::EPB;
    hipLaunchKernelGGL(adjoint_stage_kernel<N>, dim3((unsigned)blocks), dim3(S::T), 0, s, m.verts, m.side_nb, G.tab,
                       A.lam, A.mu, A.chi_in, A.chi_out, A.qs, A.cpart, G.K, A.flags, A.gs, A.as, A.gnext,
                       m.wave_speed, A.rec_elem, A.rec_w, A.rec_res, A.n_rec);
    return hipGetLastError();
}

#define ADJ_CASES(F, ...)                                                                                    \
    case 1: return F<1>(__VA_ARGS__); case 2: return F<2>(__VA_ARGS__); case 3: return F<3>(__VA_ARGS__);   \
    case 4: return F<4>(__VA_ARGS__); case 5: return F<5>(__VA_ARGS__); case 6: return F<6>(__VA_ARGS__);   \
    case 7: return F<7>(__VA_ARGS__); case 8: return F<8>(__VA_ARGS__); case 9: return F<9>(__VA_ARGS__);   \
    case 10: return F<10>(__VA_ARGS__); case 11: return F<11>(__VA_ARGS__); case 12: return F<12>(__VA_ARGS__); \
    case 13: return F<13>(__VA_ARGS__); case 14: return F<14>(__VA_ARGS__); case 15: return F<15>(__VA_ARGS__); \
    case 16: return F<16>(__VA_ARGS__);

hipError_t launch_adjoint_stage(DeviceMesh &m, const AdjointArgs &A, hipStream_t s) {
    if (m.groups.size() != 1) return hipErrorInvalidValue;
    switch (m.groups[0].N) { ADJ_CASES(adjoint_n, m, A, s) }
    return hipErrorInvalidValue;
}

hipError_t launch_jacobian_scale(const DeviceMesh &m, double *x, const int64_t *rec_elem, int64_t n_rec, int power,
                                 hipStream_t s) {
    if (m.groups.size() != 1) return hipErrorInvalidValue;
    const int64_t n = rec_elem ? n_rec : m.groups[0].K;
    if (n <= 0) return hipSuccess;
    switch (m.groups[0].N) { ADJ_CASES(jacobian_scale_n, m, x, n, rec_elem, power, s) }
    return hipErrorInvalidValue;
}

int64_t adjoint_blocks(const DeviceMesh &m) {
    const int N = m.groups[0].N;
    const int epb = 64 / (N + 1);
    return 3 * ((m.groups[0].K + epb - 1) / epb);  // one slot per wave
}

hipError_t launch_record(const DeviceMesh &m, const int64_t *elem, const double *w, int n, double *out,
                         hipStream_t s) {
    const int ne = (m.groups[0].N + 1) * (m.groups[0].N + 1);
    hipLaunchKernelGGL(record_kernel, dim3(n), dim3(64), 0, s, m.q, elem, w, ne, out);
    return hipGetLastError();
}

hipError_t launch_record_rk(const DeviceMesh &m, const int64_t *elem, const double *w_t, int n, double *out,
                            hipStream_t s) {
    const Group &G = m.groups[0];
    hipLaunchKernelGGL(record_rk_kernel, dim3(n), dim3(64), 0, s, m.g, elem, w_t, G.N + 1, G.K, stage_tile_epb(G),
                       out);
    return hipGetLastError();
}

hipError_t launch_source(double *lam, const DeviceMesh &m, const int64_t *elem, const double *w,
                         const double *res, int n, hipStream_t s) {
    const int ne = (m.groups[0].N + 1) * (m.groups[0].N + 1);
    hipLaunchKernelGGL(source_kernel, dim3(1), dim3(256), 0, s, lam, elem, w, res, n, ne);
    return hipGetLastError();
}

}  // namespace ndg
