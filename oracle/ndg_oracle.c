/*
 * ndg_oracle.c -- CPU restatement of the reference DG-SEM wave solver.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): loaded by the tests, the smoke check and the CPU-baseline
 * leg of the benchmark, as the checker. Nothing in the product links or calls it.
 *
 * Every routine follows one reference routine (cited as file:line into the reference checkout) in
 * its operation order, so that an x86-64 build without FMA contraction reproduces the reference
 * host path's arithmetic. Build with -ffp-contract=off.
 *
 * Parity pinning (how each part is checked; see tests/test_oracle_*.py):
 *  - tables (nodes, weights, barycentric weights, l(+-1), l'(+-1), D, D^, G^, interpolation,
 *    modal matrix): bitwise against fixtures emitted by the compiled reference sources
 *    (oracle/_ref), plus numpy leggauss, the reference's N=16 known answers and identities.
 *  - element traces, element/boundary geometry, face geometry, estimator, Hilbert order:
 *    bitwise against compiled-reference fixtures.
 *  - upwind flux: closed form (characteristic decomposition) on random states.
 *  - volume + lift (compute_rhs): free-stream preservation on jittered meshes, exactness of the
 *    discrete operator on polynomial fields, linearity, antisymmetry of the face coupling.
 *  - stage + time loop: the reference plane wave with inflow boundaries against its exact
 *    solution, spectral convergence in N, the periodic Fourier-series exact solution for radial
 *    pulses, energy decay, and the RK3 amplification factor on a decoupled mode.
 *  - mortar / p-nonconforming branches: L2-projection exactness for polynomials of degree
 *    <= min(N_face, N_element) and conservation of the face flux integral.
 *  The reference's mesh reader and stage phases (src/meshes/Mesh2D_t.cpp, src/solvers/
 *  Solver2D_t.cpp) cannot be compiled here (they include VTK headers absent from the image), so
 *  the stage-level functions are pinned by the mathematics above rather than by fixtures.
 */
#include "ndg_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* include/helpers/constants.h:15-19 */
static const double C_WAVE = 1.0;
static const double XY0_X = -0.2, XY0_Y = -0.2;
static const double K_X = 0.707106781186548, K_Y = 0.707106781186548;
static const double D_WIDTH = 0.120112240878645;
static const int N_LSQ_MAX = 4;
static const double PI_REF = 3.14159265358979323846;

/* ------------------------------------------------------------------------------------------ */
/* Polynomials: src/polynomials/LegendrePolynomial_t.cpp                                       */
/* ------------------------------------------------------------------------------------------ */

/* LegendrePolynomial_t.cpp:8-32 (Kopriva algorithm 22) */
static void legendre_and_derivative(int N, double x, double *L_N, double *L_N_prime) {
    if (N == 0) {
        *L_N = 1.0;
        *L_N_prime = 0.0;
    } else if (N == 1) {
        *L_N = x;
        *L_N_prime = 1.0;
    } else {
        double L_N_2 = 1.0, L_N_1 = x, L_N_2_prime = 0.0, L_N_1_prime = 1.0;
        for (int k = 2; k <= N; ++k) {
            *L_N = (2 * k - 1) * x * L_N_1 / k - (k - 1) * L_N_2 / k;
            *L_N_prime = L_N_2_prime + (2 * k - 1) * L_N_1;
            L_N_2 = L_N_1;
            L_N_1 = *L_N;
            L_N_2_prime = L_N_1_prime;
            L_N_1_prime = *L_N_prime;
        }
    }
}

/* LegendrePolynomial_t.cpp:34-58 */
static double legendre(int N, double x) {
    if (N == 0) return 1.0;
    if (N == 1) return x;
    double L_N_2 = 1.0, L_N_1 = x, L_N = 0.0;
    for (int k = 2; k <= N; ++k) {
        L_N = (2 * k - 1) * x * L_N_1 / k - (k - 1) * L_N_2 / k;
        L_N_2 = L_N_1;
        L_N_1 = L_N;
    }
    return L_N;
}

/* LegendrePolynomial_t.cpp:60-101 (Kopriva algorithm 23): Newton on L_{N+1}, at most 1000
 * iterations, symmetric fill, centre node 0 for even N. */
static void nodes_and_weights(int N, double *nodes, double *weights) {
    for (int i = 0; i < (N + 1) / 2; ++i) {
        if (N == 1) {
            nodes[0] = -sqrt(1.0 / 3.0);
            weights[0] = 1.0;
            nodes[1] = -nodes[0];
            weights[1] = weights[0];
        } else {
            nodes[i] = -cos(PI_REF * (2 * i + 1) / (2 * N + 2));
            for (int k = 0; k < 1000; ++k) {
                double L, Lp;
                legendre_and_derivative(N + 1, nodes[i], &L, &Lp);
                const double delta = -L / Lp;
                nodes[i] += delta;
                if (fabs(delta) <= DBL_MIN * 2 * fabs(nodes[i])) break;
            }
            double dummy, Lp_final;
            legendre_and_derivative(N + 1, nodes[i], &dummy, &Lp_final);
            nodes[N - i] = -nodes[i];
            weights[i] = 2.0 / ((1 - pow(nodes[i], 2)) * pow(Lp_final, 2));
            weights[N - i] = weights[i];
        }
    }
    if (N == 0) {
        nodes[0] = 0.0;
        weights[0] = 2.0;
    }
    if (N % 2 == 0) {
        double dummy, Lp_final;
        legendre_and_derivative(N + 1, 0.0, &dummy, &Lp_final);
        nodes[N / 2] = 0.0;
        weights[N / 2] = 2 / pow(Lp_final, 2);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* NDG tables: include/entities/NDG_t.tpp                                                      */
/* ------------------------------------------------------------------------------------------ */

struct o_tables {
    int N_max, n_interp;
    double **tab[12];
};

/* NDG_t.tpp:193-199 and src/meshes/Mesh2D_t.cpp:3784-3790 (cppreference almost_equal, 2 ulp) */
static int almost_equal(double x, double y) {
    return fabs(x - y) <= DBL_EPSILON * fabs(x + y) * 2 || fabs(x - y) < DBL_MIN;
}

static int tab_len(const o_tables *t, int which, int N) {
    const int nf = N + 1;
    if (which == O_D || which == O_GHAT || which == O_DHAT || which == O_POLY) return nf * nf;
    if (which == O_INTERP) return nf * t->n_interp;
    return nf;
}

o_tables *o_tables_create(int N_max, int n_interp) {
    o_tables *t = (o_tables *)calloc(1, sizeof(o_tables));
    t->N_max = N_max;
    t->n_interp = n_interp;
    for (int w = 0; w < 12; ++w) {
        t->tab[w] = (double **)calloc((size_t)N_max + 1, sizeof(double *));
        for (int N = 0; N <= N_max; ++N) t->tab[w][N] = (double *)calloc((size_t)tab_len(t, w, N), sizeof(double));
    }
    /* NDG_t.tpp:42-58: the constructor's per-N sequence */
    for (int N = 0; N <= N_max; ++N) {
        const int nf = N + 1;
        double *x = t->tab[O_NODES][N], *w = t->tab[O_WEIGHTS][N], *lam = t->tab[O_BARY][N];
        nodes_and_weights(N, x, w);
        /* NDG_t.tpp:175-188 calculate_barycentric_weights */
        for (int j = 0; j <= N; ++j) {
            double xjxi = 1.0;
            for (int i = 0; i < j; ++i) xjxi *= x[j] - x[i];
            for (int i = j + 1; i <= N; ++i) xjxi *= x[j] - x[i];
            lam[j] = 1.0 / xjxi;
        }
        /* NDG_t.tpp:327-334 calculate_polynomials */
        double *M = t->tab[O_POLY][N];
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N; ++j) M[i * nf + j] = legendre(i, x[j]) * w[j] * (2 * i + 1) * 0.5;
        /* NDG_t.tpp:251-259 polynomial_derivative_matrices (off-diagonal) */
        double *D = t->tab[O_D][N];
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N; ++j)
                if (i != j) D[i * nf + j] = lam[j] / (lam[i] * (x[i] - x[j]));
        /* NDG_t.tpp:302-325 create_interpolation_matrices */
        double *I = t->tab[O_INTERP][N];
        for (int j = 0; j < n_interp; ++j) {
            int row_has_match = 0;
            const double xc = 2.0 * j / (n_interp - 1) - 1.0;
            for (int k = 0; k <= N; ++k) {
                I[j * nf + k] = 0.0;
                if (almost_equal(xc, x[k])) {
                    I[j * nf + k] = 1.0;
                    row_has_match = 1;
                }
            }
            if (!row_has_match) {
                double total = 0.0;
                for (int k = 0; k <= N; ++k) {
                    I[j * nf + k] = lam[k] / (xc - x[k]);
                    total += I[j * nf + k];
                }
                for (int k = 0; k <= N; ++k) I[j * nf + k] /= total;
            }
        }
        /* NDG_t.tpp:209-249 lagrange interpolants at -1/+1 and their derivatives, normalised */
        double *ll = t->tab[O_LAG_LEFT][N], *lr = t->tab[O_LAG_RIGHT][N];
        double *dl = t->tab[O_DLAG_LEFT][N], *dr = t->tab[O_DLAG_RIGHT][N];
        for (int i = 0; i <= N; ++i) ll[i] = lam[i] / (-1.0 - x[i]);
        for (int i = 0; i <= N; ++i) lr[i] = lam[i] / (1.0 - x[i]);
        for (int i = 0; i <= N; ++i) dl[i] = lam[i] / ((-1.0 - x[i]) * (-1.0 - x[i]));
        for (int i = 0; i <= N; ++i) dr[i] = lam[i] / ((1.0 - x[i]) * (1.0 - x[i]));
        double s = 0.0;
        for (int i = 0; i <= N; ++i) s += ll[i];
        for (int i = 0; i <= N; ++i) ll[i] /= s;
        s = 0.0;
        for (int i = 0; i <= N; ++i) s += lr[i];
        for (int i = 0; i <= N; ++i) lr[i] /= s;
        s = 0.0;
        for (int i = 0; i <= N; ++i) s += lam[i] / (-1.0 - x[i]);
        for (int i = 0; i <= N; ++i) dl[i] /= s;
        s = 0.0;
        for (int i = 0; i <= N; ++i) s += lam[i] / (1.0 - x[i]);
        for (int i = 0; i <= N; ++i) dr[i] /= s;
        /* NDG_t.tpp:262-274 diagonal by negative row sum */
        for (int i = 0; i <= N; ++i) {
            D[i * (N + 2)] = 0.0;
            for (int j = 0; j < i; ++j) D[i * (N + 2)] -= D[i * nf + j];
            for (int j = i + 1; j <= N; ++j) D[i * (N + 2)] -= D[i * nf + j];
        }
        /* NDG_t.tpp:286-298 polynomial_cg_derivative_matrices */
        double *Gh = t->tab[O_GHAT][N];
        for (int j = 0; j <= N; ++j)
            for (int n = 0; n <= N; ++n) {
                double sum = 0.0;
                for (int k = 0; k <= N; ++k) sum += D[k * nf + n] * D[k * nf + j] * w[k];
                Gh[j * nf + n] = sum / w[j];
            }
        /* NDG_t.tpp:276-283 polynomial_derivative_matrices_hat */
        double *Dh = t->tab[O_DHAT][N];
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N; ++j) Dh[i * nf + j] = -D[j * nf + i] * w[j] / w[i];
    }
    return t;
}

void o_tables_destroy(o_tables *t) {
    if (!t) return;
    for (int w = 0; w < 12; ++w) {
        for (int N = 0; N <= t->N_max; ++N) free(t->tab[w][N]);
        free(t->tab[w]);
    }
    free(t);
}

int o_tables_get(const o_tables *t, int which, int N, double *out) {
    if (which < 0 || which > 11 || N < 0 || N > t->N_max) return -1;
    const int n = tab_len(t, which, N);
    memcpy(out, t->tab[which][N], (size_t)n * sizeof(double));
    return n;
}

/* ------------------------------------------------------------------------------------------ */
/* Geometry functions: include/functions/{quad_map,quad_metrics,analytical_solution}.tpp        */
/* ------------------------------------------------------------------------------------------ */

typedef struct { double x, y; } v2;

/* quad_map.tpp:3-8 */
static v2 quad_map(double xi, double eta, const v2 *P) {
    v2 r;
    r.x = (P[0].x * (1 - xi) * (1 - eta) + P[1].x * (xi + 1) * (1 - eta) + P[2].x * (xi + 1) * (eta + 1) +
           P[3].x * (1 - xi) * (eta + 1)) / 4;
    r.y = (P[0].y * (1 - xi) * (1 - eta) + P[1].y * (xi + 1) * (1 - eta) + P[2].y * (xi + 1) * (eta + 1) +
           P[3].y * (1 - xi) * (eta + 1)) / 4;
    return r;
}

/* quad_metrics.tpp:3-8: m[0] = {dx/dxi, dx/deta}, m[1] = {dy/dxi, dy/deta} */
static void quad_metrics(double xi, double eta, const v2 *P, v2 *m) {
    m[0].x = ((1 - eta) * (P[1].x - P[0].x) + (1 + eta) * (P[2].x - P[3].x)) / 4;
    m[0].y = ((1 - xi) * (P[3].x - P[0].x) + (1 + xi) * (P[2].x - P[1].x)) / 4;
    m[1].x = ((1 - eta) * (P[1].y - P[0].y) + (1 + eta) * (P[2].y - P[3].y)) / 4;
    m[1].y = ((1 - xi) * (P[3].y - P[0].y) + (1 + xi) * (P[2].y - P[1].y)) / 4;
}

/* analytical_solution.tpp:5-11 */
static void g_exact(double x, double y, double t, double *q) {
    const double a = K_X * (x - XY0_X) + K_Y * (y - XY0_Y) - C_WAVE * t;
    const double p = exp(-a * a / (D_WIDTH * D_WIDTH));
    q[0] = p;
    q[1] = p * K_X / C_WAVE;
    q[2] = p * K_Y / C_WAVE;
}

/* The plane wave of the speed-c system (the reference's g at c = 1): argument k.(x - x0) - c t,
 * velocity p k. */
static void g_exact_c(double x, double y, double t, double c, double *q) {
    if (c == 1.0) {
        g_exact(x, y, t, q);
        return;
    }
    const double a = K_X * (x - XY0_X) + K_Y * (y - XY0_Y) - c * t;
    const double p = exp(-a * a / (D_WIDTH * D_WIDTH));
    q[0] = p;
    q[1] = p * K_X;
    q[2] = p * K_Y;
}

static double mag(v2 a) { return sqrt(a.x * a.x + a.y * a.y); }
static v2 sub(v2 a, v2 b) { v2 r = {a.x - b.x, a.y - b.y}; return r; }

/* ------------------------------------------------------------------------------------------ */
/* Mesh model: Element2D_t / Face2D_t (include/entities/Element2D_t.h:12-127,                   */
/* include/entities/Face2D_t.h:18-83)                                                          */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int N;
    int64_t nodes[4];
    int nf[4];             /* number of faces per side */
    int64_t *faces[4];
    double *block;         /* one allocation for all arrays below */
    double *dxi_dx, *deta_dx, *dxi_dy, *deta_dy, *jac; /* Ne each */
    double *sf[4];
    double dmin, cx, cy;
    double *p, *u, *v, *Gp, *Gu, *Gv, *ip, *iu, *iv;   /* Ne each */
    double *pe[4], *ue[4], *ve[4];                    /* traces */
    double *pf[4], *uf[4], *vf[4];                    /* flux traces */
    double *pflux, *uflux, *vflux, *pder, *uder, *vder; /* line scratch */
    double error, sigma;
    int refine, coarsen;
    int status, rotation, split_level; /* Hilbert status/rotation (Element2D_t.h:62-66) */
} o_elem;

typedef struct {
    int N;
    int64_t nodes[2];
    int64_t elements[2];
    int64_t sides[2];
    double nx, ny, tx, ty, length, offset[2], scale[2];
    double *block;
    double *p[2], *u[2], *v[2], *pflux, *uflux, *vflux;
} o_face;

struct o_mesh {
    const o_tables *t;
    int64_t n_nodes;
    v2 *xy;
    int64_t K, n_elements_total;
    o_elem *el;
    int64_t F;
    o_face *fa;
    int64_t n_wall, n_sym, n_in, n_out, n_if;
    int64_t *wall, *sym, *in, *out, *if_origin, *if_side, *if_dest;
    double tol_min, tol_max;
    double c; /* wave-speed scale of the right-hand side (1 = the reference) */
    int initial_N;
    int32_t *bar_type;  /* per ghost (bar order): kind and periodic donor bar */
    int64_t *bar_donor;
    int64_t *bar_origin; /* per periodic ghost: the element and side whose trace it copies */
};

static void elem_alloc_domain(o_elem *e) {
    const int nf = e->N + 1, ne = nf * nf;
    const size_t total = (size_t)ne * 14 + (size_t)nf * (4 + 12 + 12 + 6);
    e->block = (double *)calloc(total, sizeof(double));
    double *b = e->block;
    e->dxi_dx = b; b += ne; e->deta_dx = b; b += ne; e->dxi_dy = b; b += ne; e->deta_dy = b; b += ne;
    e->jac = b; b += ne;
    e->p = b; b += ne; e->u = b; b += ne; e->v = b; b += ne;
    e->Gp = b; b += ne; e->Gu = b; b += ne; e->Gv = b; b += ne;
    e->ip = b; b += ne; e->iu = b; b += ne; e->iv = b; b += ne;
    for (int s = 0; s < 4; ++s) { e->sf[s] = b; b += nf; }
    for (int s = 0; s < 4; ++s) { e->pe[s] = b; b += nf; e->ue[s] = b; b += nf; e->ve[s] = b; b += nf; }
    for (int s = 0; s < 4; ++s) { e->pf[s] = b; b += nf; e->uf[s] = b; b += nf; e->vf[s] = b; b += nf; }
    e->pflux = b; b += nf; e->uflux = b; b += nf; e->vflux = b; b += nf;
    e->pder = b; b += nf; e->uder = b; b += nf; e->vder = b; b += nf;
}

/* Element2D_t.cpp:711-733 allocate_boundary_storage: side 0 only */
static void elem_alloc_ghost(o_elem *e) {
    const int nf = e->N + 1;
    e->block = (double *)calloc((size_t)nf * 4, sizeof(double));
    double *b = e->block;
    e->sf[0] = b; b += nf;
    e->pe[0] = b; b += nf; e->ue[0] = b; b += nf; e->ve[0] = b; b += nf;
}

static void face_alloc(o_face *f) {
    const int nf = f->N + 1;
    f->block = (double *)calloc((size_t)nf * 9, sizeof(double));
    double *b = f->block;
    for (int s = 0; s < 2; ++s) { f->p[s] = b; b += nf; f->u[s] = b; b += nf; f->v[s] = b; b += nf; }
    f->pflux = b; b += nf; f->uflux = b; b += nf; f->vflux = b;
}

/* Element2D_t.cpp:743-777 compute_geometry */
static void element_geometry(o_elem *e, const v2 *P, const double *x) {
    const int N = e->N, nf = N + 1;
    for (int i = 0; i <= N; ++i) {
        for (int j = 0; j <= N; ++j) {
            v2 m[2];
            quad_metrics(x[i], x[j], P, m);
            e->dxi_dx[i * nf + j] = m[0].x;
            e->deta_dx[i * nf + j] = m[0].y;
            e->dxi_dy[i * nf + j] = m[1].x;
            e->deta_dy[i * nf + j] = m[1].y;
            e->jac[i * nf + j] = m[0].x * m[1].y - m[0].y * m[1].x;
        }
        v2 mb[2], mr[2], mt[2], ml[2];
        quad_metrics(x[i], -1.0, P, mb);
        quad_metrics(1.0, x[i], P, mr);
        quad_metrics(x[i], 1.0, P, mt);
        quad_metrics(-1.0, x[i], P, ml);
        e->sf[0][i] = sqrt(mb[0].x * mb[0].x + mb[1].x * mb[1].x);
        e->sf[1][i] = sqrt(mr[0].y * mr[0].y + mr[1].y * mr[1].y);
        e->sf[2][i] = sqrt(mt[0].x * mt[0].x + mt[1].x * mt[1].x);
        e->sf[3][i] = sqrt(ml[0].y * ml[0].y + ml[1].y * ml[1].y);
    }
    const double a = fmin(mag(sub(P[1], P[0])), mag(sub(P[2], P[3])));
    const double b = fmin(mag(sub(P[1], P[2])), mag(sub(P[0], P[3])));
    const double c = fmin(mag(sub(P[1], P[3])), mag(sub(P[2], P[0])));
    e->dmin = fmin(fmin(a, b), c);
    e->cx = (P[0].x + P[1].x + P[2].x + P[3].x) / 4;
    e->cy = (P[0].y + P[1].y + P[2].y + P[3].y) / 4;
}

/* Element2D_t.cpp:779-790 compute_boundary_geometry */
static void boundary_geometry(o_elem *e, const v2 *P, const double *x) {
    for (int i = 0; i <= e->N; ++i) {
        v2 mb[2];
        quad_metrics(x[i], -1.0, P, mb);
        e->sf[0][i] = sqrt(mb[0].x * mb[0].x + mb[1].x * mb[1].x);
    }
    e->dmin = mag(sub(P[1], P[0]));
    e->cx = (P[0].x + P[1].x) / 2;
    e->cy = (P[0].y + P[1].y) / 2;
}

/* Face2D_t.cpp:56-81 compute_geometry. en = {L side start, L side end, R side start, R side end} */
static void face_geometry(o_face *f, const v2 *fn, const v2 *en) {
    v2 t = sub(fn[1], fn[0]);
    f->length = mag(t);
    t.x /= f->length;
    t.y /= f->length;
    v2 n = {t.y, -t.x};
    const v2 lt = sub(en[1], en[0]);
    const double sign = copysign(1.0, t.x * lt.x + t.y * lt.y);
    n.x *= sign; n.y *= sign;
    t.x *= sign; t.y *= sign;
    f->nx = n.x; f->ny = n.y; f->tx = t.x; f->ty = t.y;
    const v2 d0 = sub(en[1], en[0]), d1 = sub(en[3], en[2]);
    const v2 fd = sub(fn[1], fn[0]);
    const v2 ed0 = sub(fn[0], en[0]), ed1 = sub(fn[1], en[2]);
    f->offset[0] = mag(ed0) / mag(d0);
    f->offset[1] = mag(ed1) / mag(d1);
    f->scale[0] = mag(fd) / mag(d0);
    f->scale[1] = mag(fd) / mag(d1);
}

static void face_add_to_node(int64_t **lists, int *counts, int *caps, int64_t node, int64_t face) {
    if (counts[node] == caps[node]) {
        caps[node] = caps[node] ? 2 * caps[node] : 4;
        lists[node] = (int64_t *)realloc(lists[node], (size_t)caps[node] * sizeof(int64_t));
    }
    lists[node][counts[node]++] = face;
}

static void push(int64_t **arr, int64_t *n, int64_t v) {
    *arr = (int64_t *)realloc(*arr, (size_t)(*n + 1) * sizeof(int64_t));
    (*arr)[(*n)++] = v;
}

/* ------------------------------------------------------------------------------------------ */
/* Hilbert curve statuses (src/functions/Hilbert_splitting.cpp; Status H, A, R, B = 0..3)       */
/* ------------------------------------------------------------------------------------------ */

/* Hilbert_splitting.cpp:3-6, :8-11, :13-21 */
static int first_status(int outgoing) { static const int t[4] = {3, 0, 1, 2}; return t[outgoing]; }
static int last_status(int incoming) { static const int t[4] = {1, 2, 3, 0}; return t[incoming]; }
static int mid_status(int incoming, int outgoing) {
    static const int t[4][4] = {{0, 0, 1, 1}, {3, 3, 2, 2}, {3, 3, 2, 2}, {0, 0, 1, 1}};
    return t[incoming][outgoing];
}
/* Hilbert_splitting.cpp:23-49 child_order[rotation][status] */
static const int CHILD_ORDER[4][4][4] = {
    {{0, 3, 2, 1}, {0, 1, 2, 3}, {2, 1, 0, 3}, {2, 3, 0, 1}},
    {{3, 2, 1, 0}, {1, 2, 3, 0}, {1, 0, 3, 2}, {3, 0, 1, 2}},
    {{2, 1, 0, 3}, {2, 3, 0, 1}, {0, 3, 2, 1}, {0, 1, 2, 3}},
    {{1, 0, 3, 2}, {3, 0, 1, 2}, {3, 2, 1, 0}, {1, 2, 3, 0}}};
/* Hilbert_splitting.cpp:51-81 child_statuses[rotation][status] (H A R B = 0 1 2 3) */
static const int CHILD_STATUS[4][4][4] = {
    {{1, 3, 0, 0}, {0, 1, 1, 2}, {2, 2, 3, 1}, {3, 0, 2, 3}},
    {{3, 0, 0, 1}, {1, 1, 2, 0}, {2, 3, 1, 2}, {0, 2, 3, 3}},
    {{0, 0, 1, 3}, {1, 2, 0, 1}, {3, 1, 2, 2}, {2, 3, 3, 0}},
    {{0, 1, 3, 0}, {2, 0, 1, 1}, {1, 2, 2, 3}, {3, 3, 0, 2}}};

/* The tables above, copied out for the fixture comparison: child order and child statuses
 * [rotation][status][4], first / last element status [side], inner status [incoming][outgoing]. */
void o_hilbert_tables(int *child_order, int *child_status, int *first, int *last, int *inner) {
    for (int r = 0; r < 4; ++r)
        for (int s = 0; s < 4; ++s)
            for (int k = 0; k < 4; ++k) {
                child_order[(r * 4 + s) * 4 + k] = CHILD_ORDER[r][s][k];
                child_status[(r * 4 + s) * 4 + k] = CHILD_STATUS[r][s][k];
            }
    for (int s = 0; s < 4; ++s) {
        first[s] = first_status(s);
        last[s] = last_status(s);
        for (int o = 0; o < 4; ++o) inner[s * 4 + o] = mid_status(s, o);
    }
}

/* side through which the neighbouring centre d lies (Mesh2D_t.cpp:4083-4100) */
static int exit_side(double dx, double dy) {
    if (fabs(dy) > fabs(dx)) return dy < 0 ? 0 : 2;
    return dx < 0 ? 3 : 1;
}

/* src/meshes/Mesh2D_t.cpp:4055-4178 compute_element_status */
static void element_status(o_mesh *m) {
    static const double PI = 3.14159265358979323846;
    const double targets[4] = {-3 * PI / 4, -PI / 4, PI / 4, 3 * PI / 4};
    const int64_t K = m->K;
    for (int64_t i = 0; i < K; ++i) {
        o_elem *e = &m->el[i];
        long rotation = 0;
        for (int k = 0; k < 4; ++k) {
            const double dx = m->xy[e->nodes[k]].x - e->cx, dy = m->xy[e->nodes[k]].y - e->cy;
            const double angle = atan2(dy, dx);
            const double offset = atan2(sin(angle - targets[k]), cos(angle - targets[k]));
            const long n_turns = lround(offset * 2 / PI + 4 * (offset < 0));
            rotation += n_turns * (n_turns < 4);
        }
        rotation /= 4;
        e->rotation = (int)rotation;
        if (K == 1) {
            e->status = 0;
        } else if (i == 0) {
            e->status = first_status(exit_side(m->el[1].cx - e->cx, m->el[1].cy - e->cy));
        } else if (i == K - 1) {
            e->status = last_status(exit_side(m->el[i - 1].cx - e->cx, m->el[i - 1].cy - e->cy));
        } else {
            const int in = exit_side(m->el[i - 1].cx - e->cx, m->el[i - 1].cy - e->cy);
            const int out = exit_side(m->el[i + 1].cx - e->cx, m->el[i + 1].cy - e->cy);
            e->status = mid_status(in, out);
        }
        e->split_level = 0;
    }
}

/* Mesh creation from nodes, quads and boundary bars. A periodic bar names its donor bar (the bar
 * on the other end of the period, at mesh creation) or, when bar_origin is given, the element and
 * side whose trace it copies (after a split, where one whole bar faces two half bars). A periodic
 * ghost may be half of its element's side or cover two elements' sides (2:1 across the period). */
static o_mesh *mesh_create(const o_tables *t, int initial_N, int64_t n_nodes, const double *xy, int64_t n_quads,
                           const int64_t *quads, const int32_t *quad_N, int64_t n_bars, const int64_t *bars,
                           const int32_t *bar_type, const int64_t *bar_donor, const int64_t *bar_origin,
                           double tol_min, double tol_max) {
    o_mesh *m = (o_mesh *)calloc(1, sizeof(o_mesh));
    m->t = t;
    m->tol_min = tol_min;
    m->tol_max = tol_max;
    m->c = 1.0;
    m->initial_N = initial_N;
    m->bar_type = (int32_t *)malloc((size_t)(n_bars + 1) * sizeof(int32_t));
    m->bar_donor = (int64_t *)malloc((size_t)(n_bars + 1) * sizeof(int64_t));
    m->bar_origin = (int64_t *)malloc((size_t)(2 * n_bars + 2) * sizeof(int64_t));
    for (int64_t b = 0; b < n_bars; ++b) {
        m->bar_type[b] = bar_type[b];
        m->bar_donor[b] = bar_donor ? bar_donor[b] : -1;
        m->bar_origin[2 * b] = bar_origin ? bar_origin[2 * b] : -1;
        m->bar_origin[2 * b + 1] = bar_origin ? bar_origin[2 * b + 1] : -1;
    }
    m->n_nodes = n_nodes;
    m->xy = (v2 *)malloc((size_t)n_nodes * sizeof(v2));
    for (int64_t i = 0; i < n_nodes; ++i) { m->xy[i].x = xy[2 * i]; m->xy[i].y = xy[2 * i + 1]; }
    m->K = n_quads;
    m->n_elements_total = n_quads + n_bars;
    m->el = (o_elem *)calloc((size_t)m->n_elements_total, sizeof(o_elem));
    /* src/meshes/Mesh2D_t.cpp:676-707: domain elements, then ghost elements with nodes {a,b,b,a} */
    for (int64_t i = 0; i < n_quads; ++i) {
        o_elem *e = &m->el[i];
        e->N = quad_N ? quad_N[i] : initial_N;
        for (int k = 0; k < 4; ++k) e->nodes[k] = quads[4 * i + k];
    }
    for (int64_t b = 0; b < n_bars; ++b) {
        o_elem *e = &m->el[n_quads + b];
        e->N = initial_N;
        e->nodes[0] = bars[2 * b];
        e->nodes[1] = bars[2 * b + 1];
        e->nodes[2] = bars[2 * b + 1];
        e->nodes[3] = bars[2 * b];
    }
    /* src/meshes/Mesh2D_t.cpp:1110-1190 build_faces */
    int64_t **n2f = (int64_t **)calloc((size_t)n_nodes, sizeof(int64_t *));
    int *n2f_count = (int *)calloc((size_t)n_nodes, sizeof(int));
    int *n2f_cap = (int *)calloc((size_t)n_nodes, sizeof(int));
    int64_t cap = 2 * n_quads + n_bars + 4;
    m->fa = (o_face *)calloc((size_t)cap, sizeof(o_face));
    int64_t (*e2f)[4] = (int64_t(*)[4])malloc((size_t)m->n_elements_total * sizeof(int64_t[4]));
    m->F = 0;
    for (int64_t i = 0; i < n_quads; ++i) {
        for (int j = 0; j < 4; ++j) {
            const int64_t a = m->el[i].nodes[j], b = m->el[i].nodes[(j + 1) % 4];
            int found = 0;
            for (int k = 0; k < n2f_count[a]; ++k) {
                o_face *f = &m->fa[n2f[a][k]];
                if (f->nodes[0] == b && f->nodes[1] == a) {
                    found = 1;
                    f->elements[1] = i;
                    f->sides[1] = j;
                    e2f[i][j] = n2f[a][k];
                    break;
                }
            }
            if (!found) {
                if (m->F == cap) {
                    cap *= 2;
                    m->fa = (o_face *)realloc(m->fa, (size_t)cap * sizeof(o_face));
                }
                e2f[i][j] = m->F;
                face_add_to_node(n2f, n2f_count, n2f_cap, a, m->F);
                if (b != a) face_add_to_node(n2f, n2f_count, n2f_cap, b, m->F);
                o_face *f = &m->fa[m->F++];
                memset(f, 0, sizeof(*f));
                f->N = initial_N;
                f->nodes[0] = a; f->nodes[1] = b;
                f->elements[0] = i; f->elements[1] = -1;
                f->sides[0] = j; f->sides[1] = -1;
            }
        }
    }
    for (int64_t i = n_quads; i < m->n_elements_total; ++i) {
        o_elem *e = &m->el[i];
        const int64_t a = e->nodes[0], b = e->nodes[1];
        e2f[i][0] = -1;
        for (int k = 0; k < n2f_count[a]; ++k) {
            o_face *f = &m->fa[n2f[a][k]];
            if (f->elements[0] >= n_quads) continue;
            if (f->nodes[0] == b && f->nodes[1] == a) {
                f->elements[1] = i; f->sides[1] = 0; e2f[i][0] = n2f[a][k];
                break;
            }
            if (f->nodes[0] == a && f->nodes[1] == b) {
                f->elements[1] = i; f->sides[1] = 0; e2f[i][0] = n2f[a][k];
                e->nodes[0] = b; e->nodes[1] = a; /* "inside out" boundary element */
                break;
            }
        }
        if (e2f[i][0] < 0 && bar_type[i - n_quads] == O_BC_PERIODIC) {
            /* half of an element side, or two element sides: a face of its own, matched below */
            if (m->F == cap) {
                cap *= 2;
                m->fa = (o_face *)realloc(m->fa, (size_t)cap * sizeof(o_face));
            }
            e2f[i][0] = m->F;
            face_add_to_node(n2f, n2f_count, n2f_cap, a, m->F);
            face_add_to_node(n2f, n2f_count, n2f_cap, b, m->F);
            o_face *f = &m->fa[m->F++];
            memset(f, 0, sizeof(*f));
            f->N = initial_N;
            f->nodes[0] = a; f->nodes[1] = b;
            f->elements[0] = i; f->elements[1] = -1;
            f->sides[0] = 0; f->sides[1] = -1;
        }
    }
    /* Hanging sides: a side a -> b without a neighbour or a ghost (an element side, or a periodic
     * ghost's own face), covered by a chain of unmatched faces b -> m1 -> ... -> a along it, made
     * by the sides of smaller elements or half ghosts (the configuration h-refinement leaves behind,
     * src/meshes/Mesh2D_t.cpp:5661-5985, any number of faces per side). Those faces take the larger
     * side's element as their right element (its side then lists them from a to b) and the larger
     * side's own face is dropped. */
    int64_t *hpool = NULL, hpool_n = 0, hpool_cap = 0;
    int64_t *hstart = (int64_t *)malloc((size_t)(4 * n_quads + n_bars + 1) * sizeof(int64_t));
    int *hcount = (int *)calloc((size_t)(4 * n_quads + n_bars + 1), sizeof(int));
    char *dropped = (char *)calloc((size_t)m->F + 1, 1);
    for (int64_t q = 0; q < 4 * n_quads + n_bars; ++q) hstart[q] = -1;
    for (int64_t i = 0; i < m->n_elements_total; ++i) {
        const int ghost = i >= n_quads;
        if (ghost && bar_type[i - n_quads] != O_BC_PERIODIC) continue;
        for (int j = 0; j < (ghost ? 1 : 4); ++j) {
            const int64_t fb = e2f[i][j];
            if (fb < 0 || m->fa[fb].elements[1] != -1 || m->fa[fb].elements[0] != i) continue;
            const int64_t a = m->fa[fb].nodes[0], b = m->fa[fb].nodes[1];
            const v2 pa = m->xy[a], pb = m->xy[b], ab = sub(pb, pa);
            const double len2 = ab.x * ab.x + ab.y * ab.y, len = sqrt(len2);
            int64_t chain[64];
            int n = 0;
            int64_t x = b;
            double tx = 1.0;
            while (x != a && n < 64) {
                int64_t found = -1;
                double ty = 0.0;
                for (int k = 0; k < n2f_count[x] && found < 0; ++k) {
                    const int64_t fk = n2f[x][k];
                    const o_face *c = &m->fa[fk];
                    if (fk == fb || c->nodes[0] != x || c->elements[1] != -1 || dropped[fk]) continue;
                    const int64_t y = c->nodes[1];
                    const v2 ay = sub(m->xy[y], pa);
                    const double t = y == a ? 0.0 : (ay.x * ab.x + ay.y * ab.y) / len2;
                    const double off = fabs(ay.x * ab.y - ay.y * ab.x) / len;
                    if (y != a && (off > 1e-12 * len || t <= 1e-12 || t >= tx - 1e-12)) continue;
                    found = fk;
                    ty = t;
                }
                if (found < 0) break;
                chain[n++] = found;
                x = m->fa[found].nodes[1];
                tx = ty;
            }
            if (x != a || n < 2) continue;
            if (hpool_n + n > hpool_cap) {
                hpool_cap = 2 * (hpool_n + n) + 64;
                hpool = (int64_t *)realloc(hpool, (size_t)hpool_cap * sizeof(int64_t));
            }
            const int64_t key = ghost ? 4 * n_quads + (i - n_quads) : 4 * i + j;
            hstart[key] = hpool_n;
            hcount[key] = n;
            for (int q = n - 1; q >= 0; --q) {  /* from a to b */
                m->fa[chain[q]].elements[1] = i;
                m->fa[chain[q]].sides[1] = j;
                hpool[hpool_n++] = chain[q];
            }
            dropped[fb] = 1;
        }
    }
    {   /* compact the face array */
        int64_t *remap = (int64_t *)malloc((size_t)(m->F + 1) * sizeof(int64_t));
        int64_t nf = 0;
        for (int64_t f = 0; f < m->F; ++f) {
            remap[f] = dropped[f] ? -1 : nf;
            if (!dropped[f]) m->fa[nf++] = m->fa[f];
        }
        m->F = nf;
        for (int64_t i = 0; i < m->n_elements_total; ++i)
            for (int j = 0; j < (i < n_quads ? 4 : 1); ++j)
                if (e2f[i][j] >= 0) e2f[i][j] = remap[e2f[i][j]];
        for (int64_t q = 0; q < hpool_n; ++q) hpool[q] = remap[hpool[q]];
        free(remap);
    }
    free(dropped);
    for (int64_t i = 0; i < n_nodes; ++i) free(n2f[i]);
    free(n2f); free(n2f_count); free(n2f_cap);

    /* boundary lists (src/meshes/Mesh2D_t.cpp:715-791) and self interfaces (:820-875) */
    for (int64_t b = 0; b < n_bars; ++b) {
        const int64_t g = n_quads + b;
        switch (bar_type[b]) {
            case O_BC_WALL: push(&m->wall, &m->n_wall, g); break;
            case O_BC_SYMMETRY: push(&m->sym, &m->n_sym, g); break;
            case O_BC_INFLOW: push(&m->in, &m->n_in, g); break;
            case O_BC_OUTFLOW: push(&m->out, &m->n_out, g); break;
            case O_BC_PERIODIC: {
                if (bar_origin) {
                    int64_t tmp = m->n_if;
                    push(&m->if_origin, &tmp, bar_origin[2 * b]);
                    tmp = m->n_if;
                    push(&m->if_side, &tmp, bar_origin[2 * b + 1]);
                    push(&m->if_dest, &m->n_if, g);
                    break;
                }
                const int64_t donor_ghost = n_quads + bar_donor[b];
                const int64_t fi = e2f[donor_ghost][0];
                const int side_index = m->fa[fi].elements[0] == donor_ghost;
                int64_t tmp = m->n_if;
                push(&m->if_origin, &tmp, m->fa[fi].elements[side_index]);
                tmp = m->n_if;
                push(&m->if_side, &tmp, m->fa[fi].sides[side_index]);
                push(&m->if_dest, &m->n_if, g);
                break;
            }
            default: break;
        }
    }

    /* orders after (p-)adaptation semantics: ghost N follows its source (src/meshes/Mesh2D_t.cpp:
     * 7482-7520), face N = max(N_L, N_R) (:7522-7538) */
    for (int64_t b = 0; b < n_bars; ++b) {
        const int64_t g = n_quads + b;
        const int64_t fi = e2f[g][0];
        if (fi < 0) continue;
        const int64_t src = m->fa[fi].elements[0];
        int Ng = m->el[g].N;
        if (bar_type[b] == O_BC_PERIODIC) continue;
        if (m->el[src].N > Ng) Ng = m->el[src].N;
        m->el[g].N = Ng;
    }
    for (int64_t k = 0; k < m->n_if; ++k) m->el[m->if_dest[k]].N = m->el[m->if_origin[k]].N;
    for (int64_t f = 0; f < m->F; ++f) {
        o_face *fc = &m->fa[f];
        const int NL = m->el[fc->elements[0]].N;
        const int NR = fc->elements[1] >= 0 ? m->el[fc->elements[1]].N : NL;
        fc->N = NL > NR ? NL : NR;
    }

    /* storage, face lists, geometry (Mesh2D_t.cpp:1017-1030, 4026-4052, 4200-4235) */
    for (int64_t i = 0; i < m->n_elements_total; ++i) {
        o_elem *e = &m->el[i];
        if (i < n_quads) {
            elem_alloc_domain(e);
            for (int s = 0; s < 4; ++s) {
                if (hcount[4 * i + s] > 0) {
                    const int n = hcount[4 * i + s];
                    e->nf[s] = n;
                    e->faces[s] = (int64_t *)malloc((size_t)n * sizeof(int64_t));
                    for (int q = 0; q < n; ++q) e->faces[s][q] = hpool[hstart[4 * i + s] + q];
                    continue;
                }
                e->nf[s] = 1;
                e->faces[s] = (int64_t *)malloc(sizeof(int64_t));
                e->faces[s][0] = e2f[i][s];
            }
            v2 P[4];
            for (int k = 0; k < 4; ++k) P[k] = m->xy[e->nodes[k]];
            element_geometry(e, P, t->tab[O_NODES][e->N]);
        } else {
            elem_alloc_ghost(e);
            const int64_t key = 4 * n_quads + (i - n_quads);
            if (hcount[key] > 0) {
                const int n = hcount[key];
                e->nf[0] = n;
                e->faces[0] = (int64_t *)malloc((size_t)n * sizeof(int64_t));
                for (int q = 0; q < n; ++q) e->faces[0][q] = hpool[hstart[key] + q];
            } else {
                e->nf[0] = 1;
                e->faces[0] = (int64_t *)malloc(sizeof(int64_t));
                e->faces[0][0] = e2f[i][0];
            }
            v2 P[4];
            for (int k = 0; k < 4; ++k) P[k] = m->xy[e->nodes[k]];
            boundary_geometry(e, P, t->tab[O_NODES][e->N]);
        }
    }
    element_status(m);
    for (int64_t f = 0; f < m->F; ++f) {
        o_face *fc = &m->fa[f];
        face_alloc(fc);
        const o_elem *L = &m->el[fc->elements[0]];
        const o_elem *R = &m->el[fc->elements[1]];
        const int sL = (int)fc->sides[0], sR = (int)fc->sides[1];
        v2 fn[2] = {m->xy[fc->nodes[0]], m->xy[fc->nodes[1]]};
        v2 en[4] = {m->xy[L->nodes[sL]], m->xy[L->nodes[(sL + 1) % 4]], m->xy[R->nodes[sR]],
                    m->xy[R->nodes[(sR + 1) % 4]]};
        face_geometry(fc, fn, en);
    }
    free(e2f);
    free(hpool);
    free(hstart);
    free(hcount);
    for (int64_t k = 0; k < m->n_if; ++k) {
        m->bar_origin[2 * (m->if_dest[k] - n_quads)] = m->if_origin[k];
        m->bar_origin[2 * (m->if_dest[k] - n_quads) + 1] = m->if_side[k];
    }
    return m;
}

o_mesh *o_mesh_create(const o_tables *t, int initial_N, int64_t n_nodes, const double *xy,
                      int64_t n_quads, const int64_t *quads, const int32_t *quad_N, int64_t n_bars,
                      const int64_t *bars, const int32_t *bar_type, const int64_t *bar_donor,
                      double tol_min, double tol_max) {
    return mesh_create(t, initial_N, n_nodes, xy, n_quads, quads, quad_N, n_bars, bars, bar_type, bar_donor, NULL,
                       tol_min, tol_max);
}

void o_mesh_destroy(o_mesh *m) {
    if (!m) return;
    for (int64_t i = 0; i < m->n_elements_total; ++i) {
        free(m->el[i].block);
        for (int s = 0; s < 4; ++s) free(m->el[i].faces[s]);
    }
    for (int64_t f = 0; f < m->F; ++f) free(m->fa[f].block);
    free(m->el); free(m->fa); free(m->xy);
    free(m->wall); free(m->sym); free(m->in); free(m->out);
    free(m->if_origin); free(m->if_side); free(m->if_dest);
    free(m->bar_type); free(m->bar_donor); free(m->bar_origin);
    free(m);
}

int64_t o_mesh_n_elements(const o_mesh *m) { return m->K; }
int64_t o_mesh_n_faces(const o_mesh *m) { return m->F; }
void o_mesh_side_faces(const o_mesh *m, int32_t *nf) {
    for (int64_t i = 0; i < m->K; ++i)
        for (int s = 0; s < 4; ++s) nf[4 * i + s] = m->el[i].nf[s];
}
int64_t o_mesh_n_ghosts(const o_mesh *m) { return m->n_elements_total - m->K; }
int64_t o_mesh_n_nodes_total(const o_mesh *m) {
    int64_t n = 0;
    for (int64_t i = 0; i < m->K; ++i) n += (int64_t)(m->el[i].N + 1) * (m->el[i].N + 1);
    return n;
}
void o_mesh_get_N(const o_mesh *m, int32_t *N_out) {
    for (int64_t i = 0; i < m->K; ++i) N_out[i] = m->el[i].N;
}

/* ------------------------------------------------------------------------------------------ */
/* Initial conditions: src/meshes/Mesh2D_t.cpp:1192-1216                                       */
/* ------------------------------------------------------------------------------------------ */

static void ic_common(o_mesh *m, int kind, int n, const double *pulses, double Lx, double Ly) {
    for (int64_t ei = 0; ei < m->K; ++ei) {
        o_elem *e = &m->el[ei];
        const int N = e->N, nf = N + 1;
        const double *x = m->t->tab[O_NODES][N];
        v2 P[4];
        for (int k = 0; k < 4; ++k) P[k] = m->xy[e->nodes[k]];
        for (int i = 0; i <= N; ++i)
            for (int j = 0; j <= N; ++j) {
                const v2 X = quad_map(x[i], x[j], P);
                double q[3];
                if (kind == 0) {
                    g_exact(X.x, X.y, 0.0, q);
                } else if (kind == 2) {
                    q[0] = 0.0; q[1] = 0.0; q[2] = 0.0;
                    for (int k = 0; k < n; ++k) {
                        const double w = pulses[4 * k + 2];
                        for (int mx = -1; mx <= 1; ++mx)
                            for (int my = -1; my <= 1; ++my) {
                                const double dx = X.x - (pulses[4 * k] + mx * Lx);
                                const double dy = X.y - (pulses[4 * k + 1] + my * Ly);
                                q[0] += pulses[4 * k + 3] * exp(-M_LN2 * (dx * dx + dy * dy) / (w * w));
                            }
                    }
                } else {
                    q[0] = 0.0; q[1] = 0.0; q[2] = 0.0;
                    for (int k = 0; k < n; ++k) {
                        const double dx = X.x - pulses[4 * k], dy = X.y - pulses[4 * k + 1];
                        const double w = pulses[4 * k + 2];
                        q[0] += pulses[4 * k + 3] * exp(-M_LN2 * (dx * dx + dy * dy) / (w * w));
                    }
                }
                e->p[i * nf + j] = q[0];
                e->u[i * nf + j] = q[1];
                e->v[i * nf + j] = q[2];
                e->Gp[i * nf + j] = 0; e->Gu[i * nf + j] = 0; e->Gv[i * nf + j] = 0;
                e->ip[i * nf + j] = 0; e->iu[i * nf + j] = 0; e->iv[i * nf + j] = 0;
            }
    }
}

void o_initial_conditions_plane(o_mesh *m) { ic_common(m, 0, 0, NULL, 0.0, 0.0); }
void o_initial_conditions_pulses(o_mesh *m, int n, const double *pulses) { ic_common(m, 1, n, pulses, 0.0, 0.0); }
/* The pulses periodised with periods (Lx, Ly): each pulse and its 8 nearest images (no reference
 * counterpart; pinned against the Fourier series of the periodised pulses). */
void o_initial_conditions_periodic_pulses(o_mesh *m, int n, const double *pulses, double Lx, double Ly) {
    ic_common(m, 2, n, pulses, Lx, Ly);
}

void o_set_state(o_mesh *m, const double *p, const double *u, const double *v) {
    int64_t off = 0;
    for (int64_t i = 0; i < m->K; ++i) {
        o_elem *e = &m->el[i];
        const int ne = (e->N + 1) * (e->N + 1);
        memcpy(e->p, p + off, (size_t)ne * sizeof(double));
        memcpy(e->u, u + off, (size_t)ne * sizeof(double));
        memcpy(e->v, v + off, (size_t)ne * sizeof(double));
        off += ne;
    }
}

static void get3(const o_mesh *m, int which, double *a, double *b, double *c) {
    int64_t off = 0;
    for (int64_t i = 0; i < m->K; ++i) {
        const o_elem *e = &m->el[i];
        const int ne = (e->N + 1) * (e->N + 1);
        memcpy(a + off, which ? e->Gp : e->p, (size_t)ne * sizeof(double));
        memcpy(b + off, which ? e->Gu : e->u, (size_t)ne * sizeof(double));
        memcpy(c + off, which ? e->Gv : e->v, (size_t)ne * sizeof(double));
        off += ne;
    }
}
void o_get_state(const o_mesh *m, double *p, double *u, double *v) { get3(m, 0, p, u, v); }
void o_get_rhs(const o_mesh *m, double *Gp, double *Gu, double *Gv) { get3(m, 1, Gp, Gu, Gv); }

/* Mesh2D_t::get_solution (src/meshes/Mesh2D_t.cpp:4237-4249) and Element2D_t::interpolate_solution
 * (src/entities/Element2D_t.cpp:581-604): per element, x and y by quad_map at the uniform points
 * (i/(n-1))*2-1 and each field as sum_m sum_k q[m][k] I[i][m] I[j][k]. */
void o_get_solution(const o_mesh *m, double *x, double *y, double *p, double *u, double *v) {
    const int n = m->t->n_interp;
    for (int64_t k = 0; k < m->K; ++k) {
        const o_elem *e = &m->el[k];
        const int nf = e->N + 1;
        const double *I = m->t->tab[O_INTERP][e->N];
        const v2 P[4] = {m->xy[e->nodes[0]], m->xy[e->nodes[1]], m->xy[e->nodes[2]], m->xy[e->nodes[3]]};
        const int64_t off = k * (int64_t)n * n;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const v2 g = quad_map((double)i / (double)(n - 1) * 2 - 1, (double)j / (double)(n - 1) * 2 - 1, P);
                const int64_t at = off + (int64_t)i * n + j;
                x[at] = g.x;
                y[at] = g.y;
                double sp = 0.0, su = 0.0, sv = 0.0;
                for (int a = 0; a <= e->N; ++a)
                    for (int b = 0; b <= e->N; ++b) {
                        sp += e->p[a * nf + b] * I[i * nf + a] * I[j * nf + b];
                        su += e->u[a * nf + b] * I[i * nf + a] * I[j * nf + b];
                        sv += e->v[a * nf + b] * I[i * nf + a] * I[j * nf + b];
                    }
                p[at] = sp;
                u[at] = su;
                v[at] = sv;
            }
    }
}

/* The point fields of Mesh2D_t::get_complete_solution (src/meshes/Mesh2D_t.cu:5096-5155) with
 * Element2D_t::interpolate_complete_solution (src/entities/Element2D_t.cu:623-665): those of
 * o_get_solution, the element's G (what the last right-hand side evaluation left there, the
 * reference's G_p_) and |g(x_mk, time) - q_mk| of the plane wave at the solution nodes, each
 * interpolated with I[i][m] I[j][k]. out holds 11 planes of K n n values:
 * x, y, p, u, v, Gp, Gu, Gv, |dp|, |du|, |dv|. Pinned (tests/test_oracle_output.py): the first
 * five planes equal o_get_solution's; the error planes vanish on the plane-wave state at t = 0
 * and approach g itself on a zero state; the G planes hold the constant right-hand side of a
 * linear field on interior elements. */
void o_get_complete_solution(const o_mesh *m, double time, double *out) {
    const int n = m->t->n_interp;
    const int64_t plane = m->K * (int64_t)n * n;
    for (int64_t k = 0; k < m->K; ++k) {
        const o_elem *e = &m->el[k];
        const int nf = e->N + 1;
        const double *I = m->t->tab[O_INTERP][e->N];
        const double *xn = m->t->tab[O_NODES][e->N];
        const v2 P[4] = {m->xy[e->nodes[0]], m->xy[e->nodes[1]], m->xy[e->nodes[2]], m->xy[e->nodes[3]]};
        const double *src[6] = {e->p, e->u, e->v, e->Gp, e->Gu, e->Gv};
        const int64_t off = k * (int64_t)n * n;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const v2 g = quad_map((double)i / (double)(n - 1) * 2 - 1, (double)j / (double)(n - 1) * 2 - 1, P);
                const int64_t at = off + (int64_t)i * n + j;
                double f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
                for (int a = 0; a <= e->N; ++a)
                    for (int b = 0; b <= e->N; ++b) {
                        const double wa = I[i * nf + a], wb = I[j * nf + b];
                        for (int c = 0; c < 6; ++c) f[c] += src[c][a * nf + b] * wa * wb;
                        const v2 xg = quad_map(xn[a], xn[b], P);
                        double q[3];
                        g_exact_c(xg.x, xg.y, time, m->c, q);
                        for (int c = 0; c < 3; ++c) f[6 + c] += fabs(q[c] - src[c][a * nf + b]) * wa * wb;
                    }
                out[at] = g.x;
                out[plane + at] = g.y;
                for (int c = 0; c < 9; ++c) out[(2 + c) * plane + at] = f[c];
            }
    }
}

static void get_sides(const o_mesh *m, int flux, double *out) {
    int64_t off = 0;
    for (int64_t i = 0; i < m->K; ++i) {
        const o_elem *e = &m->el[i];
        const int nf = e->N + 1;
        for (int s = 0; s < 4; ++s) {
            memcpy(out + off, flux ? e->pf[s] : e->pe[s], (size_t)nf * sizeof(double)); off += nf;
            memcpy(out + off, flux ? e->uf[s] : e->ue[s], (size_t)nf * sizeof(double)); off += nf;
            memcpy(out + off, flux ? e->vf[s] : e->ve[s], (size_t)nf * sizeof(double)); off += nf;
        }
    }
}
void o_get_traces(const o_mesh *m, double *out) { get_sides(m, 0, out); }
void o_get_flux_traces(const o_mesh *m, double *out) { get_sides(m, 1, out); }

/* ------------------------------------------------------------------------------------------ */
/* Solver: src/solvers/Solver2D_t.cpp                                                          */
/* ------------------------------------------------------------------------------------------ */

/* Solver2D_t.cpp:19-29 (single rank: the MPI_Allreduce(MIN) is the identity) */
double o_get_delta_t(const o_mesh *m, double cfl) {
    double dt = INFINITY;
    for (int64_t i = 0; i < m->K; ++i) {
        const o_elem *e = &m->el[i];
        dt = fmin(cfl * e->dmin / (e->N * e->N * C_WAVE * m->c), dt);
    }
    return dt;
}

/* Element2D_t.cpp:86-125 interpolate_to_boundaries */
static void element_traces(o_elem *e, const double *lm, const double *lp) {
    const int N = e->N, nf = N + 1;
    for (int i = 0; i <= N; ++i) {
        e->pe[0][i] = 0; e->pe[2][N - i] = 0; e->pe[1][i] = 0; e->pe[3][N - i] = 0;
        e->ue[0][i] = 0; e->ue[2][N - i] = 0; e->ue[1][i] = 0; e->ue[3][N - i] = 0;
        e->ve[0][i] = 0; e->ve[2][N - i] = 0; e->ve[1][i] = 0; e->ve[3][N - i] = 0;
        for (int j = 0; j <= N; ++j) {
            e->pe[0][i] += lm[j] * e->p[i * nf + j];
            e->pe[2][N - i] += lp[j] * e->p[i * nf + j];
            e->pe[1][i] += lp[j] * e->p[j * nf + i];
            e->pe[3][N - i] += lm[j] * e->p[j * nf + i];
            e->ue[0][i] += lm[j] * e->u[i * nf + j];
            e->ue[2][N - i] += lp[j] * e->u[i * nf + j];
            e->ue[1][i] += lp[j] * e->u[j * nf + i];
            e->ue[3][N - i] += lm[j] * e->u[j * nf + i];
            e->ve[0][i] += lm[j] * e->v[i * nf + j];
            e->ve[2][N - i] += lp[j] * e->v[i * nf + j];
            e->ve[1][i] += lp[j] * e->v[j * nf + i];
            e->ve[3][N - i] += lm[j] * e->v[j * nf + i];
        }
    }
}

/* Accumulate the mortar projection of a neighbour trace onto ghost side 0, the shared body of
 * compute_wall_boundaries / compute_symmetry_boundaries / compute_outflow_boundaries
 * (src/meshes/Mesh2D_t.cpp:4383-4418, 4656-4691). The "only one neighbour" branch of those
 * functions tests faces_[0].size() == 0 and is never taken (a ghost has exactly one face). */
static void ghost_project(const o_mesh *m, o_elem *g, int64_t gi) {
    const o_tables *t = m->t;
    const int Ng = g->N;
    for (int i = 0; i <= Ng; ++i) { g->pe[0][i] = 0; g->ue[0][i] = 0; g->ve[0][i] = 0; }
    for (int fk = 0; fk < g->nf[0]; ++fk) {
        const o_face *f = &m->fa[g->faces[0][fk]];
        const int face_side = f->elements[0] == gi;
        const o_elem *nb = &m->el[f->elements[face_side]];
        const int ns = (int)f->sides[face_side];
        const int Nn = nb->N;
        const double *xn = t->tab[O_NODES][Nn], *xg = t->tab[O_NODES][Ng];
        const double *wn = t->tab[O_WEIGHTS][Nn], *wg = t->tab[O_WEIGHTS][Ng];
        const double *lg = t->tab[O_BARY][Ng];
        for (int j = 0; j <= Nn; ++j) {
            const double coord = (xn[Nn - j] + 1) * f->scale[face_side] + 2 * f->offset[face_side] - 1;
            int found = 0;
            for (int i = 0; i <= Ng; ++i) {
                if (almost_equal(coord, xg[i])) {
                    g->pe[0][i] += wn[j] / wg[i] * nb->pe[ns][j] * g->sf[0][i];
                    g->ue[0][i] += wn[j] / wg[i] * nb->ue[ns][j] * g->sf[0][i];
                    g->ve[0][i] += wn[j] / wg[i] * nb->ve[ns][j] * g->sf[0][i];
                    found = 1;
                    break;
                }
            }
            if (!found) {
                double s = 0.0;
                for (int i = 0; i <= Ng; ++i) s += lg[i] / (coord - xg[i]);
                for (int i = 0; i <= Ng; ++i) {
                    const double T = lg[i] / ((coord - xg[i]) * s);
                    g->pe[0][i] += T * wn[j] / wg[i] * nb->pe[ns][j] * g->sf[0][i];
                    g->ue[0][i] += T * wn[j] / wg[i] * nb->ue[ns][j] * g->sf[0][i];
                    g->ve[0][i] += T * wn[j] / wg[i] * nb->ve[ns][j] * g->sf[0][i];
                }
            }
        }
    }
}

/* src/meshes/Mesh2D_t.cpp:4298-4436 compute_wall_boundaries (symmetry :4438-4573 is identical) */
static void wall_boundary(const o_mesh *m, int64_t gi) {
    o_elem *g = &m->el[gi];
    ghost_project(m, g, gi);
    const o_face *f = &m->fa[g->faces[0][0]];
    for (int k = 0; k <= g->N; ++k) {
        const double ux = g->ue[0][k], uy = g->ve[0][k];
        const double l0 = -(ux * f->nx + uy * f->ny);
        const double l1 = ux * f->tx + uy * f->ty;
        g->ue[0][k] = f->nx * l0 + f->tx * l1;
        g->ve[0][k] = f->ny * l0 + f->ty * l1;
    }
}

/* src/meshes/Mesh2D_t.cpp:4575-4599 compute_inflow_boundaries */
static void inflow_boundary(const o_mesh *m, int64_t gi, double time) {
    o_elem *g = &m->el[gi];
    const v2 a = m->xy[g->nodes[0]], b = m->xy[g->nodes[1]];
    const double *x = m->t->tab[O_NODES][g->N];
    for (int k = 0; k <= g->N; ++k) {
        const double interp = (x[k] + 1) / 2;
        const double X = b.x * interp + a.x * (1 - interp);
        const double Y = b.y * interp + a.y * (1 - interp);
        double q[3];
        g_exact_c(X, Y, time, m->c, q);
        g->pe[0][k] = q[0];
        g->ue[0][k] = q[1];
        g->ve[0][k] = q[2];
    }
}

/* src/meshes/Mesh2D_t.cpp:3731-3782 boundary_conditions (+ :4712-4730 local_interfaces) */
static void boundary_conditions(o_mesh *m, double time) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < m->n_wall; ++k) wall_boundary(m, m->wall[k]);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < m->n_sym; ++k) wall_boundary(m, m->sym[k]);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < m->n_in; ++k) inflow_boundary(m, m->in[k], time);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < m->n_out; ++k) ghost_project(m, &m->el[m->out[k]], m->out[k]);
    for (int64_t k = 0; k < m->n_if; ++k) {
        const o_elem *src = &m->el[m->if_origin[k]];
        o_elem *dst = &m->el[m->if_dest[k]];
        const int s = (int)m->if_side[k];
        for (int i = 0; i <= src->N; ++i) {
            dst->pe[0][i] = src->pe[s][i];
            dst->ue[0][i] = src->ue[s][i];
            dst->ve[0][i] = src->ve[s][i];
        }
    }
}

static int next_side(int s) { return (s + 1) * !(s + 1 == 4); }

/* project_to_faces' barycentric interpolation of one element trace (order Ne) to the nodes of one
 * side of a face (order Nf) at xi = (x_face + 1) scale + 2 offset - 1, with the exact-node shortcut
 * (src/meshes/Mesh2D_t.cpp:3820-3890); the right side (side 1) walks the face nodes backwards. */
static void face_side_values(const o_tables *t, int Ne, const double *tp, const double *tu, const double *tv,
                             int Nf, int side, double offset, double scale, double *op, double *ou, double *ov) {
    const double *xf = t->tab[O_NODES][Nf];
    const double *xe = t->tab[O_NODES][Ne], *le = t->tab[O_BARY][Ne];
    for (int i = 0; i <= Nf; ++i) {
        const double coord = (xf[side == 0 ? i : Nf - i] + 1) * scale + 2 * offset - 1;
        double pn = 0.0, un = 0.0, vn = 0.0, den = 0.0;
        for (int j = 0; j <= Ne; ++j) {
            if (almost_equal(coord, xe[j])) {
                pn = tp[j]; un = tu[j]; vn = tv[j];
                den = 1.0;
                break;
            }
            const double tt = le[j] / (coord - xe[j]);
            pn += tt * tp[j];
            un += tt * tu[j];
            vn += tt * tv[j];
            den += tt;
        }
        op[i] = pn / den;
        ou[i] = un / den;
        ov[i] = vn / den;
    }
}

/* src/meshes/Mesh2D_t.cpp:3803-3893 project_to_faces */
static void project_to_faces(o_mesh *m) {
    const o_tables *t = m->t;
#pragma omp parallel for schedule(static)
    for (int64_t fi = 0; fi < m->F; ++fi) {
        o_face *f = &m->fa[fi];
        const int Nf = f->N;
        for (int side = 0; side < 2; ++side) {
            const o_elem *e = &m->el[f->elements[side]];
            const int es = (int)f->sides[side];
            const int conforming = side == 0
                ? (f->N == e->N && f->nodes[0] == e->nodes[es] && f->nodes[1] == e->nodes[next_side(es)])
                : (f->N == e->N && f->nodes[1] == e->nodes[es] && f->nodes[0] == e->nodes[next_side(es)]);
            if (conforming) {
                for (int i = 0; i <= Nf; ++i) {
                    const int o = side == 0 ? i : Nf - i;
                    f->p[side][o] = e->pe[es][i];
                    f->u[side][o] = e->ue[es][i];
                    f->v[side][o] = e->ve[es][i];
                }
            } else {
                face_side_values(t, e->N, e->pe[es], e->ue[es], e->ve[es], Nf, side, f->offset[side], f->scale[side],
                                 f->p[side], f->u[side], f->v[side]);
            }
        }
    }
}

/* Solver2D_t.cpp:48-63 calculate_wave_fluxes */
static void wave_fluxes(o_mesh *m) {
#pragma omp parallel for schedule(static)
    for (int64_t fi = 0; fi < m->F; ++fi) {
        o_face *f = &m->fa[fi];
        for (int i = 0; i <= f->N; ++i) {
            const double wL = f->p[0][i] + C_WAVE * (f->u[0][i] * f->nx + f->v[0][i] * f->ny);
            const double wR = f->p[1][i] - C_WAVE * (f->u[1][i] * f->nx + f->v[1][i] * f->ny);
            f->pflux[i] = C_WAVE * (wL - wR) / 2;
            f->uflux[i] = f->nx * (wL + wR) / 2;
            f->vflux[i] = f->ny * (wL + wR) / 2;
        }
    }
}

void o_wave_flux(const double *in, double *out) {
    const double nx = in[6], ny = in[7];
    const double wL = in[0] + C_WAVE * (in[1] * nx + in[2] * ny);
    const double wR = in[3] - C_WAVE * (in[4] * nx + in[5] * ny);
    out[0] = C_WAVE * (wL - wR) / 2;
    out[1] = nx * (wL + wR) / 2;
    out[2] = ny * (wL + wR) / 2;
}

/* project_to_elements' Gauss-quadrature L2 projection of one face's flux (order Nf) onto an element
 * side (order Ne), accumulated into out: T_i w_j / w_i F_j sf_i scale with the exact-node shortcut,
 * negated and walked backwards for the face's right side (src/meshes/Mesh2D_t.cpp:3935-4003). */
static void side_flux_from_face(const o_tables *t, int Ne, const double *sf, int Nf, int fwd, double offset,
                                double scale, const double *fp, const double *fu, const double *fv, double *op,
                                double *ou, double *ov) {
    const double *xe = t->tab[O_NODES][Ne], *we = t->tab[O_WEIGHTS][Ne], *le = t->tab[O_BARY][Ne];
    const double *xf = t->tab[O_NODES][Nf], *wf = t->tab[O_WEIGHTS][Nf];
    for (int j = 0; j <= Nf; ++j) {
        const double coord = (xf[fwd ? j : Nf - j] + 1) * scale + 2 * offset - 1;
        int found = 0;
        for (int i = 0; i <= Ne; ++i) {
            if (almost_equal(coord, xe[i])) {
                if (fwd) {
                    op[i] += wf[j] / we[i] * fp[j] * sf[i] * scale;
                    ou[i] += wf[j] / we[i] * fu[j] * sf[i] * scale;
                    ov[i] += wf[j] / we[i] * fv[j] * sf[i] * scale;
                } else {
                    op[i] += -wf[j] / we[i] * fp[j] * sf[i] * scale;
                    ou[i] += -wf[j] / we[i] * fu[j] * sf[i] * scale;
                    ov[i] += -wf[j] / we[i] * fv[j] * sf[i] * scale;
                }
                found = 1;
                break;
            }
        }
        if (!found) {
            double ssum = 0.0;
            for (int i = 0; i <= Ne; ++i) ssum += le[i] / (coord - xe[i]);
            for (int i = 0; i <= Ne; ++i) {
                const double T = le[i] / ((coord - xe[i]) * ssum);
                if (fwd) {
                    op[i] += T * wf[j] / we[i] * fp[j] * sf[i] * scale;
                    ou[i] += T * wf[j] / we[i] * fu[j] * sf[i] * scale;
                    ov[i] += T * wf[j] / we[i] * fv[j] * sf[i] * scale;
                } else {
                    op[i] += -T * wf[j] / we[i] * fp[j] * sf[i] * scale;
                    ou[i] += -T * wf[j] / we[i] * fu[j] * sf[i] * scale;
                    ov[i] += -T * wf[j] / we[i] * fv[j] * sf[i] * scale;
                }
            }
        }
    }
}

/* src/meshes/Mesh2D_t.cpp:3895-4006 project_to_elements */
static void project_to_elements(o_mesh *m) {
    const o_tables *t = m->t;
#pragma omp parallel for schedule(static)
    for (int64_t ei = 0; ei < m->K; ++ei) {
        o_elem *e = &m->el[ei];
        const int Ne = e->N;
        for (int s = 0; s < 4; ++s) {
            const o_face *f0 = &m->fa[e->faces[s][0]];
            if (e->nf[s] == 1 && f0->N == Ne && f0->nodes[0] == e->nodes[s] && f0->nodes[1] == e->nodes[next_side(s)]) {
                for (int j = 0; j <= f0->N; ++j) {
                    e->pf[s][j] = f0->pflux[j] * e->sf[s][j];
                    e->uf[s][j] = f0->uflux[j] * e->sf[s][j];
                    e->vf[s][j] = f0->vflux[j] * e->sf[s][j];
                }
            } else if (e->nf[s] == 1 && f0->N == Ne && f0->nodes[1] == e->nodes[s] && f0->nodes[0] == e->nodes[next_side(s)]) {
                for (int j = 0; j <= f0->N; ++j) {
                    e->pf[s][f0->N - j] = -f0->pflux[j] * e->sf[s][j];
                    e->uf[s][f0->N - j] = -f0->uflux[j] * e->sf[s][j];
                    e->vf[s][f0->N - j] = -f0->vflux[j] * e->sf[s][j];
                }
            } else {
                for (int i = 0; i <= Ne; ++i) { e->pf[s][i] = 0.0; e->uf[s][i] = 0.0; e->vf[s][i] = 0.0; }
                for (int fk = 0; fk < e->nf[s]; ++fk) {
                    const o_face *f = &m->fa[e->faces[s][fk]];
                    const int fwd = ei == f->elements[0];
                    const int fs = fwd ? 0 : 1;
                    side_flux_from_face(t, Ne, e->sf[s], f->N, fwd, f->offset[fs], f->scale[fs], f->pflux, f->uflux,
                                        f->vflux, e->pf[s], e->uf[s], e->vf[s]);
                }
            }
        }
    }
}

/* Solver2D_t.cpp:39-46 matrix_vector_multiply */
static void matvec(int N, const double *A, const double *x, double *y) {
    for (int i = 0; i <= N; ++i) {
        y[i] = 0.0;
        for (int j = 0; j <= N; ++j) y[i] += A[i * (N + 1) + j] * x[j];
    }
}

/* Solver2D_t.cpp:66-136 compute_dg_wave_derivative (x_flux = {c u, p, 0}, y_flux = {c v, 0, p},
 * :31-37) */
static void dg_derivative(o_mesh *m) {
    const o_tables *t = m->t;
#pragma omp parallel for schedule(static)
    for (int64_t ei = 0; ei < m->K; ++ei) {
        o_elem *e = &m->el[ei];
        const int N = e->N, nf = N + 1;
        const double *w = t->tab[O_WEIGHTS][N], *Dh = t->tab[O_DHAT][N];
        const double *ll = t->tab[O_LAG_LEFT][N], *lr = t->tab[O_LAG_RIGHT][N];
        for (int j = 0; j <= N; ++j) {
            for (int i = 0; i <= N; ++i) {
                const int k = i * nf + j;
                const double fx0 = C_WAVE * e->u[k], fx1 = e->p[k], fx2 = 0.0;
                const double fy0 = C_WAVE * e->v[k], fy1 = 0.0, fy2 = e->p[k];
                e->pflux[i] = e->deta_dy[k] * fx0 - e->deta_dx[k] * fy0;
                e->uflux[i] = e->deta_dy[k] * fx1 - e->deta_dx[k] * fy1;
                e->vflux[i] = e->deta_dy[k] * fx2 - e->deta_dx[k] * fy2;
            }
            matvec(N, Dh, e->pflux, e->pder);
            matvec(N, Dh, e->uflux, e->uder);
            matvec(N, Dh, e->vflux, e->vder);
            for (int i = 0; i <= N; ++i) {
                e->pder[i] += (e->pf[1][j] * lr[i] + e->pf[3][N - j] * ll[i]) / w[i];
                e->uder[i] += (e->uf[1][j] * lr[i] + e->uf[3][N - j] * ll[i]) / w[i];
                e->vder[i] += (e->vf[1][j] * lr[i] + e->vf[3][N - j] * ll[i]) / w[i];
            }
            for (int i = 0; i <= N; ++i) {
                e->Gp[i * nf + j] = -e->pder[i];
                e->Gu[i * nf + j] = -e->uder[i];
                e->Gv[i * nf + j] = -e->vder[i];
            }
        }
        for (int i = 0; i <= N; ++i) {
            for (int j = 0; j <= N; ++j) {
                const int k = i * nf + j;
                const double fx0 = C_WAVE * e->u[k], fx1 = e->p[k], fx2 = 0.0;
                const double fy0 = C_WAVE * e->v[k], fy1 = 0.0, fy2 = e->p[k];
                e->pflux[j] = -e->dxi_dy[k] * fx0 + e->dxi_dx[k] * fy0;
                e->uflux[j] = -e->dxi_dy[k] * fx1 + e->dxi_dx[k] * fy1;
                e->vflux[j] = -e->dxi_dy[k] * fx2 + e->dxi_dx[k] * fy2;
            }
            matvec(N, Dh, e->pflux, e->pder);
            matvec(N, Dh, e->uflux, e->uder);
            matvec(N, Dh, e->vflux, e->vder);
            for (int j = 0; j <= N; ++j) {
                e->pder[j] += (e->pf[2][N - i] * lr[j] + e->pf[0][i] * ll[j]) / w[j];
                e->uder[j] += (e->uf[2][N - i] * lr[j] + e->uf[0][i] * ll[j]) / w[j];
                e->vder[j] += (e->vf[2][N - i] * lr[j] + e->vf[0][i] * ll[j]) / w[j];
            }
            for (int j = 0; j <= N; ++j) {
                const int k = i * nf + j;
                e->Gp[k] = (e->Gp[k] - e->pder[j]) / e->jac[k];
                e->Gu[k] = (e->Gu[k] - e->uder[j]) / e->jac[k];
                e->Gv[k] = (e->Gv[k] - e->vder[j]) / e->jac[k];
            }
        }
        /* wave speed c: the symmetric speed-c system's operator is c times the c = 1 operator */
        for (int k = 0; k < nf * nf; ++k) {
            e->Gp[k] *= m->c;
            e->Gu[k] *= m->c;
            e->Gv[k] *= m->c;
        }
    }
}

void o_mesh_set_wave_speed(o_mesh *m, double c) { m->c = c; }

/* Threads of the element and face loops of a stage (OpenMP; every loop writes disjoint elements,
 * faces or ghosts, so any thread count gives the same results bit for bit). 0 = OpenMP's default.
 * Returns the count in use. */
int o_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* Solver2D_t.cpp:138-172 rk3_first_step / rk3_step (a = 0 marks the first stage) */
static void rk3_update(o_mesh *m, int first, double dt, double a, double g) {
#pragma omp parallel for schedule(static)
    for (int64_t ei = 0; ei < m->K; ++ei) {
        o_elem *e = &m->el[ei];
        const int ne = (e->N + 1) * (e->N + 1);
        for (int k = 0; k < ne; ++k) {
            if (first) {
                e->ip[k] = e->Gp[k]; e->iu[k] = e->Gu[k]; e->iv[k] = e->Gv[k];
            } else {
                e->ip[k] = a * e->ip[k] + e->Gp[k];
                e->iu[k] = a * e->iu[k] + e->Gu[k];
                e->iv[k] = a * e->iv[k] + e->Gv[k];
            }
            e->p[k] += g * dt * e->ip[k];
            e->u[k] += g * dt * e->iu[k];
            e->v[k] += g * dt * e->iv[k];
        }
    }
}

void o_compute_rhs(o_mesh *m, double time) {
#pragma omp parallel for schedule(static)
    for (int64_t ei = 0; ei < m->K; ++ei) {
        o_elem *e = &m->el[ei];
        element_traces(e, m->t->tab[O_LAG_LEFT][e->N], m->t->tab[O_LAG_RIGHT][e->N]);
    }
    boundary_conditions(m, time);
    project_to_faces(m);
    wave_fluxes(m);
    project_to_elements(m);
    dg_derivative(m);
}

/* include/solvers/Solver2D_t.tpp:16-18 (Williamson 2N-storage RK3) and :44-72 */
static const double AM[3] = {0, -5.0 / 9.0, -153.0 / 128.0};
static const double BM[3] = {0, 1.0 / 3.0, 0.75};
static const double GM[3] = {1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0};

void o_step(o_mesh *m, double time, double dt) {
    for (int s = 0; s < 3; ++s) {
        o_compute_rhs(m, time + BM[s] * dt);
        rk3_update(m, s == 0, dt, AM[s], GM[s]);
    }
}

int64_t o_solve(o_mesh *m, double cfl, double t_end, double *t_final) {
    double time = 0.0;
    int64_t steps = 0;
    while (time < t_end) {
        ++steps;
        double dt = o_get_delta_t(m, cfl);
        if (time + dt > t_end) dt = t_end - time;
        o_step(m, time, dt);
        time += dt;
    }
    if (t_final) *t_final = time;
    return steps;
}

void o_run_steps(o_mesh *m, double cfl, int64_t n_steps, double *time_io) {
    double time = time_io ? *time_io : 0.0;
    for (int64_t k = 0; k < n_steps; ++k) {
        const double dt = o_get_delta_t(m, cfl);
        o_step(m, time, dt);
        time += dt;
    }
    if (time_io) *time_io = time;
}

/* ------------------------------------------------------------------------------------------ */
/* Estimator: src/entities/Element2D_t.cpp:274-343                                             */
/* ------------------------------------------------------------------------------------------ */

static void estimate_p(int N, const double *p, const double *poly, double tol_min, double tol_max,
                       double *err_out, double *sigma_out, int *refine, int *coarsen, double *spectrum) {
    const int nf = N + 1;
    const int n = (N + 1) < N_LSQ_MAX ? (N + 1) : N_LSQ_MAX;
    for (int ni = 0; ni < n; ++ni) {
        spectrum[ni] = 0.0;
        const double pd = N + ni + 1 - n;
        const int pi_ = (int)pd;
        for (int i = 0; i < pd; ++i) {
            double lx = 0.0, ly = 0.0;
            for (int k = 0; k <= N; ++k)
                for (int l = 0; l <= N; ++l) {
                    lx += p[k * nf + l] * poly[i * nf + k] * poly[pi_ * nf + l];
                    ly += p[k * nf + l] * poly[pi_ * nf + k] * poly[i * nf + l];
                }
            spectrum[ni] += fabs(lx) + fabs(ly);
        }
        double lx = 0.0;
        for (int k = 0; k <= N; ++k)
            for (int l = 0; l <= N; ++l) lx += p[k * nf + l] * poly[pi_ * nf + k] * poly[pi_ * nf + l];
        spectrum[ni] += fabs(lx);
    }
    /* Element2D_t.cpp:319-343 exponential_decay */
    double x_avg = 0.0, y_avg = 0.0;
    for (int i = 0; i < n; ++i) {
        x_avg += N + i + 1 - n;
        y_avg += log(spectrum[i]);
    }
    x_avg /= n;
    y_avg /= n;
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n; ++i) {
        const int pp = N + i + 1 - n;
        num += (pp - x_avg) * (log(spectrum[i]) - y_avg);
        den += (pp - x_avg) * (pp - x_avg);
    }
    const double sigma_signed = num / den;
    const double C = exp(y_avg - sigma_signed * x_avg);
    const double sigma = fabs(sigma_signed);
    const double err = sqrt(spectrum[n - 1] * spectrum[n - 1] + C * C * 0.5 / sigma * exp(-2 * sigma * (N + 1)));
    *err_out = err;
    *sigma_out = sigma;
    *refine = err > tol_min;
    *coarsen = err < tol_max;
}

void o_estimate_error(o_mesh *m, double *error, double *sigma, int32_t *refine, int32_t *coarsen) {
    for (int64_t ei = 0; ei < m->K; ++ei) {
        o_elem *e = &m->el[ei];
        double spec[16];
        estimate_p(e->N, e->p, m->t->tab[O_POLY][e->N], m->tol_min, m->tol_max, &e->error, &e->sigma,
                   &e->refine, &e->coarsen, spec);
        if (error) error[ei] = e->error;
        if (sigma) sigma[ei] = e->sigma;
        if (refine) refine[ei] = e->refine;
        if (coarsen) coarsen[ei] = e->coarsen;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* p-adaptation: src/meshes/Mesh2D_t.cpp:1644-1706 (the branch without splitting), :5026-5049,  */
/* :7482-7538; Element2D_t.cpp:464-579, :735-741, :802-812                                      */
/* ------------------------------------------------------------------------------------------ */

/* Element2D_t.cpp:464-579 interpolate_from (same geometry, other order): at each new node, the
 * old element's exact node when both coordinates match (almost_equal, the last matching index
 * wins), a 1D barycentric interpolation along the other coordinate when one matches, else the
 * tensor barycentric interpolation. p, u, v and G are interpolated; the RK register starts at 0. */
static void interpolate_p(const o_tables *t, const o_elem *o, o_elem *e) {
    const int No = o->N, N = e->N, nfo = No + 1, nf = N + 1;
    const double *xo = t->tab[O_NODES][No], *lo = t->tab[O_BARY][No], *xn = t->tab[O_NODES][N];
    const double *src[6] = {o->p, o->u, o->v, o->Gp, o->Gu, o->Gv};
    double *dst[6] = {e->p, e->u, e->v, e->Gp, e->Gu, e->Gv};
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) {
            const double X = xn[i], Y = xn[j];
            int col = -1, row = -1;
            for (int m = 0; m <= No; ++m) {
                if (almost_equal(X, xo[m])) col = m;
                if (almost_equal(Y, xo[m])) row = m;
            }
            const int k = i * nf + j;
            if (row != -1 && col != -1) {
                for (int f = 0; f < 6; ++f) dst[f][k] = src[f][col * nfo + row];
            } else if (row != -1) {
                double num[6] = {0, 0, 0, 0, 0, 0}, den = 0.0;
                for (int m = 0; m <= No; ++m) {
                    const double tt = lo[m] / (X - xo[m]);
                    for (int f = 0; f < 6; ++f) num[f] += tt * src[f][m * nfo + row];
                    den += tt;
                }
                for (int f = 0; f < 6; ++f) dst[f][k] = num[f] / den;
            } else if (col != -1) {
                double num[6] = {0, 0, 0, 0, 0, 0}, den = 0.0;
                for (int n = 0; n <= No; ++n) {
                    const double tt = lo[n] / (Y - xo[n]);
                    for (int f = 0; f < 6; ++f) num[f] += tt * src[f][col * nfo + n];
                    den += tt;
                }
                for (int f = 0; f < 6; ++f) dst[f][k] = num[f] / den;
            } else {
                for (int f = 0; f < 6; ++f) dst[f][k] = 0.0;
                double dx = 0.0, dy = 0.0;
                for (int q = 0; q <= No; ++q) {
                    dx += lo[q] / (X - xo[q]);
                    dy += lo[q] / (Y - xo[q]);
                }
                for (int m = 0; m <= No; ++m) {
                    const double tx = lo[m] / ((X - xo[m]) * dx);
                    for (int n = 0; n <= No; ++n) {
                        const double ty = lo[n] / ((Y - xo[n]) * dy);
                        for (int f = 0; f < 6; ++f) dst[f][k] += src[f][m * nfo + n] * tx * ty;
                    }
                }
            }
        }
}

/* Element2D_t.cpp:735-741 resize_boundary_storage: the ghost's traces and edge scaling factor
 * become zero vectors of the new length; nothing recomputes the scaling factor on this path. */
static void ghost_resize(o_elem *g, int N) {
    free(g->block);
    g->N = N;
    elem_alloc_ghost(g);
}

static int64_t adapt_split(o_mesh *m, int N_max, int max_split_level);

int64_t o_adapt(o_mesh *m, int N_max, int max_split_level) {
    /* Element2D_t.cpp:802-812 would_h_refine: refine && sigma < 1 && split_level < max */
    for (int64_t i = 0; i < m->K; ++i)
        if (m->el[i].refine && m->el[i].sigma < 1.0 && m->el[i].split_level < max_split_level)
            return adapt_split(m, N_max, max_split_level);
    int64_t count = 0;
    for (int64_t i = 0; i < m->K; ++i) {
        o_elem *o = &m->el[i];
        if (!(o->refine && o->sigma >= 1.0 && o->N + 2 <= N_max)) continue;
        o_elem e = *o;
        e.N = o->N + 2;
        elem_alloc_domain(&e);
        interpolate_p(m->t, o, &e);
        v2 P[4];
        for (int k = 0; k < 4; ++k) P[k] = m->xy[e.nodes[k]];
        element_geometry(&e, P, m->t->tab[O_NODES][e.N]);
        free(o->block);
        *o = e;
        ++count;
    }
    /* adjust_boundaries (Mesh2D_t.cpp:7482-7504) for the wall, symmetry, inflow, outflow lists */
    int64_t *lists[4] = {m->wall, m->sym, m->in, m->out};
    const int64_t counts[4] = {m->n_wall, m->n_sym, m->n_in, m->n_out};
    for (int l = 0; l < 4; ++l)
        for (int64_t k = 0; k < counts[l]; ++k) {
            o_elem *g = &m->el[lists[l][k]];
            int N = g->N;
            for (int fk = 0; fk < g->nf[0]; ++fk) {
                const o_face *f = &m->fa[g->faces[0][fk]];
                const int side = f->elements[0] == lists[l][k];
                const int Ns = m->el[f->elements[side]].N;
                if (Ns > N) N = Ns;
            }
            if (N != g->N) ghost_resize(g, N);
        }
    /* adjust_interfaces (:7506-7520) */
    for (int64_t k = 0; k < m->n_if; ++k) {
        o_elem *g = &m->el[m->if_dest[k]];
        const int N = m->el[m->if_origin[k]].N;
        if (N != g->N) ghost_resize(g, N);
    }
    /* adjust_faces (:7522-7538) */
    for (int64_t fi = 0; fi < m->F; ++fi) {
        o_face *f = &m->fa[fi];
        const int NL = m->el[f->elements[0]].N, NR = m->el[f->elements[1]].N;
        const int N = NL > NR ? NL : NR;
        if (N != f->N) {
            free(f->block);
            f->N = N;
            face_alloc(f);
        }
    }
    return count;
}

/* quad_map.tpp:3-8 */
static v2 quad_map_pt(double xi, double eta, const v2 *P) {
    v2 r;
    r.x = (P[0].x * (1 - xi) * (1 - eta) + P[1].x * (xi + 1) * (1 - eta) + P[2].x * (xi + 1) * (eta + 1) +
           P[3].x * (1 - xi) * (eta + 1)) / 4;
    r.y = (P[0].y * (1 - xi) * (1 - eta) + P[1].y * (xi + 1) * (1 - eta) + P[2].y * (xi + 1) * (eta + 1) +
           P[3].y * (1 - xi) * (eta + 1)) / 4;
    return r;
}

static double cross2(v2 a, v2 b) { return a.x * b.y - a.y * b.x; }

/* inverse_quad_map.tpp:46-88 (closed-form inverse of the bilinear map) */
static v2 inverse_quad_map(v2 g, const v2 *P) {
    const v2 e = sub(P[1], P[0]), f = sub(P[3], P[0]);
    const v2 gg = {P[0].x - P[1].x + P[2].x - P[3].x, P[0].y - P[1].y + P[2].y - P[3].y};
    const v2 h = sub(g, P[0]);
    const double k2 = cross2(gg, f);
    const double k1 = cross2(e, f) + cross2(h, gg);
    const double k0 = cross2(h, e);
    v2 r;
    if (fabs(k2) < DBL_MIN) {
        const double ux = e.x * k1 - gg.x * k0, uy = e.y * k1 - gg.y * k0;
        const double u = fabs(ux) > fabs(uy) ? (h.x * k1 + f.x * k0) / ux : (h.y * k1 + f.y * k0) / uy;
        r.x = u * 2 - 1;
        r.y = -k0 / k1 * 2 - 1;
        return r;
    }
    double w = k1 * k1 - 4 * k0 * k2;
    if (w < 0.0) { r.x = -1; r.y = -1; return r; }
    w = sqrt(w);
    const double ik2 = 1 / (2 * k2);
    double v = (-k1 - w) * ik2;
    double ux = e.x + gg.x * v, uy = e.y + gg.y * v;
    double u = fabs(ux) > fabs(uy) ? (h.x - f.x * v) / ux : (h.y - f.y * v) / uy;
    if (u < 0.0 || u > 1.0 || v < 0.0 || v > 1.0) {
        v = (-k1 + w) * ik2;
        ux = e.x + gg.x * v;
        uy = e.y + gg.y * v;
        u = fabs(ux) > fabs(uy) ? (h.x - f.x * v) / ux : (h.y - f.y * v) / uy;
    }
    r.x = u * 2 - 1;
    r.y = v * 2 - 1;
    return r;
}

/* Element2D_t.cpp:345-462 interpolate_from (other geometry): each new node is mapped to the other
 * element's reference coordinates (quad_map, inverse_quad_map), then the same exact-node /
 * row / column / tensor barycentric branches as interpolate_p. */
static void interpolate_h(const o_tables *t, const o_elem *o, const v2 *Po, o_elem *e, const v2 *Pe) {
    const int No = o->N, N = e->N, nfo = No + 1, nf = N + 1;
    const double *xo = t->tab[O_NODES][No], *lo = t->tab[O_BARY][No], *xn = t->tab[O_NODES][N];
    const double *src[6] = {o->p, o->u, o->v, o->Gp, o->Gu, o->Gv};
    double *dst[6] = {e->p, e->u, e->v, e->Gp, e->Gu, e->Gv};
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j) {
            const v2 gpt = quad_map_pt(xn[i], xn[j], Pe);
            const v2 loc = inverse_quad_map(gpt, Po);
            const double X = loc.x, Y = loc.y;
            int col = -1, row = -1;
            for (int mm = 0; mm <= No; ++mm) {
                if (almost_equal(X, xo[mm])) col = mm;
                if (almost_equal(Y, xo[mm])) row = mm;
            }
            const int k = i * nf + j;
            if (row != -1 && col != -1) {
                for (int f = 0; f < 6; ++f) dst[f][k] = src[f][col * nfo + row];
            } else if (row != -1) {
                double num[6] = {0, 0, 0, 0, 0, 0}, den = 0.0;
                for (int mm = 0; mm <= No; ++mm) {
                    const double tt = lo[mm] / (X - xo[mm]);
                    for (int f = 0; f < 6; ++f) num[f] += tt * src[f][mm * nfo + row];
                    den += tt;
                }
                for (int f = 0; f < 6; ++f) dst[f][k] = num[f] / den;
            } else if (col != -1) {
                double num[6] = {0, 0, 0, 0, 0, 0}, den = 0.0;
                for (int n = 0; n <= No; ++n) {
                    const double tt = lo[n] / (Y - xo[n]);
                    for (int f = 0; f < 6; ++f) num[f] += tt * src[f][col * nfo + n];
                    den += tt;
                }
                for (int f = 0; f < 6; ++f) dst[f][k] = num[f] / den;
            } else {
                for (int f = 0; f < 6; ++f) dst[f][k] = 0.0;
                double dx = 0.0, dy = 0.0;
                for (int q = 0; q <= No; ++q) {
                    dx += lo[q] / (X - xo[q]);
                    dy += lo[q] / (Y - xo[q]);
                }
                for (int mm = 0; mm <= No; ++mm) {
                    const double tx = lo[mm] / ((X - xo[mm]) * dx);
                    for (int n = 0; n <= No; ++n) {
                        const double ty = lo[n] / ((Y - xo[n]) * dy);
                        for (int f = 0; f < 6; ++f) dst[f][k] += src[f][mm * nfo + n] * tx * ty;
                    }
                }
            }
        }
}

static int64_t edge_mid_find(const int64_t *keys, int64_t n, int64_t a, int64_t b) {
    const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
    for (int64_t k = 0; k < n; ++k)
        if (keys[3 * k] == lo && keys[3 * k + 1] == hi) return keys[3 * k + 2];
    return -1;
}

/* The splitting branch of Mesh2D_t::adapt (src/meshes/Mesh2D_t.cpp:1993-2268, hp_adapt
 * :5186-5659, split_boundaries :6232-6343, split_interfaces :6343-6756): elements with
 * would_h_refine become four children (corner s: {corner s, midpoint of side s, centre, midpoint
 * of side s-1} in local positions s, s+1, s+2, s+3), placed at position child_order[s] of the
 * Hilbert child order, with the parent's order, split level + 1, child status and the parent's
 * rotation, state interpolated from the parent; elements with would_p_refine are raised by 2
 * orders; a boundary ghost on a split side becomes two ghosts (geometry recomputed), a ghost whose
 * order changes is resized and its geometry recomputed; a periodic ghost splits with its origin
 * (the element across the period). The new topology (faces, 2:1 hanging sides, also across the
 * period, any number of faces per side) is built from the new element and ghost lists as at mesh
 * creation. A split that would leave a face smaller than 1/16 of its larger side is refused
 * (returns -2, nothing changed), as the device path does. Returns the number of split plus raised
 * elements. */
static int64_t adapt_split(o_mesh *m, int N_max, int max_split_level) {
    const int64_t K = m->K, NB = m->n_elements_total - m->K;
    char *split = (char *)calloc((size_t)K + 1, 1);
    char *raise = (char *)calloc((size_t)K + 1, 1);
    int64_t n_split = 0, n_raise = 0;
    for (int64_t i = 0; i < K; ++i) {
        const o_elem *e = &m->el[i];
        if (e->refine && e->sigma < 1.0 && e->split_level < max_split_level) { split[i] = 1; ++n_split; }
        else if (e->refine && e->sigma >= 1.0 && e->N + 2 <= N_max) { raise[i] = 1; ++n_raise; }
    }
    /* limit shared with the device path: a small side covering 1/16 of a larger side (a face scale
     * of 1/16 on the larger side), across the interior or the period, is not split again while the
     * larger element stays */
    for (int64_t i = 0; i < K; ++i) {
        if (!split[i]) continue;
        const o_elem *e = &m->el[i];
        for (int s = 0; s < 4; ++s) {
            if (e->nf[s] != 1) continue;
            const o_face *f = &m->fa[e->faces[s][0]];
            const int side = f->elements[0] == i ? 1 : 0;
            const int64_t other = f->elements[side];
            int64_t larger = other;
            if (other >= K) {
                if (m->bar_type[other - K] != O_BC_PERIODIC) continue;
                larger = m->bar_origin[2 * (other - K)];
            }
            if (f->scale[side] < 1.5 / 16 && !split[larger]) { free(split); free(raise); return -2; }
        }
    }
    /* new nodes: existing, then per split element its side midpoints (shared) and centre, then the
     * middles of periodic ghosts that split with their origin and had none */
    int64_t cap_nodes = m->n_nodes + 5 * n_split + NB + 1;
    double *xy = (double *)malloc((size_t)cap_nodes * 2 * sizeof(double));
    for (int64_t k = 0; k < m->n_nodes; ++k) { xy[2 * k] = m->xy[k].x; xy[2 * k + 1] = m->xy[k].y; }
    int64_t n_nodes = m->n_nodes;
    int64_t *mids = (int64_t *)malloc((size_t)(4 * K + 4 * n_split + 2 * NB + 1) * 3 * sizeof(int64_t));
    int64_t n_mids = 0;
    /* midpoints already present: the node at the middle of a side covered by several faces, of
     * elements and of periodic ghosts */
    for (int64_t i = 0; i < m->n_elements_total; ++i)
        for (int s = 0; s < (i < K ? 4 : 1); ++s)
            if (m->el[i].nf[s] >= 2) {
                const int64_t a = m->el[i].nodes[s], b = m->el[i].nodes[(s + 1) % 4];
                const double mx = (m->xy[a].x + m->xy[b].x) / 2, my = (m->xy[a].y + m->xy[b].y) / 2;
                const double len = mag(sub(m->xy[b], m->xy[a]));
                int64_t md = -1;
                for (int q = 0; q < m->el[i].nf[s] && md < 0; ++q)
                    for (int r = 0; r < 2; ++r) {
                        const int64_t v = m->fa[m->el[i].faces[s][q]].nodes[r];
                        if (fabs(m->xy[v].x - mx) <= 1e-12 * len && fabs(m->xy[v].y - my) <= 1e-12 * len) md = v;
                    }
                if (md < 0) continue;
                mids[3 * n_mids] = a < b ? a : b; mids[3 * n_mids + 1] = a < b ? b : a;
                mids[3 * n_mids + 2] = md;
                ++n_mids;
            }
    const int64_t Knew = K + 3 * n_split;
    int64_t *quads = (int64_t *)malloc((size_t)Knew * 4 * sizeof(int64_t));
    int32_t *qN = (int32_t *)malloc((size_t)Knew * sizeof(int32_t));
    int64_t *parent = (int64_t *)malloc((size_t)Knew * sizeof(int64_t));
    int *child_status = (int *)malloc((size_t)Knew * sizeof(int));
    int64_t *newpos = (int64_t *)malloc((size_t)(K + 1) * sizeof(int64_t));
    int64_t pos = 0;
    for (int64_t i = 0; i < K; ++i) {
        const o_elem *e = &m->el[i];
        newpos[i] = pos;
        if (!split[i]) {
            for (int k = 0; k < 4; ++k) quads[4 * pos + k] = e->nodes[k];
            qN[pos] = e->N + (raise[i] ? 2 : 0);
            parent[pos] = i;
            child_status[pos] = -1;
            ++pos;
            continue;
        }
        int64_t mid[4];
        for (int s = 0; s < 4; ++s) {
            const int64_t a = e->nodes[s], b = e->nodes[(s + 1) % 4];
            int64_t md = edge_mid_find(mids, n_mids, a, b);
            if (md < 0) {
                md = n_nodes++;
                xy[2 * md] = (xy[2 * a] + xy[2 * b]) / 2;
                xy[2 * md + 1] = (xy[2 * a + 1] + xy[2 * b + 1]) / 2;
                mids[3 * n_mids] = a < b ? a : b; mids[3 * n_mids + 1] = a < b ? b : a; mids[3 * n_mids + 2] = md;
                ++n_mids;
            }
            mid[s] = md;
        }
        const int64_t c = n_nodes++;
        double cx = 0.0, cy = 0.0;
        for (int k = 0; k < 4; ++k) { cx += xy[2 * e->nodes[k]]; cy += xy[2 * e->nodes[k] + 1]; }
        xy[2 * c] = cx / 4;
        xy[2 * c + 1] = cy / 4;
        const int *order = CHILD_ORDER[e->rotation][e->status];
        const int *stat = CHILD_STATUS[e->rotation][e->status];
        for (int s = 0; s < 4; ++s) {
            const int64_t at = pos + order[s];
            quads[4 * at + s] = e->nodes[s];
            quads[4 * at + (s + 1) % 4] = mid[s];
            quads[4 * at + (s + 2) % 4] = c;
            quads[4 * at + (s + 3) % 4] = mid[(s + 3) % 4];
            qN[at] = e->N;
            parent[at] = i;
            child_status[at] = stat[s];
        }
        pos += 4;
    }
    /* ghosts: split on split sides (two consecutive ghosts), else kept; resized ones and new ones
     * get fresh geometry, the others keep a zero scaling factor an earlier resize left. A periodic
     * ghost follows its origin instead (split_interfaces, src/meshes/Mesh2D_t.cpp:6343-6756): it
     * splits when the origin splits, its halves copying the origin's children at the side's first
     * and second corner, and stays whole otherwise, also when its own element splits (then it
     * covers two element sides). */
    int64_t *bars = (int64_t *)malloc((size_t)(2 * NB + 1) * 2 * sizeof(int64_t));
    int32_t *btype = (int32_t *)malloc((size_t)(2 * NB + 1) * sizeof(int32_t));
    int64_t *borig = (int64_t *)malloc((size_t)(2 * NB + 1) * 2 * sizeof(int64_t));
    char *bzero = (char *)calloc((size_t)(2 * NB + 1), 1);
    int64_t nb = 0;
    for (int64_t b = 0; b < NB; ++b) {
        const o_elem *g = &m->el[K + b];
        if (m->bar_type[b] == O_BC_PERIODIC) {
            const int64_t o = m->bar_origin[2 * b], so = m->bar_origin[2 * b + 1];
            if (split[o]) {
                int64_t md = edge_mid_find(mids, n_mids, g->nodes[0], g->nodes[1]);
                if (md < 0) {
                    const int64_t a = g->nodes[0], c = g->nodes[1];
                    md = n_nodes++;
                    xy[2 * md] = (xy[2 * a] + xy[2 * c]) / 2;
                    xy[2 * md + 1] = (xy[2 * a + 1] + xy[2 * c + 1]) / 2;
                    mids[3 * n_mids] = a < c ? a : c; mids[3 * n_mids + 1] = a < c ? c : a; mids[3 * n_mids + 2] = md;
                    ++n_mids;
                }
                const int *order = CHILD_ORDER[m->el[o].rotation][m->el[o].status];
                for (int h = 0; h < 2; ++h) {
                    bars[2 * nb] = h == 0 ? g->nodes[0] : md;
                    bars[2 * nb + 1] = h == 0 ? md : g->nodes[1];
                    btype[nb] = O_BC_PERIODIC;
                    borig[2 * nb] = newpos[o] + order[(so + h) % 4];
                    borig[2 * nb + 1] = so;
                    ++nb;
                }
            } else {
                bars[2 * nb] = g->nodes[0]; bars[2 * nb + 1] = g->nodes[1];
                btype[nb] = O_BC_PERIODIC;
                borig[2 * nb] = newpos[o];
                borig[2 * nb + 1] = so;
                bzero[nb] = !raise[o] && g->sf[0][0] == 0.0;
                ++nb;
            }
            continue;
        }
        const o_face *f = &m->fa[g->faces[0][0]];
        const int64_t src = f->elements[0];
        if (split[src]) {
            const int64_t md = edge_mid_find(mids, n_mids, g->nodes[0], g->nodes[1]);
            for (int h = 0; h < 2; ++h) {
                bars[2 * nb] = h == 0 ? g->nodes[0] : md;
                bars[2 * nb + 1] = h == 0 ? md : g->nodes[1];
                btype[nb] = m->bar_type[b];
                borig[2 * nb] = borig[2 * nb + 1] = -1;
                ++nb;
            }
        } else {
            bars[2 * nb] = g->nodes[0]; bars[2 * nb + 1] = g->nodes[1];
            btype[nb] = m->bar_type[b];
            borig[2 * nb] = borig[2 * nb + 1] = -1;
            bzero[nb] = !raise[src] && g->sf[0][0] == 0.0;
            ++nb;
        }
    }
    o_mesh *nm = mesh_create(m->t, m->initial_N, n_nodes, xy, Knew, quads, qN, nb, bars, btype, NULL, borig,
                             m->tol_min, m->tol_max);
    nm->c = m->c;
    for (int64_t k = 0; k < Knew; ++k) {
        o_elem *e = &nm->el[k];
        const o_elem *o = &m->el[parent[k]];
        v2 Pe[4], Po[4];
        for (int q = 0; q < 4; ++q) { Pe[q] = nm->xy[e->nodes[q]]; Po[q] = m->xy[o->nodes[q]]; }
        if (child_status[k] >= 0) {
            e->status = child_status[k];
            e->rotation = o->rotation;
            e->split_level = o->split_level + 1;
            interpolate_h(m->t, o, Po, e, Pe);
        } else {
            e->status = o->status;
            e->rotation = o->rotation;
            e->split_level = o->split_level;
            if (e->N != o->N) {
                interpolate_p(m->t, o, e);
            } else {
                const int ne = (e->N + 1) * (e->N + 1);
                memcpy(e->p, o->p, ne * sizeof(double)); memcpy(e->u, o->u, ne * sizeof(double));
                memcpy(e->v, o->v, ne * sizeof(double)); memcpy(e->Gp, o->Gp, ne * sizeof(double));
                memcpy(e->Gu, o->Gu, ne * sizeof(double)); memcpy(e->Gv, o->Gv, ne * sizeof(double));
            }
        }
    }
    for (int64_t b = 0; b < nb; ++b)
        if (bzero[b]) {
            o_elem *g = &nm->el[Knew + b];
            for (int k = 0; k <= g->N; ++k) g->sf[0][k] = 0.0;
        }
    /* swap the new mesh in */
    o_mesh tmp = *m;
    *m = *nm;
    *nm = tmp;
    o_mesh_destroy(nm);
    free(xy); free(mids); free(quads); free(qN); free(parent); free(child_status);
    free(bars); free(btype); free(borig); free(bzero); free(newpos); free(split); free(raise);
    return n_split + n_raise;
}

void o_mesh_get_elements(const o_mesh *m, int64_t *nodes, int32_t *status, int32_t *rotation, double *xy) {
    for (int64_t i = 0; i < m->K; ++i) {
        for (int k = 0; k < 4; ++k) nodes[4 * i + k] = m->el[i].nodes[k];
        if (status) status[i] = m->el[i].status;
        if (rotation) rotation[i] = m->el[i].rotation;
    }
    if (xy)
        for (int64_t k = 0; k < m->n_nodes; ++k) { xy[2 * k] = m->xy[k].x; xy[2 * k + 1] = m->xy[k].y; }
}
int64_t o_mesh_n_nodes(const o_mesh *m) { return m->n_nodes; }
void o_mesh_set_flags(o_mesh *m, const int32_t *refine, const double *sigma) {
    for (int64_t i = 0; i < m->K; ++i) {
        m->el[i].refine = refine[i];
        m->el[i].sigma = sigma[i];
    }
}

/* include/solvers/Solver2D_t.tpp:37-100 with adaptation: Δt from get_delta_t every step, the last
 * step clipped to t_end, and every adapt_interval steps estimate_error + adapt. Returns the step
 * count, or -1 if an adaptation would split an element. */
int64_t o_solve_adaptive(o_mesh *m, double cfl, double t_end, int64_t adapt_interval, int N_max,
                         int max_split_level, double *t_final) {
    double time = 0.0;
    int64_t steps = 0;
    while (time < t_end) {
        ++steps;
        double dt = o_get_delta_t(m, cfl);
        if (time + dt > t_end) dt = t_end - time;
        o_step(m, time, dt);
        time += dt;
        if (adapt_interval > 0 && steps % adapt_interval == 0) {
            o_estimate_error(m, NULL, NULL, NULL, NULL);
            if (o_adapt(m, N_max, max_split_level) < 0) return -1;
        }
    }
    if (t_final) *t_final = time;
    return steps;
}

/* ------------------------------------------------------------------------------------------ */
/* Element / face level entry points                                                           */
/* ------------------------------------------------------------------------------------------ */

void o_element_traces(const o_tables *t, int N, const double *p, const double *u, const double *v,
                      double *out) {
    o_elem e;
    memset(&e, 0, sizeof(e));
    e.N = N;
    elem_alloc_domain(&e);
    const int nf = N + 1;
    memcpy(e.p, p, (size_t)nf * nf * sizeof(double));
    memcpy(e.u, u, (size_t)nf * nf * sizeof(double));
    memcpy(e.v, v, (size_t)nf * nf * sizeof(double));
    element_traces(&e, t->tab[O_LAG_LEFT][N], t->tab[O_LAG_RIGHT][N]);
    for (int s = 0; s < 4; ++s) {
        memcpy(out + (s * 3 + 0) * nf, e.pe[s], (size_t)nf * sizeof(double));
        memcpy(out + (s * 3 + 1) * nf, e.ue[s], (size_t)nf * sizeof(double));
        memcpy(out + (s * 3 + 2) * nf, e.ve[s], (size_t)nf * sizeof(double));
    }
    free(e.block);
}

void o_element_geometry(const o_tables *t, int N, const double *pts, double *out) {
    o_elem e;
    memset(&e, 0, sizeof(e));
    e.N = N;
    elem_alloc_domain(&e);
    v2 P[4];
    for (int k = 0; k < 4; ++k) { P[k].x = pts[2 * k]; P[k].y = pts[2 * k + 1]; }
    element_geometry(&e, P, t->tab[O_NODES][N]);
    const int nf = N + 1, ne = nf * nf;
    double *o = out;
    memcpy(o, e.dxi_dx, ne * sizeof(double)); o += ne;
    memcpy(o, e.deta_dx, ne * sizeof(double)); o += ne;
    memcpy(o, e.dxi_dy, ne * sizeof(double)); o += ne;
    memcpy(o, e.deta_dy, ne * sizeof(double)); o += ne;
    memcpy(o, e.jac, ne * sizeof(double)); o += ne;
    for (int s = 0; s < 4; ++s) { memcpy(o, e.sf[s], nf * sizeof(double)); o += nf; }
    o[0] = e.dmin; o[1] = e.cx; o[2] = e.cy;
    free(e.block);
}

void o_boundary_geometry(const o_tables *t, int N, const double *pts, double *out) {
    o_elem e;
    memset(&e, 0, sizeof(e));
    e.N = N;
    elem_alloc_ghost(&e);
    v2 P[4];
    for (int k = 0; k < 4; ++k) { P[k].x = pts[2 * k]; P[k].y = pts[2 * k + 1]; }
    boundary_geometry(&e, P, t->tab[O_NODES][N]);
    memcpy(out, e.sf[0], (size_t)(N + 1) * sizeof(double));
    out[N + 1] = e.dmin; out[N + 2] = e.cx; out[N + 3] = e.cy;
    free(e.block);
}

void o_face_geometry(const double *face_nodes, const double *element_nodes, double *out) {
    o_face f;
    memset(&f, 0, sizeof(f));
    v2 fn[2], en[4];
    for (int k = 0; k < 2; ++k) { fn[k].x = face_nodes[2 * k]; fn[k].y = face_nodes[2 * k + 1]; }
    for (int k = 0; k < 4; ++k) { en[k].x = element_nodes[2 * k]; en[k].y = element_nodes[2 * k + 1]; }
    face_geometry(&f, fn, en);
    out[0] = f.nx; out[1] = f.ny; out[2] = f.tx; out[3] = f.ty; out[4] = f.length;
    out[5] = f.offset[0]; out[6] = f.offset[1]; out[7] = f.scale[0]; out[8] = f.scale[1];
}

void o_element_estimate(const o_tables *t, int N, const double *p, double tol_min, double tol_max,
                        double *out) {
    double spec[16] = {0};
    double err, sigma;
    int refine, coarsen;
    estimate_p(N, p, t->tab[O_POLY][N], tol_min, tol_max, &err, &sigma, &refine, &coarsen, spec);
    out[0] = err; out[1] = sigma; out[2] = refine; out[3] = coarsen;
    const int n = (N + 1) < N_LSQ_MAX ? (N + 1) : N_LSQ_MAX;
    for (int i = 0; i < n; ++i) out[4 + i] = spec[i];
}

/* src/functions/Hilbert.cpp:16-40 */
static void hilbert_rot(int n, int *x, int *y, int rx, int ry) {
    if (ry == 0) {
        if (rx == 1) { *x = n - 1 - *x; *y = n - 1 - *y; }
        const int tmp = *x; *x = *y; *y = tmp;
    }
}

void o_hilbert_d2xy(int n, int d, int *xy) {
    int x = 0, y = 0;
    for (int s = 1; s < n; s *= 2) {
        const int rx = 1 & (d / 2);
        const int ry = 1 & (d ^ rx);
        hilbert_rot(s, &x, &y, rx, ry);
        x += s * rx;
        y += s * ry;
        d /= 4;
    }
    xy[0] = x; xy[1] = y;
}

/* src/functions/Hilbert.cpp:3-14 */
int o_hilbert_xy2d(int n, int x, int y) {
    int d = 0;
    for (int s = n / 2; s > 0; s /= 2) {
        const int rx = (x & s) > 0, ry = (y & s) > 0;
        d += s * s * ((3 * rx) ^ ry);
        hilbert_rot(n, &x, &y, rx, ry);
    }
    return d;
}


/* ------------------------------------------------------------------------------------------ */
/* Single-operation entry points for the reference's own known-answer tests                    */
/* (tests/element_to_face_projection_test.cu, face_to_element_projection_test.cu,              */
/* element_to_element_interpolation_test.cu, inverse_quad_mapping_test.cu): the same helpers the */
/* mesh code above uses, on caller-given traces, fluxes and elements.                          */
/* ------------------------------------------------------------------------------------------ */

#define O_MAX_NF 64  /* face / element orders up to 63 */

void o_kat_face_values(const o_tables *t, int Ne, int Nf, int side, double offset, double scale,
                       const double *trace, double *out) {
    double u[O_MAX_NF], v[O_MAX_NF];
    face_side_values(t, Ne, trace, trace, trace, Nf, side, offset, scale, out, u, v);
}

void o_kat_side_flux(const o_tables *t, int Ne, int Nf, int fwd, double offset, double scale, double sf,
                     const double *flux, double *out) {
    double sfv[O_MAX_NF], u[O_MAX_NF], v[O_MAX_NF];
    for (int i = 0; i <= Ne; ++i) { sfv[i] = sf; u[i] = 0.0; v[i] = 0.0; }
    side_flux_from_face(t, Ne, sfv, Nf, fwd, offset, scale, flux, flux, flux, out, u, v);
}

void o_kat_interpolate(const o_tables *t, int No, const double *Po, int N, const double *Pn, const double *q_old,
                       double *q_new) {
    const int neo = (No + 1) * (No + 1), ne = (N + 1) * (N + 1);
    double *zo = (double *)calloc((size_t)neo, sizeof(double));
    double *zn = (double *)calloc((size_t)ne * 5, sizeof(double));
    o_elem o, e;
    memset(&o, 0, sizeof o);
    memset(&e, 0, sizeof e);
    o.N = No;
    o.p = (double *)q_old; o.u = zo; o.v = zo; o.Gp = zo; o.Gu = zo; o.Gv = zo;
    e.N = N;
    e.p = q_new; e.u = zn; e.v = zn + ne; e.Gp = zn + 2 * ne; e.Gu = zn + 3 * ne; e.Gv = zn + 4 * ne;
    if (Po && Pn) {
        v2 PO[4], PN[4];
        for (int k = 0; k < 4; ++k) {
            PO[k].x = Po[2 * k]; PO[k].y = Po[2 * k + 1];
            PN[k].x = Pn[2 * k]; PN[k].y = Pn[2 * k + 1];
        }
        interpolate_h(t, &o, PO, &e, PN);
    } else {
        interpolate_p(t, &o, &e);
    }
    free(zo);
    free(zn);
}

void o_kat_inverse_quad_map(const double *P, double x, double y, double *xi_eta) {
    v2 PP[4];
    for (int k = 0; k < 4; ++k) { PP[k].x = P[2 * k]; PP[k].y = P[2 * k + 1]; }
    v2 g; g.x = x; g.y = y;
    const v2 r = inverse_quad_map(g, PP);
    xi_eta[0] = r.x;
    xi_eta[1] = r.y;
}
