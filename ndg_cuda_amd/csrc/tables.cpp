// tables.cpp -- FP64 Legendre-Gauss tables (host side, computed once and uploaded).
//
// Same tables and the same per-N construction sequence as the reference NDG_t
// (include/entities/NDG_t.tpp:9-59) so that the device operator uses identical coefficients:
// Legendre-Gauss nodes/weights by Newton iteration on L_{N+1}
// (src/polynomials/LegendrePolynomial_t.cpp:60-101), barycentric weights (NDG_t.tpp:175-188),
// normalised Lagrange interpolants at -1/+1 and their derivatives (:209-249), the derivative
// matrix with its diagonal by negative row sum (:251-274), D^ = -D^T W / w (:276-283), the CG
// matrix (:286-298), uniform-grid interpolation (:302-325) and the modal matrix (:327-334).
// Compiled with -ffp-contract=off: the coefficients are bitwise those of the reference host.
#include <cfloat>
#include <cmath>

#include "internal.hpp"

namespace ndg {
namespace {

constexpr double kPi = 3.14159265358979323846;

// Bonnet recurrence for L_n and the derivative recurrence L'_k = L'_{k-2} + (2k-1) L_{k-1}.
void legendre_pair(int n, double x, double &L, double &dL) {
    if (n == 0) { L = 1.0; dL = 0.0; return; }
    if (n == 1) { L = x; dL = 1.0; return; }
    double Lm2 = 1.0, Lm1 = x, dLm2 = 0.0, dLm1 = 1.0;
    for (int k = 2; k <= n; ++k) {
        L = (2 * k - 1) * x * Lm1 / k - (k - 1) * Lm2 / k;
        dL = dLm2 + (2 * k - 1) * Lm1;
        Lm2 = Lm1; Lm1 = L;
        dLm2 = dLm1; dLm1 = dL;
    }
}

double legendre_value(int n, double x) {
    double L = 0.0, dL = 0.0;
    legendre_pair(n, x, L, dL);
    return L;
}

void gauss_nodes(int N, std::vector<double> &x, std::vector<double> &w) {
    if (N == 1) {
        x[0] = -std::sqrt(1.0 / 3.0); w[0] = 1.0;
        x[1] = -x[0]; w[1] = w[0];
    } else {
        for (int i = 0; i < (N + 1) / 2; ++i) {
            double r = -std::cos(kPi * (2 * i + 1) / (2 * N + 2));
            for (int it = 0; it < 1000; ++it) {
                double L, dL;
                legendre_pair(N + 1, r, L, dL);
                const double step = -L / dL;
                r += step;
                if (std::fabs(step) <= DBL_MIN * 2 * std::fabs(r)) break;
            }
            double L, dL;
            legendre_pair(N + 1, r, L, dL);
            x[i] = r;
            x[N - i] = -r;
            w[i] = 2.0 / ((1 - std::pow(r, 2)) * std::pow(dL, 2));
            w[N - i] = w[i];
        }
    }
    if (N == 0) { x[0] = 0.0; w[0] = 2.0; }
    if (N % 2 == 0) {
        double L, dL;
        legendre_pair(N + 1, 0.0, L, dL);
        x[N / 2] = 0.0;
        w[N / 2] = 2 / std::pow(dL, 2);
    }
}

bool ulp2_equal(double a, double b) {
    return std::fabs(a - b) <= DBL_EPSILON * std::fabs(a + b) * 2 || std::fabs(a - b) < DBL_MIN;
}

}  // namespace

void build_tables(int N_max, int64_t n_interp, Tables &t) {
    t.N_max = N_max;
    t.n_interp = n_interp;
    for (auto &tab : t.tab) tab.assign(N_max + 1, {});
    for (int N = 0; N <= N_max; ++N) {
        const int nf = N + 1;
        auto &x = t.tab[NDG_TABLE_NODES][N];
        auto &w = t.tab[NDG_TABLE_WEIGHTS][N];
        auto &lam = t.tab[NDG_TABLE_BARYCENTRIC_WEIGHTS][N];
        auto &D = t.tab[NDG_TABLE_DERIVATIVE][N];
        auto &I = t.tab[NDG_TABLE_INTERPOLATION][N];
        auto &M = t.tab[NDG_TABLE_POLYNOMIALS][N];
        auto &lm = t.tab[NDG_TABLE_LAGRANGE_LEFT][N];
        auto &lp = t.tab[NDG_TABLE_LAGRANGE_RIGHT][N];
        auto &dlm = t.tab[NDG_TABLE_LAGRANGE_DERIVATIVE_LEFT][N];
        auto &dlp = t.tab[NDG_TABLE_LAGRANGE_DERIVATIVE_RIGHT][N];
        auto &Gh = t.tab[NDG_TABLE_G_HAT_DERIVATIVE][N];
        auto &Dh = t.tab[NDG_TABLE_DERIVATIVE_HAT][N];
        x.assign(nf, 0.0); w.assign(nf, 0.0); lam.assign(nf, 0.0);
        D.assign(nf * nf, 0.0); M.assign(nf * nf, 0.0); Gh.assign(nf * nf, 0.0); Dh.assign(nf * nf, 0.0);
        I.assign(nf * n_interp, 0.0);
        lm.assign(nf, 0.0); lp.assign(nf, 0.0); dlm.assign(nf, 0.0); dlp.assign(nf, 0.0);

        gauss_nodes(N, x, w);
        for (int j = 0; j < nf; ++j) {
            double prod = 1.0;
            for (int i = 0; i < nf; ++i)
                if (i != j) prod *= x[j] - x[i];
            lam[j] = 1.0 / prod;
        }
        for (int i = 0; i < nf; ++i)
            for (int j = 0; j < nf; ++j) M[i * nf + j] = legendre_value(i, x[j]) * w[j] * (2 * i + 1) * 0.5;
        for (int i = 0; i < nf; ++i)
            for (int j = 0; j < nf; ++j)
                if (i != j) D[i * nf + j] = lam[j] / (lam[i] * (x[i] - x[j]));
        for (int64_t r = 0; r < n_interp; ++r) {
            const double xc = 2.0 * r / (n_interp - 1) - 1.0;
            bool hit = false;
            for (int k = 0; k < nf; ++k) {
                I[r * nf + k] = 0.0;
                if (ulp2_equal(xc, x[k])) { I[r * nf + k] = 1.0; hit = true; }
            }
            if (!hit) {
                double total = 0.0;
                for (int k = 0; k < nf; ++k) { I[r * nf + k] = lam[k] / (xc - x[k]); total += I[r * nf + k]; }
                for (int k = 0; k < nf; ++k) I[r * nf + k] /= total;
            }
        }
        for (int i = 0; i < nf; ++i) lm[i] = lam[i] / (-1.0 - x[i]);
        for (int i = 0; i < nf; ++i) lp[i] = lam[i] / (1.0 - x[i]);
        for (int i = 0; i < nf; ++i) dlm[i] = lam[i] / ((-1.0 - x[i]) * (-1.0 - x[i]));
        for (int i = 0; i < nf; ++i) dlp[i] = lam[i] / ((1.0 - x[i]) * (1.0 - x[i]));
        auto normalise = [&](std::vector<double> &v, double at) {
            double s = 0.0;
            for (int i = 0; i < nf; ++i) s += lam[i] / (at - x[i]);
            for (int i = 0; i < nf; ++i) v[i] /= s;
        };
        {
            double s = 0.0;
            for (int i = 0; i < nf; ++i) s += lm[i];
            for (int i = 0; i < nf; ++i) lm[i] /= s;
            s = 0.0;
            for (int i = 0; i < nf; ++i) s += lp[i];
            for (int i = 0; i < nf; ++i) lp[i] /= s;
        }
        normalise(dlm, -1.0);
        normalise(dlp, 1.0);
        for (int i = 0; i < nf; ++i) {
            double d = 0.0;
            for (int j = 0; j < nf; ++j)
                if (j != i) d -= D[i * nf + j];
            D[i * nf + i] = d;
        }
        for (int j = 0; j < nf; ++j)
            for (int n = 0; n < nf; ++n) {
                double s = 0.0;
                for (int k = 0; k < nf; ++k) s += D[k * nf + n] * D[k * nf + j] * w[k];
                Gh[j * nf + n] = s / w[j];
            }
        for (int i = 0; i < nf; ++i)
            for (int j = 0; j < nf; ++j) Dh[i * nf + j] = -D[j * nf + i] * w[j] / w[i];
    }
}

std::vector<double> pack_tables(const Tables &t, int N) {
    const int nf = N + 1;
    std::vector<double> out;
    out.reserve(table_pack_len(N));
    const auto &Dh = t.tab[NDG_TABLE_DERIVATIVE_HAT][N];
    out.insert(out.end(), Dh.begin(), Dh.end());
    for (int which : {NDG_TABLE_LAGRANGE_LEFT, NDG_TABLE_LAGRANGE_RIGHT, NDG_TABLE_WEIGHTS, NDG_TABLE_NODES})
        out.insert(out.end(), t.tab[which][N].begin(), t.tab[which][N].end());
    for (int i = 0; i < nf; ++i) out.push_back(1.0 / t.tab[NDG_TABLE_WEIGHTS][N][i]);
    const auto &M = t.tab[NDG_TABLE_POLYNOMIALS][N];
    out.insert(out.end(), M.begin(), M.end());
    for (int a = 0; a < nf; ++a)
        for (int i = 0; i < nf; ++i) out.push_back(Dh[i * nf + a]);
    return out;
}

}  // namespace ndg
