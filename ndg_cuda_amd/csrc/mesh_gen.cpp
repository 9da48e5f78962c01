// mesh_gen.cpp -- square structured mesh generator, the in-memory equivalent of the reference's
// CGNS mesh generator (extras/mesh_generator/main.cpp:161-382).
//
// Layout kept identical to that generator so that meshes are interchangeable:
//  - node (i, j) has id i*(res+1)+j and coordinates x_min + i*dx, y_min + j*dy (:189-196);
//  - element k sits at the Hilbert-curve cell d2xy(res, k) (:236-242, src/functions/Hilbert.cpp:
//    16-28) with local nodes bottom-right, top-right, top-left, bottom-left, so that the local
//    xi axis runs along +y and eta along -x;
//  - boundary bars in the order bottom, right, top, left with the generator's node order
//    (:253-305); periodic sides pair bar i of one side with bar res-1-i of the opposite side
//    (:355-382).
// Added here (no reference counterpart): a seeded jitter of interior vertices by a uniform
// fraction of the spacing, drawn with splitmix64 so that it is identical on every platform.
#include <cmath>

#include "internal.hpp"

namespace ndg {
namespace {

void hilbert_rotate(int n, int &x, int &y, int rx, int ry) {
    if (ry == 0) {
        if (rx == 1) {
            x = n - 1 - x;
            y = n - 1 - y;
        }
        const int tmp = x;
        x = y;
        y = tmp;
    }
}

// Cell of the d-th point of the Hilbert curve on an n x n grid.
void hilbert_cell(int n, int64_t d, int &x, int &y) {
    x = 0;
    y = 0;
    for (int s = 1; s < n; s *= 2) {
        const int rx = 1 & (int)(d / 2);
        const int ry = 1 & (int)(d ^ rx);
        hilbert_rotate(s, x, y, rx, ry);
        x += s * rx;
        y += s * ry;
        d /= 4;
    }
}

struct SplitMix64 {
    uint64_t state;
    double next() {
        uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-53;
    }
};

bool power_of_two(int v) { return v > 0 && (v & (v - 1)) == 0; }

}  // namespace
}  // namespace ndg

using namespace ndg;

extern "C" ndg_status ndg_square_mesh_sizes(const ndg_square_mesh_params *p, int64_t *n_nodes,
                                            int64_t *n_quads, int64_t *n_bars) {
    if (!p || !n_nodes || !n_quads || !n_bars) return set_error(NDG_ERR_INVALID, "ndg_square_mesh_sizes: null argument");
    if (!power_of_two(p->resolution))
        return set_error(NDG_ERR_INVALID, "ndg_square_mesh_sizes: resolution must be a power of two (Hilbert order)");
    const int64_t r = p->resolution;
    *n_nodes = (r + 1) * (r + 1);
    *n_quads = r * r;
    *n_bars = 4 * r;
    return NDG_OK;
}

extern "C" ndg_status ndg_square_mesh_fill(const ndg_square_mesh_params *p, double *xy, int64_t *quads,
                                           int64_t *bars, int32_t *bar_type, int64_t *bar_donor) {
    int64_t nn, nq, nb;
    ndg_status st = ndg_square_mesh_sizes(p, &nn, &nq, &nb);
    if (st != NDG_OK) return st;
    if (!xy || !quads || !bars || !bar_type || !bar_donor) return set_error(NDG_ERR_INVALID, "ndg_square_mesh_fill: null array");
    if (!(p->jitter >= 0.0 && p->jitter < 0.5)) return set_error(NDG_ERR_INVALID, "ndg_square_mesh_fill: jitter must be in [0, 0.5)");
    const int64_t res = p->resolution, xnr = res + 1, ynr = res + 1;
    const double dx = (p->x_max - p->x_min) / (xnr - 1), dy = (p->y_max - p->y_min) / (ynr - 1);
    for (int64_t i = 0; i < xnr; ++i)
        for (int64_t j = 0; j < ynr; ++j) {
            xy[2 * (i * ynr + j)] = p->x_min + i * (p->x_max - p->x_min) / (xnr - 1);
            xy[2 * (i * ynr + j) + 1] = p->y_min + j * (p->y_max - p->y_min) / (ynr - 1);
        }
    if (p->jitter > 0.0) {
        SplitMix64 rng{p->seed};
        for (int64_t i = 1; i < xnr - 1; ++i)
            for (int64_t j = 1; j < ynr - 1; ++j) {
                const double ux = rng.next(), uy = rng.next();
                xy[2 * (i * ynr + j)] += (2 * ux - 1) * p->jitter * dx;
                xy[2 * (i * ynr + j) + 1] += (2 * uy - 1) * p->jitter * dy;
            }
    }
    for (int64_t k = 0; k < nq; ++k) {
        int X, Y;
        hilbert_cell((int)res, k, X, Y);
        quads[4 * k + 0] = (X + 1) * ynr + Y;
        quads[4 * k + 1] = (X + 1) * ynr + Y + 1;
        quads[4 * k + 2] = X * ynr + Y + 1;
        quads[4 * k + 3] = X * ynr + Y;
    }
    // bars: bottom [0, res), right [res, 2res), top [2res, 3res), left [3res, 4res)
    for (int64_t i = 0; i < res; ++i) {
        int64_t *b = bars + 2 * i;
        b[0] = (i + 1) * ynr;
        b[1] = i * ynr;
    }
    for (int64_t j = 0; j < res; ++j) {
        int64_t *b = bars + 2 * (res + j);
        b[0] = ynr * (xnr - 1) + j + 1;
        b[1] = ynr * (xnr - 1) + j;
    }
    for (int64_t i = 0; i < res; ++i) {
        int64_t *b = bars + 2 * (2 * res + i);
        b[0] = (res - i) * ynr - 1;
        b[1] = (res - i + 1) * ynr - 1;
    }
    for (int64_t j = 0; j < res; ++j) {
        int64_t *b = bars + 2 * (3 * res + j);
        b[0] = res - j - 1;
        b[1] = res - j;
    }
    for (int side = 0; side < 4; ++side) {
        const bool periodic = (side % 2 == 0) ? p->y_periodic : p->x_periodic;
        for (int64_t k = 0; k < res; ++k) {
            const int64_t bi = side * res + k;
            if (periodic) {
                bar_type[bi] = NDG_BC_PERIODIC;
                const int opposite = (side + 2) % 4;
                bar_donor[bi] = opposite * res + (res - 1 - k);
            } else {
                bar_type[bi] = p->bc[side];
                bar_donor[bi] = -1;
            }
        }
    }
    return NDG_OK;
}
