// device.hpp -- device helpers shared by the stage kernels (kernels.hip) and the adjoint kernels
// (adjoint.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace ndg {

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each XCD walks one
// contiguous stretch of the Hilbert-ordered elements (neighbours share its L2).
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t G) {
    const int64_t per = G / 8, rem = G % 8, x = b % 8, i = b / 8;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

// Bilinear metrics (include/functions/quad_metrics.tpp:3-8): x_xi, x_eta, y_xi, y_eta.
__device__ __forceinline__ void metrics(const double *P, double xi, double eta, double &xxi,
                                        double &xeta, double &yxi, double &yeta) {
    xxi = ((1 - eta) * (P[2] - P[0]) + (1 + eta) * (P[4] - P[6])) / 4;
    xeta = ((1 - xi) * (P[6] - P[0]) + (1 + xi) * (P[4] - P[2])) / 4;
    yxi = ((1 - eta) * (P[3] - P[1]) + (1 + eta) * (P[5] - P[7])) / 4;
    yeta = ((1 - xi) * (P[7] - P[1]) + (1 + xi) * (P[5] - P[3])) / 4;
}

// Edge Jacobian of side s (src/entities/Element2D_t.cpp:756-769): constant along a side of a
// bilinear element, so it is evaluated once at the side's first node.
__device__ __forceinline__ double side_scale(const double *P, int s, double x0) {
    double xxi, xeta, yxi, yeta;
    if (s == 0) { metrics(P, x0, -1.0, xxi, xeta, yxi, yeta); return sqrt(xxi * xxi + yxi * yxi); }
    if (s == 1) { metrics(P, 1.0, x0, xxi, xeta, yxi, yeta); return sqrt(xeta * xeta + yeta * yeta); }
    if (s == 2) { metrics(P, x0, 1.0, xxi, xeta, yxi, yeta); return sqrt(xxi * xxi + yxi * yxi); }
    metrics(P, -1.0, x0, xxi, xeta, yxi, yeta);
    return sqrt(xeta * xeta + yeta * yeta);
}

// Thread layout of the line-per-lane kernels: a workgroup is three waves, one per field (p, u, v);
// lane (element e, line l) of the wave for field f handles one (N+1)-point line of that field for
// EPB = floor(64/(N+1)) elements.
template <int N>
struct LineShape {
    static constexpr int NF = N + 1;
    static constexpr int NE = NF * NF;
    static constexpr int T = 192;                            // one wave per field
    static constexpr int EPB = 64 / NF;                       // elements per workgroup
    static constexpr int FPT = (EPB * 4 * NF + T - 1) / T;    // flux points per thread
    static constexpr int QPT = (EPB * 3 * NE + T - 1) / T;    // state values per thread (tile load)
    static constexpr int VPT = (EPB * 8 + T - 1) / T;         // vertex coordinates per thread
};

// 1/x from the hardware reciprocal estimate and two Newton steps (quadratic convergence from its
// ~2^-29 relative error to the rounding error of the last step).
__device__ __forceinline__ double recip(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}

// Element2D_t::almost_equal with 2 ulp (src/meshes/Mesh2D_t.cpp:3785-3792).
__device__ __forceinline__ bool almost_equal2(double a, double b) {
    return fabs(a - b) <= 2.220446049250313e-16 * fabs(a + b) * 2 || fabs(a - b) < 2.2250738585072014e-308;
}

// Element2D_t::interpolate_from (src/entities/Element2D_t.cpp:345-579) at reference point (X, Y) of
// the old element of order No, from one field sf[(No+1)^2] of it: the old element's node when
// both coordinates match a node (almost_equal, the last matching index wins), a 1D barycentric
// interpolation along the other coordinate when one matches, else the tensor barycentric
// interpolation. xo, lo: nodes and barycentric weights of order No.
__device__ inline double interpolate_at(const double *sf, int No, double X, double Y, const double *xo,
                                        const double *lo) {
    const int nfo = No + 1;
    int col = -1, row = -1;
    for (int m = 0; m <= No; ++m) {
        if (almost_equal2(X, xo[m])) col = m;
        if (almost_equal2(Y, xo[m])) row = m;
    }
    if (row >= 0 && col >= 0) return sf[col * nfo + row];
    if (row >= 0) {
        double num = 0.0, den = 0.0;
        for (int m = 0; m <= No; ++m) {
            const double t = lo[m] / (X - xo[m]);
            num += t * sf[m * nfo + row];
            den += t;
        }
        return num / den;
    }
    if (col >= 0) {
        double num = 0.0, den = 0.0;
        for (int n = 0; n <= No; ++n) {
            const double t = lo[n] / (Y - xo[n]);
            num += t * sf[col * nfo + n];
            den += t;
        }
        return num / den;
    }
    double dx = 0.0, dy = 0.0;
    for (int k = 0; k <= No; ++k) {
        dx += lo[k] / (X - xo[k]);
        dy += lo[k] / (Y - xo[k]);
    }
    double v = 0.0;
    for (int m = 0; m <= No; ++m) {
        const double tx = lo[m] / ((X - xo[m]) * dx);
        for (int n = 0; n <= No; ++n) v += sf[m * nfo + n] * tx * (lo[n] / ((Y - xo[n]) * dy));
    }
    return v;
}

}  // namespace ndg
