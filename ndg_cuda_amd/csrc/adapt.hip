// adapt.hip -- device-resident hp-adaptation.
//
// The mesh's element-side topology lives in device memory in the caller's element order
// (LogicalMesh), so an adaptation runs as a chain of kernels with two small read-backs of counts:
//   1. decisions from the estimator (Element2D_t.cpp:802-812): keep, raise by 2, or split;
//   2. the new logical mesh: every split element becomes four children placed in the Hilbert child
//      order (Hilbert_splitting.cpp:23-81), raised elements take N + 2, and every new element side
//      finds its exterior side from the old topology alone (conforming, 2:1 hanging, periodic or
//      ghost), a ghost on a split side becoming two (src/meshes/Mesh2D_t.cpp:5186-5985, :6232-6343;
//      the reference's device form is src/meshes/Mesh2D_t.cu:6206-7232, :7498, :8832-8902);
//   3. the order-grouped layout: a stable counting sort by order (dev_of and the group spans) and
//      exclusive scans that number the mortar records, side-flux slots and mortar faces;
//   4. the device mesh arrays: vertices, neighbour codes, boundary and mortar records, mortar faces
//      with their geometry and matrix offsets;
//   5. the solution transfer from each new element's parent (copy, new order, or a child's nodes
//      through the inverse bilinear map).
// Geometry (midpoints, face frames, edge Jacobians, the inverse map) is computed with contraction
// off, as the host path computes it, so both paths give the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>

#include "device.hpp"
#include "internal.hpp"

namespace ndg {

namespace {

__constant__ int dChildOrder[4][4][4] = NDG_HILBERT_CHILD_ORDER;
__constant__ int dChildStatus[4][4][4] = NDG_HILBERT_CHILD_STATUS;

constexpr int kThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;
constexpr int kOrders = kMaxOrder + 1;

unsigned grid_for(int64_t n, int per = kThreads) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

struct v2 {
    double x, y;
};

__device__ __forceinline__ v2 corner(const LogicalMesh &L, int64_t e, int k) {
    const int kk = k & 3;
    return v2{L.verts[8 * e + 2 * kk], L.verts[8 * e + 2 * kk + 1]};
}

// Adds v over the wave's lanes, one atomic per wave (every lane of the wave must call it).
__device__ void wave_add(unsigned long long *dst, unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, v);
}

// ---------------------------------------------------------------------------------------------
// Geometry with the host path's operation order (contraction off).

__device__ double mag(double x, double y) {
#pragma clang fp contract(off)
    return sqrt(x * x + y * y);
}

// Face geometry with the domain element on the left (src/entities/Face2D_t.cpp:56-66).
__device__ void face_frame(v2 f0, v2 f1, v2 l0, v2 l1, double &nx, double &ny, double &tx, double &ty) {
#pragma clang fp contract(off)
    tx = f1.x - f0.x;
    ty = f1.y - f0.y;
    const double len = mag(tx, ty);
    tx /= len;
    ty /= len;
    nx = ty;
    ny = -tx;
    const double lx = l1.x - l0.x, ly = l1.y - l0.y;
    const double sign = copysign(1.0, tx * lx + ty * ly);
    nx *= sign;
    ny *= sign;
    tx *= sign;
    ty *= sign;
}

// Face offsets and scales of the two sides (src/entities/Face2D_t.cpp:68-81).
__device__ void face_offsets(v2 f0, v2 f1, const v2 en[4], double offset[2], double scale[2]) {
#pragma clang fp contract(off)
    const double d0 = mag(en[1].x - en[0].x, en[1].y - en[0].y);
    const double d1 = mag(en[3].x - en[2].x, en[3].y - en[2].y);
    const double fd = mag(f1.x - f0.x, f1.y - f0.y);
    offset[0] = mag(f0.x - en[0].x, f0.y - en[0].y) / d0;
    offset[1] = mag(f1.x - en[2].x, f1.y - en[2].y) / d1;
    scale[0] = fd / d0;
    scale[1] = fd / d1;
}

// Edge Jacobian of side s of the bilinear quad P (quad_metrics.tpp:3-8 on the side).
__device__ double side_scale(const v2 P[4], int s) {
#pragma clang fp contract(off)
    if (s == 0 || s == 2) {
        const double eta = s == 0 ? -1.0 : 1.0;
        const double mx = ((1 - eta) * (P[1].x - P[0].x) + (1 + eta) * (P[2].x - P[3].x)) / 4;
        const double my = ((1 - eta) * (P[1].y - P[0].y) + (1 + eta) * (P[2].y - P[3].y)) / 4;
        return sqrt(mx * mx + my * my);
    }
    const double xi = s == 1 ? 1.0 : -1.0;
    const double mx = ((1 - xi) * (P[3].x - P[0].x) + (1 + xi) * (P[2].x - P[1].x)) / 4;
    const double my = ((1 - xi) * (P[3].y - P[0].y) + (1 + xi) * (P[2].y - P[1].y)) / 4;
    return sqrt(mx * mx + my * my);
}

// quad_map (include/functions/quad_map.tpp:3-8).
__device__ v2 quad_point(double xi, double eta, const v2 P[4]) {
#pragma clang fp contract(off)
    return v2{(P[0].x * (1 - xi) * (1 - eta) + P[1].x * (xi + 1) * (1 - eta) + P[2].x * (xi + 1) * (eta + 1) +
               P[3].x * (1 - xi) * (eta + 1)) / 4,
              (P[0].y * (1 - xi) * (1 - eta) + P[1].y * (xi + 1) * (1 - eta) + P[2].y * (xi + 1) * (eta + 1) +
               P[3].y * (1 - xi) * (eta + 1)) / 4};
}

__device__ double cross(v2 a, v2 b) {
#pragma clang fp contract(off)
    return a.x * b.y - a.y * b.x;
}

// inverse_quad_map (include/functions/inverse_quad_map.tpp:46-88): the root of the bilinear map's
// quadratic inside the unit square.
__device__ v2 reference_point(v2 g, const v2 P[4]) {
#pragma clang fp contract(off)
    const v2 e{P[1].x - P[0].x, P[1].y - P[0].y}, f{P[3].x - P[0].x, P[3].y - P[0].y};
    const v2 gg{P[0].x - P[1].x + P[2].x - P[3].x, P[0].y - P[1].y + P[2].y - P[3].y};
    const v2 h{g.x - P[0].x, g.y - P[0].y};
    const double k2 = cross(gg, f), k1 = cross(e, f) + cross(h, gg), k0 = cross(h, e);
    if (fabs(k2) < DBL_MIN) {
        const double ux = e.x * k1 - gg.x * k0, uy = e.y * k1 - gg.y * k0;
        const double u = fabs(ux) > fabs(uy) ? (h.x * k1 + f.x * k0) / ux : (h.y * k1 + f.y * k0) / uy;
        return v2{u * 2 - 1, -k0 / k1 * 2 - 1};
    }
    double w = k1 * k1 - 4 * k0 * k2;
    if (w < 0.0) return v2{-1, -1};
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
    return v2{u * 2 - 1, v * 2 - 1};
}

__device__ v2 midpoint(v2 a, v2 b) {
#pragma clang fp contract(off)
    return v2{(a.x + b.x) / 2, (a.y + b.y) / 2};
}

// ---------------------------------------------------------------------------------------------
// Exclusive scan of int64 counts in place: x[0..n) becomes its exclusive prefix sums and x[n] the
// total. Tiles of kScanTile values (kScanItems consecutive values per thread), three launches.

__device__ int64_t block_exclusive(int64_t v, int64_t *lds, int64_t *total) {
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kThreads; o <<= 1) {
        const int64_t t = threadIdx.x >= (unsigned)o ? lds[threadIdx.x - o] : 0;
        __syncthreads();
        lds[threadIdx.x] += t;
        __syncthreads();
    }
    const int64_t incl = lds[threadIdx.x];
    *total = lds[kThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ void __launch_bounds__(kThreads) scan_tiles_kernel(const int64_t *__restrict__ x, int64_t n,
                                                               int64_t *__restrict__ tile_sum) {
    __shared__ int64_t lds[kThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v = 0;
    for (int i = 0; i < kScanItems; ++i)
        if (base + i < n) v += x[base + i];
    int64_t total;
    (void)block_exclusive(v, lds, &total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// One workgroup: the tile sums become their exclusive prefixes; x[n] = tile_sum[n_tiles] = total.
__global__ void __launch_bounds__(kThreads) scan_tile_sums_kernel(int64_t *__restrict__ tile_sum, int64_t n_tiles,
                                                                   int64_t *__restrict__ x, int64_t n) {
    __shared__ int64_t lds[kThreads];
    const int64_t per = (n_tiles + kThreads - 1) / kThreads;
    const int64_t b0 = (int64_t)threadIdx.x * per;
    int64_t v = 0;
    for (int64_t i = 0; i < per; ++i)
        if (b0 + i < n_tiles) v += tile_sum[b0 + i];
    int64_t total;
    int64_t run = block_exclusive(v, lds, &total);
    for (int64_t i = 0; i < per; ++i)
        if (b0 + i < n_tiles) {
            const int64_t t = tile_sum[b0 + i];
            tile_sum[b0 + i] = run;
            run += t;
        }
    if (threadIdx.x == 0) {
        tile_sum[n_tiles] = total;
        x[n] = total;
    }
}

__global__ void __launch_bounds__(kThreads) scan_apply_kernel(int64_t *__restrict__ x, int64_t n,
                                                               const int64_t *__restrict__ tile_sum) {
    __shared__ int64_t lds[kThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems], sum = 0;
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = base + i < n ? x[base + i] : 0;
        sum += v[i];
    }
    int64_t total;
    int64_t run = block_exclusive(sum, lds, &total) + tile_sum[blockIdx.x];
    for (int i = 0; i < kScanItems; ++i)
        if (base + i < n) {
            x[base + i] = run;
            run += v[i];
        }
}

// x must hold n + 1 values; blocks (work arena) holds the tile sums.
hipError_t exclusive_scan(int64_t *x, int64_t n, Arena &blocks, hipStream_t s) {
    const int64_t n_tiles = std::max<int64_t>(1, (n + kScanTile - 1) / kScanTile);
    hipError_t e = arena_fit(blocks, sizeof(int64_t) * (n_tiles + 1));
    if (e != hipSuccess) return e;
    int64_t *ts = (int64_t *)blocks.p;
    hipLaunchKernelGGL(scan_tiles_kernel, dim3((unsigned)n_tiles), dim3(kThreads), 0, s, x, n, ts);
    hipLaunchKernelGGL(scan_tile_sums_kernel, dim3(1), dim3(kThreads), 0, s, ts, n_tiles, x, n);
    hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)n_tiles), dim3(kThreads), 0, s, x, n, ts);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// 1. Decisions.

__global__ void __launch_bounds__(kThreads) decide_kernel(LogicalMesh L, const double *__restrict__ est, int N_max,
                                                           int tables_N_max, int max_split_level,
                                                           int8_t *__restrict__ flags, AdaptCounts *c) {
    const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    unsigned long long split = 0, raise = 0, dofs = 0, nodes = 0;
    bool bad = false;
    if (e < L.K) {
        const double *ee = est + 4 * L.dev_of[e];
        const double err = ee[0], sig = ee[1], ref = ee[2], coa = ee[3];
        L.est[4 * e + 0] = err;
        L.est[4 * e + 1] = sig;
        L.est[4 * e + 2] = ref;
        L.est[4 * e + 3] = coa;
        const int N = L.order[e];
        const bool refine = ref != 0.0;
        int f = 0;
        if (refine && sig < 1.0 && L.level[e] < max_split_level) f = 2;       // would_h_refine
        else if (refine && sig >= 1.0 && N + 2 <= N_max) f = 1;               // would_p_refine
        flags[e] = (int8_t)f;
        const int Nn = N + (f == 1 ? 2 : 0);
        const unsigned long long k = f == 2 ? 4 : 1;
        split = f == 2;
        raise = f == 1;
        dofs = k * (unsigned long long)((Nn + 1) * (Nn + 1));
        nodes = k * (unsigned long long)(Nn + 1);
        bad = f == 1 && (Nn > kMaxOrder || Nn > tables_N_max);
    }
    wave_add(&c->n_split, split);
    wave_add(&c->n_raise, raise);
    wave_add(&c->dofs, dofs);
    wave_add(&c->nodes1d, nodes);
    if (bad) atomicOr(&c->refused, (unsigned long long)kRefuseOrder);
}

// Refusals of the split branch (a small hanging element already covering 1 / 2^kMaxSpanLevel of a
// larger side, across the interior or a period, that splits while the larger element does not)
// and the ghosts that split with their element.
__global__ void __launch_bounds__(kThreads) check_kernel(LogicalMesh L, const int8_t *__restrict__ flags,
                                                          AdaptCounts *c) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    unsigned long long refused = 0, gs = 0;
    if (i < L.K && flags[i] == 2)
        for (int s = 0; s < 4; ++s) {
            const int64_t es = 4 * i + s;
            if (L.hang[es] == 1 && class_level(L.span[es]) >= kMaxSpanLevel && flags[L.link[es] / 4] != 2)
                refused |= kRefuseSpan;
        }
    if (i < L.n_ghost && flags[L.ghost_side[i] / 4] == 2) gs = 1;
    if (refused) atomicOr(&c->refused, refused);
    wave_add(&c->ghost_split, gs);
}

// ---------------------------------------------------------------------------------------------
// 2. The new logical mesh.

__global__ void count_kernel(const int8_t *__restrict__ flags, int64_t K, int64_t *__restrict__ newpos,
                             const int64_t *__restrict__ ghost_side, int64_t n_ghost, int64_t *__restrict__ newghost) {
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < K) newpos[i] = flags[i] == 2 ? 4 : 1;
    if (i < n_ghost) newghost[i] = flags[ghost_side[i] / 4] == 2 ? 2 : 1;
}

// Small sides along a larger side: the chain from its link through hang2, in the larger side's
// direction (at most 2^kMaxSpanLevel of them; the bound only guards the walk).
constexpr int kMaxChain = 1 << kMaxSpanLevel;

__device__ int chain_length(const LogicalMesh &L, int64_t large) {
    int n = 0;
    for (int64_t c = L.link[large]; c >= 0 && n < kMaxChain; c = L.hang2[c]) ++n;
    return n;
}

// Position of small side t in its larger side's chain.
__device__ int chain_position(const LogicalMesh &L, int64_t t) {
    int pos = 0;
    for (int64_t c = L.link[L.link[t]]; c != t && c >= 0 && pos < kMaxChain; c = L.hang2[c]) ++pos;
    return pos;
}

// Half (0 at the larger side's first corner) of the larger side that small side t lies in.
__device__ int span_half(const LogicalMesh &L, int64_t t) {
    const int cls = L.span[t], d = class_level(cls);
    return class_index(cls) >> (d - 1);
}

// The middle of side s of old element e: on a larger interior hanging side the node its small
// sides share there (the start of the small side whose piece ends at the middle: small sides run
// against the larger side), else the midpoint of the side's corners (the small sides of a larger
// periodic side lie at the other end of the period).
__device__ v2 side_middle(const LogicalMesh &L, int64_t e, int s) {
    const int64_t es = 4 * e + s;
    if (L.hang[es] == 2 && !L.periodic[es]) {
        int n = 0;
        for (int64_t f = L.link[es]; f >= 0 && n < kMaxChain; f = L.hang2[f], ++n) {
            const int cls = L.span[f];
            if ((class_index(cls) + 1) * 2 == (1 << class_level(cls))) return corner(L, f / 4, (int)(f % 4));
        }
    }
    return midpoint(corner(L, e, s), corner(L, e, s + 1));
}

__global__ void __launch_bounds__(kThreads) elements_kernel(LogicalMesh L, LogicalMesh Ln,
                                                             const int8_t *__restrict__ flags,
                                                             const int64_t *__restrict__ newpos,
                                                             int64_t *__restrict__ parent, int8_t *__restrict__ corner_of) {
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= L.K) return;
    const int f = flags[e];
    const int64_t np = newpos[e];
    const int N = L.order[e];
    if (f != 2) {
        for (int k = 0; k < 8; ++k) Ln.verts[8 * np + k] = L.verts[8 * e + k];
        Ln.order[np] = N + (f == 1 ? 2 : 0);
        Ln.status[np] = L.status[e];
        Ln.rotation[np] = L.rotation[e];
        Ln.level[np] = L.level[e];
        for (int k = 0; k < 4; ++k) Ln.est[4 * np + k] = f == 1 ? 0.0 : L.est[4 * e + k];
        parent[np] = e;
        corner_of[np] = -1;
        return;
    }
    v2 P[4], md[4];
    for (int k = 0; k < 4; ++k) P[k] = corner(L, e, k);
    for (int s = 0; s < 4; ++s) md[s] = side_middle(L, e, s);
    double cx = 0.0, cy = 0.0;
    for (int k = 0; k < 4; ++k) {
        cx += P[k].x;
        cy += P[k].y;
    }
    const v2 c{cx / 4, cy / 4};
    const int rot = L.rotation[e], st = L.status[e];
    for (int s = 0; s < 4; ++s) {
        const int64_t at = np + dChildOrder[rot][st][s];
        const v2 pos[4] = {P[s], md[s], c, md[(s + 3) & 3]};
        for (int k = 0; k < 4; ++k) {
            Ln.verts[8 * at + 2 * ((s + k) & 3)] = pos[k].x;
            Ln.verts[8 * at + 2 * ((s + k) & 3) + 1] = pos[k].y;
        }
        Ln.order[at] = N;
        Ln.status[at] = dChildStatus[rot][st][s];
        Ln.rotation[at] = rot;
        Ln.level[at] = L.level[e] + 1;
        for (int k = 0; k < 4; ++k) Ln.est[4 * at + k] = 0.0;
        parent[at] = e;
        corner_of[at] = (int8_t)s;
    }
}

struct SideMap {
    const LogicalMesh &L;
    const int64_t *newpos;
    const int8_t *flags;
    // Side rs of unsplit old element r in the new numbering.
    __device__ int64_t same(int64_t r, int rs) const { return 4 * newpos[r] + rs; }
    // The child of split old element r covering half h of its side rs (h = 0: the half at the
    // side's first corner): the child at corner rs + h, whose side rs it is.
    __device__ int64_t half(int64_t r, int rs, int h) const {
        return 4 * (newpos[r] + dChildOrder[L.rotation[r]][L.status[r]][(rs + h) & 3]) + rs;
    }
    // The first new side of old small side n in its larger side's direction (n runs the other way,
    // so a split n's half at its second corner comes first); -1 for n < 0.
    __device__ int64_t first_piece(int64_t n) const {
        if (n < 0) return -1;
        const int64_t r = n / 4;
        const int rs = (int)(n % 4);
        return flags[r] == 2 ? half(r, rs, 1) : same(r, rs);
    }
};

// The exterior of every new side from the old topology alone, one thread per side. A small side
// covers the piece span = (d, j) of its larger side (in the larger side's direction) and links to
// the next small side along it; a larger side links to its first small side. Through a period the
// partner's side runs against this one, as an interior neighbour's does, so periodic sides follow
// the same rules and keep their flag.
__global__ void __launch_bounds__(kThreads) links_kernel(LogicalMesh L, LogicalMesh Ln, const int8_t *__restrict__ flags,
                                                          const int64_t *__restrict__ newpos,
                                                          const int64_t *__restrict__ newghost,
                                                          const int64_t *__restrict__ parent,
                                                          const int8_t *__restrict__ corner_of) {
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= 4 * Ln.K) return;
    const SideMap M{L, newpos, flags};
    const int64_t k = t / 4, e = parent[k];
    const int s = (int)(t % 4), c = corner_of[k];
    int64_t link = 0, next = -1;
    int hang = 0, per = 0, span = 0;
    // A small side (old side es, piece (d, j) of larger side l of old element r) whose larger
    // element splits: it lies in half H of l, covered by r's child there, in which its piece is
    // (d - 1, j mod 2^(d-1)) refined by (dd, jj) (for a half of es: one level more).
    auto under_split_large = [&](int64_t es, int64_t l, int dd, int jj, int64_t nxt) {
        const int cls = L.span[es], d = class_level(cls), j = class_index(cls), H = j >> (d - 1);
        const int d2 = d - 1 + dd, j2 = ((j & ((1 << (d - 1)) - 1)) << dd) + jj;
        link = M.half(l / 4, (int)(l % 4), H);
        if (d2 == 0) return;  // the whole side of r's child: conforming
        hang = 1;
        span = make_class(d2, j2);
        next = nxt >= 0 && span_half(L, nxt) == H ? M.first_piece(nxt) : -1;
    };
    if (c < 0) {  // unsplit element: side s of e
        const int64_t es = 4 * e + s, l = L.link[es];
        const int hk = L.hang[es];
        per = L.periodic[es];
        if (l < 0) {
            link = -(1 + newghost[-l - 1]);
        } else if (hk == 0) {
            const int64_t r = l / 4;
            const int rs = (int)(l % 4);
            if (flags[r] != 2) {
                link = M.same(r, rs);
            } else {  // the neighbour splits: this side becomes the larger side of its two halves
                hang = 2;
                link = M.half(r, rs, 1);
            }
        } else if (hk == 2) {  // still the larger side; its small elements may have split
            hang = 2;
            link = M.first_piece(l);
        } else {  // small side of the larger side l
            if (flags[l / 4] != 2) {
                hang = 1;
                link = M.same(l / 4, (int)(l % 4));
                span = L.span[es];
                next = M.first_piece(L.hang2[es]);
            } else {
                under_split_large(es, l, 0, 0, L.hang2[es]);
            }
        }
    } else {  // child at corner c of split element e
        const int rel = (s - c) & 3;
        const int64_t base = newpos[e];
        const int *co = dChildOrder[L.rotation[e]][L.status[e]];
        if (rel == 1) {         // centre line shared with the child at corner c + 1
            link = 4 * (base + co[(c + 1) & 3]) + ((c + 3) & 3);
        } else if (rel == 2) {  // centre line shared with the child at corner c - 1
            link = 4 * (base + co[(c + 3) & 3]) + c;
        } else {
            // side c is the first half of the parent's side c, side c + 3 the second half of the
            // parent's side c + 3
            const int ps = rel == 0 ? c : (c + 3) & 3, h = rel == 0 ? 0 : 1;
            const int64_t es = 4 * e + ps, l = L.link[es];
            const int hk = L.hang[es];
            per = L.periodic[es];
            // in the exterior side's direction this half is piece 1 - h of the parent side, the
            // half h = 1 first, followed by its sibling half
            const int64_t sibling = M.half(e, ps, 0);
            if (l < 0) {  // the ghost splits: ghost g + 1 covers the first half, g the second
                link = -(1 + newghost[-l - 1] + (h == 0 ? 1 : 0));
            } else if (hk == 0) {
                const int64_t r = l / 4;
                const int rs = (int)(l % 4);
                if (flags[r] == 2) {  // both split: the neighbour's side runs the other way
                    link = M.half(r, rs, 1 - h);
                } else {
                    hang = 1;
                    link = M.same(r, rs);
                    span = make_class(1, 1 - h);
                    next = h == 1 ? sibling : -1;
                }
            } else if (hk == 1) {  // a piece of the larger side l
                const int cls = L.span[es], d = class_level(cls), j = class_index(cls);
                const int64_t nxt = h == 1 ? -2 : L.hang2[es];
                if (flags[l / 4] != 2) {  // one level finer (refused beyond kMaxSpanLevel)
                    hang = 1;
                    link = M.same(l / 4, (int)(l % 4));
                    span = make_class(d + 1, 2 * j + 1 - h);
                    next = h == 1 ? sibling : M.first_piece(nxt);
                } else {
                    under_split_large(es, l, 1, 1 - h, nxt >= 0 ? nxt : -1);
                    if (h == 1) next = sibling;
                }
            } else {  // the larger side: the small sides on this half
                int64_t f = l;
                for (int n = 0; f >= 0 && n < kMaxChain && span_half(L, f) != h; ++n) f = L.hang2[f];
                const int cls = f >= 0 ? L.span[f] : 0;
                if (f >= 0 && class_level(cls) == 1 && flags[f / 4] != 2) {  // one small side, unsplit
                    link = M.same(f / 4, (int)(f % 4));
                } else {
                    hang = 2;
                    link = M.first_piece(f);
                }
            }
        }
    }
    Ln.link[t] = link;
    Ln.hang2[t] = next;
    Ln.hang[t] = (int8_t)hang;
    Ln.periodic[t] = (int8_t)per;
    Ln.span[t] = (int8_t)span;
}

// Ghosts: a ghost on a split side becomes the ghosts of bars (P(s+1), m) and (m, P(s)), on the
// children's sides m -> P(s+1) and P(s) -> m, with fresh geometry (split_boundaries,
// src/meshes/Mesh2D_t.cpp:6232-6343); the others keep their record. A ghost keeps a zero scale
// if its storage was resized: in the no-split branch once its element is raised (the reference's
// adjust_boundaries resize, Element2D_t.cpp:735-741), in the split branch until then.
__global__ void __launch_bounds__(kThreads) ghosts_kernel(LogicalMesh L, LogicalMesh Ln, const int8_t *__restrict__ flags,
                                                           const int64_t *__restrict__ newpos,
                                                           const int64_t *__restrict__ newghost, int p_only) {
    const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (g >= L.n_ghost) return;
    const int64_t es = L.ghost_side[g], e = es / 4, ng = newghost[g];
    const int s = (int)(es % 4), f = flags[e];
    if (f != 2) {
        Ln.ghost[ng] = L.ghost[g];
        Ln.ghost_side[ng] = 4 * newpos[e] + s;
        const bool z = L.ghost_zero[g] != 0;
        Ln.ghost_zero[ng] = (int8_t)(p_only ? (z || f == 1) : (z && f != 1));
        return;
    }
    const SideMap M{L, newpos, flags};
    const v2 u0 = corner(L, e, s), u1 = corner(L, e, s + 1), md = midpoint(u0, u1);
    for (int h2 = 0; h2 < 2; ++h2) {
        const v2 a0 = h2 == 0 ? md : u0, a1 = h2 == 0 ? u1 : md;  // the element side a0 -> a1
        BcSide rec = L.ghost[g];
        rec.ax = a1.x;
        rec.ay = a1.y;
        rec.bx = a0.x;
        rec.by = a0.y;
        const v2 G[4] = {a1, a0, a0, a1};
        rec.sf = side_scale(G, 0);
        face_frame(a0, a1, a0, a1, rec.nx, rec.ny, rec.tx, rec.ty);
        rec.slot = -1;
        Ln.ghost[ng + h2] = rec;
        Ln.ghost_side[ng + h2] = M.half(e, s, h2 == 0 ? 1 : 0);
        Ln.ghost_zero[ng + h2] = 0;
    }
}

// ---------------------------------------------------------------------------------------------
// 3. Layout: stable counting sort by order, mortar numbering.

__global__ void __launch_bounds__(kThreads) order_hist_kernel(const int32_t *__restrict__ order, int64_t K,
                                                               int64_t *__restrict__ hist) {
    __shared__ int cnt[kOrders];
    if (threadIdx.x < kOrders) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (e < K) atomicAdd(&cnt[order[e]], 1);
    __syncthreads();
    if (threadIdx.x < kOrders) hist[(int64_t)blockIdx.x * kOrders + threadIdx.x] = cnt[threadIdx.x];
}

// One workgroup per order: the column of block counts becomes its exclusive prefixes.
__global__ void __launch_bounds__(kThreads) order_scan_kernel(int64_t *__restrict__ hist, int64_t nb, AdaptCounts *c) {
    __shared__ int64_t lds[kThreads];
    const int n = blockIdx.x;
    const int64_t per = (nb + kThreads - 1) / kThreads, b0 = (int64_t)threadIdx.x * per;
    int64_t v = 0;
    for (int64_t i = 0; i < per; ++i)
        if (b0 + i < nb) v += hist[(b0 + i) * kOrders + n];
    int64_t total;
    int64_t run = block_exclusive(v, lds, &total);
    for (int64_t i = 0; i < per; ++i)
        if (b0 + i < nb) {
            int64_t &h = hist[(b0 + i) * kOrders + n];
            const int64_t t = h;
            h = run;
            run += t;
        }
    if (threadIdx.x == 0) c->order_count[n] = (unsigned long long)total;
}

__global__ void spans_kernel(const AdaptCounts *c, GroupSpan *spans) {
    if (threadIdx.x != 0) return;
    int64_t e_begin = 0, q_off = 0, tr_off = 0;
    for (int n = 0; n < kOrders; ++n) {
        const int64_t K = (int64_t)c->order_count[n];
        spans[n] = GroupSpan{K, e_begin, q_off, tr_off};
        e_begin += K;
        q_off += 3 * (int64_t)(n + 1) * (n + 1) * K;
        tr_off += 12 * (int64_t)(n + 1) * K;
    }
}

__global__ void __launch_bounds__(kThreads) order_place_kernel(const int32_t *__restrict__ order, int64_t K,
                                                                const int64_t *__restrict__ hist,
                                                                const GroupSpan *__restrict__ spans,
                                                                int64_t *__restrict__ dev_of) {
    __shared__ int8_t ord[kThreads];
    const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int o = e < K ? order[e] : -1;
    ord[threadIdx.x] = (int8_t)o;
    __syncthreads();
    if (e >= K) return;
    int rank = 0;
    for (unsigned j = 0; j < threadIdx.x; ++j) rank += ord[j] == o;
    dev_of[e] = spans[o].e_begin + hist[(int64_t)blockIdx.x * kOrders + o] + rank;
}

// A side is a mortar side (its flux comes from the mortar kernel) when it is hanging or its
// exterior has another order. The mortar faces a side owns: an interior small hanging side owns
// the face it shares with the larger side, an interior side of full length the face of a pair of
// different orders when it is the lower-numbered element's; a periodic mortar side owns its own
// faces against its partner's trace, one per face on the side (two on a larger side), as the
// reference's periodic ghosts give each end of the period its own faces.
__device__ void side_roles(const LogicalMesh &L, int64_t t, bool &mortar, int &owned) {
    const int64_t k = t / 4, l = L.link[t];
    const int hk = L.hang[t];
    const bool other = l >= 0 && L.order[l / 4] != L.order[k];
    mortar = hk != 0 || other;
    if (L.periodic[t]) owned = mortar ? (hk == 2 ? chain_length(L, t) : 1) : 0;
    else owned = hk == 1 || (hk == 0 && other && l / 4 > k);
}

__global__ void __launch_bounds__(kThreads) side_count_kernel(LogicalMesh L, int64_t *__restrict__ cnt_rec,
                                                               int64_t *__restrict__ cnt_slot,
                                                               int64_t *__restrict__ cnt_face, AdaptCounts *c) {
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    unsigned long long pairs = 0, hfaces = 0, pfaces = 0;
    if (t < 4 * L.K) {
        bool mortar;
        int owned;
        side_roles(L, t, mortar, owned);
        const int64_t k = t / 4, l = L.link[t];
        const int N = L.order[k], hk = L.hang[t];
        const bool per = L.periodic[t] != 0;
        const int n_faces = hk == 2 ? chain_length(L, t) : 1;
        cnt_rec[t] = mortar;
        cnt_slot[t] = mortar ? 3 * (int64_t)(N + 1) * n_faces : 0;
        cnt_face[t] = owned;
        pairs = l >= 0 && hk == 0 && !per && l > t;
        hfaces = hk == 2 && !per ? n_faces : 0;
        pfaces = per ? n_faces : 0;
        if (owned) {
            int Nr = L.order[l / 4];
            if (hk == 2) {
                int n = 0;
                for (int64_t f = L.hang2[l]; f >= 0 && n < kMaxChain; f = L.hang2[f], ++n)
                    Nr = max(Nr, (int)L.order[f / 4]);
            }
            atomicMax(&c->mortar_nodes, (unsigned long long)(max(N, Nr) + 1));
        }
    }
    wave_add(&c->conforming_pairs, pairs);
    wave_add(&c->hanging_faces, hfaces);
    wave_add(&c->periodic_faces, pfaces);
}

__global__ void totals_kernel(const int64_t *rec, const int64_t *slot, const int64_t *face, int64_t n, AdaptCounts *c) {
    if (threadIdx.x != 0) return;
    c->n_rec = (unsigned long long)rec[n];
    c->n_slot = (unsigned long long)slot[n];
    c->n_face = (unsigned long long)face[n];
}

// ---------------------------------------------------------------------------------------------
// 4. Device mesh arrays.

struct FillArgs {
    const GroupSpan *spans;
    const int64_t *key_off, *rec_idx, *slot_off, *face_idx;
    double *verts;
    int32_t *side_nb;
    BcSide *bc;
    MortarFace *faces;
    int64_t n_ghost;
    AdaptCounts *c;
};

// The matrix offsets of a mortar face's two sides from their offset/scale classes (-1: a plain
// copy); a face whose sides fit no class is counted as bad.
__device__ void face_matrices(MortarFace &fc, const int64_t *key_off, AdaptCounts *c) {
    for (int side = 0; side < 2; ++side) {
        const int Ne = side ? fc.NR : fc.NL;
        const int cls = mortar_class(fc.offset[side], fc.scale[side]);
        if (cls < 0) {
            atomicAdd(&c->bad_face, 1ull);
            fc.mat[side] = fc.mat[2 + side] = -1;
            continue;
        }
        const bool copy = Ne == fc.Nf && cls == 0;
        fc.mat[side] = copy ? -1 : key_off[mortar_matrix_key(Ne, fc.Nf, side, cls, 0)];
        fc.mat[2 + side] = copy ? -1 : key_off[mortar_matrix_key(Ne, fc.Nf, side, cls, 1)];
    }
}

__global__ void __launch_bounds__(kThreads) side_fill_kernel(LogicalMesh L, FillArgs A) {
    const int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (t >= 4 * L.K) return;
    const int64_t k = t / 4, l = L.link[t];
    const int s = (int)(t % 4), N = L.order[k], nf = N + 1, hk = L.hang[t];
    const int64_t dv = L.dev_of[k];
    const GroupSpan &G = A.spans[N];
    const int64_t local = dv - G.e_begin;
    if (s == 0)
        for (int i = 0; i < 8; ++i) A.verts[8 * dv + i] = L.verts[8 * k + i];
    bool mortar;
    int owned;
    side_roles(L, t, mortar, owned);
    int64_t code;
    if (mortar) {
        code = -(1 + A.n_ghost + A.rec_idx[t]);
        BcSide rec{};
        rec.kind = kKindMortar;
        rec.slot = A.slot_off[t];
        rec.n_faces = hk == 2 ? chain_length(L, t) : 1;
        A.bc[A.n_ghost + A.rec_idx[t]] = rec;
    } else if (l < 0) {
        code = l;
    } else {
        const int64_t r = l / 4;
        code = (L.dev_of[r] - A.spans[L.order[r]].e_begin) * 4 + l % 4;
    }
    A.side_nb[4 * dv + s] = (int32_t)code;
    if (!owned) return;
    auto trace_of = [&](int64_t side) {
        const int64_t e = side / 4;
        const int n = L.order[e] + 1;
        const GroupSpan &S = A.spans[n - 1];
        return S.tr_off + (L.dev_of[e] - S.e_begin) * 12 * n + (side % 4) * 3 * n;
    };
    const v2 c0 = corner(L, k, s), c1 = corner(L, k, s + 1);
    const v2 PL[4] = {corner(L, k, 0), corner(L, k, 1), corner(L, k, 2), corner(L, k, 3)};
    const double sfL = side_scale(PL, s);
    if (L.periodic[t]) {
        // One face per face of this side, against a ghost holding the partner side's trace (in
        // the partner's direction, which runs against this side through the period): a whole side
        // against a whole partner or against its piece of a larger partner, or each piece of a
        // larger side against its whole small partner, in the side's direction.
        double tx, ty, nx, ny;
        face_frame(c0, c1, c0, c1, nx, ny, tx, ty);
        int64_t lr = l;
        for (int h = 0; h < owned; ++h, lr = L.hang2[lr]) {
            const int NR = L.order[lr / 4];
            MortarFace fc{};
            fc.NL = N;
            fc.NR = NR;
            fc.Nf = max(N, NR);
            fc.trL = G.tr_off + local * 12 * nf + s * 3 * nf;
            fc.trR = trace_of(lr);
            fc.nx = nx;
            fc.ny = ny;
            const int own = hk == 2 ? L.span[lr] : 0, ghost = hk == 1 ? L.span[t] : 0;
            fc.offset[0] = class_offset(own);
            fc.scale[0] = class_scale(own);
            fc.offset[1] = class_offset(ghost);
            fc.scale[1] = class_scale(ghost);
            fc.sfL = sfL;
            fc.slotL = A.slot_off[t] + h * 3 * nf;
            fc.slotR = -1;
            face_matrices(fc, A.key_off, A.c);
            A.faces[A.face_idx[t] + h] = fc;
        }
        return;
    }
    const int64_t r = l / 4;
    const int rs = (int)(l % 4), NR = L.order[r], nfr = NR + 1;
    MortarFace fc{};
    fc.NL = N;
    fc.NR = NR;
    fc.Nf = max(N, NR);
    fc.trL = G.tr_off + local * 12 * nf + s * 3 * nf;
    fc.trR = trace_of(l);
    double tx, ty;
    face_frame(c0, c1, c0, c1, fc.nx, fc.ny, tx, ty);
    const v2 en[4] = {c0, c1, corner(L, r, rs), corner(L, r, rs + 1)};
    face_offsets(c0, c1, en, fc.offset, fc.scale);
    fc.sfL = sfL;
    fc.slotL = A.slot_off[t];
    const v2 PR[4] = {corner(L, r, 0), corner(L, r, 1), corner(L, r, 2), corner(L, r, 3)};
    fc.sfR = side_scale(PR, rs);
    fc.slotR = A.slot_off[l] + (hk == 1 ? chain_position(L, t) * 3 * (int64_t)nfr : 0);
    face_matrices(fc, A.key_off, A.c);
    A.faces[A.face_idx[t]] = fc;
}

__global__ void ghost_fill_kernel(LogicalMesh L, BcSide *__restrict__ bc) {
    const int64_t g = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (g >= L.n_ghost) return;
    BcSide rec = L.ghost[g];
    if (L.ghost_zero[g]) rec.sf = 0.0;
    bc[g] = rec;
}

// ---------------------------------------------------------------------------------------------
// 5. Transfer: one workgroup per new element.

__global__ void __launch_bounds__(kThreads) transfer_dev_kernel(LogicalMesh L, const GroupSpan *__restrict__ so,
                                                                 const double *__restrict__ q_old, LogicalMesh Ln,
                                                                 const GroupSpan *__restrict__ sn,
                                                                 double *__restrict__ q_new,
                                                                 const int64_t *__restrict__ parent,
                                                                 const int8_t *__restrict__ corner_of,
                                                                 const double *__restrict__ tab1d) {
    const int64_t k = blockIdx.x, e = parent[k];
    const int No = L.order[e], Nn = Ln.order[k], nfo = No + 1, nf = Nn + 1, neo = nfo * nfo, ne = nf * nf;
    const double *src = q_old + so[No].q_off + (L.dev_of[e] - so[No].e_begin) * 3 * neo;
    double *dst = q_new + sn[Nn].q_off + (Ln.dev_of[k] - sn[Nn].e_begin) * 3 * ne;
    const bool child = corner_of[k] >= 0;
    if (No == Nn && !child) {
        for (int r = threadIdx.x; r < 3 * ne; r += kThreads) dst[r] = src[r];
        return;
    }
    const double *xo = tab1d + (3 * No + 0) * kTab1dStride, *lo = tab1d + (3 * No + 2) * kTab1dStride;
    const double *xn = tab1d + (3 * Nn + 0) * kTab1dStride;
    v2 Pc[4], Pp[4];
    for (int q = 0; q < 4; ++q) {
        Pc[q] = corner(Ln, k, q);
        Pp[q] = corner(L, e, q);
    }
    for (int node = threadIdx.x; node < ne; node += kThreads) {
        const int i = node / nf, j = node % nf;
        double X = xn[i], Y = xn[j];
        if (child) {
            const v2 r = reference_point(quad_point(xn[i], xn[j], Pc), Pp);
            X = r.x;
            Y = r.y;
        }
        for (int f = 0; f < 3; ++f) dst[f * ne + node] = interpolate_at(src + f * neo, No, X, Y, xo, lo);
    }
}

__global__ void copy_estimate_kernel(LogicalMesh L, const double *__restrict__ est) {
    const int64_t e = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (e >= L.K) return;
    for (int k = 0; k < 4; ++k) L.est[4 * e + k] = est[4 * L.dev_of[e] + k];
}

template <typename T>
hipError_t fit(Arena &a, int64_t n, T **p) {
    const hipError_t e = arena_fit(a, sizeof(T) * (size_t)std::max<int64_t>(n, 1));
    *p = (T *)a.p;
    return e;
}

#define TRY(x)                                  \
    do {                                        \
        const hipError_t _e = (x);              \
        if (_e != hipSuccess) return _e;        \
    } while (0)

}  // namespace

hipError_t adapt_decide(const LogicalMesh &L, const double *est, int N_max, int tables_N_max, int max_split_level,
                        AdaptWork &w, hipStream_t s) {
    int8_t *flags;
    AdaptCounts *c;
    TRY(fit(w.flags, L.K, &flags));
    TRY(fit(w.counts, 1, &c));
    TRY(hipMemsetAsync(c, 0, sizeof(AdaptCounts), s));
    hipLaunchKernelGGL(decide_kernel, dim3(grid_for(L.K)), dim3(kThreads), 0, s, L, est, N_max, tables_N_max,
                       max_split_level, flags, c);
    hipLaunchKernelGGL(check_kernel, dim3(grid_for(std::max(L.K, L.n_ghost))), dim3(kThreads), 0, s, L, flags, c);
    return hipGetLastError();
}

hipError_t adapt_rebuild(const LogicalMesh &L, LogicalMesh &Ln, bool p_only, AdaptWork &w, hipStream_t s) {
    const int8_t *flags = (const int8_t *)w.flags.p;
    int64_t *newpos, *newghost, *parent;
    int8_t *corner_of;
    TRY(fit(w.newpos, L.K + 1, &newpos));
    TRY(fit(w.newghost, L.n_ghost + 1, &newghost));
    TRY(fit(w.parent, Ln.K, &parent));
    TRY(fit(w.corner, Ln.K, &corner_of));
    hipLaunchKernelGGL(count_kernel, dim3(grid_for(std::max(L.K, L.n_ghost))), dim3(kThreads), 0, s, flags, L.K,
                       newpos, L.ghost_side, L.n_ghost, newghost);
    TRY(exclusive_scan(newpos, L.K, w.blocks, s));
    if (L.n_ghost > 0) TRY(exclusive_scan(newghost, L.n_ghost, w.blocks, s));
    hipLaunchKernelGGL(elements_kernel, dim3(grid_for(L.K)), dim3(kThreads), 0, s, L, Ln, flags, newpos, parent,
                       corner_of);
    hipLaunchKernelGGL(links_kernel, dim3(grid_for(4 * Ln.K)), dim3(kThreads), 0, s, L, Ln, flags, newpos, newghost,
                       parent, corner_of);
    if (L.n_ghost > 0)
        hipLaunchKernelGGL(ghosts_kernel, dim3(grid_for(L.n_ghost)), dim3(kThreads), 0, s, L, Ln, flags, newpos,
                           newghost, p_only ? 1 : 0);
    return hipGetLastError();
}

hipError_t adapt_layout_count(LogicalMesh &Ln, GroupSpan *spans, AdaptWork &w, hipStream_t s) {
    const int64_t nb = grid_for(Ln.K), n_sides = 4 * Ln.K;
    int64_t *hist, *rec, *slot, *face;
    AdaptCounts *c = (AdaptCounts *)w.counts.p;
    TRY(fit(w.hist, nb * kOrders, &hist));
    TRY(fit(w.cnt_rec, n_sides + 1, &rec));
    TRY(fit(w.cnt_slot, n_sides + 1, &slot));
    TRY(fit(w.cnt_face, n_sides + 1, &face));
    hipLaunchKernelGGL(order_hist_kernel, dim3((unsigned)nb), dim3(kThreads), 0, s, Ln.order, Ln.K, hist);
    hipLaunchKernelGGL(order_scan_kernel, dim3(kOrders), dim3(kThreads), 0, s, hist, nb, c);
    hipLaunchKernelGGL(spans_kernel, dim3(1), dim3(64), 0, s, c, spans);
    hipLaunchKernelGGL(order_place_kernel, dim3((unsigned)nb), dim3(kThreads), 0, s, Ln.order, Ln.K, hist, spans,
                       Ln.dev_of);
    hipLaunchKernelGGL(side_count_kernel, dim3(grid_for(n_sides)), dim3(kThreads), 0, s, Ln, rec, slot, face, c);
    TRY(exclusive_scan(rec, n_sides, w.blocks, s));
    TRY(exclusive_scan(slot, n_sides, w.blocks, s));
    TRY(exclusive_scan(face, n_sides, w.blocks, s));
    hipLaunchKernelGGL(totals_kernel, dim3(1), dim3(64), 0, s, rec, slot, face, n_sides, c);
    return hipGetLastError();
}

hipError_t adapt_layout_fill(const LogicalMesh &Ln, const GroupSpan *spans, const int64_t *mortar_key_off,
                             DeviceMesh &d, AdaptWork &w, hipStream_t s) {
    FillArgs A;
    A.spans = spans;
    A.key_off = mortar_key_off;
    A.rec_idx = (const int64_t *)w.cnt_rec.p;
    A.slot_off = (const int64_t *)w.cnt_slot.p;
    A.face_idx = (const int64_t *)w.cnt_face.p;
    A.verts = d.verts;
    A.side_nb = d.side_nb;
    A.bc = d.bc;
    A.faces = d.faces;
    A.n_ghost = Ln.n_ghost;
    A.c = (AdaptCounts *)w.counts.p;
    hipLaunchKernelGGL(side_fill_kernel, dim3(grid_for(4 * Ln.K)), dim3(kThreads), 0, s, Ln, A);
    if (Ln.n_ghost > 0)
        hipLaunchKernelGGL(ghost_fill_kernel, dim3(grid_for(Ln.n_ghost)), dim3(kThreads), 0, s, Ln, d.bc);
    return hipGetLastError();
}

hipError_t adapt_transfer(const LogicalMesh &L, const GroupSpan *spans_old, const double *q_old,
                          const LogicalMesh &Ln, const GroupSpan *spans_new, double *q_new, const double *tab1d,
                          AdaptWork &w, hipStream_t s) {
    if (Ln.K <= 0) return hipSuccess;
    hipLaunchKernelGGL(transfer_dev_kernel, dim3((unsigned)Ln.K), dim3(kThreads), 0, s, L, spans_old, q_old, Ln,
                       spans_new, q_new, (const int64_t *)w.parent.p, (const int8_t *)w.corner.p, tab1d);
    return hipGetLastError();
}

hipError_t adapt_copy_estimate(const LogicalMesh &L, const double *est, hipStream_t s) {
    if (L.K <= 0) return hipSuccess;
    hipLaunchKernelGGL(copy_estimate_kernel, dim3(grid_for(L.K)), dim3(kThreads), 0, s, L, est);
    return hipGetLastError();
}

}  // namespace ndg
