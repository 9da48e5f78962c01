// internal.hpp -- structures shared by the host side of libndg_amd (mesh build, C ABI) and the
// HIP kernels. Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "ndg_amd.h"

namespace ndg {

constexpr double kWaveSpeed = 1.0;  // include/helpers/constants.h:15
constexpr int kMaxOrder = 16;       // largest order with a compiled kernel instance

// Boundary (ghost) side record. One per element side whose neighbour is a wall, symmetry,
// inflow or outflow ghost element. n/t are the face normal and tangent of the reference
// (Face2D_t::compute_geometry with the domain element on the left); a/b the ghost element's
// first two nodes after face building; sf the ghost's edge scaling factor (half the edge
// length, src/entities/Element2D_t.cpp:779-790).
struct BcSide {
    double nx, ny, tx, ty;
    double ax, ay, bx, by;
    double sf;
    int32_t kind;
    int32_t pad;
};

// Per-order constant tables in device memory, packed as Dhat[Nf*Nf], l(-1)[Nf], l(+1)[Nf], w[Nf],
// x[Nf], 1/w[Nf], then the modal matrix poly[Nf*Nf] used by the estimator.
inline int64_t table_pack_len(int N) {
    const int nf = N + 1;
    return (int64_t)nf * nf * 2 + 5 * nf;
}
inline int64_t table_poly_offset(int N) {
    const int nf = N + 1;
    return (int64_t)nf * nf + 5 * nf;
}

struct Tables {
    int N_max = 0;
    int64_t n_interp = 0;
    std::vector<std::vector<double>> tab[12];
};

void build_tables(int N_max, int64_t n_interp, Tables &t);
// Packs the device table block for order N (layout above).
std::vector<double> pack_tables(const Tables &t, int N);

// Device-side description of a uniform-order conforming mesh.
struct DeviceMesh {
    int N = 0;
    int64_t K = 0;
    double *verts = nullptr;     // [K][4][2]
    int32_t *side_nb = nullptr;  // [K][4]: nb*4+side (>=0) or -(1+bc index)
    BcSide *bc = nullptr;        // [n_bc]
    int64_t n_bc = 0;
    double *tab = nullptr;       // packed tables for N
    double *q = nullptr;         // [K][3][Ne]
    double *g = nullptr;         // [K][3][Ne]
    double *trace[2] = {nullptr, nullptr};  // [K][4][3][Nf]
    double *rhs = nullptr;       // [K][3][Ne], compute_rhs only (allocated on demand)
    double *scratch = nullptr;   // reductions
    int64_t scratch_len = 0;
    int cur = 0;                 // trace buffer holding the traces of the current state
    // FP32 form with compensated state (allocated by ndg_mesh_set_precision): the state is the
    // unevaluated sum qh + ql of two floats per value, the RK register and traces are floats.
    int precision = NDG_PRECISION_FP64;
    float *qh = nullptr, *ql = nullptr, *g32 = nullptr;  // [K][3][Ne]
    float *trace32[2] = {nullptr, nullptr};              // [K][4][3][Nf]
    float *tab32 = nullptr;                              // packed tables for N, as floats
};

// Kernel launchers (kernels.hip). All return hipSuccess or the first launch error.
hipError_t launch_traces(const DeviceMesh &m, hipStream_t s);
hipError_t launch_stage(DeviceMesh &m, int stage, double t_stage, double a, double gdt,
                        hipStream_t s);
hipError_t launch_rhs(DeviceMesh &m, double t, hipStream_t s);
// FP32 compensated form: conversions between the FP64 state (q, g) and (qh, ql, g32), traces of
// qh, and the stage.
hipError_t launch_to_c32(DeviceMesh &m, hipStream_t s);
hipError_t launch_from_c32(DeviceMesh &m, hipStream_t s);
hipError_t launch_traces_c32(const DeviceMesh &m, hipStream_t s);
hipError_t launch_stage_c32(DeviceMesh &m, int stage, double t_stage, double a, double gdt,
                            hipStream_t s);
hipError_t launch_ic(const DeviceMesh &m, int kind, int n_pulses, const double *pulses_dev,
                     hipStream_t s);
hipError_t launch_delta_t(const DeviceMesh &m, double cfl, double *dt_dev, hipStream_t s);
hipError_t launch_estimate(const DeviceMesh &m, double tol_min, double tol_max, double *out_dev,
                           hipStream_t s);
bool order_supported(int N);

// Records msg as the calling thread's last error and returns code.
ndg_status set_error(ndg_status code, const char *msg);

}  // namespace ndg
