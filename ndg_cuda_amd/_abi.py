"""ctypes binding of include/ndg_amd.h (libndg_amd.so, built in-tree by ndg_cuda_amd/build.py).

There is no fallback: if the library is missing or a call fails, an exception is raised.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libndg_amd.so")

NDG_OK = 0
NDG_ERR_INVALID = 1
NDG_ERR_HIP = 2
NDG_ERR_UNSUPPORTED = 3
NDG_ERR_NO_DEVICE = 4

BC_NONE, BC_WALL, BC_SYMMETRY, BC_INFLOW, BC_OUTFLOW, BC_PERIODIC = range(6)
IC_PLANE_WAVE, IC_PULSES, IC_PERIODIC_PULSES = 0, 1, 2
ADAPT_DEVICE, ADAPT_HOST = 0, 1
PRECISION_FP64, PRECISION_FP32_COMPENSATED = 0, 1
GRAPH_OFF, GRAPH_SEQUENTIAL, GRAPH_PARALLEL = 0, 1, 2
ADJOINT_C_RECOMPUTE, ADJOINT_C_STAGE_RECORDS = 0, 1

# ndg_output_fn: status (mesh, time, user)
OUTPUT_FN = ctypes.CFUNCTYPE(ctypes.c_int32, ctypes.c_void_p, ctypes.c_double, ctypes.c_void_p)

TABLE_NAMES = ["nodes", "weights", "barycentric_weights", "lagrange_interpolant_left",
               "lagrange_interpolant_right", "lagrange_interpolant_derivative_left",
               "lagrange_interpolant_derivative_right", "derivative_matrices",
               "g_hat_derivative_matrices", "derivative_matrices_hat", "interpolation_matrices",
               "polynomials"]

# Every symbol include/ndg_amd.h declares (checked by the CPU tests).
EXPORTED = ["ndg_last_error", "ndg_device_count", "ndg_tables_create", "ndg_tables_destroy",
            "ndg_tables_get", "ndg_square_mesh_sizes", "ndg_square_mesh_fill", "ndg_su2_sizes", "ndg_su2_fill", "ndg_su2_sizes_periodic", "ndg_su2_fill_periodic",
            "ndg_mesh_create",
            "ndg_mesh_destroy", "ndg_mesh_set_stream", "ndg_mesh_info", "ndg_mesh_set_precision", "ndg_set_ic",
            "ndg_set_state", "ndg_get_state", "ndg_get_solution", "ndg_get_complete_solution", "ndg_get_traces", "ndg_compute_rhs",
            "ndg_get_delta_t", "ndg_step", "ndg_solve", "ndg_solve_output", "ndg_run_steps", "ndg_pre_condition", "ndg_synchronize",
            "ndg_estimate_error", "ndg_mesh_set_graph", "ndg_mesh_autotune", "ndg_mesh_tuning",
            "ndg_mesh_set_wave_speed", "ndg_record_receivers", "ndg_adjoint_set_store_budget", "ndg_adjoint_set_c_derivative", "ndg_adjoint_gradient", "ndg_apply_rhs_transpose",
            "ndg_adapt", "ndg_mesh_set_adaptivity", "ndg_mesh_orders", "ndg_hilbert_child_order",
            "ndg_hilbert_element_status", "ndg_record_receivers_noalloc", "ndg_mesh_set_adapt_mode",
            "ndg_mesh_sides", "ndg_inverse_quad_map"]


class NDGError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[ndg status {code}] {message}")
        self.code = code


class MeshDesc(ctypes.Structure):
    _fields_ = [("n_nodes", ctypes.c_int64), ("xy", ctypes.c_void_p),
                ("n_quads", ctypes.c_int64), ("quads", ctypes.c_void_p), ("quad_N", ctypes.c_void_p),
                ("n_bars", ctypes.c_int64), ("bars", ctypes.c_void_p), ("bar_type", ctypes.c_void_p),
                ("bar_donor", ctypes.c_void_p), ("initial_N", ctypes.c_int32),
                ("maximum_N", ctypes.c_int32), ("max_split_level", ctypes.c_int32),
                ("tolerance_min", ctypes.c_double), ("tolerance_max", ctypes.c_double)]


class SquareMeshParams(ctypes.Structure):
    _fields_ = [("resolution", ctypes.c_int32), ("x_min", ctypes.c_double), ("x_max", ctypes.c_double),
                ("y_min", ctypes.c_double), ("y_max", ctypes.c_double), ("bc", ctypes.c_int32 * 4),
                ("x_periodic", ctypes.c_int32), ("y_periodic", ctypes.c_int32),
                ("jitter", ctypes.c_double), ("seed", ctypes.c_uint64)]


class Receivers(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("element", ctypes.c_void_p), ("xi_eta", ctypes.c_void_p)]


COMPLETE_POINT_FIELDS = ("x", "y", "p", "u", "v", "dp_dt", "du_dt", "dv_dt", "p_analytical_error",
                         "u_analytical_error", "v_analytical_error")
COMPLETE_ELEMENT_FIELDS = (("N", np.int32), ("p_error", np.float64), ("u_error", np.float64),
                           ("v_error", np.float64), ("p_sigma", np.float64), ("u_sigma", np.float64),
                           ("v_sigma", np.float64), ("refine", np.int32), ("coarsen", np.int32),
                           ("split_level", np.int32), ("status", np.int32), ("rotation", np.int32))


class CompleteSolution(ctypes.Structure):
    """ndg_complete_solution: one pointer per output field, in the header's order."""
    _fields_ = [(name, ctypes.c_void_p) for name in COMPLETE_POINT_FIELDS] + \
               [(name, ctypes.c_void_p) for name, _ in COMPLETE_ELEMENT_FIELDS]


class SU2Periodic(ctypes.Structure):
    """ndg_su2_periodic: n_pairs marker-tag pairs joined by a translation."""
    _fields_ = [("n_pairs", ctypes.c_int32), ("marker_a", ctypes.POINTER(ctypes.c_char_p)),
                ("marker_b", ctypes.POINTER(ctypes.c_char_p))]


class ICDesc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("n_pulses", ctypes.c_int32), ("pulses", ctypes.c_void_p),
                ("period_x", ctypes.c_double), ("period_y", ctypes.c_double)]


_lib = None


def lib():
    """Loads libndg_amd.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or python -m "
                              "ndg_cuda_amd.build) first; there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        sigs = {
            "ndg_last_error": [ctypes.c_char_p, ctypes.c_size_t],
            "ndg_device_count": [vp],
            "ndg_tables_create": [i32, i64, vp],
            "ndg_tables_destroy": [vp],
            "ndg_tables_get": [vp, i32, i32, vp, i64, vp],
            "ndg_square_mesh_sizes": [vp, vp, vp, vp],
            "ndg_square_mesh_fill": [vp, vp, vp, vp, vp, vp],
            "ndg_su2_sizes": [ctypes.c_char_p, vp, vp, vp],
            "ndg_su2_fill": [ctypes.c_char_p, vp, vp, vp, vp],
            "ndg_su2_sizes_periodic": [ctypes.c_char_p, vp, vp, vp, vp],
            "ndg_su2_fill_periodic": [ctypes.c_char_p, vp, vp, vp, vp, vp, vp],
            "ndg_mesh_create": [vp, vp, vp],
            "ndg_mesh_destroy": [vp],
            "ndg_mesh_set_stream": [vp, vp],
            "ndg_mesh_info": [vp, vp, vp, vp, vp, vp],
            "ndg_mesh_set_precision": [vp, i32],
            "ndg_set_ic": [vp, vp],
            "ndg_set_state": [vp, vp, vp, vp],
            "ndg_get_state": [vp, vp, vp, vp],
            "ndg_get_solution": [vp, vp, vp, vp, vp, vp],
            "ndg_get_complete_solution": [vp, f64, vp],
            "ndg_get_traces": [vp, vp],
            "ndg_compute_rhs": [vp, f64, vp, vp, vp],
            "ndg_get_delta_t": [vp, f64, vp],
            "ndg_step": [vp, f64, f64],
            "ndg_solve": [vp, f64, f64, vp, vp],
            "ndg_solve_output": [vp, f64, vp, i64, vp, vp, vp, vp],
            "ndg_run_steps": [vp, f64, i64, vp],
            "ndg_pre_condition": [vp, f64, i64, i64, i32],
            "ndg_synchronize": [vp],
            "ndg_estimate_error": [vp, vp, vp, vp, vp],
            "ndg_mesh_set_graph": [vp, i32],
            "ndg_mesh_autotune": [vp, i32],
            "ndg_mesh_tuning": [vp, vp, vp, vp, i32, vp],
            "ndg_mesh_set_wave_speed": [vp, f64],
            "ndg_record_receivers": [vp, vp, f64, i64, vp],
            "ndg_adjoint_gradient": [vp, vp, f64, i64, i64, vp, vp, vp, vp, vp, vp],
            "ndg_adjoint_set_store_budget": [vp, i64],
            "ndg_adjoint_set_c_derivative": [vp, i32],
            "ndg_apply_rhs_transpose": [vp, vp, vp, vp, vp, vp, vp],
            "ndg_adapt": [vp, i32, i32, vp],
            "ndg_mesh_set_adaptivity": [vp, i64],
            "ndg_mesh_orders": [vp, vp],
            "ndg_mesh_set_adapt_mode": [vp, i32],
            "ndg_inverse_quad_map": [vp, i64, vp, vp],
            "ndg_mesh_sides": [vp, vp, vp, vp, vp, vp],
            "ndg_hilbert_child_order": [i32, i32, vp, vp],
            "ndg_hilbert_element_status": [i32, i32, vp],
            "ndg_record_receivers_noalloc": [vp, vp, f64, i64, vp],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def last_error():
    buf = ctypes.create_string_buffer(1024)
    lib().ndg_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status):
    if status != NDG_OK:
        raise NDGError(status, last_error())


def ptr(a):
    """Address of a C-contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data


def device_count():
    n = ctypes.c_int32(0)
    check(lib().ndg_device_count(ctypes.byref(n)))
    return n.value


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def as_i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
