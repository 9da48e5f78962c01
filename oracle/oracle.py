"""ctypes wrapper of oracle/liboracle.so -- THE ORACLE (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module, and only
as the checker (or as the timed CPU restatement). The product (ndg_cuda_amd) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

O_NODES, O_WEIGHTS, O_BARY, O_LAG_LEFT, O_LAG_RIGHT, O_DLAG_LEFT, O_DLAG_RIGHT, O_D, O_GHAT, O_DHAT, \
    O_INTERP, O_POLY = range(12)
TABLE_KEYS = ["nodes", "weights", "bary", "lag_left", "lag_right", "dlag_left", "dlag_right", "D",
              "Ghat", "Dhat", "interp", "poly"]

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
        sig = {
            "o_tables_create": ([i32, i32], vp), "o_tables_destroy": ([vp], None),
            "o_tables_get": ([vp, i32, i32, vp], i32),
            "o_mesh_create": ([vp, i32, i64, vp, i64, vp, vp, i64, vp, vp, vp, f64, f64], vp),
            "o_mesh_destroy": ([vp], None), "o_mesh_n_elements": ([vp], i64),
            "o_mesh_n_faces": ([vp], i64), "o_mesh_side_faces": ([vp, vp], None), "o_mesh_n_ghosts": ([vp], i64),
            "o_mesh_n_nodes_total": ([vp], i64), "o_mesh_get_N": ([vp, vp], None),
            "o_initial_conditions_plane": ([vp], None),
            "o_initial_conditions_pulses": ([vp, i32, vp], None),
            "o_initial_conditions_periodic_pulses": ([vp, i32, vp, f64, f64], None),
            "o_kat_face_values": ([vp, i32, i32, i32, f64, f64, vp, vp], None),
            "o_kat_side_flux": ([vp, i32, i32, i32, f64, f64, f64, vp, vp], None),
            "o_kat_interpolate": ([vp, i32, vp, i32, vp, vp, vp], None),
            "o_kat_inverse_quad_map": ([vp, f64, f64, vp], None),
            "o_set_state": ([vp, vp, vp, vp], None), "o_get_state": ([vp, vp, vp, vp], None),
            "o_get_rhs": ([vp, vp, vp, vp], None), "o_get_traces": ([vp, vp], None),
            "o_get_solution": ([vp, vp, vp, vp, vp, vp], None),
            "o_get_complete_solution": ([vp, ctypes.c_double, vp], None),
            "o_get_flux_traces": ([vp, vp], None), "o_get_delta_t": ([vp, f64], f64),
            "o_compute_rhs": ([vp, f64], None), "o_mesh_set_wave_speed": ([vp, f64], None),
            "o_adapt": ([vp, ctypes.c_int, ctypes.c_int], ctypes.c_int64),
            "o_mesh_get_elements": ([vp, vp, vp, vp, vp], None), "o_mesh_n_nodes": ([vp], ctypes.c_int64),
            "o_mesh_set_flags": ([vp, vp, vp], None),
            "o_solve_adaptive": ([vp, f64, f64, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double)], ctypes.c_int64), "o_step": ([vp, f64, f64], None),
            "o_solve": ([vp, f64, f64, vp], i64), "o_run_steps": ([vp, f64, i64, vp], None),
            "o_estimate_error": ([vp, vp, vp, vp, vp], None),
            "o_element_traces": ([vp, i32, vp, vp, vp, vp], None),
            "o_element_geometry": ([vp, i32, vp, vp], None),
            "o_boundary_geometry": ([vp, i32, vp, vp], None),
            "o_face_geometry": ([vp, vp, vp], None),
            "o_element_estimate": ([vp, i32, vp, f64, f64, vp], None),
            "o_wave_flux": ([vp, vp], None), "o_hilbert_d2xy": ([i32, i32, vp], None),
            "o_hilbert_xy2d": ([i32, i32, i32], i32),
            "o_hilbert_tables": ([vp, vp, vp, vp, vp], None),
            "o_set_threads": ([i32], i32),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Tables:
    def __init__(self, N_max, n_interp):
        self.N_max, self.n_interp = N_max, n_interp
        self.h = lib().o_tables_create(N_max, n_interp)

    def get(self, which, N):
        nf = N + 1
        n = nf * nf if which in (O_D, O_GHAT, O_DHAT, O_POLY) else (nf * self.n_interp if which == O_INTERP else nf)
        out = np.empty(n)
        lib().o_tables_get(self.h, which, N, _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().o_tables_destroy(self.h)
            self.h = None


class Mesh:
    """Reference-model mesh built from the same description the product takes."""

    def __init__(self, tables, desc, initial_N, tol_min=1e-6, tol_max=1e-14):
        self.tables = tables
        xy = _f(desc.xy)
        quads = np.ascontiguousarray(desc.quads, dtype=np.int64)
        bars = np.ascontiguousarray(desc.bars, dtype=np.int64)
        bt = np.ascontiguousarray(desc.bar_type, dtype=np.int32)
        bd = np.ascontiguousarray(desc.bar_donor, dtype=np.int64)
        qn = None if desc.quad_N is None else np.ascontiguousarray(desc.quad_N, dtype=np.int32)
        self._keep = (xy, quads, bars, bt, bd, qn)
        self.h = lib().o_mesh_create(tables.h, initial_N, xy.shape[0], _p(xy), quads.shape[0], _p(quads), _p(qn),
                                     bars.shape[0], _p(bars), _p(bt), _p(bd), tol_min, tol_max)
        self.K = lib().o_mesh_n_elements(self.h)
        self.n = lib().o_mesh_n_nodes_total(self.h)
        self.N = np.empty(self.K, dtype=np.int32)
        lib().o_mesh_get_N(self.h, _p(self.N))

    @property
    def n_faces(self):
        return lib().o_mesh_n_faces(self.h)

    def side_faces(self):
        """Faces per element side, shape (K, 4)."""
        out = np.empty((self.K, 4), dtype=np.int32)
        lib().o_mesh_side_faces(self.h, _p(out))
        return out

    def initial_conditions(self, pulses=None, period=None):
        if pulses is None:
            lib().o_initial_conditions_plane(self.h)
        elif period is not None:
            arr = _f(np.asarray(pulses, dtype=np.float64).reshape(-1, 4))
            lib().o_initial_conditions_periodic_pulses(self.h, arr.shape[0], _p(arr), float(period[0]),
                                                       float(period[1]))
        else:
            arr = _f(np.asarray(pulses, dtype=np.float64).reshape(-1, 4))
            lib().o_initial_conditions_pulses(self.h, arr.shape[0], _p(arr))

    def set_state(self, p, u, v):
        p, u, v = _f(p), _f(u), _f(v)
        lib().o_set_state(self.h, _p(p), _p(u), _p(v))

    def get_state(self):
        out = [np.empty(self.n) for _ in range(3)]
        lib().o_get_state(self.h, *(_p(a) for a in out))
        return tuple(out)

    def get_rhs(self):
        out = [np.empty(self.n) for _ in range(3)]
        lib().o_get_rhs(self.h, *(_p(a) for a in out))
        return tuple(out)

    def _sides(self, fn):
        out = np.empty(int(np.sum(12 * (self.N.astype(np.int64) + 1))))
        fn(self.h, _p(out))
        return out

    def get_traces(self):
        return self._sides(lib().o_get_traces)

    def get_flux_traces(self):
        return self._sides(lib().o_get_flux_traces)

    def set_wave_speed(self, c):
        lib().o_mesh_set_wave_speed(self.h, float(c))

    def delta_t(self, cfl):
        return lib().o_get_delta_t(self.h, cfl)

    def compute_rhs(self, t=0.0):
        lib().o_compute_rhs(self.h, t)
        return self.get_rhs()

    def step(self, t, dt):
        lib().o_step(self.h, t, dt)

    def solve(self, cfl, t_end):
        tf = ctypes.c_double()
        n = lib().o_solve(self.h, cfl, t_end, ctypes.byref(tf))
        return n, tf.value

    def run_steps(self, cfl, n_steps, time=0.0):
        t = ctypes.c_double(time)
        lib().o_run_steps(self.h, cfl, n_steps, ctypes.byref(t))
        return t.value

    def adapt(self, N_max, max_split_level=0):
        """One adaptation from the last estimate; returns the number of raised elements (-1: an
        element would split) and refreshes the element orders."""
        n = lib().o_adapt(self.h, int(N_max), int(max_split_level))
        self._refresh()
        return n

    def solve_adaptive(self, cfl, t_end, adapt_interval, N_max, max_split_level=0):
        tf = ctypes.c_double()
        n = lib().o_solve_adaptive(self.h, cfl, t_end, int(adapt_interval), int(N_max), int(max_split_level),
                                   ctypes.byref(tf))
        self._refresh()
        return n, tf.value

    def pre_condition(self, cfl, n_adaptivity_steps, interval, N_max, max_split_level, pulses, iterative=False,
                      period=None):
        """Solver2D_t::pre_condition (include/solvers/Solver2D_t.tpp:138-236) and
        pre_condition_iterative (:238-336): from time 0, steps of get_delta_t with estimate_error +
        adapt every `interval` steps until n_adaptivity_steps adaptations are done; the initial
        condition (pulses, or the plane wave for None) is applied again at the end, or after every
        adaptation with the time set back to 0 (iterative)."""
        time, timestep, n = 0.0, 0, 0
        while n < n_adaptivity_steps:
            timestep += 1
            dt = self.delta_t(cfl)
            self.step(time, dt)
            time += dt
            if timestep % interval == 0:
                self.estimate_error()
                if self.adapt(N_max, max_split_level) < 0:
                    raise RuntimeError("oracle pre_condition: adaptation refused")
                n += 1
                if iterative:
                    self.initial_conditions(pulses, period)
                    time = 0.0
        if not iterative:
            self.initial_conditions(pulses, period)

    def set_flags(self, refine, sigma):
        r = np.ascontiguousarray(refine, dtype=np.int32)
        sg = _f(sigma)
        lib().o_mesh_set_flags(self.h, _p(r), _p(sg))

    def elements(self):
        """(element nodes (K, 4), Hilbert status (K,), rotation (K,), node coordinates (n, 2))."""
        nodes = np.empty((self.K, 4), dtype=np.int64)
        st, rot = np.empty(self.K, dtype=np.int32), np.empty(self.K, dtype=np.int32)
        xy = np.empty((lib().o_mesh_n_nodes(self.h), 2))
        lib().o_mesh_get_elements(self.h, _p(nodes), _p(st), _p(rot), _p(xy))
        return nodes, st, rot, xy

    def _refresh(self):
        self.K = lib().o_mesh_n_elements(self.h)
        self.N = np.empty(self.K, dtype=np.int32)
        self.n = lib().o_mesh_n_nodes_total(self.h)
        lib().o_mesh_get_N(self.h, _p(self.N))

    def get_solution(self, n):
        """x, y, p, u, v at the n x n uniform output points of every element, shape (K, n, n)
        (n: the tables' interpolation point count)."""
        out = [np.empty((self.K, n, n)) for _ in range(5)]
        lib().o_get_solution(self.h, *(_p(a) for a in out))
        return tuple(out)

    def get_complete_solution(self, n, time):
        """The 11 point fields of get_complete_solution, shape (11, K, n, n): x, y, p, u, v, the
        right-hand side left by the last compute_rhs / step, and the plane-wave errors at time."""
        out = np.empty((11, self.K, n, n))
        lib().o_get_complete_solution(self.h, float(time), _p(out))
        return out

    def estimate_error(self):
        err, sig = np.empty(self.K), np.empty(self.K)
        ref, coa = np.empty(self.K, dtype=np.int32), np.empty(self.K, dtype=np.int32)
        lib().o_estimate_error(self.h, _p(err), _p(sig), _p(ref), _p(coa))
        return err, sig, ref.astype(bool), coa.astype(bool)

    def __del__(self):
        if getattr(self, "h", None):
            lib().o_mesh_destroy(self.h)
            self.h = None


def element_traces(tables, N, p, u, v):
    out = np.empty(12 * (N + 1))
    p, u, v = _f(p), _f(u), _f(v)
    lib().o_element_traces(tables.h, N, _p(p), _p(u), _p(v), _p(out))
    return out


def element_geometry(tables, N, pts):
    nf = N + 1
    out = np.empty(5 * nf * nf + 4 * nf + 3)
    pts = _f(pts)
    lib().o_element_geometry(tables.h, N, _p(pts), _p(out))
    return out


def boundary_geometry(tables, N, pts):
    out = np.empty(N + 4)
    pts = _f(pts)
    lib().o_boundary_geometry(tables.h, N, _p(pts), _p(out))
    return out


def face_geometry(face_nodes, element_nodes):
    out = np.empty(9)
    a, b = _f(face_nodes), _f(element_nodes)
    lib().o_face_geometry(_p(a), _p(b), _p(out))
    return out


def element_estimate(tables, N, p, tol_min=1e-6, tol_max=1e-14):
    out = np.zeros(8)
    p = _f(p)
    lib().o_element_estimate(tables.h, N, _p(p), tol_min, tol_max, _p(out))
    return out


def wave_flux(state8):
    out = np.empty(3)
    s = _f(state8)
    lib().o_wave_flux(_p(s), _p(out))
    return out


def set_threads(n):
    """Threads of the oracle's element and face loops (OpenMP); 0 keeps the default. Returns the
    thread count in use. Results are the same bit for bit for any count."""
    return int(lib().o_set_threads(int(n)))


def hilbert_tables():
    """(child_order [rotation][status][4], child_status [rotation][status][4], first status [side],
    last status [side], inner status [incoming][outgoing]) as int arrays."""
    co = np.zeros((4, 4, 4), dtype=np.int32)
    cs = np.zeros((4, 4, 4), dtype=np.int32)
    fi = np.zeros(4, dtype=np.int32)
    la = np.zeros(4, dtype=np.int32)
    inner = np.zeros((4, 4), dtype=np.int32)
    lib().o_hilbert_tables(_p(co), _p(cs), _p(fi), _p(la), _p(inner))
    return co, cs, fi, la, inner


def hilbert_d2xy(n, d):
    xy = np.zeros(2, dtype=np.int32)
    lib().o_hilbert_d2xy(n, d, _p(xy))
    return int(xy[0]), int(xy[1])


# Single operations for the reference's known-answer tests.
def kat_face_values(tables, Ne, Nf, side, offset, scale, trace):
    """project_to_faces on one face side: the element trace (order Ne) at the Nf face nodes."""
    tr = _f(np.asarray(trace, dtype=np.float64))
    out = np.empty(Nf + 1)
    lib().o_kat_face_values(tables.h, Ne, Nf, side, offset, scale, _p(tr), _p(out))
    return out


def kat_side_flux(tables, Ne, Nf, fwd, offset, scale, sf, flux, out=None):
    """project_to_elements of one face's flux (order Nf) into an order-Ne side, accumulated into out."""
    fl = _f(np.asarray(flux, dtype=np.float64))
    out = np.zeros(Ne + 1) if out is None else out
    lib().o_kat_side_flux(tables.h, Ne, Nf, int(fwd), offset, scale, sf, _p(fl), _p(out))
    return out


def kat_interpolate(tables, No, q_old, N, P_old=None, P_new=None):
    """Element2D_t::interpolate_from: the old element's field (order No) at the new element's nodes
    (same geometry, or a child P_new inside P_old, vertices (4, 2))."""
    q = _f(np.asarray(q_old, dtype=np.float64))
    out = np.empty((N + 1) ** 2)
    Po = None if P_old is None else _f(np.asarray(P_old, dtype=np.float64).ravel())
    Pn = None if P_new is None else _f(np.asarray(P_new, dtype=np.float64).ravel())
    lib().o_kat_interpolate(tables.h, No, _p(Po), N, _p(Pn), _p(q), _p(out))
    return out


def kat_inverse_quad_map(P, x, y):
    """inverse_quad_map: reference coordinates of (x, y) in the quad P (4, 2)."""
    PP = _f(np.asarray(P, dtype=np.float64).ravel())
    out = np.empty(2)
    lib().o_kat_inverse_quad_map(_p(PP), float(x), float(y), _p(out))
    return out
