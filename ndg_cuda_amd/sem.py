"""Python mirror of the reference host API (SEM::Host / SEM::Device classes of NDG_cuda) over the
C ABI of libndg_amd. Names, argument meaning and call order follow the reference:

    NDG_t(N_max, n_interpolation_points)                   include/entities/NDG_t.h:11-26
    Mesh2D_t(mesh, initial_N, maximum_N, n_interpolation_points, max_split_level,
             adaptivity_interval, load_balancing_interval, tolerance_min, tolerance_max,
             load_balancing_threshold, NDG)                 include/meshes/Mesh2D_t.h:21
    mesh.initial_conditions(...)                             src/meshes/Mesh2D_t.cpp:1192-1216
    Solver2D_t(CFL, output_times, viscosity)                 include/solvers/Solver2D_t.h:14-49
    solver.solve(NDG, mesh)                                  include/solvers/Solver2D_t.tpp:7-135

Where the reference reads a CGNS/SU2 file, Mesh2D_t takes a MeshDescription (the same content in
memory); square_mesh() is the equivalent of the reference's mesh generator. Errors raise NDGError
instead of exiting the process.
"""
import ctypes
import os
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _abi
from ._abi import (BC_INFLOW, BC_NONE, BC_OUTFLOW, BC_PERIODIC, BC_SYMMETRY, BC_WALL,  # noqa: F401
                   IC_PERIODIC_PULSES, IC_PLANE_WAVE, IC_PULSES, NDGError, TABLE_NAMES, check, lib, ptr)


class NDG_t:
    """Legendre-Gauss tables for every order 0..N_max (FP64, computed on the host)."""

    def __init__(self, N_max: int, n_interpolation_points: int):
        self.N_max_ = int(N_max)
        self.N_interpolation_points_ = int(n_interpolation_points)
        h = ctypes.c_void_p()
        check(lib().ndg_tables_create(self.N_max_, self.N_interpolation_points_, ctypes.byref(h)))
        self._h = h
        for which, name in enumerate(TABLE_NAMES):
            setattr(self, name + "_", [self._get(which, N) for N in range(self.N_max_ + 1)])

    def _get(self, which, N):
        count = ctypes.c_int64()
        check(lib().ndg_tables_get(self._h, which, N, None, 0, ctypes.byref(count)))
        out = np.empty(count.value, dtype=np.float64)
        check(lib().ndg_tables_get(self._h, which, N, ptr(out), out.size, None))
        return out

    @property
    def handle(self):
        return self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _abi._lib is not None:
            _abi._lib.ndg_tables_destroy(h)
            self._h = None


@dataclass
class MeshDescription:
    """Nodes, QUAD_4 elements, BAR_2 boundary elements with kinds, periodic donors."""
    xy: np.ndarray                       # (n_nodes, 2)
    quads: np.ndarray                    # (n_quads, 4) int64, counter-clockwise
    bars: np.ndarray                     # (n_bars, 2) int64
    bar_type: np.ndarray                 # (n_bars,) int32, BC_*
    bar_donor: np.ndarray                # (n_bars,) int64 (periodic donor bar or -1)
    quad_N: Optional[np.ndarray] = None  # (n_quads,) int32 or None

    @property
    def n_elements(self):
        return int(self.quads.shape[0])


def inverse_quad_map(vertices, xy):
    """SEM::inverse_quad_map (include/functions/inverse_quad_map.tpp:46-88): reference coordinates
    (n, 2) of the physical points xy (n, 2) in the bilinear quad of vertices (4, 2)."""
    P = _abi.as_f64(np.asarray(vertices, dtype=np.float64).reshape(4, 2))
    X = _abi.as_f64(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
    out = np.empty_like(X)
    check(lib().ndg_inverse_quad_map(ptr(P), X.shape[0], ptr(X), ptr(out)))
    return out


def square_mesh(resolution: int, x_min=-1.0, x_max=1.0, y_min=-1.0, y_max=1.0,
                bc: Sequence[int] = (BC_WALL, BC_WALL, BC_WALL, BC_WALL),
                x_periodic=False, y_periodic=False, jitter=0.0, seed=0) -> MeshDescription:
    """Square mesh in Hilbert-curve order (extras/mesh_generator/main.cpp:161-382); bc is
    {bottom, right, top, left}; jitter moves interior vertices by a seeded uniform fraction of
    the spacing."""
    p = _abi.SquareMeshParams()
    p.resolution = int(resolution)
    p.x_min, p.x_max, p.y_min, p.y_max = float(x_min), float(x_max), float(y_min), float(y_max)
    for k in range(4):
        p.bc[k] = int(bc[k])
    p.x_periodic, p.y_periodic = int(bool(x_periodic)), int(bool(y_periodic))
    p.jitter, p.seed = float(jitter), int(seed)
    nn, nq, nb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    check(lib().ndg_square_mesh_sizes(ctypes.byref(p), ctypes.byref(nn), ctypes.byref(nq), ctypes.byref(nb)))
    xy = np.empty((nn.value, 2), dtype=np.float64)
    quads = np.empty((nq.value, 4), dtype=np.int64)
    bars = np.empty((nb.value, 2), dtype=np.int64)
    bar_type = np.empty(nb.value, dtype=np.int32)
    bar_donor = np.empty(nb.value, dtype=np.int64)
    check(lib().ndg_square_mesh_fill(ctypes.byref(p), ptr(xy), ptr(quads), ptr(bars), ptr(bar_type), ptr(bar_donor)))
    return MeshDescription(xy, quads, bars, bar_type, bar_donor)


def read_su2(path, periodic=None) -> MeshDescription:
    """Mesh2D_t::read_su2 (src/meshes/Mesh2D_t.cpp:78-330): nodes, QUAD_4 elements and the
    farfield / wall / inlet / symmetry markers of an SU2 file as a MeshDescription (bars in the
    reference's ghost order: walls, symmetry, inflow, outflow). `periodic` lists (marker_a,
    marker_b) tag pairs joined by a translation, as SU2's MARKER_PERIODIC does; their bars come
    last, as periodic bars with their donors (ndg_su2_fill_periodic)."""
    p = os.fsencode(path)
    pairs = list(periodic or [])
    tags_a = (ctypes.c_char_p * max(1, len(pairs)))(*[a.encode() for a, _ in pairs])
    tags_b = (ctypes.c_char_p * max(1, len(pairs)))(*[b.encode() for _, b in pairs])
    per = _abi.SU2Periodic(len(pairs), tags_a, tags_b)
    nn, nq, nb = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    check(lib().ndg_su2_sizes_periodic(p, ctypes.byref(per), ctypes.byref(nn), ctypes.byref(nq), ctypes.byref(nb)))
    xy = np.empty((nn.value, 2), dtype=np.float64)
    quads = np.empty((nq.value, 4), dtype=np.int64)
    bars = np.empty((nb.value, 2), dtype=np.int64)
    bar_type = np.empty(nb.value, dtype=np.int32)
    bar_donor = np.empty(nb.value, dtype=np.int64)
    check(lib().ndg_su2_fill_periodic(p, ctypes.byref(per), ptr(xy), ptr(quads), ptr(bars), ptr(bar_type),
                                      ptr(bar_donor)))
    return MeshDescription(xy, quads, bars, bar_type, bar_donor)


def _mesh_from(mesh) -> MeshDescription:
    """The constructor's mesh argument: a MeshDescription, or a file path like the reference's
    (src/meshes/Mesh2D_t.cpp:56-70; .su2 only, there is no CGNS reader here)."""
    if isinstance(mesh, MeshDescription):
        return mesh
    ext = os.path.splitext(os.fsdecode(mesh))[1].lower()
    if ext == ".su2":
        return read_su2(mesh)
    raise NDGError(_abi.NDG_ERR_UNSUPPORTED, f"Mesh2D_t: mesh file extension '{ext}' is not read here (only .su2)")


class Mesh2D_t:
    """Device-resident mesh and solution (one HIP stream)."""

    def __init__(self, mesh: MeshDescription, initial_N: int, maximum_N: int, n_interpolation_points: int,
                 max_split_level: int, adaptivity_interval: int, load_balancing_interval: int,
                 tolerance_min: float, tolerance_max: float, load_balancing_threshold: float, NDG: NDG_t,
                 worker_weights=None):
        self.initial_N_ = int(initial_N)
        self.maximum_N_ = int(maximum_N)
        self.n_interpolation_points_ = int(n_interpolation_points)
        self.max_split_level_ = int(max_split_level)
        self.adaptivity_interval_ = int(adaptivity_interval)
        self.load_balancing_interval_ = int(load_balancing_interval)
        self.tolerance_min_ = float(tolerance_min)
        self.tolerance_max_ = float(tolerance_max)
        self._ndg = NDG
        mesh = _mesh_from(mesh)
        self._arrays = [_abi.as_f64(mesh.xy), _abi.as_i64(mesh.quads), _abi.as_i64(mesh.bars),
                        _abi.as_i32(mesh.bar_type), _abi.as_i64(mesh.bar_donor),
                        None if mesh.quad_N is None else _abi.as_i32(mesh.quad_N)]
        xy, quads, bars, bar_type, bar_donor, quad_N = self._arrays
        d = _abi.MeshDesc()
        d.n_nodes, d.xy = xy.shape[0], ptr(xy)
        d.n_quads, d.quads, d.quad_N = quads.shape[0], ptr(quads), ptr(quad_N)
        d.n_bars, d.bars, d.bar_type, d.bar_donor = bars.shape[0], ptr(bars), ptr(bar_type), ptr(bar_donor)
        d.initial_N, d.maximum_N, d.max_split_level = self.initial_N_, self.maximum_N_, self.max_split_level_
        d.tolerance_min, d.tolerance_max = self.tolerance_min_, self.tolerance_max_
        h = ctypes.c_void_p()
        check(lib().ndg_mesh_create(ctypes.byref(d), NDG.handle, ctypes.byref(h)))
        self._h = h
        self._refresh()
        if adaptivity_interval:
            check(lib().ndg_mesh_set_adaptivity(h, int(adaptivity_interval)))

    def _refresh(self):
        """Re-reads the sizes and element orders (they change with adaptation)."""
        ne, nf, nd = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        nmin, nmax = ctypes.c_int32(), ctypes.c_int32()
        check(lib().ndg_mesh_info(self._h, ctypes.byref(ne), ctypes.byref(nf), ctypes.byref(nd),
                                  ctypes.byref(nmin), ctypes.byref(nmax)))
        self.n_elements_ = ne.value
        self.n_faces_ = nf.value
        self.n_dofs_per_field_ = nd.value
        self.N_ = nmin.value
        self.N_max_ = nmax.value
        self.orders_ = np.empty(self.n_elements_, dtype=np.int32)
        check(lib().ndg_mesh_orders(self._h, ptr(self.orders_)))

    def set_adapt_mode(self, mode):
        """"device" (default): adaptation rebuilt on the GPU from the device-resident topology;
        "host": the topology rebuilt on the host from the node lists (refused after a device
        adaptation). Both give the same meshes and states bit for bit."""
        modes = {"device": _abi.ADAPT_DEVICE, "host": _abi.ADAPT_HOST}
        check(lib().ndg_mesh_set_adapt_mode(self._h, modes[mode]))

    def sides(self):
        """Element-side topology, 4 entries per element: link (exterior side 4r+rs, or -(1+ghost);
        a larger side's first small side), hang (0 conforming, 1 small hanging side, 2 larger
        side), a small side's next small side along its larger side (-1 after the last), the
        periodic flag, and a small side's piece of its larger side as 2^d - 1 + j for
        [j / 2^d, (j + 1) / 2^d]."""
        n = 4 * self.n_elements_
        link, nxt = np.empty(n, dtype=np.int64), np.empty(n, dtype=np.int64)
        hang, per, span = np.empty(n, dtype=np.int8), np.empty(n, dtype=np.int8), np.empty(n, dtype=np.int8)
        check(lib().ndg_mesh_sides(self._h, ptr(link), ptr(hang), ptr(nxt), ptr(per), ptr(span)))
        return link, hang, nxt, per, span

    def adapt(self, N_max=None, max_split_level=None):
        """Mesh2D_t::adapt: estimate, then either split the elements that would h-refine (when
        any would, the hp_adapt branch) or raise by 2 the order of those that would p-refine.
        Returns the number of changed elements."""
        n = ctypes.c_int64()
        check(lib().ndg_adapt(self._h, int(self._ndg.N_max_ if N_max is None else N_max),
                              int(self.max_split_level_ if max_split_level is None else max_split_level),
                              ctypes.byref(n)))
        self._refresh()
        return n.value

    @property
    def handle(self):
        return self._h

    @property
    def n_dofs(self):
        """DOF = 3 * sum over elements of (N+1)^2."""
        return 3 * self.n_dofs_per_field_

    def set_precision(self, precision: str):
        """'fp64' (the reference arithmetic) or 'fp32c' (FP32 right-hand side and RK register, state
        kept as a compensated pair of floats)."""
        code = {"fp64": _abi.PRECISION_FP64, "fp32c": _abi.PRECISION_FP32_COMPENSATED}[precision]
        check(lib().ndg_mesh_set_precision(self._h, code))
        self.precision_ = precision

    def set_graph(self, mode: str):
        """'off' (direct launches), 'sequential' or 'parallel' (one captured hipGraph per RK step,
        order groups in sequence or as concurrent branches)."""
        code = {"off": _abi.GRAPH_OFF, "sequential": _abi.GRAPH_SEQUENTIAL, "parallel": _abi.GRAPH_PARALLEL}[mode]
        check(lib().ndg_mesh_set_graph(self._h, code))

    def autotune(self, repeats: int = 5):
        """Times the launch shapes of every order group's stage kernel and keeps the fastest;
        returns [(N, elements per workgroup, XCD block order)] per group."""
        check(lib().ndg_mesh_autotune(self._h, int(repeats)))
        return self.tuning()

    def tuning(self):
        n = ctypes.c_int32()
        check(lib().ndg_mesh_tuning(self._h, None, None, None, 0, ctypes.byref(n)))
        o, e, r = (np.empty(n.value, dtype=np.int32) for _ in range(3))
        check(lib().ndg_mesh_tuning(self._h, ptr(o), ptr(e), ptr(r), n.value, ctypes.byref(n)))
        return [(int(a), int(b), int(c)) for a, b, c in zip(o, e, r)]

    def set_stream(self, hip_stream_ptr: int):
        check(lib().ndg_mesh_set_stream(self._h, ctypes.c_void_p(hip_stream_ptr)))

    def initial_conditions(self, pulses=None, period=None):
        """Reference plane wave g(x, 0) when pulses is None, else radial Gaussian pulses given as
        rows {x0, y0, width, amplitude}; with period = (Lx, Ly) the pulses are periodised (each
        with its 8 nearest images)."""
        ic = _abi.ICDesc()
        arr = None
        if pulses is None:
            ic.kind = IC_PLANE_WAVE
        else:
            arr = _abi.as_f64(np.asarray(pulses, dtype=np.float64).reshape(-1, 4))
            ic.kind, ic.n_pulses, ic.pulses = IC_PULSES, arr.shape[0], ptr(arr)
            if period is not None:
                ic.kind = IC_PERIODIC_PULSES
                ic.period_x, ic.period_y = float(period[0]), float(period[1])
        check(lib().ndg_set_ic(self._h, ctypes.byref(ic)))

    def set_solution(self, p, u, v):
        p, u, v = (_abi.as_f64(a) for a in (p, u, v))
        for a in (p, u, v):
            if a.size != self.n_dofs_per_field_:
                raise NDGError(_abi.NDG_ERR_INVALID, "set_solution: wrong field length")
        check(lib().ndg_set_state(self._h, ptr(p), ptr(u), ptr(v)))

    def get_solution(self):
        out = [np.empty(self.n_dofs_per_field_, dtype=np.float64) for _ in range(3)]
        check(lib().ndg_get_state(self._h, *(ptr(a) for a in out)))
        return tuple(out)

    def get_interpolated_solution(self):
        """Mesh2D_t::get_solution (src/meshes/Mesh2D_t.cpp:4237-4249): x, y, p, u, v at the n x n
        uniform output points of every element (n = the tables' n_interpolation_points), arrays of
        shape (n_elements, n, n). Also kept as x_output_, y_output_, p_output_, u_output_,
        v_output_ like the reference's members."""
        n = self._ndg.N_interpolation_points_
        out = [np.empty((self.n_elements_, n, n), dtype=np.float64) for _ in range(5)]
        check(lib().ndg_get_solution(self._h, *(ptr(a) for a in out)))
        self.x_output_, self.y_output_, self.p_output_, self.u_output_, self.v_output_ = out
        return tuple(out)

    def get_complete_solution(self, time: float):
        """Mesh2D_t::get_complete_solution (src/meshes/Mesh2D_t.cu:1861-1950): a dict of the point
        fields x, y, p, u, v, dp_dt, du_dt, dv_dt, p/u/v_analytical_error, each of shape
        (n_elements, n, n), and the element fields N, p/u/v_error, p/u/v_sigma, refine, coarsen,
        split_level, status, rotation, each of shape (n_elements,). dp_dt etc. are the right-hand
        side of the current state at `time`; the analytical errors are against the reference plane
        wave at `time`; the element errors are those of the last estimate_error."""
        n, K = self._ndg.N_interpolation_points_, self.n_elements_
        out = {name: np.empty((K, n, n), dtype=np.float64) for name in _abi.COMPLETE_POINT_FIELDS}
        out.update({name: np.empty(K, dtype=dt) for name, dt in _abi.COMPLETE_ELEMENT_FIELDS})
        cs = _abi.CompleteSolution(*(out[name].ctypes.data for name, _ in _abi.CompleteSolution._fields_))
        check(lib().ndg_get_complete_solution(self._h, float(time), ctypes.byref(cs)))
        self.x_output_, self.y_output_ = out["x"], out["y"]
        self.p_output_, self.u_output_, self.v_output_ = out["p"], out["u"], out["v"]
        return out

    def write_data(self, time: float, data_writer):
        """Mesh2D_t::write_data (src/meshes/Mesh2D_t.cpp:1601-1609): get_solution, then
        data_writer.write_data with the output points of every element."""
        x, y, p, u, v = self.get_interpolated_solution()
        return data_writer.write_data(self._ndg.N_interpolation_points_, self.n_elements_, time, x, y, p, u, v)

    def write_complete_data(self, time: float, data_writer):
        """Mesh2D_t::write_complete_data (src/meshes/Mesh2D_t.cu:1861-1950):
        get_complete_solution, then data_writer.write_complete_data (rank 0)."""
        c = self.get_complete_solution(time)
        return data_writer.write_complete_data(
            self._ndg.N_interpolation_points_, self.n_elements_, time, 0, c["x"], c["y"], c["p"], c["u"], c["v"],
            c["N"], c["dp_dt"], c["du_dt"], c["dv_dt"], c["p_error"], c["u_error"], c["v_error"], c["p_sigma"],
            c["u_sigma"], c["v_sigma"], c["refine"], c["coarsen"], c["split_level"], c["p_analytical_error"],
            c["u_analytical_error"], c["v_analytical_error"], c["status"], c["rotation"])

    def interpolate_to_boundaries(self):
        """Element traces of the current state: shape (K, 4, 3, N+1) when every element has the
        same order, else a list of K arrays of shape (4, 3, N_e+1)."""
        nf = self.orders_.astype(np.int64) + 1
        flat = np.empty(int(12 * nf.sum()), dtype=np.float64)
        check(lib().ndg_get_traces(self._h, ptr(flat)))
        if self.N_ == self.N_max_:
            return flat.reshape(self.n_elements_, 4, 3, self.N_ + 1)
        offs = np.concatenate([[0], np.cumsum(12 * nf)])
        return [flat[offs[e]:offs[e + 1]].reshape(4, 3, nf[e]) for e in range(self.n_elements_)]

    def compute_rhs(self, t=0.0):
        out = [np.empty(self.n_dofs_per_field_, dtype=np.float64) for _ in range(3)]
        check(lib().ndg_compute_rhs(self._h, float(t), *(ptr(a) for a in out)))
        return tuple(out)

    def estimate_error(self):
        K = self.n_elements_
        err, sig = np.empty(K), np.empty(K)
        ref, coa = np.empty(K, dtype=np.int32), np.empty(K, dtype=np.int32)
        check(lib().ndg_estimate_error(self._h, ptr(err), ptr(sig), ptr(ref), ptr(coa)))
        return err, sig, ref.astype(bool), coa.astype(bool)

    def synchronize(self):
        check(lib().ndg_synchronize(self._h))

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _abi._lib is not None:
            _abi._lib.ndg_mesh_destroy(h)
        self._h = None

    def __del__(self):
        self.close()


class Solver2D_t:
    """Low-storage RK3 time stepping on the device (include/solvers/Solver2D_t.h:14-49)."""

    def __init__(self, CFL: float, output_times: Sequence[float] = (), viscosity: float = 0.0):
        self.CFL_ = float(CFL)
        self.output_times_ = [float(t) for t in output_times]
        self.viscosity_ = float(viscosity)
        self.time_ = 0.0
        self.n_steps_ = 0

    def get_delta_t(self, mesh: Mesh2D_t) -> float:
        dt = ctypes.c_double()
        check(lib().ndg_get_delta_t(mesh.handle, self.CFL_, ctypes.byref(dt)))
        return dt.value

    def solve(self, NDG: NDG_t, mesh: Mesh2D_t, data_writer=None):
        """Advances to output_times[-1] with the reference's step clipping; returns the number of
        steps taken. At the reference's output events (include/solvers/Solver2D_t.tcu:21-29,
        :74-83, :112-127), a DataWriter_t gets what the reference writes there: estimate_error
        (after the first step) and write_complete_data. Any other callable is called as
        data_writer(time, mesh) and may read the mesh (get_solution, get_interpolated_solution,
        get_complete_solution, estimate_error)."""
        from .data_writer import DataWriter_t
        if isinstance(data_writer, DataWriter_t):
            writer = data_writer

            def data_writer(time, mesh):
                if time > 0.0:
                    mesh.estimate_error()
                mesh.write_complete_data(time, writer)
        t_final, steps = ctypes.c_double(), ctypes.c_int64()
        if data_writer is None:
            check(lib().ndg_solve(mesh.handle, self.CFL_, self.output_times_[-1], ctypes.byref(t_final),
                                  ctypes.byref(steps)))
        else:
            times = np.asarray(self.output_times_, dtype=np.float64)
            raised = []

            def event(_mesh, time, _user):
                try:
                    mesh._refresh()
                    data_writer(time, mesh)
                    return _abi.NDG_OK
                except BaseException as exc:  # re-raised after the library call returns
                    raised.append(exc)
                    return _abi.NDG_ERR_INVALID

            cb = _abi.OUTPUT_FN(event)
            status = lib().ndg_solve_output(mesh.handle, self.CFL_, ptr(times), times.size,
                                            ctypes.cast(cb, ctypes.c_void_p), None, ctypes.byref(t_final),
                                            ctypes.byref(steps))
            if raised:
                raise raised[0]
            check(status)
        self.time_, self.n_steps_ = t_final.value, steps.value
        mesh._refresh()
        return steps.value

    def pre_condition(self, NDG: NDG_t, mesh: Mesh2D_t, n_adaptivity_steps: int, pre_condition_interval: int):
        """Solver2D_t::pre_condition (include/solvers/Solver2D_t.tpp:138-236): steps with an adapt
        every pre_condition_interval steps until n_adaptivity_steps adaptations are done, then the
        mesh's initial condition again on the adapted mesh."""
        check(lib().ndg_pre_condition(mesh.handle, self.CFL_, int(n_adaptivity_steps), int(pre_condition_interval), 0))
        mesh._refresh()

    def pre_condition_iterative(self, NDG: NDG_t, mesh: Mesh2D_t, n_adaptivity_steps: int,
                                pre_condition_interval: int):
        """Solver2D_t::pre_condition_iterative (include/solvers/Solver2D_t.tpp:238-336): the same,
        with the initial condition applied again (and the time reset) after every adaptation."""
        check(lib().ndg_pre_condition(mesh.handle, self.CFL_, int(n_adaptivity_steps), int(pre_condition_interval), 1))
        mesh._refresh()

    def step(self, mesh: Mesh2D_t, t: float, dt: float):
        check(lib().ndg_step(mesh.handle, float(t), float(dt)))

    def run_steps(self, mesh: Mesh2D_t, n_steps: int, time: float = 0.0) -> float:
        t = ctypes.c_double(time)
        check(lib().ndg_run_steps(mesh.handle, self.CFL_, int(n_steps), ctypes.byref(t)))
        return t.value
