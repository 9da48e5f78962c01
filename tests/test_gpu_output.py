"""Solver2D_t::solve with output times (include/solvers/Solver2D_t.tpp:21-29, :74-83, :112-127)
through the C ABI: the writer is called at the reference's events (time 0 for an output time in
[0, dt), the first step ending in [t_o, t_o + dt), the end when nothing was written there), output
times do not shorten steps, and the state seen at each event equals the oracle's after the same
number of steps."""
import numpy as np
import pytest

import ndg_cuda_amd as nd
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def expected_events(dt, output_times):
    """(time, steps) of each event, with the reference's host arithmetic."""
    t_end = output_times[-1]
    events = [(0.0, 0) for e in output_times if 0.0 >= e and 0.0 < e + dt]
    time, steps, step_dt = 0.0, 0, dt
    while time < t_end:
        steps += 1
        step_dt = dt if time + dt <= t_end else t_end - time
        time += step_dt
        for e in output_times:
            if time >= e and time < e + step_dt:
                events.append((time, steps))
                break
    if not any(time >= e and time < e + step_dt for e in output_times):
        events.append((time, steps))
    return events


@pytest.mark.parametrize("output_times", [[0.0, 0.01, 0.025, 0.04], [0.013, 0.0301]])
def test_output_events(output_times):
    print("[gpu] output events", output_times, flush=True)
    N = 4
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True, jitter=0.1, seed=2)
    gm = nd.Mesh2D_t(desc, N, N, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, nd.NDG_t(N, 5))
    gm.initial_conditions([(0.1, -0.2, 0.2, 1.0)])
    solver = nd.Solver2D_t(0.5, output_times)
    dt = solver.get_delta_t(gm)
    seen = []
    solver.solve(None, gm, lambda t, m: seen.append((t, [a.copy() for a in m.get_solution()])))
    want = expected_events(dt, output_times)
    assert [t for t, _ in seen] == [t for t, _ in want]
    ot = O.Tables(N, 5)
    for (t, state), (_, steps) in zip(seen, want):
        om = O.Mesh(ot, desc, N)
        om.initial_conditions([(0.1, -0.2, 0.2, 1.0)])
        if steps and t == output_times[-1] and t != steps * dt:
            om.solve(0.5, t)  # the clipped last step
        else:
            om.run_steps(0.5, steps)
        for g, r in zip(state, om.get_state()):
            assert np.max(np.abs(g - r)) <= 1e-12 * np.max(np.abs(r))


def test_writer_exception_stops_the_solve():
    print("[gpu] output writer error", flush=True)
    desc = nd.square_mesh(4, x_periodic=True, y_periodic=True)
    gm = nd.Mesh2D_t(desc, 3, 3, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, nd.NDG_t(3, 5))
    gm.initial_conditions(None)

    def writer(t, m):
        if t > 0:
            raise KeyError("stop")

    with pytest.raises(KeyError):
        nd.Solver2D_t(0.5, [0.0, 0.02, 0.05]).solve(None, gm, writer)


def test_solve_with_data_writer(tmp_path):
    """Solver2D_t.solve(NDG, mesh, DataWriter_t): the reference's output events write
    write_complete_data's files (point grid and element grid) and the .series index, estimating
    the error before every write but the one before the first step."""
    import json
    from ndg_cuda_amd.data_writer import read_vtu
    print("[gpu] solve with DataWriter_t", flush=True)
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True, jitter=0.1, seed=2)
    ndg = nd.NDG_t(5, 4)
    gm = nd.Mesh2D_t(desc, 5, 5, 4, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    gm.initial_conditions([(0.1, -0.2, 0.2, 1.0)])
    writer = nd.DataWriter_t(tmp_path / "out" / "run.pvtu")
    solver = nd.Solver2D_t(0.5, [0.0, 0.01, 0.02])
    solver.solve(ndg, gm, writer)
    files = json.loads((tmp_path / "out" / "run.pvtu.series").read_text())["files"]
    assert [f["time"] for f in files][0] == 0.0 and files[-1]["time"] == solver.time_ == 0.02
    assert len(files) == 3
    last = read_vtu(tmp_path / "out" / files[-1]["name"])
    x, y, p, u, v = gm.get_interpolated_solution()
    assert np.array_equal(last["point_data"]["Pressure"], p.ravel())
    assert np.array_equal(last["points"][:, 0], x.ravel())
    elements = json.loads((tmp_path / "out" / "run_element.pvtu.series").read_text())["files"]
    first = read_vtu(tmp_path / "out" / elements[0]["name"])
    final = read_vtu(tmp_path / "out" / elements[-1]["name"])
    assert not np.any(first["cell_data"]["PressureError"])  # no estimate before the first step
    err, sig, _, _ = gm.estimate_error()
    assert np.array_equal(final["cell_data"]["PressureError"], err)
    assert np.all(final["cell_data"]["N"] == 5)


def test_output_calls_do_not_move_the_solver_clock():
    """A writer that evaluates the right-hand side or the complete solution at other times (which
    run the right-hand-side kernels) leaves the solve's device clock alone: with an inflow
    boundary, whose ghost state depends on the stage times, the final state is the same bit for
    bit as without those calls."""
    print("[gpu] output calls and the device clock", flush=True)
    desc = nd.square_mesh(8, bc=(nd.BC_INFLOW, nd.BC_OUTFLOW, nd.BC_WALL, nd.BC_INFLOW), jitter=0.1, seed=6)
    states = []
    for disturb in (False, True):
        gm = nd.Mesh2D_t(desc, 4, 4, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, nd.NDG_t(4, 5))
        gm.initial_conditions(None)

        def writer(t, m):
            if disturb:
                m.compute_rhs(123.0)
                m.get_complete_solution(45.0)

        nd.Solver2D_t(0.5, [0.01, 0.02, 0.03]).solve(None, gm, writer)
        states.append([a.copy() for a in gm.get_solution()])
    for a, b in zip(*states):
        assert np.array_equal(a, b)
