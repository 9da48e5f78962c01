"""Meshes whose elements have different orders (p-nonconforming faces): the HIP path, called
through the C ABI, against the oracle on the same inputs.

Elements are grouped by order on the device and the faces between orders go through the mortar
kernel (interpolation to the face order, upwind flux, L2 projection back). Bar: relative max error
<= 1e-12 per field, as for uniform meshes.
"""
import numpy as np
import pytest

import ndg_cuda_amd as nd
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
PULSES = [(0.1, -0.2, 0.2, 1.0), (-0.4, 0.3, 0.2, 0.5)]
MIXED_BC = (nd.BC_WALL, nd.BC_INFLOW, nd.BC_OUTFLOW, nd.BC_SYMMETRY)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def say(*args):
    print("[gpu]", *args, flush=True)


@pytest.fixture(scope="module")
def tables():
    return nd.NDG_t(16, 5), O.Tables(16, 5)


def build(tables, res, orders, initial_N, bc=(nd.BC_WALL,) * 4, periodic=False, jitter=0.0, seed=1):
    ndg, ot = tables
    desc = nd.square_mesh(res, bc=bc, x_periodic=periodic, y_periodic=periodic, jitter=jitter, seed=seed)
    desc.quad_N = np.asarray(orders, dtype=np.int32)
    gm = nd.Mesh2D_t(desc, initial_N, 16, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    om = O.Mesh(ot, desc, initial_N)
    return desc, gm, om


def random_orders(K, choices, seed):
    return np.random.default_rng(seed).choice(choices, size=K).astype(np.int32)


CASES = [
    ("periodic-jitter", dict(periodic=True, jitter=0.2), (4, 5, 6, 8)),
    ("wall-jitter", dict(jitter=0.2), (3, 4, 7)),
    ("mixed-bc", dict(bc=MIXED_BC, jitter=0.1), (2, 3, 5)),
    ("outflow", dict(bc=(nd.BC_OUTFLOW,) * 4, jitter=0.1), (6, 9)),
    ("zero-ghost-high", dict(bc=(nd.BC_NONE,) * 4), (10, 12, 16)),
]


@pytest.mark.parametrize("name,kw,choices", CASES)
def test_mixed_rhs(tables, name, kw, choices):
    say("mixed rhs", name, choices)
    orders = random_orders(64, choices, 3)
    desc, gm, om = build(tables, 8, orders, min(choices), **kw)
    assert gm.n_dofs_per_field_ == int(((orders + 1) ** 2).sum())
    rng = np.random.default_rng(5)
    q = [rng.normal(size=gm.n_dofs_per_field_) for _ in range(3)]
    gm.set_solution(*q)
    om.set_state(*q)
    for g, r in zip(gm.get_solution(), q):
        assert np.array_equal(g, r)
    t = 0.21
    ref = om.compute_rhs(t)
    got = gm.compute_rhs(t)
    for g, r in zip(got, ref):
        assert rel(g, r) < TOL, name
    # element traces come back in the caller's element order, each of its own order
    tr_ref = om.get_traces()
    tr = np.concatenate([a.ravel() for a in gm.interpolate_to_boundaries()])
    assert rel(tr, tr_ref) < TOL


@pytest.mark.parametrize("name,kw,choices", CASES)
def test_mixed_steps(tables, name, kw, choices):
    say("mixed steps", name, choices)
    orders = random_orders(64, choices, 11)
    desc, gm, om = build(tables, 8, orders, min(choices), **kw)
    if kw.get("bc") == MIXED_BC:
        gm.initial_conditions(None)
        om.initial_conditions(None)
    else:
        gm.initial_conditions(PULSES)
        om.initial_conditions(PULSES)
    solver = nd.Solver2D_t(0.5, [0.0])
    assert solver.get_delta_t(gm) == om.delta_t(0.5)
    t_gpu = solver.run_steps(gm, 9, 0.0)
    t_cpu = om.run_steps(0.5, 9, 0.0)
    assert t_gpu == t_cpu
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL, name


def test_mixed_blocks(tables):
    """Orders laid out in blocks (an order jump along whole lines of faces) on a jittered periodic
    mesh, run to t = 0.3 with the reference time loop."""
    say("mixed blocks periodic N 4/8")
    res = 16
    orders = np.where(np.arange(res * res) % 64 < 32, 4, 8).astype(np.int32)
    desc, gm, om = build(tables, res, orders, 4, periodic=True, jitter=0.1, seed=4)
    gm.initial_conditions(PULSES)
    om.initial_conditions(PULSES)
    solver = nd.Solver2D_t(0.5, [0.3])
    solver.solve(None, gm)
    steps_cpu, t_cpu = om.solve(0.5, 0.3)
    assert solver.time_ == t_cpu
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL


def test_mixed_estimator(tables):
    say("mixed estimator")
    orders = random_orders(64, (6, 8, 10), 2)
    desc, gm, om = build(tables, 8, orders, 6, jitter=0.1)
    gm.initial_conditions(PULSES)
    om.initial_conditions(PULSES)
    err, sig, refine, coarsen = gm.estimate_error()
    e_ref, s_ref, r_ref, c_ref = om.estimate_error()
    assert np.max(np.abs(err - e_ref)) < 1e-13
    big = e_ref > 1e-7
    assert np.max(np.abs(sig[big] - s_ref[big]) / np.abs(s_ref[big])) < 1e-8


def test_mixed_fp32c(tables):
    say("mixed fp32c")
    orders = random_orders(256, (4, 6, 7), 8)
    desc, gm, om = build(tables, 16, orders, 4, jitter=0.15, seed=9)
    gm.initial_conditions(PULSES)
    om.initial_conditions(PULSES)
    gm.set_precision("fp32c")
    solver = nd.Solver2D_t(0.5, [0.0])
    solver.run_steps(gm, 30)
    om.run_steps(0.5, 30)
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < 2e-5


@pytest.mark.parametrize("mode", ["sequential", "parallel"])
def test_graph_steps_after_autotune(tables, mode):
    """Captured-step replays (device clock drives the inflow ghosts' time) after autotuning, on a
    mixed-order mesh with every boundary kind; the autotuner leaves the state bit-identical."""
    say("graph", mode)
    orders = random_orders(256, (2, 4, 5, 7), 21)
    desc, gm, om = build(tables, 16, orders, 2, bc=MIXED_BC, jitter=0.1, seed=6)
    gm.initial_conditions(None)
    om.initial_conditions(None)
    before = np.concatenate(gm.get_solution())
    tuning = gm.autotune(2)
    say("tuning", tuning)
    assert [t[0] for t in tuning] == [2, 4, 5, 7]
    assert np.array_equal(np.concatenate(gm.get_solution()), before)
    gm.set_graph(mode)
    solver = nd.Solver2D_t(0.5, [0.0])
    t_gpu = solver.run_steps(gm, 6, 0.0)
    t_cpu = om.run_steps(0.5, 6, 0.0)
    assert t_gpu == t_cpu
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL
    # direct steps interleaved with replays continue from the same clock
    t_gpu = solver.run_steps(gm, 3, t_gpu)
    t_cpu = om.run_steps(0.5, 3, t_cpu)
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL


def test_graph_solve_16x16_N6(tables):
    """The reference time loop in graph mode: 288 replayed steps plus the clipped 289th launched
    directly (16 x 16 periodic, N = 6, t = 0.5)."""
    say("graph solve 16x16 periodic N=6")
    ndg, ot = tables
    desc = nd.square_mesh(16, x_periodic=True, y_periodic=True)
    gm = nd.Mesh2D_t(desc, 6, 6, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    om = O.Mesh(ot, desc, 6)
    pulse = [(0.1, -0.2, 0.2, 1.0)]
    gm.initial_conditions(pulse)
    om.initial_conditions(pulse)
    gm.set_graph("parallel")
    solver = nd.Solver2D_t(0.5, [0.5])
    solver.solve(None, gm)
    steps, t_cpu = om.solve(0.5, 0.5)
    assert solver.n_steps_ == steps == 289 and solver.time_ == t_cpu
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL


@pytest.mark.parametrize("periodic,bc", [(True, (nd.BC_WALL,) * 4), (False, MIXED_BC)])
def test_adaptive_solve_matches_oracle(tables, periodic, bc):
    """The reference time loop with p-adaptation every 4 steps (estimate, raise by 2 where
    refine && sigma >= 1, interpolate, ghosts and faces follow) against the oracle's restatement."""
    say("adaptive solve", periodic, bc)
    ndg, ot = tables
    desc = nd.square_mesh(8, bc=bc, x_periodic=periodic, y_periodic=periodic, jitter=0.1, seed=7)
    gm = nd.Mesh2D_t(desc, 3, 9, 5, 0, 4, 0, 1e-4, 1e-14, 0.0, nd.NDG_t(9, 5))
    om = O.Mesh(O.Tables(9, 5), desc, 3, tol_min=1e-4, tol_max=1e-14)
    if periodic:
        gm.initial_conditions(PULSES)
        om.initial_conditions(PULSES)
    else:
        gm.initial_conditions(None)
        om.initial_conditions(None)
    solver = nd.Solver2D_t(0.5, [0.2])
    solver.solve(None, gm)
    steps, t_cpu = om.solve_adaptive(0.5, 0.2, 4, 9, 0)
    say("steps", steps, "orders", np.bincount(om.N))
    assert solver.n_steps_ == steps and solver.time_ == t_cpu
    assert np.array_equal(gm.orders_, om.N)
    assert om.N.max() > 3
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL


@pytest.mark.parametrize("split", [0, 1])
def test_adaptive_solve_in_fp32c_form(split):
    """The adaptive solve stepping in the FP32 compensated form (each adaptation converts the state
    to FP64, rebuilds and converts back): the same step count, time, elements and orders as the
    oracle's FP64 adaptive solve on this workload, and the state within 2e-5 of it (the FP32
    form's bar against FP64, relative to the largest value of each field)."""
    say("adaptive solve in the fp32c form, split level", split)
    desc = nd.square_mesh(8, bc=(nd.BC_WALL,) * 4, jitter=0.1, seed=7)
    gm = nd.Mesh2D_t(desc, 3, 9, 5, split, 4, 0, 1e-4, 1e-14, 0.0, nd.NDG_t(9, 5))
    om = O.Mesh(O.Tables(9, 5), desc, 3, tol_min=1e-4, tol_max=1e-14)
    gm.initial_conditions(PULSES)
    om.initial_conditions(PULSES)
    gm.set_precision("fp32c")
    solver = nd.Solver2D_t(0.5, [0.2])
    solver.solve(None, gm)
    steps, t_cpu = om.solve_adaptive(0.5, 0.2, 4, 9, split)
    say("steps", steps, "K", om.K, "orders", np.bincount(om.N).tolist())
    assert solver.n_steps_ == steps and solver.time_ == t_cpu
    assert gm.n_elements_ == om.K and np.array_equal(gm.orders_, om.N)
    assert om.N.max() > 3 and (split == 0 or om.K > 64)
    gm.set_precision("fp64")
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < 2e-5


@pytest.mark.parametrize("bc,ic", [((nd.BC_WALL,) * 4, PULSES), (MIXED_BC, None),
                                   ((nd.BC_WALL,) * 4, [(0.1, -0.2, 0.08, 1.0)])])
def test_adaptive_split_solve_matches_oracle(tables, bc, ic):
    """hp-adaptation every 4 steps with max_split_level = 1: elements whose estimate asks for h
    refinement become four children in Hilbert order (state interpolated through the inverse
    bilinear map), split ghosts become two, the others are raised by 2; hanging faces appear
    between split and unsplit elements. Element count, orders, time and state against the
    oracle's restatement."""
    say("adaptive split solve", bc, ic is None)
    desc = nd.square_mesh(8, bc=bc, jitter=0.1, seed=7)
    gm = nd.Mesh2D_t(desc, 3, 9, 5, 1, 4, 0, 1e-4, 1e-14, 0.0, nd.NDG_t(9, 5))
    om = O.Mesh(O.Tables(9, 5), desc, 3, tol_min=1e-4, tol_max=1e-14)
    gm.initial_conditions(ic)
    om.initial_conditions(ic)
    solver = nd.Solver2D_t(0.5, [0.2])
    solver.solve(None, gm)
    steps, t_cpu = om.solve_adaptive(0.5, 0.2, 4, 9, 1)
    say("steps", steps, "K", om.K, "orders", np.bincount(om.N))
    assert om.K > 64
    assert gm.n_elements_ == om.K and gm.n_faces_ == om.n_faces
    assert solver.n_steps_ == steps and solver.time_ == t_cpu
    assert np.array_equal(gm.orders_, om.N)
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL


def test_complete_solution(tables):
    """Mesh2D_t::get_complete_solution on a mixed-order mesh with hanging faces, on the state 3
    steps give: the 11 point fields (state, right-hand side at the output time, plane-wave errors)
    against the oracle's restatement on the same state; element fields zero before the first
    estimate, then exactly the last estimate's."""
    say("complete solution")
    desc, gm, om = hanging_build(tables, 8, dict(jitter=0.1), (3, 5, 6), 21)
    gm.initial_conditions(PULSES)
    om.initial_conditions(PULSES)
    t = nd.Solver2D_t(0.5, [0.0]).run_steps(gm, 3)
    om.set_state(*gm.get_solution())  # the output arithmetic on one state (steps are pinned elsewhere)
    got = gm.get_complete_solution(t)
    om.compute_rhs(t)
    ref = om.get_complete_solution(5, t)
    names = nd._abi.COMPLETE_POINT_FIELDS
    for k, name in enumerate(names):
        assert got[name].shape == (om.K, 5, 5)
        if k < 2:
            assert np.max(np.abs(got[name] - ref[k])) < 1e-15, name
        else:
            assert rel(got[name], ref[k]) < TOL, name
    assert np.max(got["p_analytical_error"]) > 1e-3  # the pulses are not the plane wave
    assert np.array_equal(got["N"], om.N)
    _, status, rotation, _ = om.elements()
    assert np.array_equal(got["status"], status) and np.array_equal(got["rotation"], rotation)
    for name in ("p_error", "p_sigma", "refine", "coarsen", "u_error", "v_sigma"):
        assert not np.any(got[name]), name
    err, sig, ref_f, coa = gm.estimate_error()  # the estimator itself is pinned by the estimator tests
    got = gm.get_complete_solution(t)
    assert np.array_equal(got["p_error"], err) and np.array_equal(got["p_sigma"], sig)
    assert np.array_equal(got["refine"], ref_f) and np.array_equal(got["coarsen"], coa)
    assert not np.any(got["u_error"]) and not np.any(got["v_sigma"])


def test_complete_solution_after_adaptation():
    """Element fields across adaptations: a p-adaptation keeps the estimate of the elements it
    leaves alone and zeroes those it raises; a splitting adaptation zeroes the children. Split
    levels count the children; Hilbert statuses and rotations equal the oracle's."""
    say("complete solution after adaptation")
    desc = nd.square_mesh(8, jitter=0.1, seed=7)
    gm = nd.Mesh2D_t(desc, 3, 9, 5, 1, 0, 0, 1e-4, 1e-14, 0.0, nd.NDG_t(9, 5))
    gm.initial_conditions(PULSES)
    nd.Solver2D_t(0.5, [0.0]).run_steps(gm, 4)
    err, sig, refine, coarsen = gm.estimate_error()
    before = gm.orders_.copy()
    gm.adapt(9, 0)
    c = gm.get_complete_solution(0.0)
    raised = gm.orders_ != before
    assert raised.any() and not raised.all()
    assert np.array_equal(c["p_error"][~raised], err[~raised]) and not np.any(c["p_error"][raised])
    assert np.array_equal(c["refine"][~raised], refine[~raised].astype(np.int32))

    om = O.Mesh(O.Tables(9, 5), desc, 3, tol_min=1e-4, tol_max=1e-14)
    om.initial_conditions(PULSES)
    gm2 = nd.Mesh2D_t(desc, 3, 9, 5, 1, 4, 0, 1e-4, 1e-14, 0.0, nd.NDG_t(9, 5))
    gm2.initial_conditions(PULSES)
    solver = nd.Solver2D_t(0.5, [0.2])
    solver.solve(None, gm2)
    om.solve_adaptive(0.5, 0.2, 4, 9, 1)
    c = gm2.get_complete_solution(solver.time_)
    assert gm2.n_elements_ == om.K > 64
    _, status, rotation, _ = om.elements()
    assert np.array_equal(c["status"], status) and np.array_equal(c["rotation"], rotation)
    assert set(np.unique(c["split_level"])) <= {0, 1}
    assert 3 * int(np.sum(c["split_level"] == 1)) == 4 * (om.K - 64)


HANGING = [
    ("periodic-uniform", dict(periodic=True, jitter=0.15), (5,)),
    ("periodic-mixed", dict(periodic=True, jitter=0.1), (3, 5, 6)),
    ("walls-mixed", dict(jitter=0.1), (4, 6)),
]


def hanging_build(tables, res, kw, choices, seed):
    import meshes
    ndg, ot = tables
    desc = nd.square_mesh(res, x_periodic=kw.get("periodic", False), y_periodic=kw.get("periodic", False),
                          jitter=kw.get("jitter", 0.0), seed=seed)
    rng = np.random.default_rng(seed)
    inner = meshes.interior_elements(desc)
    desc.quad_N = rng.choice(choices, size=res * res).astype(np.int32)
    fine = meshes.split_elements(desc, rng.choice(inner, size=len(inner) // 3, replace=False))
    gm = nd.Mesh2D_t(fine, min(choices), 16, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    om = O.Mesh(ot, fine, min(choices))
    return fine, gm, om


@pytest.mark.parametrize("name,kw,choices", HANGING)
def test_hanging_rhs_and_steps(tables, name, kw, choices):
    """2:1 hanging faces (a side covered by two half faces of split neighbours): the interpolation
    to the half faces with offset/scale, and the sum of the two projected half-face fluxes."""
    say("hanging", name, choices)
    desc, gm, om = hanging_build(tables, 8, kw, choices, 13)
    assert gm.n_faces_ == om.n_faces
    rng = np.random.default_rng(4)
    q = [rng.normal(size=gm.n_dofs_per_field_) for _ in range(3)]
    gm.set_solution(*q)
    om.set_state(*q)
    for g, r in zip(gm.compute_rhs(0.0), om.compute_rhs(0.0)):
        assert rel(g, r) < TOL, name
    gm.initial_conditions(PULSES)
    om.initial_conditions(PULSES)
    solver = nd.Solver2D_t(0.5, [0.0])
    t_gpu = solver.run_steps(gm, 8, 0.0)
    t_cpu = om.run_steps(0.5, 8, 0.0)
    assert t_gpu == t_cpu
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL, name


def test_hanging_fp32c(tables):
    say("hanging fp32c")
    desc, gm, om = hanging_build(tables, 8, dict(periodic=True, jitter=0.1), (4, 6), 17)
    gm.initial_conditions(PULSES)
    om.initial_conditions(PULSES)
    gm.set_precision("fp32c")
    solver = nd.Solver2D_t(0.5, [0.0])
    solver.run_steps(gm, 20)
    om.run_steps(0.5, 20)
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < 2e-5


def test_output_interpolation(tables):
    """Mesh2D_t::get_solution on a mesh with mixed orders and hanging faces after a few steps:
    x, y, p, u, v at the uniform output points against the oracle."""
    say("output interpolation")
    desc, gm, om = hanging_build(tables, 8, dict(jitter=0.1), (3, 5, 6), 21)
    gm.initial_conditions(PULSES)
    om.initial_conditions(PULSES)
    nd.Solver2D_t(0.5, [0.0]).run_steps(gm, 3)
    om.run_steps(0.5, 3)
    got = gm.get_interpolated_solution()
    ref = om.get_solution(5)
    assert got[0].shape == (om.K, 5, 5)
    for g, r in zip(got[:2], ref[:2]):
        assert np.max(np.abs(g - r)) < 1e-15
    for g, r in zip(got[2:], ref[2:]):
        assert rel(g, r) < TOL
