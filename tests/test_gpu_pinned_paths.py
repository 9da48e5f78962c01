"""GPU parity on the paths whose oracle is pinned by the wave-speed and Hilbert-rotation tests:
wave speed c != 1, splitting in every Hilbert rotation, and the hp-adaptive loop at the shape of
the batched hp workload (N0 = 4 up to N = 12, split level <= 1, 8 seeded pulses, adaptation every
20 steps). The HIP path is called through the C ABI and compared with the oracle on the same
inputs. Bar: relative max error <= 1e-12 per field; element counts, orders, statuses, rotations,
step counts and times equal.
"""
import numpy as np
import pytest

import ndg_cuda_amd as nd
from ndg_cuda_amd import adjoint as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
MIXED_BC = (nd.BC_WALL, nd.BC_INFLOW, nd.BC_OUTFLOW, nd.BC_SYMMETRY)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def say(*args):
    print("[gpu]", *args, flush=True)


@pytest.fixture(scope="module")
def tables():
    return nd.NDG_t(16, 5), O.Tables(16, 5)


SPEED_CASES = [
    ("periodic-jitter", dict(periodic=True, jitter=0.2), 6),
    ("mixed-bc", dict(bc=MIXED_BC, jitter=0.1), 5),
    ("inflow", dict(bc=(nd.BC_INFLOW,) * 4, jitter=0.1), 7),
]


@pytest.mark.parametrize("c", [1.05, 0.8])
@pytest.mark.parametrize("name,kw,N", SPEED_CASES)
def test_wave_speed_rhs_and_steps(tables, c, name, kw, N):
    say("wave speed", c, name, "N", N)
    ndg, ot = tables
    desc = nd.square_mesh(8, bc=kw.get("bc", (nd.BC_WALL,) * 4), x_periodic=kw.get("periodic", False),
                          y_periodic=kw.get("periodic", False), jitter=kw.get("jitter", 0.0), seed=5)
    gm = nd.Mesh2D_t(desc, N, N, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    om = O.Mesh(ot, desc, N)
    A.set_wave_speed(gm, c)
    om.set_wave_speed(c)
    rng = np.random.default_rng(41)
    q = [rng.normal(size=gm.n_dofs_per_field_) for _ in range(3)]
    gm.set_solution(*q)
    om.set_state(*q)
    for g, r in zip(gm.compute_rhs(0.07), om.compute_rhs(0.07)):
        assert rel(g, r) < TOL
    solver = nd.Solver2D_t(0.5, [0.0])
    assert solver.get_delta_t(gm) == om.delta_t(0.5)
    if "inflow" in name or "mixed" in name:
        gm.initial_conditions(None)
        om.initial_conditions(None)
    t_gpu = solver.run_steps(gm, 9, 0.0)
    t_cpu = om.run_steps(0.5, 9, 0.0)
    assert t_gpu == t_cpu
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL


@pytest.mark.parametrize("c", [1.05, 0.8])
def test_wave_speed_full_solve(tables, c):
    """16 x 16 periodic, N = 6, CFL 0.5, to t = 0.5 at wave speed c: step count, time and state."""
    say("wave speed full solve", c)
    ndg, ot = tables
    desc = nd.square_mesh(16, x_periodic=True, y_periodic=True)
    gm = nd.Mesh2D_t(desc, 6, 6, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    om = O.Mesh(ot, desc, 6)
    A.set_wave_speed(gm, c)
    om.set_wave_speed(c)
    pulse = [(0.1, -0.2, 0.2, 1.0)]
    gm.initial_conditions(pulse)
    om.initial_conditions(pulse)
    solver = nd.Solver2D_t(0.5, [0.5])
    solver.solve(None, gm)
    steps, t_cpu = om.solve(0.5, 0.5)
    assert solver.n_steps_ == steps and solver.time_ == t_cpu
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL


def rotated(desc, shift):
    """The same mesh with every element's local nodes renumbered from node `shift` on (the
    generator's rotation 1 becomes rotation (1 + shift) % 4)."""
    return nd.MeshDescription(desc.xy, np.roll(np.asarray(desc.quads), -shift, axis=1), desc.bars,
                              desc.bar_type, desc.bar_donor)


def assert_same_elements(gm, om, t):
    cs = gm.get_complete_solution(t)
    _, st, rot, _ = om.elements()
    assert np.array_equal(cs["status"], st)
    assert np.array_equal(cs["rotation"], rot)
    assert np.array_equal(cs["N"], om.N)


@pytest.mark.parametrize("shift", [0, 1, 2, 3])
def test_split_solve_every_rotation(shift):
    """hp-adaptation every 4 steps, split level <= 1, on meshes whose elements have rotation
    0..3: element count, faces, orders, Hilbert statuses and rotations, time and state."""
    say("split solve rotation", (1 + shift) % 4)
    desc = rotated(nd.square_mesh(8, bc=(nd.BC_WALL,) * 4, jitter=0.1, seed=7), shift)
    ic = [(0.1, -0.2, 0.08, 1.0)]
    gm = nd.Mesh2D_t(desc, 3, 9, 5, 1, 4, 0, 1e-4, 1e-14, 0.0, nd.NDG_t(9, 5))
    om = O.Mesh(O.Tables(9, 5), desc, 3, tol_min=1e-4, tol_max=1e-14)
    gm.initial_conditions(ic)
    om.initial_conditions(ic)
    solver = nd.Solver2D_t(0.5, [0.2])
    solver.solve(None, gm)
    steps, t_cpu = om.solve_adaptive(0.5, 0.2, 4, 9, 1)
    say("steps", steps, "K", om.K)
    assert om.K > 64
    assert gm.n_elements_ == om.K and gm.n_faces_ == om.n_faces
    assert solver.n_steps_ == steps and solver.time_ == t_cpu
    assert np.array_equal(gm.orders_, om.N)
    assert_same_elements(gm, om, t_cpu)
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL


def batched_hp_pulses():
    """The 8 seeded pulses of the batched hp workload (seed 77: centres in [-0.7, 0.7]^2, widths
    0.05-0.2, amplitudes 0.2-1.0)."""
    rng = np.random.default_rng(77)
    return np.column_stack([rng.uniform(-0.7, 0.7, 8), rng.uniform(-0.7, 0.7, 8),
                            rng.uniform(0.05, 0.2, 8), rng.uniform(0.2, 1.0, 8)])


def test_hp_adaptive_at_workload_shape():
    """The batched hp workload's adaptive loop on a 32 x 32 base: walls, N0 = 4, N_max = 12,
    max_split_level = 1, tolerance 1e-6, the 8 seeded pulses, adaptation every 20 steps, to
    t = 0.08 (145 steps; the mesh grows to 1,729 elements of orders 4..10). Element and face
    counts, orders, statuses, rotations, step count, time and state against the oracle."""
    say("hp adaptive at workload shape")
    pulses = batched_hp_pulses()
    desc = nd.square_mesh(32, bc=(nd.BC_WALL,) * 4)
    gm = nd.Mesh2D_t(desc, 4, 12, 5, 1, 20, 0, 1e-6, 1e-14, 0.0, nd.NDG_t(12, 5))
    om = O.Mesh(O.Tables(12, 5), desc, 4, tol_min=1e-6, tol_max=1e-14)
    gm.initial_conditions(pulses)
    om.initial_conditions(pulses)
    solver = nd.Solver2D_t(0.5, [0.08])
    solver.solve(None, gm)
    steps, t_cpu = om.solve_adaptive(0.5, 0.08, 20, 12, 1)
    say("steps", steps, "K", om.K, "orders", np.bincount(om.N).tolist())
    assert om.K > 1024 and om.N.max() == 10 and om.N.min() == 4
    assert gm.n_elements_ == om.K and gm.n_faces_ == om.n_faces
    assert solver.n_steps_ == steps and solver.time_ == t_cpu
    assert np.array_equal(gm.orders_, om.N)
    assert_same_elements(gm, om, t_cpu)
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL


def test_periodised_pulses_initial_condition_and_pre_condition():
    """The periodised-pulse initial condition (pulses and their 8 nearest images) on a mixed-order
    periodic mesh against the oracle, and the iterative pre-conditioning that re-applies it after
    each p-adaptation (the equal-error workload's start)."""
    say("periodised pulses")
    pulses = np.array([[0.1, -0.2, 0.2, 1.0], [-0.6, 0.7, 0.15, 0.5], [0.8, 0.85, 0.1, 0.7]])
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True, jitter=0.1, seed=3)
    desc.quad_N = np.array([4, 6, 8, 5] * 16, dtype=np.int32)
    gm = nd.Mesh2D_t(desc, 4, 12, 5, 0, 0, 0, 1e-5, 1e-14, 0.0, nd.NDG_t(12, 5))
    om = O.Mesh(O.Tables(12, 5), desc, 4, tol_min=1e-5, tol_max=1e-14)
    gm.initial_conditions(pulses, period=(2.0, 2.0))
    om.initial_conditions(pulses, period=(2.0, 2.0))
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL
    solver = nd.Solver2D_t(0.5, [0.0])
    solver.pre_condition_iterative(None, gm, 2, 3)
    om.pre_condition(0.5, 2, 3, 12, 0, pulses, iterative=True, period=(2.0, 2.0))
    assert np.array_equal(gm.orders_, om.N)
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL
