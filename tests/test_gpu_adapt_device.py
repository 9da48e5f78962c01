"""GPU tests of the device-resident adaptation (ndg_mesh_set_adapt_mode NDG_ADAPT_DEVICE): the
same adaptive solves through the device path and through the host path (topology rebuilt on the
host from the node lists) give the same element counts, face counts, orders, element-side links,
Hilbert statuses, rotations, split levels, estimates and states bit for bit; refused adaptations
change nothing; and the device path's adaptive loop matches the oracle at the batched workload's
shape. The oracle comparisons of the other adaptive tests run through the device path too (it is
the default).
"""
import numpy as np
import pytest

import ndg_cuda_amd as nd
from oracle import oracle as O
from test_gpu_pinned_paths import batched_hp_pulses, rotated

pytestmark = pytest.mark.gpu

MIXED_BC = (nd.BC_WALL, nd.BC_INFLOW, nd.BC_OUTFLOW, nd.BC_SYMMETRY)
PULSES = [(0.1, -0.2, 0.2, 1.0), (-0.4, 0.3, 0.15, 0.5)]


def say(*args):
    print("[gpu]", *args, flush=True)


def pair(desc, N0, N_max, split, interval, tol, ic, period=None):
    """Two meshes of one description, one adapting on the device and one on the host, with the
    same initial condition."""
    out = []
    for mode in ("device", "host"):
        m = nd.Mesh2D_t(desc, N0, N_max, 5, split, interval, 0, tol, 1e-14, 0.0, nd.NDG_t(N_max, 5))
        m.set_adapt_mode(mode)
        m.initial_conditions(ic, period=period) if period else m.initial_conditions(ic)
        out.append(m)
    return out


def assert_identical(dm, hm, t):
    assert dm.n_elements_ == hm.n_elements_ and dm.n_faces_ == hm.n_faces_
    assert dm.n_dofs_per_field_ == hm.n_dofs_per_field_
    assert np.array_equal(dm.orders_, hm.orders_)
    for a, b, name in zip(dm.sides(), hm.sides(), ("link", "hang", "hang_second", "periodic")):
        assert np.array_equal(a, b), name
    cd, ch = dm.get_complete_solution(t), hm.get_complete_solution(t)
    for name in ("N", "status", "rotation", "split_level", "p_error", "p_sigma", "refine", "coarsen", "x", "y"):
        assert np.array_equal(cd[name], ch[name]), name
    for a, b in zip(dm.get_solution(), hm.get_solution()):
        assert np.array_equal(a, b)


CASES = [
    ("p periodic walls", dict(x_periodic=True, y_periodic=True, jitter=0.1, seed=7), 3, 9, 0, 4, 1e-4, PULSES),
    ("p mixed bc", dict(bc=MIXED_BC, jitter=0.1, seed=7), 3, 9, 0, 4, 1e-4, None),
    ("hp walls", dict(bc=(nd.BC_WALL,) * 4, jitter=0.1, seed=7), 3, 9, 1, 4, 1e-4, PULSES),
    ("hp mixed bc", dict(bc=MIXED_BC, jitter=0.1, seed=7), 3, 9, 1, 4, 1e-4, None),
    ("hp narrow pulse", dict(bc=(nd.BC_WALL,) * 4, jitter=0.1, seed=7), 3, 9, 1, 4, 1e-4, [(0.1, -0.2, 0.08, 1.0)]),
]


@pytest.mark.parametrize("name,mesh,N0,N_max,split,interval,tol,ic", CASES, ids=[c[0] for c in CASES])
def test_device_adaptation_equals_host_path(name, mesh, N0, N_max, split, interval, tol, ic):
    say("device vs host adaptation", name)
    desc = nd.square_mesh(8, **mesh)
    dm, hm = pair(desc, N0, N_max, split, interval, tol, ic)
    solver = nd.Solver2D_t(0.5, [0.2])
    solver.solve(None, dm)
    steps, t = solver.n_steps_, solver.time_
    solver.solve(None, hm)
    assert solver.n_steps_ == steps and solver.time_ == t
    say("K", dm.n_elements_, "orders", np.bincount(dm.orders_).tolist())
    assert dm.orders_.max() > N0
    assert_identical(dm, hm, t)


@pytest.mark.parametrize("shift", [0, 1, 2, 3])
def test_device_splits_in_every_rotation_equal_host_path(shift):
    say("device vs host adaptation, rotation", (1 + shift) % 4)
    desc = rotated(nd.square_mesh(8, bc=(nd.BC_WALL,) * 4, jitter=0.1, seed=7), shift)
    dm, hm = pair(desc, 3, 9, 1, 4, 1e-4, [(0.1, -0.2, 0.08, 1.0)])
    solver = nd.Solver2D_t(0.5, [0.2])
    solver.solve(None, dm)
    t = solver.time_
    solver.solve(None, hm)
    assert dm.n_elements_ > 64
    assert_identical(dm, hm, t)


def test_device_adaptation_twice_split_equals_host_path():
    """Split level 2 (a wide pulse, adaptation every 2 steps): children split again, next to
    neighbours that split in the same pass or earlier, so hanging faces are created and resolved
    (the 8 x 8 mesh grows to 616 elements)."""
    say("device vs host adaptation, split level 2")
    desc = nd.square_mesh(8, bc=(nd.BC_WALL,) * 4)
    dm, hm = pair(desc, 3, 9, 2, 2, 1e-4, [(0.0, 0.0, 0.5, 1.0)])
    solver = nd.Solver2D_t(0.5, [0.1])
    solver.solve(None, hm)
    t = solver.time_
    solver.solve(None, dm)
    levels = np.bincount(dm.get_complete_solution(t)["split_level"].astype(int).ravel()).tolist()
    say("K", dm.n_elements_, "levels", levels)
    assert len(levels) == 3 and dm.n_elements_ > 256
    assert_identical(dm, hm, t)


def test_device_adaptation_refusal_changes_nothing():
    """A split of an element with a periodic side is refused by the device path before anything
    changes: same orders, element count and state as before the call."""
    say("device refusal")
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True, jitter=0.1, seed=7)
    m = nd.Mesh2D_t(desc, 3, 9, 5, 1, 0, 0, 1e-4, 1e-14, 0.0, nd.NDG_t(9, 5))
    m.initial_conditions([(0.9, 0.9, 0.06, 1.0)])  # a narrow pulse on the periodic corner: h-refinement there
    before = m.get_solution()
    orders = m.orders_.copy()
    with pytest.raises(nd.NDGError) as info:
        m.adapt()
    assert info.value.code == nd._abi.NDG_ERR_UNSUPPORTED
    m._refresh()
    assert np.array_equal(m.orders_, orders) and m.n_elements_ == 64
    for a, b in zip(m.get_solution(), before):
        assert np.array_equal(a, b)


def test_device_adaptation_at_workload_shape_matches_oracle():
    """The batched hp workload's adaptive loop on a 32 x 32 base through the device path (walls,
    N0 = 4, N_max = 12, split level <= 1, 8 seeded pulses, adaptation every 20 steps) against the
    oracle, then against the host path bit for bit."""
    say("device adaptation at workload shape")
    pulses = batched_hp_pulses()
    desc = nd.square_mesh(32, bc=(nd.BC_WALL,) * 4)
    dm, hm = pair(desc, 4, 12, 1, 20, 1e-6, pulses)
    om = O.Mesh(O.Tables(12, 5), desc, 4, tol_min=1e-6, tol_max=1e-14)
    om.initial_conditions(pulses)
    solver = nd.Solver2D_t(0.5, [0.08])
    solver.solve(None, dm)
    steps, t = solver.n_steps_, solver.time_
    s_o, t_o = om.solve_adaptive(0.5, 0.08, 20, 12, 1)
    assert steps == s_o and t == t_o
    assert dm.n_elements_ == om.K and dm.n_faces_ == om.n_faces
    assert np.array_equal(dm.orders_, om.N)
    for g, r in zip(dm.get_solution(), om.get_state()):
        assert float(np.max(np.abs(g - r)) / np.max(np.abs(r))) < 1e-12
    solver.solve(None, hm)
    assert_identical(dm, hm, t)
    with pytest.raises(nd.NDGError):  # the node lists are gone after a device adaptation
        dm.set_adapt_mode("host")
