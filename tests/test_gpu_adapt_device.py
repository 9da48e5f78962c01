"""GPU tests of the device-resident adaptation (ndg_mesh_set_adapt_mode NDG_ADAPT_DEVICE): the
same adaptive solves through the device path and through the host path (topology rebuilt on the
host from the node lists) give the same element counts, face counts, orders, element-side links,
Hilbert statuses, rotations, split levels, estimates and states bit for bit; refused adaptations
change nothing; and the device path's adaptive loop matches the oracle at the batched workload's
shape. The oracle comparisons of the other adaptive tests run through the device path too (it is
the default). Splits across periodic sides and sides covered by more than two faces are compared
with the oracle, face count per side included, with the same solve shifted by half a period, and
with the host path bit for bit.
"""
import numpy as np
import pytest

import ndg_cuda_amd as nd
from oracle import oracle as O
from test_gpu_pinned_paths import batched_hp_pulses, rotated
from test_oracle_adapt import PERIODIC_SPLITS, SHIFT_CASES

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
    for a, b, name in zip(dm.sides(), hm.sides(), ("link", "hang", "next_small", "periodic", "span")):
        assert np.array_equal(a, b), name
    cd, ch = dm.get_complete_solution(t), hm.get_complete_solution(t)
    for name in ("N", "status", "rotation", "split_level", "p_error", "p_sigma", "refine", "coarsen", "x", "y"):
        # the estimator gives NaN on an element whose state is zero, on both paths
        assert np.array_equal(cd[name], ch[name], equal_nan=True), name
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


def rough_state(om, pred, rng):
    """Noise on the elements whose centre satisfies pred (they h-refine: at N = 6 a noise field's
    spectral decay rate is below 1), zero elsewhere (no refinement), packed in element order."""
    nodes, _, _, xy = om.elements()
    out = []
    for f in range(3):
        parts = []
        for e, (x, y) in enumerate(xy[nodes].mean(axis=1)):
            n = (int(om.N[e]) + 1) ** 2
            parts.append(rng.normal(size=n) if pred(x, y) else np.zeros(n))
        out.append(np.concatenate(parts))
    return out


def periodic_pair(N0=6, N_max=8, jitter=0.1):
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True, jitter=jitter, seed=7)
    m = nd.Mesh2D_t(desc, N0, N_max, 5, 2, 0, 0, 1e-4, 1e-14, 0.0, nd.NDG_t(N_max, 5))
    om = O.Mesh(O.Tables(N_max, 5), desc, N0, tol_min=1e-4, tol_max=1e-14)
    return m, om


def split_both(m, om, pred, rng, level=5):
    q = rough_state(om, pred, rng)
    m.set_solution(*q)
    om.set_state(*q)
    om.estimate_error()
    return om.adapt(8, level), q


def faces_per_side(m):
    """Faces on each element side from the device topology: a larger side's small sides (its chain
    of next links), one on any other side."""
    link, hang, nxt, _, _ = m.sides()
    out = np.ones(len(link), dtype=np.int64)
    for t in np.nonzero(hang == 2)[0]:
        n, c = 0, link[t]
        while c >= 0:
            n, c = n + 1, nxt[c]
        out[t] = n
    return out.reshape(-1, 4)


@pytest.mark.parametrize("name", list(PERIODIC_SPLITS))
def test_device_splits_across_the_period_match_oracle(name):
    """Splits next to the periods (2:1 across them, at split levels 1 and 2, and conforming again
    when the partners split) and small elements split again next to unsplit ones (sides with up to
    five faces, pieces down to 1/16): the device adaptation splits the same elements as the oracle
    (split_interfaces, src/meshes/Mesh2D_t.cpp:6343-6756), with the same element and face counts,
    faces on every side and orders; then the right-hand side of a periodised two-pulse state
    (relative to its largest field) and the state after 20 RK3 steps match the oracle to 1e-12."""
    say("periodic splits", name)
    # the 1/16 pieces are picked by their centres near the corner (0, 0), which jitter would move
    m, om = periodic_pair(jitter=0.0 if name == "sixteenth" else 0.1)
    rng = np.random.default_rng(11)
    for pred in PERIODIC_SPLITS[name]:
        n_o, _ = split_both(m, om, pred, rng)
        assert n_o > 0
        assert m.adapt(8, 5) == n_o
        assert m.n_elements_ == om.K and m.n_faces_ == om.n_faces
        assert np.array_equal(m.orders_, om.N)
        assert np.array_equal(faces_per_side(m), om.side_faces())
    link, hang, nxt, per, span = m.sides()
    assert ((hang != 0) & (per != 0)).any() or name in ("then the partners", "four faces", "sixteenth")
    say("K", m.n_elements_, "faces", m.n_faces_, "periodic hanging sides", int(((hang != 0) & (per != 0)).sum()),
        "most faces on a side", int(faces_per_side(m).max()))
    pulses = [(0.9, 0.85, 0.2, 1.0), (-0.6, 0.2, 0.3, 0.5)]
    m.initial_conditions(pulses, period=(2.0, 2.0))
    om.initial_conditions(pulses, period=(2.0, 2.0))
    G, R = m.compute_rhs(0.0), om.compute_rhs(0.0)
    scale = max(np.abs(r).max() for r in R)  # the pressure rate of a state at rest is tiny
    for g, r in zip(G, R):
        assert float(np.max(np.abs(g - r))) < 1e-12 * scale
    solver = nd.Solver2D_t(0.5)
    t = solver.run_steps(m, 20)
    assert t == om.run_steps(0.5, 20)
    for g, r in zip(m.get_solution(), om.get_state()):
        assert float(np.max(np.abs(g - r)) / np.max(np.abs(r))) < 1e-12


# interior splits that leave sides with several faces: four quarters, pieces down to 1/16, and
# the four-face configuration's larger neighbours split afterwards (their midpoints are nodes of
# the pieces' chain)
MULTI_FACE_SPLITS = {
    "four faces": PERIODIC_SPLITS["four faces"],
    "sixteenth": PERIODIC_SPLITS["sixteenth"],
    "four faces then the neighbours": PERIODIC_SPLITS["four faces"] + (
        (lambda x, y: -0.25 < x < 0 and 0 < y < 0.25 or 0 < x < 0.25 and -0.25 < y < 0),),
}
# the most faces on one side each leaves (after the neighbours split, their children's sides carry
# at most two faces; the upper neighbour's side keeps two quarters and a half)
MULTI_FACE_MOST = {"four faces": 4, "sixteenth": 5, "four faces then the neighbours": 3}


@pytest.mark.parametrize("name", list(MULTI_FACE_SPLITS))
def test_multi_face_splits_equal_host_path_and_oracle(name):
    """Sides covered by up to five faces on a walls mesh: the device path, the host path (topology
    rebuilt from the node lists, pieces found along each unmatched side) and the oracle split the
    same elements; device and host give the same element-side links, pieces, statuses and states
    bit for bit, before and after 10 RK3 steps, and match the oracle's face counts per side."""
    say("multi-face splits, device vs host", name)
    desc = nd.square_mesh(8, bc=(nd.BC_WALL,) * 4, jitter=0.0 if name == "sixteenth" else 0.1, seed=7)
    dm, hm = pair(desc, 6, 8, 5, 0, 1e-4, PULSES)
    om = O.Mesh(O.Tables(8, 5), desc, 6, tol_min=1e-4, tol_max=1e-14)
    rng = np.random.default_rng(11)
    for pred in MULTI_FACE_SPLITS[name]:
        q = rough_state(om, pred, rng)
        for m in (dm, hm):
            m.set_solution(*q)
        om.set_state(*q)
        om.estimate_error()
        n_o = om.adapt(8, 5)
        assert n_o > 0
        assert dm.adapt(8, 5) == n_o and hm.adapt(8, 5) == n_o
        assert_identical(dm, hm, 0.0)
        assert dm.n_elements_ == om.K and dm.n_faces_ == om.n_faces
        assert np.array_equal(faces_per_side(dm), om.side_faces())
    say("K", dm.n_elements_, "most faces on a side", int(faces_per_side(dm).max()))
    assert faces_per_side(dm).max() == MULTI_FACE_MOST[name]
    solver = nd.Solver2D_t(0.5)
    for m in (dm, hm):
        m.initial_conditions(PULSES)
    t = solver.run_steps(dm, 10)
    assert solver.run_steps(hm, 10) == t
    assert_identical(dm, hm, t)


@pytest.mark.parametrize("name", list(PERIODIC_SPLITS))
def test_periodic_splits_equal_host_path(name):
    """The splits next to the periods and the multi-face splits of the oracle comparison above,
    through the host path too (periodic bars split with their translation, donors found again
    at the partner end, a periodic side covered by shorter partner bars hanging as an interior
    side would): the same element-side links, pieces, periodic flags, statuses and states as the
    device path bit for bit after every adaptation and after 20 RK3 steps."""
    say("periodic splits, device vs host", name)
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True, jitter=0.0 if name == "sixteenth" else 0.1, seed=7)
    dm, hm = pair(desc, 6, 8, 5, 0, 1e-4, PULSES)
    om = O.Mesh(O.Tables(8, 5), desc, 6, tol_min=1e-4, tol_max=1e-14)
    rng = np.random.default_rng(11)
    for pred in PERIODIC_SPLITS[name]:
        q = rough_state(om, pred, rng)
        for m in (dm, hm):
            m.set_solution(*q)
        om.set_state(*q)
        om.estimate_error()
        n_o = om.adapt(8, 5)
        assert n_o > 0
        assert dm.adapt(8, 5) == n_o and hm.adapt(8, 5) == n_o
        assert_identical(dm, hm, 0.0)
        assert np.array_equal(faces_per_side(hm), om.side_faces())
    pulses = [(0.9, 0.85, 0.2, 1.0), (-0.6, 0.2, 0.3, 0.5)]
    for m in (dm, hm):
        m.initial_conditions(pulses, period=(2.0, 2.0))
    solver = nd.Solver2D_t(0.5)
    t = solver.run_steps(dm, 20)
    assert solver.run_steps(hm, 20) == t
    assert_identical(dm, hm, t)


@pytest.mark.parametrize("name", ["four faces", "sixteenth"])
def test_multi_face_description_matches_oracle(name):
    """A mesh description that already holds sides with several faces (the oracle's mesh after
    the splits, given as nodes and quads): the mesh built from it finds every chain of pieces
    (faces per side as the oracle's) and its right-hand side of a random state matches the
    oracle's to 1e-12 relative to the largest field."""
    say("multi-face description", name)
    desc = nd.square_mesh(8, bc=(nd.BC_WALL,) * 4, jitter=0.0 if name == "sixteenth" else 0.1, seed=7)
    om = O.Mesh(O.Tables(8, 5), desc, 6, tol_min=1e-4, tol_max=1e-14)
    rng = np.random.default_rng(11)
    for pred in MULTI_FACE_SPLITS[name]:
        om.set_state(*rough_state(om, pred, rng))
        om.estimate_error()
        assert om.adapt(8, 5) > 0
    nodes, _, _, xy = om.elements()
    fine = nd.MeshDescription(xy, nodes, desc.bars, desc.bar_type, desc.bar_donor)
    fine.quad_N = np.asarray(om.N, dtype=np.int32)
    m = nd.Mesh2D_t(fine, 6, 8, 5, 0, 0, 0, 1e-4, 1e-14, 0.0, nd.NDG_t(8, 5))
    assert m.n_elements_ == om.K and m.n_faces_ == om.n_faces
    assert np.array_equal(faces_per_side(m), om.side_faces()) and om.side_faces().max() == MULTI_FACE_MOST[name]
    q = [rng.normal(size=om.n) for _ in range(3)]
    m.set_solution(*q)
    om.set_state(*q)
    G, R = m.compute_rhs(0.0), om.compute_rhs(0.0)
    scale = max(np.abs(r).max() for r in R)
    for g, r in zip(G, R):
        assert float(np.max(np.abs(g - r))) < 1e-12 * scale


@pytest.mark.parametrize("mode", ["device", "host"])
def test_refuses_faces_below_a_sixteenth(mode):
    """A small element covering 1/16 of a larger side that splits while the larger element does
    not would leave a face of 1/32 of the side: the device path and the host path refuse before
    anything changes (same orders, element count and state), as the oracle does."""
    say("refusal below 1/16", mode)
    m, om = periodic_pair(jitter=0.0)
    m.set_adapt_mode(mode)
    rng = np.random.default_rng(11)
    for pred in PERIODIC_SPLITS["sixteenth"]:
        n_o, _ = split_both(m, om, pred, rng, level=6)
        assert m.adapt(8, 6) == n_o > 0
    n_o, q = split_both(m, om, lambda x, y: 0 < x < 0.01 and 0 < y < 0.01, rng, level=6)
    assert n_o == -2
    K, faces, orders = m.n_elements_, m.n_faces_, m.orders_.copy()
    with pytest.raises(nd.NDGError) as info:
        m.adapt(8, 6)
    assert info.value.code == nd._abi.NDG_ERR_UNSUPPORTED
    m._refresh()
    assert m.n_elements_ == K and m.n_faces_ == faces and np.array_equal(m.orders_, orders)
    for a, b in zip(m.get_solution(), q):
        assert np.array_equal(a, b)


def corner_pulse_solve(x0, y0, split, tol, width, mode="device"):
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True)
    m = nd.Mesh2D_t(desc, 4, 10, 5, split, 5, 0, tol, 1e-14, 0.0, nd.NDG_t(10, 5))
    m.set_adapt_mode(mode)
    m.initial_conditions([(x0, y0, width, 1.0)], period=(2.0, 2.0))
    solver = nd.Solver2D_t(0.5, [0.3])
    solver.solve(None, m)
    return m, solver


@pytest.mark.parametrize("split,tol,width,most", SHIFT_CASES)
def test_device_periodic_split_solve_matches_oracle_and_the_shifted_solve(split, tol, width, most):
    """The hp-adaptive solve with the pulse on the periodic corner (its elements split across both
    periods; at split level 2, sides with up to four faces) against the oracle (step count, time,
    element and face counts, faces per side, orders, state <= 1e-12), then against the same device
    solve with the pulse moved by half a period to the centre (split in the interior), shifted back
    element by element (<= 1e-12)."""
    say("periodic corner pulse solve, split level", split)
    a, sa = corner_pulse_solve(0.95, 0.95, split, tol, width)
    om = O.Mesh(O.Tables(10, 5), nd.square_mesh(8, x_periodic=True, y_periodic=True), 4, tol_min=tol,
                tol_max=1e-14)
    om.initial_conditions([(0.95, 0.95, width, 1.0)], period=(2.0, 2.0))
    steps, t = om.solve_adaptive(0.5, 0.3, 5, 10, split)
    assert sa.n_steps_ == steps and sa.time_ == t
    assert a.n_elements_ == om.K > 64 and a.n_faces_ == om.n_faces
    assert np.array_equal(faces_per_side(a), om.side_faces()) and om.side_faces().max() == most
    assert np.array_equal(a.orders_, om.N)
    for g, r in zip(a.get_solution(), om.get_state()):
        assert float(np.max(np.abs(g - r)) / np.max(np.abs(r))) < 1e-12
    b, sb = corner_pulse_solve(-0.05, -0.05, split, tol, width)
    assert sb.n_steps_ == steps and b.n_elements_ == a.n_elements_

    def centres(m):
        c = m.get_complete_solution(t)
        return np.stack([c["x"].mean(axis=(1, 2)), c["y"].mean(axis=(1, 2))], axis=1)

    key = lambda c: tuple(np.round(c, 9))
    index_b = {key(c): k for k, c in enumerate(centres(b))}
    off_a = np.concatenate([[0], np.cumsum((a.orders_.astype(np.int64) + 1) ** 2)])
    off_b = np.concatenate([[0], np.cumsum((b.orders_.astype(np.int64) + 1) ** 2)])
    qa, qb = a.get_solution(), b.get_solution()
    scale = max(np.abs(q).max() for q in qb)
    for e, c in enumerate(centres(a)):
        j = index_b[key((c - 1.0 + 1.0) % 2.0 - 1.0)]
        assert a.orders_[e] == b.orders_[j]
        for f in range(3):
            assert np.abs(qa[f][off_a[e]:off_a[e + 1]] - qb[f][off_b[j]:off_b[j + 1]]).max() < 1e-12 * scale


@pytest.mark.parametrize("split,tol,width,most", SHIFT_CASES)
def test_periodic_split_solve_equals_host_path(split, tol, width, most):
    """The corner-pulse hp solve above through the host path: the same steps, time, topology and
    state as the device path bit for bit."""
    say("periodic corner pulse solve, device vs host, split level", split)
    a, sa = corner_pulse_solve(0.95, 0.95, split, tol, width)
    b, sb = corner_pulse_solve(0.95, 0.95, split, tol, width, mode="host")
    assert sa.n_steps_ == sb.n_steps_ and sa.time_ == sb.time_
    assert faces_per_side(b).max() == most
    assert_identical(a, b, sa.time_)


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
