"""The oracle's adaptation (src/meshes/Mesh2D_t.cpp:1644-2268, :5026-5049, :5186-5659,
:6232-6343, :7482-7538; Element2D_t.cpp:345-579, :735-741, :802-812; Hilbert_splitting.cpp)
pinned by mathematics (CPU only).

- interpolate_from to order N + 2 reproduces the old element polynomial at the new nodes (checked
  against a numpy Legendre-Vandermonde evaluation, independent of the barycentric formula);
- the decisions follow would_p_refine / would_h_refine (refine and sigma >= 1 and N + 2 <= N_max;
  a split request makes the no-split branch refuse);
- after adaptation the mesh is the same operator as a mesh built directly at the new orders with
  the same state, except that ghosts whose order changed have zero storage (the reference
  resizes them without recomputing the edge scaling factor): with every element raised, walls act
  as zero ghosts, which a directly built mesh with BC_NONE reproduces exactly; periodic meshes are
  unchanged in kind;
- splitting keeps the Hilbert curve continuous (element statuses, child order and child status
  tables), carries polynomial fields exactly to the children, and keeps the operator's exactness
  on the resulting hanging faces.
"""
import numpy as np
import pytest

import exact
import ndg_cuda_amd as nd
from oracle import oracle as O


def smooth_state(desc, tables, N):
    X, Y = exact.node_coordinates(desc, tables.get(O.O_NODES, N))
    return [np.exp(0.7 * X) * np.cos(1.3 * Y), np.sin(X + 0.3 * Y), np.cos(2 * X * Y)]


def legendre_eval(x_old, vals, x_new):
    """Evaluates the degree-N tensor interpolant of vals (nodes x_old) at x_new x x_new through a
    Legendre Vandermonde solve."""
    V = np.polynomial.legendre.legvander(x_old, len(x_old) - 1)
    W = np.polynomial.legendre.legvander(x_new, len(x_old) - 1)
    C = np.linalg.solve(V, np.linalg.solve(V, vals).T).T  # coefficients c[a][b]
    return W @ C @ W.T


def mesh(tables, res, N, bc=(nd.BC_WALL,) * 4, periodic=False, jitter=0.1, tol_min=1e-14):
    desc = nd.square_mesh(res, bc=bc, x_periodic=periodic, y_periodic=periodic, jitter=jitter, seed=3)
    return desc, O.Mesh(tables, desc, N, tol_min=tol_min, tol_max=1e-300)


def test_p_refinement_interpolates_the_element_polynomial(oracle_tables):
    N = 5
    desc, om = mesh(oracle_tables, 4, N)
    q = smooth_state(desc, oracle_tables, N)
    om.set_state(*q)
    err, sigma, refine, _ = om.estimate_error()
    assert refine.all() and (sigma >= 1).all()
    assert om.adapt(N_max=N + 2) == om.K
    assert (om.N == N + 2).all()
    x_old, x_new = oracle_tables.get(O.O_NODES, N), oracle_tables.get(O.O_NODES, N + 2)
    new = om.get_state()
    ne_old, ne_new = (N + 1) ** 2, (N + 3) ** 2
    for f in range(3):
        for e in range(om.K):
            old = q[f][e * ne_old:(e + 1) * ne_old].reshape(N + 1, N + 1)
            ref = legendre_eval(x_old, old, x_new).ravel()
            got = new[f][e * ne_new:(e + 1) * ne_new]
            assert np.max(np.abs(got - ref)) < 1e-13 * max(1.0, np.abs(ref).max())


def test_decisions_follow_the_reference_rules(oracle_tables):
    N = 4
    desc, om = mesh(oracle_tables, 4, N)
    q = smooth_state(desc, oracle_tables, N)
    om.set_state(*q)
    om.estimate_error()
    assert om.adapt(N_max=N + 1) == 0                  # N + 2 > N_max: nothing changes
    assert (om.N == N).all()
    # a rough field (sigma < 1 somewhere): with max_split_level = 0 the estimate only raises the
    # refine & sigma >= 1 elements (the no-split branch)
    rng = np.random.default_rng(0)
    om.set_state(*(rng.normal(size=om.n) for _ in range(3)))
    _, sigma, refine, _ = om.estimate_error()
    assert (refine & (sigma < 1)).any()
    expect = int((refine & (sigma >= 1)).sum())
    assert om.adapt(N_max=12, max_split_level=0) == expect
    assert int((om.N == N + 2).sum()) == expect


@pytest.mark.parametrize("periodic", [False, True])
def test_adapted_mesh_equals_directly_built_mesh(oracle_tables, periodic):
    """Every element raised: the adapted operator equals the directly built one at N + 2 with the
    same state, walls turned into zero ghosts (their storage was resized to zero)."""
    N = 4
    bc = (nd.BC_WALL,) * 4
    desc, om = mesh(oracle_tables, 4, N, bc=bc, periodic=periodic)
    om.set_state(*smooth_state(desc, oracle_tables, N))
    om.estimate_error()
    assert om.adapt(N_max=N + 2) == om.K
    state = om.get_state()
    desc2 = nd.square_mesh(4, bc=(nd.BC_NONE,) * 4 if not periodic else bc, x_periodic=periodic,
                           y_periodic=periodic, jitter=0.1, seed=3)
    fresh = O.Mesh(oracle_tables, desc2, N + 2)
    fresh.set_state(*state)
    for a, b in zip(om.compute_rhs(0.0), fresh.compute_rhs(0.0)):
        assert np.array_equal(a, b)


def test_solve_adaptive_runs_the_reference_loop(oracle_tables):
    """Adaptation every 3 steps changes Δt (recomputed each step from the new orders) and keeps
    the total pressure integral (the periodic upwind scheme conserves it, interpolation to a
    higher order preserves the element polynomial)."""
    N = 3
    desc, om = mesh(oracle_tables, 8, N, periodic=True, jitter=0.0, tol_min=1e-4)
    om.initial_conditions([(0.1, -0.2, 0.2, 1.0)])

    def mass():
        total = 0.0
        p = om.get_state()[0]
        off = 0
        for e in range(om.K):
            n = om.N[e]
            w = oracle_tables.get(O.O_WEIGHTS, n)
            ne = (n + 1) ** 2
            # uniform mesh: J = (h/2)^2 everywhere
            total += float(np.outer(w, w).ravel() @ p[off:off + ne]) * (0.25 / 2) ** 2
            off += ne
        return total

    m0 = mass()
    steps, t = om.solve_adaptive(0.5, 0.05, 3, N_max=9)
    assert steps > 0 and t == pytest.approx(0.05, abs=0)
    assert om.N.max() > N
    assert abs(mass() - m0) < 1e-12 * abs(m0)


# ---------------------------------------------------------------------------------------------
# Splitting (hp_adapt, Mesh2D_t.cpp:5186-5659; Hilbert child order, Hilbert_splitting.cpp)

def split_all(om, level_cap):
    om.set_flags(np.ones(om.K, dtype=np.int32), np.zeros(om.K))
    return om.adapt(N_max=16, max_split_level=level_cap)


def assert_hilbert_path(om):
    """Consecutive elements share an edge: the element order is still a continuous curve."""
    nodes, _, _, _ = om.elements()
    for k in range(om.K - 1):
        assert len(set(nodes[k].tolist()) & set(nodes[k + 1].tolist())) == 2, k


@pytest.mark.parametrize("res", [4, 8])
def test_split_keeps_the_hilbert_curve(oracle_tables, res):
    """Statuses from compute_element_status and the child order / child status tables: splitting
    every element once and then again keeps the element sequence a continuous path."""
    desc, om = mesh(oracle_tables, res, 3, jitter=0.0)
    assert_hilbert_path(om)
    assert split_all(om, 2) == res * res
    assert om.K == 4 * res * res
    assert_hilbert_path(om)
    assert split_all(om, 2) == 4 * res * res
    assert om.K == 16 * res * res
    assert_hilbert_path(om)
    assert split_all(om, 2) == 0 or om.K == 16 * res * res


def rotated(desc, shift):
    """The same mesh with every element's local nodes renumbered from node `shift` on, which turns
    the generator's rotation 1 into rotation (1 + shift) % 4 (compute_element_status)."""
    out = nd.MeshDescription(desc.xy, np.roll(np.asarray(desc.quads), -shift, axis=1), desc.bars,
                             desc.bar_type, desc.bar_donor)
    return out


def curve_statuses(nodes, xy, fx):
    """compute_element_status (Mesh2D_t.cpp:4055-4178) from the element centres, with the status
    tables of the compiled Hilbert_splitting.cpp (fixtures): the side towards the previous element
    (incoming) and towards the next one (outgoing)."""
    c = xy[nodes].mean(axis=1)

    def side(d):
        return (0 if d[1] < 0 else 2) if abs(d[1]) > abs(d[0]) else (3 if d[0] < 0 else 1)

    K = len(nodes)
    out = np.empty(K, dtype=np.int32)
    out[0] = fx["hilbert_first_status"][side(c[1] - c[0])]
    out[-1] = fx["hilbert_last_status"][side(c[-2] - c[-1])]
    for k in range(1, K - 1):
        out[k] = fx["hilbert_element_status"][4 * side(c[k - 1] - c[k]) + side(c[k + 1] - c[k])]
    return out


@pytest.mark.parametrize("shift", [0, 1, 2, 3])
def test_split_keeps_the_hilbert_curve_every_rotation(oracle_tables, ref_fixtures, shift):
    """Every rotation row of the child order / child status tables: with the element frames
    turned (rotations 0..3), splitting every element twice keeps the curve continuous, and each
    child's status from the child status table is the one compute_element_status gives it from
    its curve neighbours."""
    desc = rotated(nd.square_mesh(8, bc=(nd.BC_WALL,) * 4, jitter=0.0, seed=3), shift)
    om = O.Mesh(oracle_tables, desc, 3, tol_min=1e-14, tol_max=1e-300)
    nodes, st, rot, xy = om.elements()
    assert (rot == (1 + shift) % 4).all()
    assert np.array_equal(st, curve_statuses(nodes, xy, ref_fixtures))
    assert_hilbert_path(om)
    for level in (1, 2):
        assert split_all(om, 2) > 0
        assert om.K == 64 * 4 ** level
        assert_hilbert_path(om)
        nodes, st, rot, xy = om.elements()
        assert (rot == (1 + shift) % 4).all()
        assert np.array_equal(st, curve_statuses(nodes, xy, ref_fixtures)), level


def test_split_interpolates_and_conserves(oracle_tables):
    """Children get the parent polynomial at their nodes (inverse_quad_map + barycentric
    interpolation), so a polynomial field of degree <= N is carried exactly, and the operator on
    the split mesh (hanging faces where a neighbour is not split) keeps its polynomial exactness."""
    N = 4
    desc, om = mesh(oracle_tables, 8, N, periodic=True, jitter=0.0)
    import test_oracle_math as tm
    X, Y = tm.packed_nodes(desc, oracle_tables, om.N)
    f = lambda x, y: (x**2 + 3 * x * y - y**3 + 0.5, y**2 - x * y + 2 * x**3, x * y + x**2 * y)
    om.set_state(*f(X, Y))
    inner = np.zeros(om.K, dtype=np.int32)
    inner[[e for e in range(om.K) if e % 3 == 0]] = 1
    # splitting must avoid the periodic boundary: keep only interior elements
    import meshes
    interior = set(meshes.interior_elements(desc))
    inner = np.array([1 if (e in interior and e % 3 == 0) else 0 for e in range(om.K)], dtype=np.int32)
    om.set_flags(inner, np.zeros(om.K))
    assert om.adapt(N_max=16, max_split_level=1) == int(inner.sum())
    nodes, _, _, xy = om.elements()
    fine = nd.MeshDescription(xy, nodes, desc.bars, desc.bar_type, desc.bar_donor)
    Xf, Yf = tm.packed_nodes(fine, oracle_tables, om.N)
    for got, ref in zip(om.get_state(), f(Xf, Yf)):
        assert np.max(np.abs(got - ref)) < 1e-12
    Gp, Gu, Gv = om.compute_rhs(0.0)
    ex = (-((-Yf + 6 * Xf**2) + (Xf + Xf**2)), -(2 * Xf + 3 * Yf), -(3 * Xf - 3 * Yf**2))
    mask = tm.interior_mask(fine, om.N, 8)
    for got, ref in zip((Gp, Gu, Gv), ex):
        assert np.max(np.abs(got[mask] - ref[mask])) < 1e-9


def test_split_walls_and_limits(oracle_tables):
    """Splitting an element on a wall splits its ghost in two (geometry recomputed); splitting
    every element of a periodic mesh splits every periodic ghost with its origin."""
    desc, om = mesh(oracle_tables, 4, 3, jitter=0.1)
    flags = np.zeros(om.K, dtype=np.int32)
    flags[0] = 1  # the first Hilbert element is a corner element: two wall sides
    om.set_flags(flags, np.zeros(om.K))
    assert om.adapt(N_max=16, max_split_level=1) == 1
    assert om.K == 16 + 3
    om.set_state(*(np.ones(om.n) for _ in range(3)))
    assert all(np.isfinite(g).all() for g in om.compute_rhs(0.0))
    # with max level 1 the four children stay, every other element splits (their sides facing the
    # children become conforming again)
    om.set_flags(np.ones(om.K, dtype=np.int32), np.zeros(om.K))
    assert om.adapt(N_max=16, max_split_level=1) == 15
    assert om.K == 64
    assert_hilbert_path(om)
    om.set_state(*(np.ones(om.n) for _ in range(3)))
    assert all(np.isfinite(g).all() for g in om.compute_rhs(0.0))
    # every element of a periodic mesh: conforming again across the period
    d2, om2 = mesh(oracle_tables, 4, 3, periodic=True)
    assert om2.n_faces == 2 * 4 * 3 + 4 * 4  # interior faces, then one ghost face per periodic side
    om2.set_flags(np.ones(om2.K, dtype=np.int32), np.zeros(om2.K))
    assert om2.adapt(N_max=16, max_split_level=1) == 16
    assert om2.K == 64 and om2.n_faces == 2 * 8 * 7 + 4 * 8
    assert_hilbert_path(om2)


# ---------------------------------------------------------------------------------------------
# Splitting across the period (split_interfaces, Mesh2D_t.cpp:6343-6756): a periodic ghost splits
# with its origin, so a split element faces an unsplit partner through two half ghosts on one end
# and one ghost covering two sides on the other (2:1 across the period). Sides covered by more than
# two faces (a small element split again next to an unsplit one), in the interior and across the
# period.

def centre_flags(om, pred):
    nodes, _, _, xy = om.elements()
    return np.array([1 if pred(x, y) else 0 for x, y in xy[nodes].mean(axis=1)], dtype=np.int32)


def split_where(om, *preds, level=5):
    for pred in preds:
        om.set_flags(centre_flags(om, pred), np.zeros(om.K))
        n = om.adapt(N_max=16, max_split_level=level)
        if n < 0:
            return n
    return n


def fine_description(om, desc):
    nodes, _, _, xy = om.elements()
    return nd.MeshDescription(xy, nodes, desc.bars, desc.bar_type, desc.bar_donor)


PERIODIC_SPLITS = {
    # elements at the periodic corner: 2:1 across both periods
    "corner block": ((lambda x, y: x > 0.5 and y > 0.5),),
    # one column and one row next to the periods (their corner element split on two sides)
    "column and row": ((lambda x, y: x > 0.75 or y < -0.75),),
    # the right column, then its partners on the left: conforming again across the period
    "then the partners": ((lambda x, y: x > 0.75), (lambda x, y: x < -0.75)),
    # the right column, then its partners, then the children on the period at level 2 (2:1 across
    # the period between levels 1 and 2)
    "level two": ((lambda x, y: x > 0.75), (lambda x, y: x < -0.75), (lambda x, y: x > 0.9 and y > 0.5)),
    # the right column, then its children on the period again: four quarter faces on the unsplit
    # partners' sides
    "four faces across the period": ((lambda x, y: x > 0.75), (lambda x, y: x > 0.9 and y > 0.5)),
    # one element, then its two children on its left side: four quarter faces on the left
    # neighbour's side, and two quarters and a half on the lower neighbour's
    "four faces": ((lambda x, y: 0 < x < 0.25 and 0 < y < 0.25), (lambda x, y: 0 < x < 0.1 and 0 < y < 0.25)),
    # one element, then its corner child three more times: pieces down to 1/16 of the
    # neighbours' sides
    "sixteenth": ((lambda x, y: 0 < x < 0.25 and 0 < y < 0.25), (lambda x, y: 0 < x < 0.1 and 0 < y < 0.1),
                  (lambda x, y: 0 < x < 0.05 and 0 < y < 0.05), (lambda x, y: 0 < x < 0.02 and 0 < y < 0.02)),
    # the same, then their partners too
    "two levels": ((lambda x, y: x > 0.75), (lambda x, y: x < -0.75),
                   (lambda x, y: x > 0.9 and y > 0.5), (lambda x, y: x < -0.9 and y > 0.5)),
}


# the most faces on one element side each configuration leaves
MOST_FACES = {"corner block": 2, "column and row": 2, "then the partners": 2, "level two": 2,
              "four faces across the period": 4, "four faces": 4, "sixteenth": 5, "two levels": 2}


@pytest.mark.parametrize("name", list(PERIODIC_SPLITS))
def test_periodic_split_keeps_the_operator_exact_and_conservative(oracle_tables, name):
    """On meshes split next to the periods (2:1 across them) or with sides covered by up to four
    faces (pieces down to 1/16 of a side): a constant state stays at rest, the
    mass integrals of a random state's right-hand side vanish, and the right-hand side of a
    smooth periodic field matches its exact derivative to the order-10 interpolation error
    (observed 1.4e-12; a periodic ghost copying the wrong child of its origin gives 20)."""
    N = 10
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True)
    om = O.Mesh(oracle_tables, desc, N, tol_min=1e-14, tol_max=1e-300)
    assert split_where(om, *PERIODIC_SPLITS[name]) > 0
    assert om.K > 64
    assert om.side_faces().max() == MOST_FACES[name]
    import test_oracle_math as tm
    fine = fine_description(om, desc)
    X, Y = tm.packed_nodes(fine, oracle_tables, om.N)
    W = tm.quadrature_weights(fine, oracle_tables, om.N)
    om.set_state(np.ones(om.n), 0.3 * np.ones(om.n), -0.2 * np.ones(om.n))
    assert max(np.abs(g).max() for g in om.compute_rhs(0.0)) < 1e-11
    rng = np.random.default_rng(1)
    om.set_state(*(rng.normal(size=om.n) for _ in range(3)))
    for g in om.compute_rhs(0.0):
        assert abs(float(W @ g)) < 1e-13
    k = np.pi
    om.set_state(np.sin(k * X) * np.cos(k * Y), np.cos(k * X + 0.3), np.sin(k * Y) * np.cos(k * X))
    exact_rhs = (k * np.sin(k * X + 0.3) - k * np.cos(k * Y) * np.cos(k * X), -k * np.cos(k * X) * np.cos(k * Y),
                 k * np.sin(k * X) * np.sin(k * Y))
    for g, e in zip(om.compute_rhs(0.0), exact_rhs):
        assert np.abs(g - e).max() < 1e-10


def test_split_refuses_faces_below_a_sixteenth(oracle_tables):
    """A small element covering 1/16 of a larger side that splits while the larger element does
    not would leave a face of 1/32 of the side: refused, nothing changed (the limit the device
    path shares)."""
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True)
    om = O.Mesh(oracle_tables, desc, 4, tol_min=1e-14, tol_max=1e-300)
    assert split_where(om, *PERIODIC_SPLITS["sixteenth"]) > 0
    K, faces = om.K, om.n_faces
    assert split_where(om, lambda x, y: 0 < x < 0.01 and 0 < y < 0.01) == -2
    assert om.K == K and om.n_faces == faces


def periodic_pulse_solve(oracle_tables, x0, y0, split=1, tol=1e-5, width=0.15):
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True)
    om = O.Mesh(oracle_tables, desc, 4, tol_min=tol, tol_max=1e-14)
    om.initial_conditions([(x0, y0, width, 1.0)], period=(2.0, 2.0))
    steps, t = om.solve_adaptive(0.5, 0.3, 5, 10, split)
    return om, steps, t


# (max split level, tolerance_min, pulse width, most faces on one side)
SHIFT_CASES = [(1, 1e-5, 0.15, 2), (2, 1e-4, 0.08, 4)]


@pytest.mark.parametrize("split,tol,width,most", SHIFT_CASES)
def test_periodic_split_solve_equals_the_shifted_solve(oracle_tables, split, tol, width, most):
    """Translation invariance: an hp-adaptive solve with the pulse on the periodic corner (its
    elements split across both periods; at split level 2 sides with up to four faces appear)
    equals, element by element, the same solve with the pulse moved by half a period to the centre
    (split in the interior), shifted back: same step count, element count and orders, states equal
    to rounding (observed 4e-15)."""
    a, steps_a, t_a = periodic_pulse_solve(oracle_tables, 0.95, 0.95, split, tol, width)
    b, steps_b, t_b = periodic_pulse_solve(oracle_tables, -0.05, -0.05, split, tol, width)
    assert steps_a == steps_b > 0 and t_a == t_b
    assert a.K == b.K > 64
    assert a.side_faces().max() == most
    assert np.array_equal(np.bincount(a.N), np.bincount(b.N))
    assert_same_up_to_shift(a, b, 1.0, 1.0, 1e-12)


def assert_same_up_to_shift(a, b, dx, dy, tol):
    """States of meshes a and b equal element by element, element e of a matching the element of
    b whose centre is e's centre moved by (-dx, -dy) modulo the period 2."""
    def centres(m):
        nodes, _, _, xy = m.elements()
        return xy[nodes].mean(axis=1)
    key = lambda c: tuple(np.round(c, 9))
    index_b = {key(c): k for k, c in enumerate(centres(b))}
    off_a = np.concatenate([[0], np.cumsum((a.N.astype(np.int64) + 1) ** 2)])
    off_b = np.concatenate([[0], np.cumsum((b.N.astype(np.int64) + 1) ** 2)])
    sa, sb = a.get_state(), b.get_state()
    scale = max(np.abs(s).max() for s in sb)
    for e, c in enumerate(centres(a)):
        shifted = (c - (dx, dy) + 1.0) % 2.0 - 1.0
        j = index_b[key(shifted)]
        assert a.N[e] == b.N[j]
        for f in range(3):
            d = np.abs(sa[f][off_a[e]:off_a[e + 1]] - sb[f][off_b[j]:off_b[j + 1]]).max()
            assert d < tol * scale, (e, f, d)
