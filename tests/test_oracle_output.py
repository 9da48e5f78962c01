"""The oracle's output interpolation, Mesh2D_t::get_solution (src/meshes/Mesh2D_t.cpp:4237-4249,
Element2D_t::interpolate_solution src/entities/Element2D_t.cpp:581-604), pinned by mathematics
(CPU only):

- the output coordinates are the bilinear map at the uniform points (i / (n - 1)) * 2 - 1: the
  four corner points are the element's vertices, and every point lies on the straight line
  between the points of the first and last rows of its column (bilinear maps are affine along a
  reference line);
- a field that is a polynomial of degree <= N in the reference coordinates, with different powers
  in xi and eta (so a transposed interpolation fails), is reproduced at the output points;
- a physical polynomial x^a y^b with a + b <= N on a jittered mesh with mixed orders is reproduced
  at the output coordinates.
"""
import numpy as np

import ndg_cuda_amd as nd
from oracle import oracle as O

import test_oracle_math as tm


def build(res, N, n, jitter=0.15, orders=None):
    T = O.Tables(12, n)
    desc = nd.square_mesh(res, bc=(nd.BC_WALL,) * 4, jitter=jitter, seed=9)
    if orders is not None:
        desc.quad_N = np.asarray(orders, dtype=np.int32)
    return T, desc, O.Mesh(T, desc, N)


def test_output_coordinates_are_the_bilinear_map():
    n = 6
    T, desc, om = build(4, 3, n)
    om.set_state(*(np.zeros(om.n) for _ in range(3)))
    x, y, _, _, _ = om.get_solution(n)
    P = desc.xy[desc.quads]  # (K, 4, 2)
    for (i, j), k in (((0, 0), 0), ((n - 1, 0), 1), ((n - 1, n - 1), 2), ((0, n - 1), 3)):
        assert np.max(np.abs(x[:, i, j] - P[:, k, 0])) < 1e-15
        assert np.max(np.abs(y[:, i, j] - P[:, k, 1])) < 1e-15
    t = np.arange(n) / (n - 1)
    for a in (x, y):  # affine along xi at fixed eta index and along eta at fixed xi index
        assert np.max(np.abs(a - (a[:, :1, :] * (1 - t)[None, :, None] + a[:, -1:, :] * t[None, :, None]))) < 1e-15
        assert np.max(np.abs(a - (a[:, :, :1] * (1 - t)[None, None, :] + a[:, :, -1:] * t[None, None, :]))) < 1e-15


def test_reference_polynomials_are_reproduced():
    n, N = 7, 4
    T, desc, om = build(2, N, n)
    x = T.get(O.O_NODES, N)
    fields = [lambda a, b: a**4 * b, lambda a, b: a * b**3 - 2 * a**2, lambda a, b: 1.0 + b**4 - a**3 * b]
    XI, ETA = np.meshgrid(x, x, indexing="ij")  # node (i, j) at (x_i, x_j)
    om.set_state(*(np.tile(f(XI, ETA).ravel(), om.K) for f in fields))
    out = om.get_solution(n)[2:]
    r = np.arange(n) / (n - 1) * 2 - 1
    RI, RJ = np.meshgrid(r, r, indexing="ij")
    for f, got in zip(fields, out):
        assert np.max(np.abs(got - f(RI, RJ)[None])) < 1e-13


def test_physical_polynomials_on_mixed_orders():
    n = 5
    orders = np.random.default_rng(3).choice([3, 4, 6], size=16)
    T, desc, om = build(4, 3, n, orders=orders)
    X, Y = tm.packed_nodes(desc, T, om.N)
    f = [lambda a, b: a**2 * b - a + 0.25, lambda a, b: a * b + b**3, lambda a, b: np.ones_like(a) * 2.0]
    om.set_state(*(g(X, Y) for g in f))
    x, y, *vals = om.get_solution(n)
    for g, got in zip(f, vals):
        assert np.max(np.abs(got - g(x, y))) < 1e-12


def plane_wave(x, y, t, c=1.0):
    """The reference's g (include/functions/analytical_solution.tpp:5-11, include/helpers/
    constants.h:15-18): k = (sqrt(2)/2, sqrt(2)/2), x0 = (-0.2, -0.2), d = 0.2 / (2 sqrt(ln 2))."""
    k = np.sqrt(2.0) / 2
    d = 0.2 / (2.0 * np.sqrt(np.log(2.0)))
    p = np.exp(-((k * (x + 0.2) + k * (y + 0.2) - c * t) ** 2) / d**2)
    return p, p * k, p * k


def test_complete_solution_point_fields():
    """The oracle's get_complete_solution (src/entities/Element2D_t.cu:623-665):
    - its first five planes are get_solution's;
    - on the plane-wave initial state at time 0 the analytical-error planes are exactly zero;
    - with a zero state they interpolate |g| = g, so they approach g at the output points as the
      order grows;
    - after compute_rhs of a physically linear p (u = v = 0) on an affine mesh the right-hand side
      planes hold the constant (0, -a, -b) on the elements away from the walls (the interior
      upwind fluxes of a continuous field vanish)."""
    n, N = 6, 8
    T, desc, om = build(4, N, n, jitter=0.0)
    om.initial_conditions(None)  # the plane wave
    full = om.get_complete_solution(n, 0.0)
    for a, b in zip(full[:5], om.get_solution(n)):
        assert np.array_equal(a, b)
    assert not np.any(full[8:])
    om.set_state(*(np.zeros(om.n) for _ in range(3)))
    for t in (0.0, 0.3):
        full = om.get_complete_solution(n, t)
        exact = plane_wave(full[0], full[1], t)
        for got, want in zip(full[8:], exact):
            assert np.max(np.abs(got - want)) < 5e-3 * np.max(want)
    X, Y = tm.packed_nodes(desc, T, om.N)
    a, b = 0.7, -1.3
    om.set_state(a * X + b * Y + 0.5, np.zeros(om.n), np.zeros(om.n))
    om.compute_rhs(0.0)
    full = om.get_complete_solution(n, 0.0)
    cx, cy = full[0].mean(axis=(1, 2)), full[1].mean(axis=(1, 2))
    inner = (np.abs(cx) < 0.5) & (np.abs(cy) < 0.5)
    assert inner.sum() == 4
    for plane, want in zip(full[5:8], (0.0, -a, -b)):
        assert np.max(np.abs(plane[inner] - want)) < 1e-11
