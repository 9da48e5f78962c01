"""CPU tests of the host exact-solution utilities (ndg_cuda_amd/exact.py): the factorised mode sum
against the direct per-point Fourier sum of tests/exact.py, the node layout against the mesh
generator's geometry, and the quadrature L2 norm against closed forms."""
import types

import numpy as np

import ndg_cuda_amd as nd
from ndg_cuda_amd import exact as E
import exact as X


PULSES = np.array([[0.1, -0.2, 0.12, 1.0], [-0.4, 0.3, 0.08, 0.5]])


def test_tensor_sum_matches_direct_sum():
    xs = np.linspace(-1, 1, 9)
    ys = np.linspace(-0.9, 0.95, 7)
    GX, GY = np.meshgrid(xs, ys, indexing="ij")
    for c, t in ((1.0, 0.0), (1.0, 0.37), (1.05, 0.21)):
        got = E.periodic_pulses(GX, GY, PULSES, t, c=c)
        ref = X.periodic_pulse(GX, GY, t, PULSES, n_modes=E._modes_needed(PULSES), c=c)
        for g, r in zip(got, ref):
            assert np.max(np.abs(g - r)) < 1e-13


def test_initial_pressure_is_the_periodised_gaussian():
    xs = np.linspace(-1, 1, 11)
    GX, GY = np.meshgrid(xs, xs, indexing="ij")
    p, u, v = E.periodic_pulses(GX, GY, PULSES, 0.0)
    ref = np.zeros_like(GX)
    for x0, y0, w, a in PULSES:
        for i in range(-2, 3):
            for j in range(-2, 3):
                r2 = (GX - x0 - 2 * i) ** 2 + (GY - y0 - 2 * j) ** 2
                ref += a * np.exp(-np.log(2.0) * r2 / w**2)
    assert np.max(np.abs(p - ref)) < 1e-13
    assert np.max(np.abs(u)) == 0.0 and np.max(np.abs(v)) == 0.0


def test_mesh_nodes_layout_and_quadrature_weights():
    desc = nd.square_mesh(4, x_periodic=True, y_periodic=True)
    ndg = nd.NDG_t(6, 5)
    orders = np.array([2, 4, 6, 3] * 4, dtype=np.int32)

    def nodes_of(n):
        return ndg.nodes_[n], ndg.weights_[n]

    Xn, Yn, W = E.mesh_nodes(desc, orders, nodes_of)
    assert Xn.size == int(((orders + 1) ** 2).sum())
    # the weights integrate 1 over the square, x^2 exactly (degree 2 <= every order)
    assert abs(W.sum() - 4.0) < 1e-13
    assert abs(float(np.sum(W * Xn**2)) - 4.0 / 3.0) < 1e-13
    assert abs(float(np.sum(W * Xn * Yn))) < 1e-14
    # the same coordinates as the per-order node map used by the oracle tests
    off = 0
    for e, n in enumerate(orders):
        sub = types.SimpleNamespace(xy=desc.xy, quads=desc.quads[e:e + 1])
        xr, yr = X.node_coordinates(sub, ndg.nodes_[n])
        sz = (n + 1) ** 2
        assert np.array_equal(Xn[off:off + sz], xr) and np.array_equal(Yn[off:off + sz], yr)
        off += sz


def test_l2_error_closed_forms():
    W = np.array([0.5, 1.5, 2.0])
    ex = (np.array([1.0, 2.0, 0.0]), np.zeros(3), np.array([0.0, 0.0, 1.0]))
    assert E.l2_error(ex, ex, W) == 0.0
    st = (ex[0] + 0.1, ex[1], ex[2])
    # num = 0.01 * 4, den = 0.5 + 6 + 2
    assert abs(E.l2_error(st, ex, W) - np.sqrt(0.04 / 8.5)) < 1e-15
