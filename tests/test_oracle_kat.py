"""The reference's own known-answer tests for the mortar projections, the p/h solution transfer and
the inverse bilinear map, restated on the oracle's single-operation entry points (the same helpers
its mesh code runs) and, for the inverse map, on the library's host function too. Each test cites
the reference test it restates; the tolerances are the reference's.

Element-side traces, face fluxes and element fields are the reference tests' analytic functions
(sin x cos y, x, y on the element boundary; cos y, y^2, y on the face) sampled at the nodes, and the
expected values are the same functions at the output nodes. The inverse-map vectors are the
reference test's own data (tests/golden/reference_kats.json, extracted by make_reference_kats.py).
"""
import json
import os

import numpy as np
import pytest

import ndg_cuda_amd as nd
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-9  # element_to_face / face_to_element / element_to_element tests (max_error = 1e-9)


@pytest.fixture(scope="module")
def T():
    return O.Tables(16, 16)


def quad_map(xi, eta, P):
    """include/functions/quad_map.tpp:3-8."""
    P = np.asarray(P, dtype=np.float64)
    return (P[0] * (1 - xi) * (1 - eta) + P[1] * (xi + 1) * (1 - eta) + P[2] * (xi + 1) * (eta + 1)
            + P[3] * (1 - xi) * (eta + 1)) / 4


def fields(g):
    return np.array([np.sin(g[0]) * np.cos(g[1]), g[0], g[1]])


def element_traces(T, N, P):
    """The reference tests' element boundary traces: side s at reference points (x_i, -1),
    (1, x_i), (x_{N-i}, 1), (-1, x_{N-i}) (counter-clockwise storage), 3 fields."""
    x = T.get(0, N)
    out = np.empty((4, 3, N + 1))
    for i in range(N + 1):
        ref = [(x[i], -1.0), (1.0, x[i]), (x[N - i], 1.0), (-1.0, x[N - i])]
        for s in range(4):
            out[s, :, i] = fields(quad_map(*ref[s], P))
    return out


def face_values(T, Ne, Nf, side, offset, scale, tr):
    return np.array([O.kat_face_values(T, Ne, Nf, side, offset, scale, tr[f]) for f in range(3)])


POINTS = [(1.0, -1.0), (1.0, 1.0), (-1.0, 1.0), (-1.0, -1.0)]


@pytest.mark.parametrize("Ne", [16, 14])
def test_element_to_face_projection(T, Ne):
    """tests/element_to_face_projection_test.cu:265 (N_e = N_f = 16) and :377 (N_e = 14 to a face
    of order 16): the bottom and right sides go to their faces as the faces' first element, the top
    and left sides as the second element (face nodes walked backwards)."""
    Nf = 16
    tr = element_traces(T, Ne, POINTS)
    xf = T.get(0, Nf)
    P = np.array(POINTS)
    ends = [(0, 1), (1, 2), (3, 2), (0, 3)]      # face node pairs of the bottom, right, top, left faces
    for s, (a, b) in enumerate(ends):
        got = face_values(T, Ne, Nf, 0 if s < 2 else 1, 0.0, 1.0, tr[s])
        for j in range(Nf + 1):
            w = (xf[j] + 1) / 2
            exp = fields(P[b] * w + P[a] * (1 - w))
            assert np.max(np.abs(got[:, j] - exp)) < TOL, (s, j)


def test_element_to_two_faces_projection(T):
    """tests/element_to_face_projection_test.cu:489: every side covered by two faces of half its
    length (offsets 0 and 1/2, scales 1/2), the first four faces with the element as first element,
    the last four as second element."""
    N = 16
    fp = np.array([(1, -1), (1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1), (0, -1)], dtype=np.float64)
    tr = element_traces(T, N, fp[[0, 2, 4, 6]])
    x = T.get(0, N)
    # (element side, face side, offset of that side, face node pair (from, to))
    faces = [(0, 0, 0.0, (0, 1)), (0, 0, 0.5, (1, 2)), (1, 0, 0.0, (2, 3)), (1, 0, 0.5, (3, 4)),
             (2, 1, 0.5, (6, 5)), (2, 1, 0.0, (5, 4)), (3, 1, 0.5, (0, 7)), (3, 1, 0.0, (7, 6))]
    for s, side, off, (a, b) in faces:
        got = face_values(T, N, N, side, off, 0.5, tr[s])
        for j in range(N + 1):
            w = (x[j] + 1) / 2
            exp = fields(fp[b] * w + fp[a] * (1 - w))
            assert np.max(np.abs(got[:, j] - exp)) < TOL, (s, off, j)


def face_flux(x, y0, y1):
    """The reference tests' face flux cos y, y^2, y along the face from y0 to y1 (x = 0)."""
    y = y1 * (x + 1) / 2 + y0 * (1 - (x + 1) / 2)
    return np.array([np.cos(y), y * y, y])


@pytest.mark.parametrize("Ne", [16, 14])
def test_face_to_element_projection(T, Ne):
    """tests/face_to_element_projection_test.cu:161 (orders 16) and :267 (a face of order 16 onto
    elements of order 14): the face from (0, -1) to (0, 1) projects onto side 0 of its first element
    and, negated and reversed, onto side 2 of its second element; both edges have length 2 (edge
    scale 1)."""
    Nf = 16
    xf, xe = T.get(0, Nf), T.get(0, Ne)
    fl = face_flux(xf, -1.0, 1.0)
    left = np.array([O.kat_side_flux(T, Ne, Nf, 1, 0.0, 1.0, 1.0, fl[f]) for f in range(3)])
    right = np.array([O.kat_side_flux(T, Ne, Nf, 0, 0.0, 1.0, 1.0, fl[f]) for f in range(3)])
    exp = face_flux(xe, -1.0, 1.0)
    assert np.max(np.abs(left - exp)) < TOL
    assert np.max(np.abs(right[:, ::-1] + exp)) < TOL


def test_faces_to_element_projection(T):
    """tests/face_to_element_projection_test.cu:373: two faces of half length, (0, -1) -> (0, 0)
    and (0, 0) -> (0, 1) (offsets 0 and 1/2 on the first element, 1/2 and 0 on the second, scales
    1/2), summed onto side 0 of the first element and side 2 of the second."""
    N = 16
    x = T.get(0, N)
    halves = [(face_flux(x, -1.0, 0.0), 0.0, 0.5), (face_flux(x, 0.0, 1.0), 0.5, 0.0)]
    left = np.zeros((3, N + 1))
    right = np.zeros((3, N + 1))
    for fl, off_l, off_r in halves:
        for f in range(3):
            O.kat_side_flux(T, N, N, 1, off_l, 0.5, 1.0, fl[f], left[f])
            O.kat_side_flux(T, N, N, 0, off_r, 0.5, 1.0, fl[f], right[f])
    exp = face_flux(x, -1.0, 1.0)
    assert np.max(np.abs(left - exp)) < TOL
    assert np.max(np.abs(right[:, ::-1] + exp)) < TOL


def element_fields(T, N, P):
    x = T.get(0, N)
    out = np.empty((3, (N + 1) ** 2))
    for i in range(N + 1):
        for j in range(N + 1):
            out[:, i * (N + 1) + j] = fields(quad_map(x[i], x[j], P))
    return out


def test_element_to_smaller_element_interpolation(T):
    """tests/element_to_element_interpolation_test.cu:115: the fields of the element on [-1, 1]^2
    carried to the child (1, -1), (1, 0), (0, 0), (0, -1) through the inverse bilinear map."""
    N = 16
    small = [(1.0, -1.0), (1.0, 0.0), (0.0, 0.0), (0.0, -1.0)]
    src = element_fields(T, N, POINTS)
    got = np.array([O.kat_interpolate(T, N, src[f], N, POINTS, small) for f in range(3)])
    assert np.max(np.abs(got - element_fields(T, N, small))) < TOL


def test_element_to_higher_order_element_interpolation(T):
    """tests/element_to_element_interpolation_test.cu:195: the fields of an order-14 element carried
    to order 16 on the same element."""
    src = element_fields(T, 14, POINTS)
    got = np.array([O.kat_interpolate(T, 14, src[f], 16) for f in range(3)])
    assert np.max(np.abs(got - element_fields(T, 16, POINTS))) < TOL


REF = json.load(open(os.path.join(HERE, "golden", "reference_kats.json")))


@pytest.mark.parametrize("case", REF["inverse_quad_map"], ids=[c["name"] for c in REF["inverse_quad_map"]])
def test_inverse_quad_map_known_answers(case):
    """tests/inverse_quad_mapping_test.cu:9, :81, :154: square, parallelogram and trapezoid, 25
    points each, within the reference's 1e-6; the oracle and the library's host function agree
    bit for bit."""
    P, loc, glo = np.array(case["points"]), np.array(case["local"]), np.array(case["global"])
    got_o = np.array([O.kat_inverse_quad_map(P, x, y) for x, y in glo])
    got_l = nd.inverse_quad_map(P, glo)
    assert np.max(np.abs(got_o - loc)) < case["tolerance"]
    assert np.array_equal(got_o, got_l)
