"""The oracle pinned bit for bit against the reference's own compiled sources.

tests/golden/ref_fixtures.json was produced by oracle/ref_fixtures.cpp linked with the
reference's LegendrePolynomial_t.cpp, Element2D_t.cpp, Face2D_t.cpp and Hilbert*.cpp (see
tests/golden/make_golden.py). Every comparison here is exact (==), not a tolerance.
"""
import numpy as np
import pytest

from oracle import oracle as O


@pytest.mark.parametrize("which", range(12))
def test_tables_bitwise(ref_fixtures, oracle_tables, which):
    key = O.TABLE_KEYS[which]
    for N in range(ref_fixtures["N_max"] + 1):
        ref = np.array(ref_fixtures[key][N])
        got = oracle_tables.get(which, N)
        assert got.shape == ref.shape, (key, N)
        assert np.array_equal(got, ref), (key, N, np.max(np.abs(got - ref)))


def test_element_traces_bitwise(ref_fixtures, oracle_tables):
    for case in ref_fixtures["elements"]:
        N = case["N"]
        got = O.element_traces(oracle_tables, N, case["p"], case["u"], case["v"])
        assert np.array_equal(got, np.array(case["traces"])), N


def test_element_geometry_bitwise(ref_fixtures, oracle_tables):
    for case in ref_fixtures["elements"]:
        got = O.element_geometry(oracle_tables, case["N"], case["points"])
        assert np.array_equal(got, np.array(case["geometry"])), case["N"]


def test_estimator_bitwise(ref_fixtures, oracle_tables):
    for case in ref_fixtures["elements"]:
        N = case["N"]
        got = O.element_estimate(oracle_tables, N, case["p_smooth"])
        ref = np.array(case["estimate"])
        assert np.array_equal(got[: ref.size], ref), (N, got[: ref.size], ref)


def test_boundary_geometry_bitwise(ref_fixtures, oracle_tables):
    for case in ref_fixtures["boundary_elements"]:
        a, b = case["bar"][:2], case["bar"][2:]
        got = O.boundary_geometry(oracle_tables, case["N"], a + b + b + a)
        assert np.array_equal(got, np.array(case["out"]))


def test_face_geometry_bitwise(ref_fixtures):
    for case in ref_fixtures["faces"]:
        c = case["in"]
        got = O.face_geometry(c[:4], c[4:])
        assert np.array_equal(got, np.array(case["out"])), c


def test_hilbert_bitwise(ref_fixtures):
    for n, d, x, y in ref_fixtures["hilbert"]:
        assert O.hilbert_d2xy(n, d) == (x, y)


def test_hilbert_splitting_tables_bitwise(ref_fixtures):
    """Child order, child statuses and the first / last / inner element status tables for every
    rotation, status and side (Hilbert_splitting.cpp:3-81, compiled)."""
    co, cs, fi, la, inner = O.hilbert_tables()
    assert np.array_equal(co.reshape(16, 4), np.array(ref_fixtures["hilbert_child_order"]))
    assert np.array_equal(cs.reshape(16, 4), np.array(ref_fixtures["hilbert_child_status"]))
    assert np.array_equal(fi, np.array(ref_fixtures["hilbert_first_status"]))
    assert np.array_equal(la, np.array(ref_fixtures["hilbert_last_status"]))
    assert np.array_equal(inner.ravel(), np.array(ref_fixtures["hilbert_element_status"]))
