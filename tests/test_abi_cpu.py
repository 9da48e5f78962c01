"""C ABI checks that need no GPU: the library loads, exports every declared entry point, its
host-side tables and mesh generator match the reference, and invalid input is reported with a
status code instead of terminating the process."""
import ctypes
import os
import re

import numpy as np
import pytest

import ndg_cuda_amd as nd
from ndg_cuda_amd import _abi
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ndg_amd.h")).read()
    return sorted(set(re.findall(r"^ndg_status\s+(ndg_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert len(names) >= 20
    assert sorted(names) == sorted(_abi.EXPORTED)
    L = _abi.lib()
    for n in names:
        assert hasattr(L, n), n


def test_device_count_without_gpu_is_not_an_error():
    assert nd.device_count() >= 0


@pytest.mark.parametrize("which", range(12))
def test_product_tables_match_reference_bitwise(ref_fixtures, which):
    ndg = nd.NDG_t(16, 5)
    key = O.TABLE_KEYS[which]
    name = _abi.TABLE_NAMES[which] + "_"
    for N in range(17):
        assert np.array_equal(getattr(ndg, name)[N], np.array(ref_fixtures[key][N])), (key, N)


def test_hilbert_tables_match_reference_bitwise(ref_fixtures):
    """The product's Hilbert splitting tables (child order, child statuses, first / last / inner
    status) for every rotation, status and side against the compiled Hilbert_splitting.cpp."""
    L = _abi.lib()
    order = np.zeros(4, dtype=np.int32)
    stat = np.zeros(4, dtype=np.int32)
    for rot in range(4):
        for st in range(4):
            _abi.check(L.ndg_hilbert_child_order(st, rot, _abi.ptr(order), _abi.ptr(stat)))
            assert order.tolist() == ref_fixtures["hilbert_child_order"][rot * 4 + st], (rot, st)
            assert stat.tolist() == ref_fixtures["hilbert_child_status"][rot * 4 + st], (rot, st)
    out = ctypes.c_int32()
    for side in range(4):
        _abi.check(L.ndg_hilbert_element_status(-1, side, ctypes.byref(out)))
        assert out.value == ref_fixtures["hilbert_first_status"][side]
        _abi.check(L.ndg_hilbert_element_status(side, -1, ctypes.byref(out)))
        assert out.value == ref_fixtures["hilbert_last_status"][side]
        for o in range(4):
            _abi.check(L.ndg_hilbert_element_status(side, o, ctypes.byref(out)))
            assert out.value == ref_fixtures["hilbert_element_status"][side * 4 + o]
    assert L.ndg_hilbert_child_order(4, 0, None, None) == _abi.NDG_ERR_INVALID
    assert L.ndg_hilbert_element_status(-1, -1, ctypes.byref(out)) == _abi.NDG_ERR_INVALID


def test_generator_layout_matches_reference_generator(ref_fixtures):
    """extras/mesh_generator/main.cpp:189-305: node ids, Hilbert element order, local node order
    BR, TR, TL, BL, and the bar sections in the order bottom, right, top, left."""
    res = 16
    m = nd.square_mesh(res, x_min=0.0, x_max=1.0, y_min=0.0, y_max=1.0)
    ynr = res + 1
    hil = {(n, d): (x, y) for n, d, x, y in ref_fixtures["hilbert"]}
    for k in range(res * res):
        X, Y = hil[(res, k)]
        assert list(m.quads[k]) == [(X + 1) * ynr + Y, (X + 1) * ynr + Y + 1, X * ynr + Y + 1, X * ynr + Y]
    i = np.arange(res)
    assert np.array_equal(m.bars[:res], np.stack([(i + 1) * ynr, i * ynr], 1))
    assert np.array_equal(m.bars[res:2 * res], np.stack([ynr * res + i + 1, ynr * res + i], 1))
    assert np.array_equal(m.bars[2 * res:3 * res], np.stack([(res - i) * ynr - 1, (res - i + 1) * ynr - 1], 1))
    assert np.array_equal(m.bars[3 * res:], np.stack([res - i - 1, res - i], 1))
    assert np.allclose(m.xy[5 * ynr + 3], [5 / 16, 3 / 16])
    # counter-clockwise elements, positive area
    P = m.xy[m.quads]
    area = 0.5 * np.sum(P[:, :, 0] * np.roll(P[:, :, 1], -1, 1) - np.roll(P[:, :, 0], -1, 1) * P[:, :, 1], 1)
    assert np.all(area > 0)


def test_generator_periodic_pairs_and_jitter():
    res = 8
    m = nd.square_mesh(res, x_periodic=True, y_periodic=True, jitter=0.15, seed=7)
    assert np.all(m.bar_type == nd.BC_PERIODIC)
    for b in range(4 * res):
        d = m.bar_donor[b]
        assert m.bar_donor[d] == b
        # donor edges are translates of each other (same length, opposite direction)
        e1 = m.xy[m.bars[b, 1]] - m.xy[m.bars[b, 0]]
        e2 = m.xy[m.bars[d, 1]] - m.xy[m.bars[d, 0]]
        assert np.allclose(e1, -e2)
    h = 2.0 / res
    base = nd.square_mesh(res, x_periodic=True, y_periodic=True)
    shift = m.xy - base.xy
    assert np.abs(shift).max() <= 0.15 * h + 1e-15
    assert np.abs(shift).max() > 0.1 * h
    edge = np.isclose(np.abs(base.xy), 1.0).any(axis=1)
    assert np.all(shift[edge] == 0.0)
    again = nd.square_mesh(res, x_periodic=True, y_periodic=True, jitter=0.15, seed=7)
    assert np.array_equal(again.xy, m.xy)


def test_invalid_input_is_a_status_not_an_exit():
    with pytest.raises(nd.NDGError) as e:
        nd.square_mesh(12)
    assert e.value.code == _abi.NDG_ERR_INVALID and "power of two" in str(e.value)
    with pytest.raises(nd.NDGError):
        nd.NDG_t(-1, 5)
    ndg = nd.NDG_t(8, 5)
    desc = nd.square_mesh(4)
    # an element below initial_N would meet a boundary ghost of higher order
    desc.quad_N = np.array([4] * 15 + [3], dtype=np.int32)
    with pytest.raises(nd.NDGError) as e:
        nd.Mesh2D_t(desc, 4, 8, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    assert e.value.code == _abi.NDG_ERR_UNSUPPORTED
    bad = nd.square_mesh(4)
    bad.quads = bad.quads.copy()
    bad.quads[0, 0] = 10 ** 6
    with pytest.raises(nd.NDGError) as e:
        nd.Mesh2D_t(bad, 4, 8, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    assert e.value.code == _abi.NDG_ERR_INVALID


@pytest.mark.skipif(nd.device_count() > 0, reason="checks the no-device path")
def test_mesh_create_without_gpu_reports_no_device():
    ndg = nd.NDG_t(8, 5)
    with pytest.raises(nd.NDGError) as e:
        nd.Mesh2D_t(nd.square_mesh(4), 4, 8, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    assert e.value.code == _abi.NDG_ERR_NO_DEVICE


@pytest.mark.skipif(nd.device_count() > 0, reason="the host topology checks run before the no-device error")
def test_mesh_topology_errors_and_hanging_faces_on_the_host():
    """ndg_mesh_create builds the topology on the host before it looks for a device: malformed
    descriptions are refused with their own message, and a mesh with 2:1 hanging faces passes
    (reaching the no-device error), so every side found a neighbour, a half face or a bar."""
    import meshes
    ndg = nd.NDG_t(8, 5)

    def create(desc):
        with pytest.raises(nd.NDGError) as e:
            nd.Mesh2D_t(desc, 4, 8, 5, 1, 0, 0, 1e-6, 1e-14, 0.0, ndg)
        return e.value

    base = nd.square_mesh(8)
    split = meshes.split_elements(base, meshes.interior_elements(base)[::3])
    assert create(split).code == _abi.NDG_ERR_NO_DEVICE
    twice = nd.square_mesh(8)
    twice.quads = np.concatenate([twice.quads, twice.quads[:1]])
    err = create(twice)
    assert err.code == _abi.NDG_ERR_INVALID and "used twice" in str(err)
    stray = nd.square_mesh(8)
    stray.bars = stray.bars.copy()
    stray.bars[0] = [stray.quads[0, 0], stray.quads[0, 2]]  # a diagonal, no element edge
    err = create(stray)
    assert err.code == _abi.NDG_ERR_INVALID and "matches no element edge" in str(err)
    open_side = nd.square_mesh(8)
    keep = np.arange(1, len(open_side.bars))
    open_side.bars, open_side.bar_type = open_side.bars[keep], open_side.bar_type[keep]
    err = create(open_side)
    assert err.code == _abi.NDG_ERR_INVALID and "neither a neighbour nor a boundary bar" in str(err)
