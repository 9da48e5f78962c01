"""Element renumbering (extras/mesh_renumberer/mesh_renumberer.py:88-155, hilbert.py) and the SU2
writer, CPU only:

- hilbert_xy2d inverts the square generator's d2xy (the reference's Hilbert.cpp, which the
  generator's element order pins bitwise), for every cell of 8 x 8 and 16 x 16 grids;
- "circular" orders the element centres by angle; "hilbert-circular" orders them by the Hilbert
  index of their mapped centre, and the map sends the disk onto the square (the four points on the
  diagonals at radius 1 go to the corners, the axes stay on the axes);
- renumbering is a relabelling: the oracle's right-hand side on the renumbered mesh is the
  permuted right-hand side of the original (bit for bit at uniform order);
- write_su2 then read_su2 returns the description (bars regrouped by kind).
"""
import numpy as np
import pytest

import meshes
import ndg_cuda_amd as nd
from ndg_cuda_amd import mesh_tools as mt
from oracle import oracle as O


@pytest.mark.parametrize("res", [8, 16])
def test_xy2d_inverts_the_generator_order(res):
    desc = nd.square_mesh(res)
    centers = mt.element_centers(desc)
    h = 2.0 / res
    for d, (cx, cy) in enumerate(centers):
        i, j = int((cx + 1) / h), int((cy + 1) / h)
        assert mt.hilbert_xy2d(res, i, j) == d


def test_disk_to_square_map():
    c = np.array([[np.sqrt(0.5), np.sqrt(0.5)], [-np.sqrt(0.5), np.sqrt(0.5)], [0.3, 0.0], [0.0, -0.6]])
    sq = mt.disk_to_square(c)
    assert np.allclose(np.abs(sq[:2]), 1.0, atol=1e-7)
    assert sq[2, 1] == 0.0 and 0.3 < sq[2, 0] < 1 and sq[3, 0] == 0.0 and -1 < sq[3, 1] < -0.6


def test_orders():
    desc = nd.square_mesh(16, jitter=0.2, seed=5)
    c = mt.element_centers(desc)
    circ = mt.renumber(desc, "circular")
    theta = np.arctan2(*mt.element_centers(circ)[:, ::-1].T)
    assert np.all(np.diff(theta) >= 0)
    hil = mt.renumber(desc, "hilbert-circular")
    order = mt.hilbert_circular_order(c)
    assert np.array_equal(hil.quads, desc.quads[order])
    n = 512  # smallest power of two above K = 256
    cells = np.rint((mt.disk_to_square(c) + 1) * n / 2).astype(np.int64) % n
    d = np.array([mt.hilbert_xy2d(n, int(a), int(b)) for a, b in cells])
    assert np.all(np.diff(d[order]) >= 0)
    with pytest.raises(ValueError):
        mt.renumber(desc, "spiral")


@pytest.mark.parametrize("algorithm", ["circular", "hilbert-circular"])
@pytest.mark.parametrize("mixed", [False, True])
def test_renumbering_is_a_relabelling(oracle_tables, algorithm, mixed):
    """Uniform order: bit for bit. Mixed orders: the side that is 'left' on a mortar face follows
    the numbering, and the face's interpolation/projection is taken from that side, so the same
    flux comes out of different (equivalent) arithmetic: 1e-12."""
    desc = nd.square_mesh(8, bc=(nd.BC_WALL, nd.BC_INFLOW, nd.BC_OUTFLOW, nd.BC_SYMMETRY), jitter=0.15, seed=2)
    orders = [3, 4, 6] if mixed else [4]
    desc.quad_N = np.random.default_rng(1).choice(orders, size=64).astype(np.int32)
    order = mt.circular_order(mt.element_centers(desc)) if algorithm == "circular" else \
        mt.hilbert_circular_order(mt.element_centers(desc))
    new = mt.renumber(desc, algorithm)
    a, b = O.Mesh(oracle_tables, desc, min(orders)), O.Mesh(oracle_tables, new, min(orders))
    rng = np.random.default_rng(3)
    ne = (desc.quad_N.astype(int) + 1) ** 2
    off = np.concatenate([[0], np.cumsum(ne)])
    q = [rng.normal(size=off[-1]) for _ in range(3)]
    perm = np.concatenate([np.arange(off[e], off[e + 1]) for e in order])
    a.set_state(*q)
    b.set_state(*(x[perm] for x in q))
    for ga, gb in zip(a.compute_rhs(0.0), b.compute_rhs(0.0)):
        if mixed:
            assert np.max(np.abs(ga[perm] - gb)) <= 1e-12 * np.max(np.abs(ga))
        else:
            assert np.array_equal(ga[perm], gb)


def test_su2_writer_round_trip(tmp_path):
    desc = mt.renumber(nd.square_mesh(8, bc=(nd.BC_WALL, nd.BC_INFLOW, nd.BC_OUTFLOW, nd.BC_SYMMETRY)),
                       "hilbert-circular")
    path = tmp_path / "r.su2"
    mt.write_su2(desc, path)
    got = nd.read_su2(path)
    assert np.array_equal(got.quads, desc.quads) and np.array_equal(got.xy, desc.xy)
    assert sorted(map(tuple, got.bars.tolist())) == sorted(map(tuple, np.asarray(desc.bars).tolist()))
    out = tmp_path / "c.su2"
    assert mt.main(["renumber", "-i", str(path), "-o", str(out), "-a", "circular"]) == 0
    assert np.array_equal(nd.read_su2(out).quads, mt.renumber(got, "circular").quads)
    bad = nd.square_mesh(4)
    bad.bar_type = np.full_like(np.asarray(bad.bar_type), nd.BC_NONE)  # no SU2 marker for this kind
    with pytest.raises(ValueError):
        mt.write_su2(bad, tmp_path / "n.su2")


def canonical_periodic(desc):
    """{element-edge node pair of a periodic bar: that of its donor}, independent of bar order."""
    bars = np.asarray(desc.bars)
    out = {}
    for b in np.nonzero(np.asarray(desc.bar_type) == nd.BC_PERIODIC)[0]:
        d = int(desc.bar_donor[b])
        out[frozenset(map(int, bars[b]))] = frozenset(map(int, bars[d]))
    return out


@pytest.mark.parametrize("xp,yp,jitter", [(True, True, 0.0), (True, False, 0.1), (False, True, 0.15)])
def test_periodic_su2_round_trip(tmp_path, xp, yp, jitter):
    """A generator mesh with periodic sides written as SU2 marker pairs and read back with those
    pairs: same nodes, elements, boundary kinds and periodic couplings (ndg_su2_fill_periodic)."""
    desc = nd.square_mesh(8, bc=(nd.BC_WALL, nd.BC_OUTFLOW, nd.BC_SYMMETRY, nd.BC_INFLOW), x_periodic=xp,
                          y_periodic=yp, jitter=jitter, seed=4)
    path = tmp_path / "p.su2"
    tags = mt.write_su2(desc, path)
    assert len(tags) == int(xp) + int(yp)
    back = nd.read_su2(path, periodic=tags)
    assert np.array_equal(back.xy, desc.xy) and np.array_equal(back.quads, desc.quads)
    assert canonical_periodic(back) == canonical_periodic(desc)
    kinds = lambda d: sorted(map(int, np.asarray(d.bar_type)))
    assert kinds(back) == kinds(desc)
    without = nd.read_su2  # without the pairs the periodic tags are unknown markers
    with pytest.raises(nd.NDGError) as e:
        without(path)
    assert "periodic_0_a" in str(e.value)


def test_periodic_su2_errors(tmp_path):
    desc = nd.square_mesh(4, x_periodic=True, y_periodic=True)
    path = tmp_path / "p.su2"
    tags = mt.write_su2(desc, path)
    with pytest.raises(nd.NDGError) as e:
        nd.read_su2(path, periodic=tags + [("ghost_a", "ghost_b")])
    assert "'ghost_a' not found" in str(e.value)
    # pairing a marker with the other pair's marker: no translated partners
    with pytest.raises(nd.NDGError) as e:
        nd.read_su2(path, periodic=[(tags[0][0], tags[1][0]), (tags[0][1], tags[1][1])])
    assert "partner" in str(e.value)


@pytest.mark.skipif(nd.device_count() > 0, reason="the host topology runs before the no-device error")
def test_periodic_su2_mesh_passes_the_host_topology(tmp_path):
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True, seed=1)
    path = tmp_path / "p.su2"
    tags = mt.write_su2(desc, path)
    with pytest.raises(nd.NDGError) as e:
        nd.Mesh2D_t(nd.read_su2(path, periodic=tags), 4, 4, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, nd.NDG_t(4, 5))
    assert e.value.code == nd._abi.NDG_ERR_NO_DEVICE


def test_renumber_cli_keeps_periodic_pairs(tmp_path, capsys):
    desc = nd.square_mesh(8, x_periodic=True, y_periodic=True)
    src, dst = tmp_path / "in.su2", tmp_path / "out.su2"
    tags = mt.write_su2(desc, src)
    args = ["renumber", "-i", str(src), "-o", str(dst), "-a", "hilbert-circular"]
    for a, b in tags:
        args += ["-p", f"{a}:{b}"]
    assert mt.main(args) == 0
    printed = [l.split(": ")[1].split(":") for l in capsys.readouterr().out.strip().splitlines()]
    back = nd.read_su2(dst, periodic=[tuple(p) for p in printed])
    assert len(canonical_periodic(back)) == len(canonical_periodic(desc))
