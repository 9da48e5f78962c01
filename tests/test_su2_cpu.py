"""SU2 mesh files (Mesh2D_t::read_su2, src/meshes/Mesh2D_t.cpp:78-330) through the C ABI, CPU only:
a written square mesh reads back as the same description (bars regrouped in the reference's ghost
order: walls, symmetry, inflow, outflow), and the reference's error cases are refused."""
import numpy as np
import pytest

import meshes
import ndg_cuda_amd as nd

MIXED_BC = (nd.BC_WALL, nd.BC_INFLOW, nd.BC_OUTFLOW, nd.BC_SYMMETRY)


def test_round_trip(tmp_path):
    desc = nd.square_mesh(8, bc=MIXED_BC, jitter=0.2, seed=4)
    path = tmp_path / "square.su2"
    meshes.write_su2(desc, path)
    got = nd.read_su2(path)
    assert np.array_equal(got.xy, desc.xy)  # repr() of a double reads back exactly
    assert np.array_equal(got.quads, desc.quads)
    order = [nd.BC_WALL, nd.BC_SYMMETRY, nd.BC_INFLOW, nd.BC_OUTFLOW]
    idx = np.concatenate([np.nonzero(desc.bar_type == k)[0] for k in order])
    assert np.array_equal(got.bars, desc.bars[idx])
    assert np.array_equal(got.bar_type, desc.bar_type[idx])
    assert (got.bar_donor == -1).all()


def test_marker_aliases_and_blank_lines(tmp_path):
    desc = nd.square_mesh(4, bc=(nd.BC_WALL,) * 4)
    path = tmp_path / "a.su2"
    meshes.write_su2(desc, path, tags={nd.BC_WALL: "AIRFOIL"})
    text = "\n\n" + path.read_text().replace("NELEM=", "\nnelem=")
    path.write_text(text)
    got = nd.read_su2(path)
    assert (got.bar_type == nd.BC_WALL).all() and len(got.bars) == 16


@pytest.mark.parametrize("edit,reason", [
    (lambda t: t.replace("NDIME= 2", "NDIME= 3"), "2 dimensions"),
    (lambda t: t.replace("NDIME= 2", "DIM= 2"), "NDIME="),
    (lambda t: t.replace("MARKER_TAG= wall", "MARKER_TAG= outlet"), "marker tag"),
    (lambda t: t.replace("\n9 ", "\n5 ", 1), "element type"),
    (lambda t: t.replace("\n3 ", "\n4 ", 1), "boundary element type"),
    (lambda t: t + "NZONE= 1\n", "expected"),
])
def test_reference_errors(tmp_path, edit, reason):
    desc = nd.square_mesh(2, bc=(nd.BC_WALL,) * 4)
    path = tmp_path / "bad.su2"
    meshes.write_su2(desc, path)
    path.write_text(edit(path.read_text()))
    with pytest.raises(nd.NDGError) as e:
        nd.read_su2(path)
    assert reason in str(e.value)


def test_missing_file_and_other_extensions(tmp_path):
    with pytest.raises(nd.NDGError):
        nd.read_su2(tmp_path / "none.su2")
    with pytest.raises(nd.NDGError) as e:
        nd.Mesh2D_t(str(tmp_path / "mesh.cgns"), 4, 4, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, nd.NDG_t(4, 5))
    assert e.value.code == 3
