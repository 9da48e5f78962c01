"""The C++ mirror of the reference host API (include/ndg_sem.hpp), driven by tests/cpp/host_api_main
the way the reference's src/main.cpp drives its classes."""
import os
import subprocess

import numpy as np
import pytest

import ndg_cuda_amd as nd
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "cpp", "host_api_main")
FLOW = os.path.join(ROOT, "tests", "cpp", "reference_flow_main")


def test_driver_built():
    assert os.path.exists(DRIVER), "run __graft_entry__.build() first"


@pytest.mark.skipif(nd.device_count() > 0, reason="checks the no-device error path")
def test_driver_reports_status_without_gpu(tmp_path):
    r = subprocess.run([DRIVER, str(tmp_path / "o.bin"), "4", "4", "0.1"], capture_output=True, text=True)
    assert r.returncode == 1 and "error 4" in r.stderr


@pytest.mark.gpu
def test_driver_matches_oracle(tmp_path):
    res, N, t_end = 16, 6, 0.5
    print("[gpu] C++ host API driver", flush=True)
    r = subprocess.run([DRIVER, str(tmp_path / "o.bin"), str(res), str(N), str(t_end)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split("\n")
    events = [float(x) for x in lines[0].split()[1:]]
    steps = int(lines[1].split()[1])
    assert len(events) == 2 and events[1] == t_end and t_end / 2 <= events[0] < t_end / 2 + 0.0018
    got = np.fromfile(tmp_path / "o.bin", dtype=np.float64).reshape(3, -1)
    desc = nd.square_mesh(res, x_periodic=True, y_periodic=True)
    om = O.Mesh(O.Tables(N, 5), desc, N)
    om.initial_conditions([(0.1, -0.2, 0.2, 1.0)])
    n_ref, _ = om.solve(0.5, t_end)
    assert steps == n_ref == 289
    for g, ref in zip(got, om.get_state()):
        assert np.max(np.abs(g - ref)) / np.max(np.abs(ref)) < 1e-12
    outs = np.fromfile(str(tmp_path / "o.bin") + ".out", dtype=np.float64).reshape(5, res * res, 5, 5)
    for g, ref in zip(outs, om.get_solution(5)):
        assert np.max(np.abs(g - ref)) <= 1e-12 * max(np.max(np.abs(ref)), 1.0)
    # DataWriter_t files: write_complete_data of the final state, then a solve with the writer
    import json
    from ndg_cuda_amd.data_writer import read_vtu
    vtk = tmp_path / "o.bin_vtk"
    pts = read_vtu(vtk / "output_t0_500000000s.pvtu")
    assert np.array_equal(pts["point_data"]["Pressure"], outs[2].ravel())
    assert np.array_equal(pts["points"][:, 0], outs[0].ravel())
    assert np.array_equal(pts["point_data"]["Velocity"][:, 1], outs[4].ravel())
    els = read_vtu(vtk / "output_element_t0_500000000s.pvtu")
    assert np.all(els["cell_data"]["N"] == N) and np.array_equal(els["cell_data"]["Index"], np.arange(res * res))
    assert not np.any(els["cell_data"]["PressureError"])  # no estimate was made
    series = json.loads((vtk / "output.pvtu.series").read_text())
    assert series["files"] == [{"name": "output_t0_500000000s.pvtu", "time": 0.5}]
    written = float(lines[2].split()[1])
    files = json.loads((tmp_path / "o.bin_vtk2" / "output.vtu.series").read_text())["files"]
    assert [f["time"] for f in files] == [0.0, written] and written >= t_end / 8
    late = read_vtu(tmp_path / "o.bin_vtk2" / files[1]["name"])
    el2 = read_vtu(tmp_path / "o.bin_vtk2" / files[1]["name"].replace("output_", "output_element_"))
    assert np.any(el2["cell_data"]["PressureError"])  # estimated before the write after a step
    assert late["point_data"]["PressureDerivative"].shape == (res * res * 25,)


@pytest.mark.gpu
def test_driver_reads_su2_file(tmp_path):
    """Mesh2D_t(path, ...) on an SU2 file (jittered square with every boundary kind the format
    has), through the C++ mirror: the solve matches the oracle on the same description."""
    import meshes
    res, N, t_end = 8, 4, 0.05
    print("[gpu] C++ host API driver, SU2 mesh", flush=True)
    desc = nd.square_mesh(res, bc=(nd.BC_WALL, nd.BC_INFLOW, nd.BC_OUTFLOW, nd.BC_SYMMETRY), jitter=0.15, seed=3)
    path = tmp_path / "mesh.su2"
    meshes.write_su2(desc, path)
    r = subprocess.run([DRIVER, str(tmp_path / "o.bin"), str(res), str(N), str(t_end), str(path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    steps = int(r.stdout.split("\n")[1].split()[1])
    got = np.fromfile(tmp_path / "o.bin", dtype=np.float64).reshape(3, -1)
    om = O.Mesh(O.Tables(N, 5), nd.read_su2(path), N)
    om.initial_conditions([(0.1, -0.2, 0.2, 1.0)])
    n_ref, _ = om.solve(0.5, t_end)
    assert steps == n_ref
    for g, ref in zip(got, om.get_state()):
        assert np.max(np.abs(g - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.gpu
def test_driver_reads_periodic_su2_file(tmp_path):
    """The periodic 16 x 16 mesh (N = 6, to t = 0.5) written as an SU2 file with periodic marker pairs and read
    back by the C++ mirror's read_su2(path, pairs): the solve to t = 0.5 matches the oracle on the
    generator's description."""
    from ndg_cuda_amd import mesh_tools
    res, N, t_end = 16, 6, 0.5
    print("[gpu] C++ host API driver, periodic SU2 mesh", flush=True)
    desc = nd.square_mesh(res, x_periodic=True, y_periodic=True)
    path = tmp_path / "periodic.su2"
    tags = mesh_tools.write_su2(desc, path)
    r = subprocess.run([DRIVER, str(tmp_path / "o.bin"), str(res), str(N), str(t_end), str(path)]
                       + [f"{a}:{b}" for a, b in tags], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    steps = int(r.stdout.split("\n")[1].split()[1])
    got = np.fromfile(tmp_path / "o.bin", dtype=np.float64).reshape(3, -1)
    om = O.Mesh(O.Tables(N, 5), desc, N)
    om.initial_conditions([(0.1, -0.2, 0.2, 1.0)])
    n_ref, _ = om.solve(0.5, t_end)
    assert steps == n_ref == 289
    for g, ref in zip(got, om.get_state()):
        assert np.max(np.abs(g - ref)) / np.max(np.abs(ref)) < 1e-12


def test_reference_flow_driver_built():
    """tests/cpp/reference_flow_main.cpp is the reference program's call sequence (src/main.cpp
    host classes, src/main.cu device classes with a stream) compiled against ndg_sem.hpp."""
    assert os.path.exists(FLOW), "run __graft_entry__.build() first"


@pytest.mark.skipif(nd.device_count() > 0, reason="checks the no-device error path")
def test_reference_flow_reports_status_without_gpu(tmp_path):
    desc = nd.square_mesh(4, bc=(nd.BC_INFLOW,) * 4)
    from ndg_cuda_amd import mesh_tools
    mesh_tools.write_su2(desc, tmp_path / "m.su2")
    r = subprocess.run([FLOW, "host", str(tmp_path / "m.su2"), str(tmp_path / "o.bin"), str(tmp_path / "out.vtu"),
                        "4", "0.1"], capture_output=True, text=True)
    assert r.returncode == 1 and "error 4" in r.stderr, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("flow", ["host", "device"])
@pytest.mark.parametrize("adaptive", [False, True])
def test_reference_flow_matches_oracle(tmp_path, flow, adaptive):
    """The reference program's sequence with the reference constructor (polynomial nodes, one
    worker weight), initial_conditions(nodes) (the plane wave) and solve(NDG, mesh, writer) on an
    SU2 file with inlets on every side: steps, time, element count and state against the oracle;
    with adaptive=True the mesh adapts every 4 steps with max_split_level 1 (N 4 up to 8)."""
    from ndg_cuda_amd import mesh_tools
    res, N, t_end = 8, 4, 0.1
    print("[gpu] reference call sequence,", flow, "classes, adaptive" if adaptive else "", flush=True)
    desc = nd.square_mesh(res, bc=(nd.BC_INFLOW,) * 4, jitter=0.1, seed=5)
    path = tmp_path / "mesh.su2"
    mesh_tools.write_su2(desc, path)
    args = [FLOW, flow, str(path), str(tmp_path / "o.bin"), str(tmp_path / "vtk" / "out.vtu"), str(N), str(t_end)]
    if adaptive:
        args += ["1", "4"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    words = r.stdout.split()
    K, steps, time = int(words[1]), int(words[3]), float(words[5])
    got = np.fromfile(tmp_path / "o.bin", dtype=np.float64).reshape(3, -1)
    N_max = N + 4 if adaptive else N
    om = O.Mesh(O.Tables(N_max, 5), nd.read_su2(path), N, tol_min=1e-6, tol_max=1e-14)
    om.initial_conditions(None)
    if adaptive:
        n_ref, t_ref = om.solve_adaptive(0.5, t_end, 4, N_max, 1)
        assert om.K > res * res or om.N.max() > N
    else:
        n_ref, t_ref = om.solve(0.5, t_end)
    assert (K, steps, time) == (om.K, n_ref, t_ref)
    for g, ref in zip(got, om.get_state()):
        assert np.max(np.abs(g - ref)) / np.max(np.abs(ref)) < 1e-12
    assert (tmp_path / "vtk" / "out.vtu.series").exists()
