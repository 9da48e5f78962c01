"""DataWriter_t (src/helpers/DataWriter_t.cpp) without VTK: file names, the .series index, the
point and cell layout, and the arrays read back from the XML (binary and ASCII, .vtu and .pvtu).
CPU only: the inputs are synthetic output grids."""
import json
import os

import numpy as np
import pytest

from ndg_cuda_amd.data_writer import DataWriter_t, quad_connectivity, read_vtu, time_string


def synthetic(K=3, n=4, seed=0):
    rng = np.random.default_rng(seed)
    s = np.linspace(-1, 1, n)
    xi, eta = np.meshgrid(s, s, indexing="ij")
    x = np.stack([xi + 2 * e for e in range(K)])
    y = np.stack([eta + 0.5 * e for e in range(K)])
    f = {name: rng.normal(size=(K, n, n)) for name in ("p", "u", "v", "dp", "du", "dv", "pa", "ua", "va")}
    return K, n, x, y, f


def test_time_string():
    """std::fixed with 9 decimals, '.' replaced by '_' (DataWriter_t.cpp:113-116)."""
    assert time_string(0.0) == "_t0_000000000s"
    assert time_string(0.5) == "_t0_500000000s"
    assert time_string(1.25) == "_t1_250000000s"
    assert time_string(12.0000000004) == "_t12_000000000s"
    assert time_string(3e-10) == "_t0_000000000s"


def test_quad_connectivity_brute_force():
    """Cell (i, j) of element e: points (i + 1, j), (i + 1, j + 1), (i, j + 1), (i, j), counter-
    clockwise in (xi, eta) (DataWriter_t.cpp:58-71)."""
    K, n = 2, 4
    c = quad_connectivity(n, K)
    want = []
    for e in range(K):
        o = e * n * n
        for i in range(n - 1):
            for j in range(n - 1):
                want.append([o + (i + 1) * n + j, o + (i + 1) * n + j + 1, o + i * n + j + 1, o + i * n + j])
    assert np.array_equal(c, np.array(want))


@pytest.mark.parametrize("ext,fmt", [(".pvtu", "binary"), (".vtu", "binary"), (".vtu", "ascii")])
def test_write_data_round_trip(tmp_path, ext, fmt):
    K, n, x, y, f = synthetic()
    w = DataWriter_t(tmp_path / "data" / ("output" + ext), fmt=fmt)
    w.write_data(n, K, 0.0, x, y, f["p"], f["u"], f["v"])
    out = w.write_data(n, K, 0.25, x, y, f["p"], f["u"], f["v"])
    assert out.name == "output_t0_250000000s" + ext
    if ext == ".pvtu":
        assert (tmp_path / "data" / "output_t0_250000000s_0.vtu").exists()
    g = read_vtu(out)
    assert g["points"].shape == (K * n * n, 3)
    assert np.array_equal(g["points"][:, 0], x.ravel()) and np.array_equal(g["points"][:, 1], y.ravel())
    assert not np.any(g["points"][:, 2])
    assert np.array_equal(g["connectivity"], quad_connectivity(n, K))
    assert np.array_equal(g["offsets"], 4 * np.arange(1, K * (n - 1) ** 2 + 1))
    assert np.all(g["types"] == 9)
    assert np.array_equal(g["point_data"]["Pressure"], f["p"].ravel())
    assert np.array_equal(g["point_data"]["Velocity"], np.stack([f["u"].ravel(), f["v"].ravel()], axis=1))
    assert list(g["point_data"]) == ["Pressure", "Velocity"] and not g["cell_data"]
    series = json.loads((tmp_path / "data" / ("output" + ext + ".series")).read_text())
    assert series == {"file-series-version": "1.0",
                      "files": [{"name": "output_t0_000000000s" + ext, "time": 0.0},
                                {"name": "output_t0_250000000s" + ext, "time": 0.25}]}
    element_series = json.loads((tmp_path / "data" / ("output_element" + ext + ".series")).read_text())
    assert element_series == {"file-series-version": "1.0", "files": []}


def test_series_file_layout(tmp_path):
    """The series files are JSON with 2-space indentation, recreated empty by the constructor."""
    w = DataWriter_t(tmp_path / "o.pvtu")
    text = (tmp_path / "o.pvtu.series").read_text()
    assert text == '{\n  "file-series-version": "1.0",\n  "files": []\n}\n'
    K, n, x, y, f = synthetic(1, 2)
    w.write_data(n, K, 0.5, x, y, f["p"], f["u"], f["v"])
    assert '      "name": "o_t0_500000000s.pvtu",\n      "time": 0.5\n' in (tmp_path / "o.pvtu.series").read_text()
    DataWriter_t(tmp_path / "o.pvtu")
    assert json.loads((tmp_path / "o.pvtu.series").read_text())["files"] == []


def test_write_complete_data(tmp_path):
    K, n, x, y, f = synthetic(K=4, n=5, seed=3)
    rng = np.random.default_rng(9)
    N = np.array([3, 5, 4, 7], dtype=np.int32)
    err, sig = rng.random(K), rng.random(K)
    zeros = np.zeros(K)
    refine, coarsen = np.array([1, 0, 1, 0]), np.array([0, 1, 0, 0])
    level, status, rotation = np.array([0, 1, 1, 0]), np.array([0, 2, 3, 1]), np.array([1, 0, 2, 3])
    w = DataWriter_t(tmp_path / "out.pvtu")
    pts, els = w.write_complete_data(n, K, 0.125, 0, x, y, f["p"], f["u"], f["v"], N, f["dp"], f["du"], f["dv"],
                                     err, zeros, zeros, sig, zeros, zeros, refine, coarsen, level, f["pa"],
                                     f["ua"], f["va"], status, rotation)
    assert pts.name == "out_t0_125000000s.pvtu" and els.name == "out_element_t0_125000000s.pvtu"
    g = read_vtu(pts)
    assert list(g["point_data"]) == ["Pressure", "Velocity", "PressureDerivative", "VelocityDerivative",
                                     "PressureAnalyticalError", "VelocityAnalyticalError"]
    assert np.array_equal(g["point_data"]["PressureDerivative"], f["dp"].ravel())
    assert np.array_equal(g["point_data"]["VelocityAnalyticalError"][:, 1], f["va"].ravel())
    e = read_vtu(els)
    assert list(e["cell_data"]) == ["N", "Index", "Status", "Rotation", "Rank", "PressureError", "VelocityError",
                                    "PressureSigma", "VelocitySigma", "Refine", "Coarsen", "SplitLevel"]
    cd = e["cell_data"]
    assert cd["N"].dtype == np.int32 and np.array_equal(cd["N"], N)
    assert cd["Index"].dtype == np.uint64 and np.array_equal(cd["Index"], np.arange(K))
    assert np.array_equal(cd["PressureError"], err) and cd["VelocityError"].shape == (K, 2)
    assert np.array_equal(cd["Status"], status) and np.array_equal(cd["SplitLevel"], level)
    assert not np.any(cd["Rank"])
    # element corners: output points (0, 0), (n-1, 0), (n-1, n-1), (0, n-1) of each element
    corners = e["points"].reshape(K, 4, 3)
    for k in range(K):
        for c, (i, j) in enumerate([(0, 0), (n - 1, 0), (n - 1, n - 1), (0, n - 1)]):
            assert corners[k, c, 0] == x[k, i, j] and corners[k, c, 1] == y[k, i, j]
    assert np.array_equal(e["connectivity"], np.arange(4 * K).reshape(K, 4))
    series = json.loads((tmp_path / "out_element.pvtu.series").read_text())
    assert series["files"] == [{"name": "out_element_t0_125000000s.pvtu", "time": 0.125}]


def test_bad_extension(tmp_path):
    with pytest.raises(ValueError):
        DataWriter_t(tmp_path / "out.vtk")


CPP_WRITER = r"""
#include <cstdlib>
#include "ndg_data_writer.hpp"
int main(int argc, char **argv) {
    const bool ascii = argc > 2 && std::atoi(argv[2]);
    SEM::Helpers::DataWriter_t w(argv[1], ascii);
    const std::size_t n = 3, K = 2;
    std::vector<double> x(K * n * n), y(K * n * n), f(K * n * n);
    for (std::size_t i = 0; i < x.size(); ++i) {
        x[i] = 0.25 * (double)i;
        y[i] = 1.0 - 0.125 * (double)i;
        f[i] = 1.0 / (1.0 + (double)i);
    }
    w.write_data(n, K, 0.125, x, y, f, y, x);
    std::vector<int> N{4, 6}, ints{1, 0};
    std::vector<double> e{1e-3, 2e-4};
    w.write_complete_data(n, K, 0.25, 0, x, y, f, y, x, N, f, y, x, e, e, e, e, e, e, ints, ints, ints, f, y, x, ints,
                          ints);
    return 0;
}
"""


@pytest.mark.parametrize("ascii_mode", [0, 1])
def test_cpp_writer_matches_python_writer(tmp_path, ascii_mode):
    """The header-only C++ DataWriter_t (include/ndg_data_writer.hpp) writes the same arrays, file
    names and series as the Python one."""
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src, exe = tmp_path / "w.cpp", tmp_path / "w"
    src.write_text(CPP_WRITER)
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + os.path.join(root, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    subprocess.run([str(exe), str(tmp_path / "cpp" / "o.pvtu"), str(ascii_mode)], check=True)
    K, n = 2, 3
    i = np.arange(K * n * n, dtype=np.float64)
    x, y, f = 0.25 * i, 1.0 - 0.125 * i, 1.0 / (1.0 + i)
    e = np.array([1e-3, 2e-4])
    ints = np.array([1, 0])
    w = DataWriter_t(tmp_path / "py" / "o.pvtu", fmt="ascii" if ascii_mode else "binary")
    w.write_data(n, K, 0.125, x, y, f, y, x)
    w.write_complete_data(n, K, 0.25, 0, x, y, f, y, x, np.array([4, 6]), f, y, x, e, e, e, e, e, e, ints, ints, ints,
                          f, y, x, ints, ints)
    for name in ("o_t0_125000000s.pvtu", "o_t0_250000000s.pvtu", "o_element_t0_250000000s.pvtu"):
        a, b = read_vtu(tmp_path / "cpp" / name), read_vtu(tmp_path / "py" / name)
        assert np.array_equal(a["points"], b["points"]) and np.array_equal(a["connectivity"], b["connectivity"])
        for kind in ("point_data", "cell_data"):
            assert list(a[kind]) == list(b[kind])
            for key in a[kind]:
                assert a[kind][key].dtype == b[kind][key].dtype and np.array_equal(a[kind][key], b[kind][key]), key
    for series in ("o.pvtu.series", "o_element.pvtu.series"):
        assert json.loads((tmp_path / "cpp" / series).read_text()) == json.loads((tmp_path / "py" / series).read_text())
