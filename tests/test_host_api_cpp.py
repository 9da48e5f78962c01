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
    steps = int(r.stdout.split()[1])
    got = np.fromfile(tmp_path / "o.bin", dtype=np.float64).reshape(3, -1)
    desc = nd.square_mesh(res, x_periodic=True, y_periodic=True)
    om = O.Mesh(O.Tables(N, 5), desc, N)
    om.initial_conditions([(0.1, -0.2, 0.2, 1.0)])
    n_ref, _ = om.solve(0.5, t_end)
    assert steps == n_ref == 289
    for g, ref in zip(got, om.get_state()):
        assert np.max(np.abs(g - ref)) / np.max(np.abs(ref)) < 1e-12
