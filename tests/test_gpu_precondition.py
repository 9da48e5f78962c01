"""GPU parity of Solver2D_t::pre_condition and pre_condition_iterative (include/solvers/
Solver2D_t.tpp:138-336) through the C ABI against the oracle's restatement: element counts,
orders and the re-applied initial state after the adaptations, then the state after a solve.
Bar: relative max error <= 1e-12 per field; counts and orders equal."""
import numpy as np
import pytest

import ndg_cuda_amd as nd
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
PULSES = [(0.1, -0.2, 0.08, 1.0), (-0.45, 0.3, 0.12, 0.5)]


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("iterative", [False, True])
@pytest.mark.parametrize("split", [0, 1])
def test_pre_condition_matches_oracle(iterative, split):
    print("[gpu] pre_condition iterative", iterative, "split", split, flush=True)
    desc = nd.square_mesh(8, bc=(nd.BC_WALL,) * 4, jitter=0.1, seed=3)
    gm = nd.Mesh2D_t(desc, 3, 9, 5, split, 0, 0, 1e-4, 1e-14, 0.0, nd.NDG_t(9, 5))
    om = O.Mesh(O.Tables(9, 5), desc, 3, tol_min=1e-4, tol_max=1e-14)
    gm.initial_conditions(PULSES)
    om.initial_conditions(PULSES)
    solver = nd.Solver2D_t(0.5, [0.05])
    if iterative:
        solver.pre_condition_iterative(None, gm, 3, 2)
    else:
        solver.pre_condition(None, gm, 3, 2)
    om.pre_condition(0.5, 3, 2, 9, split, PULSES, iterative=iterative)
    assert gm.n_elements_ == om.K and gm.n_faces_ == om.n_faces
    assert np.array_equal(gm.orders_, om.N)
    assert int(gm.orders_.max()) > 3
    if split:
        assert om.K > 64
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL
    solver.solve(None, gm)
    n, t = om.solve(0.5, 0.05)
    assert solver.n_steps_ == n and solver.time_ == t
    for g, r in zip(gm.get_solution(), om.get_state()):
        assert rel(g, r) < TOL


def test_pre_condition_without_initial_condition_is_refused():
    desc = nd.square_mesh(2, bc=(nd.BC_WALL,) * 4)
    gm = nd.Mesh2D_t(desc, 3, 5, 5, 0, 0, 0, 1e-4, 1e-14, 0.0, nd.NDG_t(5, 5))
    with pytest.raises(nd.NDGError):
        nd.Solver2D_t(0.5, [0.1]).pre_condition(None, gm, 1, 1)
    gm.initial_conditions(PULSES)
    with pytest.raises(nd.NDGError):
        nd.Solver2D_t(0.5, [0.1]).pre_condition(None, gm, 1, 0)
