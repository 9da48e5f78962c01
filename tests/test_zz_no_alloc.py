"""The allocation-free entry point (ndg_record_receivers_noalloc, no reference counterpart): called
through the C ABI, it gives the same records and the same final state as ndg_record_receivers,
bit for bit, on two calls in a row (the second with fewer receivers and steps, so the fixed
scratch array holds the first call's values where the second does not write), and refuses a
request larger than its scratch without touching the mesh."""
import numpy as np
import pytest

import ndg_cuda_amd as nd
from ndg_cuda_amd import adjoint as A

pytestmark = pytest.mark.gpu


def say(*args):
    print("[gpu]", *args, flush=True)


def make_mesh(ndg):
    desc = nd.square_mesh(16, x_periodic=True, y_periodic=True, jitter=0.1, seed=4)
    m = nd.Mesh2D_t(desc, 6, 6, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg)
    m.initial_conditions([(0.1, -0.2, 0.2, 1.0), (-0.4, 0.3, 0.15, 0.5)])
    return m


def test_noalloc_records_match_the_allocating_entry_point():
    say("no-alloc receivers: two calls, second smaller")
    ndg = nd.NDG_t(6, 5)
    ref, got = make_mesh(ndg), make_mesh(ndg)
    dt = nd.Solver2D_t(0.5, [0.0]).get_delta_t(ref)
    for n_rec, n_steps, seed in ((32, 40, 1), (5, 11, 2)):
        rec = A.Receivers.seeded(ref.n_elements_, n_rec, seed)
        a = A.record_receivers(ref, rec, dt, n_steps)
        b = A.record_receivers_noalloc(got, rec, dt, n_steps)
        assert b.shape == (n_steps, n_rec)
        assert np.array_equal(a, b), (n_rec, n_steps, np.max(np.abs(a - b)))
        for x, y in zip(ref.get_solution(), got.get_solution()):
            assert np.array_equal(x, y)


def test_noalloc_refuses_requests_beyond_its_scratch():
    say("no-alloc receivers: size limit")
    ndg = nd.NDG_t(6, 5)
    m = make_mesh(ndg)
    before = m.get_solution()
    rec = A.Receivers.seeded(m.n_elements_, 64, 3)
    with pytest.raises(nd.NDGError, match="exceed its fixed scratch"):
        A.record_receivers_noalloc(m, rec, 1e-3, 8192)   # 64 * 8192 * 8 B = 4 MiB of records alone
    for x, y in zip(before, m.get_solution()):
        assert np.array_equal(x, y)
