"""Discrete adjoint of the RK3 stage (no reference counterpart), through the C ABI.

Checks, from the operator up:
- dot-product test of the transposed right-hand side: <y, G x> = <G^T y, x> (1e-12 relative);
- misfit J and its gradient against a CPU tangent-linear model built on the oracle's right-hand
  side (J to 1e-11, directional derivatives to 1e-9), including the wave-speed derivative and a
  last checkpoint segment shorter than the others;
- the 128 x 128, N = 8, 2000-step configuration with checkpoints every 50 steps: gradient against
  central finite differences (eps = 1e-6) along 8 seeded directions, to 1e-6 relative;
- the torch.autograd.Function wrapper's backward equals the adjoint gradient.
"""
import numpy as np
import pytest

import ndg_cuda_amd as nd
from ndg_cuda_amd import adjoint as A
from oracle import oracle as O

pytestmark = pytest.mark.gpu

AM = (0.0, -5.0 / 9.0, -153.0 / 128.0)
GM = (1.0 / 3.0, 15.0 / 16.0, 8.0 / 15.0)


def say(*args):
    print("[gpu]", *args, flush=True)


@pytest.fixture(scope="module")
def tables():
    return nd.NDG_t(12, 5), O.Tables(12, 5)


def periodic_mesh(tables, res, N, jitter, seed):
    ndg, ot = tables
    desc = nd.square_mesh(res, x_periodic=True, y_periodic=True, jitter=jitter, seed=seed)
    return desc, nd.Mesh2D_t(desc, N, N, 5, 0, 0, 0, 1e-6, 1e-14, 0.0, ndg), O.Mesh(ot, desc, N)


def lagrange(x, lam, xi):
    hit = np.nonzero(x == xi)[0]
    if hit.size:
        out = np.zeros_like(x)
        out[hit[0]] = 1.0
        return out
    t = lam / (xi - x)
    return t / t.sum()


def receiver_weights(ot, N, rec):
    x, lam = ot.get(O.O_NODES, N), ot.get(O.O_BARY, N)
    return [np.outer(lagrange(x, lam, a), lagrange(x, lam, b)).ravel() for a, b in rec.xi_eta]


def test_transpose_dot_product(tables):
    say("adjoint dot product")
    desc, gm, _ = periodic_mesh(tables, 8, 5, 0.2, 3)
    A.set_wave_speed(gm, 1.3)
    rng = np.random.default_rng(1)
    n = gm.n_dofs_per_field_
    x = [rng.normal(size=n) for _ in range(3)]
    y = [rng.normal(size=n) for _ in range(3)]
    gm.set_solution(*x)
    Gx = gm.compute_rhs(0.0)
    Gty = A.apply_rhs_transpose(gm, *y)
    lhs = sum(float(a @ b) for a, b in zip(y, Gx))
    rhs = sum(float(a @ b) for a, b in zip(Gty, x))
    scale = sum(float(np.abs(a) @ np.abs(b)) for a, b in zip(y, Gx))
    assert abs(lhs - rhs) <= 1e-12 * scale


@pytest.mark.parametrize("N", range(1, 17))
def test_transpose_dot_product_every_order(N):
    """<y, G x> = <G^T y, x> for every compiled order (one transposed-stage instantiation each) on
    a jittered periodic 4 x 4 mesh, to 1e-12 of sum |y| |G x|."""
    say("adjoint dot product, N", N)
    desc, gm, _ = periodic_mesh((nd.NDG_t(N, 5), O.Tables(N, 5)), 4, N, 0.2, 40 + N)
    A.set_wave_speed(gm, 0.9)
    rng = np.random.default_rng(N)
    n = gm.n_dofs_per_field_
    x = [rng.normal(size=n) for _ in range(3)]
    y = [rng.normal(size=n) for _ in range(3)]
    gm.set_solution(*x)
    Gx = gm.compute_rhs(0.0)
    Gty = A.apply_rhs_transpose(gm, *y)
    lhs = sum(float(a @ b) for a, b in zip(y, Gx))
    rhs = sum(float(a @ b) for a, b in zip(Gty, x))
    scale = sum(float(np.abs(a) @ np.abs(b)) for a, b in zip(y, Gx))
    assert abs(lhs - rhs) <= 1e-12 * scale


def tlm_run(om, ot, N, q0, dq0, c, dc, dt, n_steps, rec, data):
    """Forward run and tangent-linear model on the oracle's right-hand side: returns (J, dJ)."""
    om.set_wave_speed(c)

    def rhs(q):
        om.set_state(*q)
        return [np.array(a) for a in om.compute_rhs(0.0)]

    ne = (N + 1) ** 2
    w = receiver_weights(ot, N, rec)
    q = [a.copy() for a in q0]
    dq = [a.copy() for a in dq0]
    g = dg = None
    J = dJ = 0.0
    for n in range(n_steps):
        for s in range(3):
            G = rhs(q)
            dG = [a + (dc / c) * b for a, b in zip(rhs(dq), G)]
            g = G if s == 0 else [AM[s] * a + b for a, b in zip(g, G)]
            dg = dG if s == 0 else [AM[s] * a + b for a, b in zip(dg, dG)]
            q = [a + (GM[s] * dt) * b for a, b in zip(q, g)]
            dq = [a + (GM[s] * dt) * b for a, b in zip(dq, dg)]
        for r, e in enumerate(rec.element):
            pr = float(w[r] @ q[0][e * ne:(e + 1) * ne])
            dpr = float(w[r] @ dq[0][e * ne:(e + 1) * ne])
            res = pr - data[n, r]
            J += 0.5 * res * res
            dJ += res * dpr
    return J, dJ


def test_gradient_vs_cpu_tangent_linear(tables):
    say("adjoint vs CPU TLM, 8x8 N=4, 30 steps, checkpoints every 7")
    ndg, ot = tables
    N, c, n_steps = 4, 1.1, 30
    desc, gm, om = periodic_mesh(tables, 8, N, 0.15, 5)
    rec = A.Receivers.seeded(gm.n_elements_, 5, 11)
    pulses_true = [(0.2, -0.1, 0.25, 1.0), (-0.3, 0.3, 0.2, 0.6)]
    gm.initial_conditions(pulses_true)
    A.set_wave_speed(gm, 1.05)
    dt = 0.5 * 0.25 / (N * N)
    data = A.record_receivers(gm, rec, dt, n_steps)
    gm.initial_conditions([(0.15, -0.1, 0.25, 0.9)])
    A.set_wave_speed(gm, c)
    q0 = [np.array(a) for a in gm.get_solution()]
    J, gp, gu, gv, gc = A.adjoint_gradient(gm, rec, dt, n_steps, 7, data)
    # the mesh state is back to the initial fields
    for a, b in zip(gm.get_solution(), q0):
        assert np.array_equal(a, b)
    rng = np.random.default_rng(2)
    for k in range(3):
        d = [rng.normal(size=q0[0].size) for _ in range(3)]
        dc = float(rng.normal()) if k > 0 else 0.0
        J_cpu, dJ_cpu = tlm_run(om, ot, N, q0, d, c, dc, dt, n_steps, rec, data)
        assert abs(J - J_cpu) <= 1e-11 * abs(J_cpu)
        dJ = float(gp @ d[0] + gu @ d[1] + gv @ d[2]) + gc * dc
        say("direction", k, dJ, dJ_cpu)
        assert abs(dJ - dJ_cpu) <= 1e-9 * max(abs(dJ_cpu), 1e-300)


def adjoint_128_setup(tables):
    N, res, n_steps, interval = 8, 128, 2000, 50
    desc, gm, _ = periodic_mesh(tables, res, N, 0.0, 1)
    rec = A.Receivers.seeded(gm.n_elements_, 32, 2024)
    dt = 0.5 * (2.0 / res) / (N * N)  # frozen: the c = 1 CFL step
    gm.initial_conditions([(0.1, -0.2, 0.2, 1.0), (-0.4, 0.3, 0.2, 0.5)])
    A.set_wave_speed(gm, 1.05)
    data = A.record_receivers(gm, rec, dt, n_steps)
    gm.initial_conditions([(0.1, -0.2, 0.2, 1.0)])
    A.set_wave_speed(gm, 1.0)
    return gm, rec, dt, n_steps, interval, data


def test_adjoint_128x128_gradient_vs_finite_differences(tables):
    say("adjoint 128x128 N=8 periodic, 2000 steps, checkpoints 50, FD along 8 directions")
    gm, rec, dt, n_steps, interval, data = adjoint_128_setup(tables)
    q0 = [np.array(a) for a in gm.get_solution()]
    J, gp, gu, gv, gc = A.adjoint_gradient(gm, rec, dt, n_steps, interval, data)
    say("J", J, "dJ/dc", gc)

    def misfit(q, c):
        gm.set_solution(*q)
        A.set_wave_speed(gm, c)
        p = A.record_receivers(gm, rec, dt, n_steps)
        return 0.5 * float(((p - data) ** 2).sum())

    assert abs(misfit(q0, 1.0) - J) <= 1e-12 * J
    rng = np.random.default_rng(8)
    eps = 1e-6
    scale = max(np.abs(a).max() for a in q0)
    for k in range(8):
        d = [rng.normal(size=q0[0].size) * scale for _ in range(3)]
        dc = float(rng.normal())
        Jp = misfit([a + eps * b for a, b in zip(q0, d)], 1.0 + eps * dc)
        Jm = misfit([a - eps * b for a, b in zip(q0, d)], 1.0 - eps * dc)
        fd = (Jp - Jm) / (2 * eps)
        ad = float(gp @ d[0] + gu @ d[1] + gv @ d[2]) + gc * dc
        say("direction", k, "adjoint", ad, "fd", fd)
        assert abs(ad - fd) <= 1e-6 * abs(fd)


def test_store_budget_gives_the_same_gradient(tables):
    """Keeping the stage states of more segments in the forward pass (fewer recomputed) changes
    nothing in the result, including a last segment shorter than the others."""
    say("adjoint store budget")
    desc, gm, _ = periodic_mesh(tables, 8, 5, 0.1, 4)
    rec = A.Receivers.seeded(gm.n_elements_, 6, 3)
    dt, n_steps, interval = 0.5 * 0.25 / 25, 23, 5
    gm.initial_conditions([(0.1, 0.0, 0.3, 1.0)])
    data = A.record_receivers(gm, rec, dt, n_steps) * 0.95
    gm.initial_conditions([(0.12, 0.0, 0.3, 1.0)])
    A.set_wave_speed(gm, 1.02)
    ref = A.adjoint_gradient(gm, rec, dt, n_steps, interval, data)
    seg_bytes = 3 * interval * 3 * gm.n_dofs_per_field_ * 8
    for extra in (1, 2, 10):
        A.set_store_budget(gm, extra * seg_bytes)
        got = A.adjoint_gradient(gm, rec, dt, n_steps, interval, data)
        assert got[0] == ref[0] and got[4] == ref[4]
        for a, b in zip(got[1:4], ref[1:4]):
            assert np.array_equal(a, b)
    A.set_store_budget(gm, 0)


@pytest.mark.parametrize("interval", [7, 30])
def test_stage_records_c_derivative(tables, interval):
    """dJ/dc from the stage right-hand sides recorded in the forward pass (no checkpoints, no
    recompute) against the recompute form; the initial-field gradients are the same bit for bit."""
    say("adjoint c-derivative from stage records")
    desc, gm, _ = periodic_mesh(tables, 8, 5, 0.15, 6)
    rec = A.Receivers.seeded(gm.n_elements_, 7, 9)
    dt, n_steps = 0.5 * 0.25 / 25, 30
    gm.initial_conditions([(0.1, 0.05, 0.25, 1.0)])
    A.set_wave_speed(gm, 1.1)
    data = A.record_receivers(gm, rec, dt, n_steps)
    gm.initial_conditions([(0.12, 0.0, 0.25, 0.9)])
    A.set_wave_speed(gm, 0.95)
    ref = A.adjoint_gradient(gm, rec, dt, n_steps, interval, data)
    A.set_c_derivative(gm, "stage_records")
    got = A.adjoint_gradient(gm, rec, dt, n_steps, interval, data)
    A.set_c_derivative(gm, "recompute")
    say("dJ/dc", ref[4], got[4])
    assert got[0] == ref[0]
    for a, b in zip(got[1:4], ref[1:4]):
        assert np.array_equal(a, b)
    assert abs(got[4] - ref[4]) <= 1e-10 * abs(ref[4])


def test_autograd_function(tables):
    torch = pytest.importorskip("torch")
    say("autograd wrapper")
    desc, gm, _ = periodic_mesh(tables, 8, 4, 0.1, 2)
    rec = A.Receivers.seeded(gm.n_elements_, 4, 5)
    dt, n_steps = 0.5 * 0.25 / 16, 12
    gm.initial_conditions([(0.0, 0.1, 0.3, 1.0)])
    data = A.record_receivers(gm, rec, dt, n_steps) * 0.9
    gm.initial_conditions([(0.05, 0.1, 0.3, 1.0)])
    q0 = [torch.tensor(np.array(a), requires_grad=True) for a in gm.get_solution()]
    c = torch.tensor(1.0, dtype=torch.float64, requires_grad=True)
    f = A.misfit_function()
    J = f.apply(q0[0], q0[1], q0[2], c, gm, rec, dt, n_steps, 5, data)
    (2.0 * J).backward()
    gm.set_solution(*(t.detach().numpy() for t in q0))
    A.set_wave_speed(gm, 1.0)
    Jr, gp, gu, gv, gc = A.adjoint_gradient(gm, rec, dt, n_steps, 5, data)
    assert float(J.detach()) == Jr
    for t, g in zip(q0, (gp, gu, gv)):
        assert np.array_equal(t.grad.numpy(), 2.0 * g)
    assert float(c.grad) == 2.0 * gc
