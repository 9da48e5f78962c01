"""The oracle's stage-level functions pinned by mathematics (CPU only).

The reference's stage phases (Mesh2D_t.cpp, Solver2D_t.cpp) cannot be compiled in this image, so
the oracle's restatement of them is checked against facts the mathematics fixes: a library
Gauss rule, the reference's own N=16 known answers, closed forms of the upwind flux and of the
wall ghost state, exactness of the discrete operator on polynomials, free-stream preservation,
conservation, linearity, the RK3 power series, exact solutions and spectral convergence.
"""
import numpy as np
import pytest

import exact
from oracle import oracle as O
import ndg_cuda_amd as nd


# ---------------------------------------------------------------------------------------------
# Tables (A1-A6)

# tests/polynomials_test.cu:187-223 of the reference (Legendre-Gauss, N = 16)
KAT_NODES_16 = [-0.9905754753144173356754, -0.9506755217687677612227, -0.880239153726985902123,
                -0.7815140038968014069252, -0.6576711592166907658503, -0.5126905370864769678863,
                -0.3512317634538763152972, -0.1784841814958478558507, 0.0]
KAT_WEIGHTS_16 = [0.0241483028685479319601, 0.0554595293739872011294, 0.0850361483171791808835,
                  0.111883847193403971095, 0.1351363684685254732863, 0.1540457610768102880814,
                  0.16800410215645004451, 0.1765627053669926463253, 0.1794464703562065254583]


def test_nodes_weights_vs_library_and_kat(oracle_tables):
    for N in range(0, 17):
        x = oracle_tables.get(O.O_NODES, N)
        w = oracle_tables.get(O.O_WEIGHTS, N)
        xr, wr = np.polynomial.legendre.leggauss(N + 1)
        assert np.max(np.abs(x - xr)) < 1e-14, N
        assert np.max(np.abs(w - wr)) < 1e-14, N
        assert np.array_equal(x, -x[::-1]) and np.array_equal(w, w[::-1]), N  # exact symmetry
    x16, w16 = oracle_tables.get(O.O_NODES, 16), oracle_tables.get(O.O_WEIGHTS, 16)
    assert np.max(np.abs(x16[:9] - KAT_NODES_16)) < 2e-16
    assert np.max(np.abs(w16[:9] - KAT_WEIGHTS_16)) < 1e-15
    assert oracle_tables.get(O.O_NODES, 1)[1] == np.sqrt(1.0 / 3.0)
    x2 = oracle_tables.get(O.O_NODES, 2)
    assert x2[1] == 0.0 and abs(x2[2] - np.sqrt(0.6)) < 1e-16


@pytest.mark.parametrize("N", [1, 2, 4, 6, 8, 12, 16])
def test_table_identities(oracle_tables, N):
    nf = N + 1
    x = oracle_tables.get(O.O_NODES, N)
    w = oracle_tables.get(O.O_WEIGHTS, N)
    lm, lp = oracle_tables.get(O.O_LAG_LEFT, N), oracle_tables.get(O.O_LAG_RIGHT, N)
    D = oracle_tables.get(O.O_D, N).reshape(nf, nf)
    Dh = oracle_tables.get(O.O_DHAT, N).reshape(nf, nf)
    M = oracle_tables.get(O.O_POLY, N).reshape(nf, nf)
    assert abs(lm.sum() - 1) < 1e-14 and abs(lp.sum() - 1) < 1e-14
    for k in range(nf):  # exact for polynomials of degree <= N
        f = x ** k
        assert abs(lm @ f - (-1.0) ** k) < 1e-12
        assert abs(lp @ f - 1.0) < 1e-12
        df = k * x ** (k - 1) if k else np.zeros(nf)
        assert np.max(np.abs(D @ f - df)) < 1e-10 * max(1, k * k)
    assert np.max(np.abs(Dh - (-(D.T * w[None, :]) / w[:, None]))) < 1e-15 * np.abs(Dh).max()
    L = np.array([np.polynomial.legendre.legval(x, np.eye(nf)[k]) for k in range(nf)])  # L_k(x_j)
    assert np.max(np.abs(M @ L.T - np.eye(nf))) < 1e-12


# ---------------------------------------------------------------------------------------------
# Upwind flux (A13): characteristic closed form F = (f_n(qL) + f_n(qR))/2 - |A_n| (qR - qL)/2

def test_wave_flux_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(200):
        pL, uL, vL, pR, uR, vR = rng.normal(size=6)
        th = rng.uniform(0, 2 * np.pi)
        nx, ny = np.cos(th), np.sin(th)
        unL, unR = uL * nx + vL * ny, uR * nx + vR * ny
        fL = np.array([unL, pL * nx, pL * ny])
        fR = np.array([unR, pR * nx, pR * ny])
        diss = np.array([pR - pL, (unR - unL) * nx, (unR - unL) * ny])
        ref = 0.5 * (fL + fR) - 0.5 * diss
        got = O.wave_flux([pL, uL, vL, pR, uR, vR, nx, ny])
        assert np.max(np.abs(got - ref)) < 1e-14


# ---------------------------------------------------------------------------------------------
# Meshes and helpers

def make(res, N, jitter=0.0, periodic=True, bc=nd.BC_WALL, quad_N=None, seed=3, tables=None):
    desc = nd.square_mesh(res, bc=(bc,) * 4, x_periodic=periodic, y_periodic=periodic, jitter=jitter, seed=seed)
    if quad_N is not None:
        desc.quad_N = np.asarray(quad_N, dtype=np.int32)
    T = tables or O.Tables(16, 5)
    m = O.Mesh(T, desc, N)
    return desc, T, m


def packed_nodes(desc, T, Ns):
    """Physical coordinates and per-element offsets for possibly mixed orders."""
    Xs, Ys = [], []
    for e, N in enumerate(Ns):
        sub = nd.MeshDescription(desc.xy, desc.quads[e:e + 1], desc.bars, desc.bar_type, desc.bar_donor)
        X, Y = exact.node_coordinates(sub, T.get(O.O_NODES, int(N)))
        Xs.append(X)
        Ys.append(Y)
    return np.concatenate(Xs), np.concatenate(Ys)


def quadrature_weights(desc, T, Ns):
    """w_i w_j J_ij per node (packed), J from the bilinear map."""
    out = []
    for e, N in enumerate(Ns):
        P = desc.xy[desc.quads[e]].ravel()
        geo = O.element_geometry(T, int(N), P)
        nf = N + 1
        J = geo[4 * nf * nf:5 * nf * nf]
        w = T.get(O.O_WEIGHTS, int(N))
        out.append(np.outer(w, w).ravel() * J)
    return np.concatenate(out)


def interior_mask(desc, Ns, res):
    """Nodes of elements that do not touch the domain boundary (packed)."""
    out = []
    xy = desc.xy
    lo, hi = xy.min(axis=0), xy.max(axis=0)
    for e, N in enumerate(Ns):
        P = xy[desc.quads[e]]
        touch = np.any(np.isclose(P, lo[None, :]) | np.isclose(P, hi[None, :]))
        out.append(np.full((N + 1) ** 2, not touch))
    return np.concatenate(out)


# ---------------------------------------------------------------------------------------------
# Volume + lift (A15), traces (A10), face projection (A12/A14)

@pytest.mark.parametrize("jitter", [0.0, 0.2])
def test_free_stream_preservation(jitter):
    desc, T, m = make(8, 6, jitter=jitter, periodic=True)
    n = m.n
    m.set_state(np.full(n, 0.75), np.zeros(n), np.zeros(n))
    Gp, Gu, Gv = m.compute_rhs(0.0)
    assert max(np.abs(Gp).max(), np.abs(Gu).max(), np.abs(Gv).max()) < 1e-11


@pytest.mark.parametrize("mixed", [False, True])
def test_polynomial_exactness_interior(mixed):
    """Continuous polynomial fields of degree <= N: interior elements see continuous states, the
    upwind flux reduces to the exact flux, and G equals -div f exactly (up to roundoff) on affine
    elements. With mixed orders this runs the non-conforming face interpolation and the L2
    mortar projection of the reference (Mesh2D_t.cpp:3830-3889, 3937-4003)."""
    res = 8
    rng = np.random.default_rng(5)
    quad_N = rng.choice([4, 6, 8], size=res * res) if mixed else None
    desc, T, m = make(res, 4 if mixed else 6, periodic=True, quad_N=quad_N)
    Ns = m.N
    X, Y = packed_nodes(desc, T, Ns)
    p = X**2 + 3 * X * Y - Y**3 + 0.5
    u = Y**2 - X * Y + 2 * X**3
    v = X * Y + X**2 * Y
    m.set_state(p, u, v)
    Gp, Gu, Gv = m.compute_rhs(0.0)
    ex_p = -((-Y + 6 * X**2) + (X + X**2))       # -(u_x + v_y)
    ex_u = -(2 * X + 3 * Y)                      # -p_x
    ex_v = -(3 * X - 3 * Y**2)                   # -p_y
    mask = interior_mask(desc, Ns, res)
    for got, ref in ((Gp, ex_p), (Gu, ex_u), (Gv, ex_v)):
        assert np.max(np.abs(got[mask] - ref[mask])) < 1e-10


@pytest.mark.parametrize("mixed,jitter", [(False, 0.2), (True, 0.0)])
def test_conservation_periodic(mixed, jitter):
    """sum_e sum_ij w_i w_j J_ij G_p = 0: D^ has zero weighted column sums, lift sums to the face
    fluxes, and each face contributes with opposite signs to its two elements."""
    res = 8
    rng = np.random.default_rng(7)
    quad_N = rng.choice([4, 6, 8], size=res * res) if mixed else None
    desc, T, m = make(res, 4 if mixed else 6, jitter=jitter, periodic=True, quad_N=quad_N)
    n = m.n
    m.set_state(rng.normal(size=n), rng.normal(size=n), rng.normal(size=n))
    Gp, Gu, Gv = m.compute_rhs(0.0)
    W = quadrature_weights(desc, T, m.N)
    scale = np.sum(W * np.abs(Gp))
    assert abs(np.sum(W * Gp)) < 1e-12 * scale


def test_linearity_walls_jittered():
    desc, T, m = make(8, 5, jitter=0.25, periodic=False, bc=nd.BC_WALL)
    rng = np.random.default_rng(11)
    n = m.n
    q1 = [rng.normal(size=n) for _ in range(3)]
    q2 = [rng.normal(size=n) for _ in range(3)]
    a, b = 0.7, -1.3
    m.set_state(*q1)
    G1 = m.compute_rhs(0.0)
    m.set_state(*q2)
    G2 = m.compute_rhs(0.0)
    m.set_state(*[a * x + b * y for x, y in zip(q1, q2)])
    G12 = m.compute_rhs(0.0)
    for g12, g1, g2 in zip(G12, G1, G2):
        assert np.max(np.abs(g12 - (a * g1 + b * g2))) < 1e-12 * np.abs(g12).max()


def test_wall_flux_closed_form():
    """Reference wall ghost (F4 quirk): p_R = s p_L, u_R.n = -s u_L.n, u_R.t = s u_L.t with s the
    ghost scaling factor (half the edge length), so w_R = s w_L and the wall flux is
    F_p = (1 - s) w_L / 2, F_u,v = n (1 + s) w_L / 2, times the element edge Jacobian."""
    res, N = 4, 5
    desc, T, m = make(res, N, jitter=0.0, periodic=False, bc=nd.BC_WALL)
    rng = np.random.default_rng(2)
    n = m.n
    m.set_state(rng.normal(size=n), rng.normal(size=n), rng.normal(size=n))
    m.compute_rhs(0.0)
    tr = m.get_traces().reshape(m.K, 4, 3, N + 1)
    fl = m.get_flux_traces().reshape(m.K, 4, 3, N + 1)
    xy = desc.xy
    checked = 0
    for e in range(m.K):
        P = xy[desc.quads[e]]
        for s in range(4):
            a, b = P[s], P[(s + 1) % 4]
            mid = 0.5 * (a + b)
            if not (np.isclose(abs(mid[0]), 1.0) or np.isclose(abs(mid[1]), 1.0)):
                continue
            t = (b - a) / np.linalg.norm(b - a)
            nrm = np.array([t[1], -t[0]])
            sf = np.linalg.norm(b - a) / 2
            s_g = sf                       # ghost scaling factor: same edge, half its length
            pL, uL, vL = tr[e, s]
            wL = pL + uL * nrm[0] + vL * nrm[1]
            exp_p = (1 - s_g) * wL / 2 * sf
            exp_u = nrm[0] * (1 + s_g) * wL / 2 * sf
            exp_v = nrm[1] * (1 + s_g) * wL / 2 * sf
            assert np.max(np.abs(fl[e, s, 0] - exp_p)) < 1e-14
            assert np.max(np.abs(fl[e, s, 1] - exp_u)) < 1e-14
            assert np.max(np.abs(fl[e, s, 2] - exp_v)) < 1e-14
            checked += 1
    assert checked == 4 * res


def test_outflow_flux_closed_form():
    """Reference outflow ghost (Mesh2D_t.cpp:4601-4710; the multi-face branch always runs): the
    element trace projected onto the ghost with the weight ratio times the ghost scaling factor,
    i.e. q_R = s q_L for all three fields on a conforming boundary edge (s = half the edge
    length). The upwind flux is then F_p = (1 + s) u_n/2 + (1 - s) p/2 and
    F_(u,v) = n ((1 + s) p/2 + (1 - s) u_n/2), times the element edge Jacobian."""
    res, N = 4, 6
    desc, T, m = make(res, N, jitter=0.0, periodic=False, bc=nd.BC_OUTFLOW)
    rng = np.random.default_rng(12)
    n = m.n
    m.set_state(rng.normal(size=n), rng.normal(size=n), rng.normal(size=n))
    m.compute_rhs(0.0)
    tr = m.get_traces().reshape(m.K, 4, 3, N + 1)
    fl = m.get_flux_traces().reshape(m.K, 4, 3, N + 1)
    xy = desc.xy
    checked = 0
    for e in range(m.K):
        P = xy[desc.quads[e]]
        for s in range(4):
            a, b = P[s], P[(s + 1) % 4]
            mid = 0.5 * (a + b)
            if not (np.isclose(abs(mid[0]), 1.0) or np.isclose(abs(mid[1]), 1.0)):
                continue
            t = (b - a) / np.linalg.norm(b - a)
            nrm = np.array([t[1], -t[0]])
            sf = np.linalg.norm(b - a) / 2
            pL, uL, vL = tr[e, s]
            un = uL * nrm[0] + vL * nrm[1]
            exp_p = ((1 + sf) * un / 2 + (1 - sf) * pL / 2) * sf
            exp_n = ((1 + sf) * pL / 2 + (1 - sf) * un / 2) * sf
            assert np.max(np.abs(fl[e, s, 0] - exp_p)) < 1e-14
            assert np.max(np.abs(fl[e, s, 1] - nrm[0] * exp_n)) < 1e-14
            assert np.max(np.abs(fl[e, s, 2] - nrm[1] * exp_n)) < 1e-14
            checked += 1
    assert checked == 4 * res


# ---------------------------------------------------------------------------------------------
# Time stepping (A9, A16, A17)

def test_rk3_is_the_third_order_power_series():
    """For the linear autonomous (periodic) system q' = L q one low-storage RK3 step equals
    q + dt L q + dt^2/2 L^2 q + dt^3/6 L^3 q: pins the Williamson coefficients and stage order."""
    desc, T, m = make(4, 4, jitter=0.15, periodic=True)
    rng = np.random.default_rng(3)
    n = m.n
    q0 = [rng.normal(size=n) for _ in range(3)]

    def L(q):
        m.set_state(*q)
        return m.compute_rhs(0.0)

    L1 = L(q0)
    L2 = L(L1)
    L3 = L(L2)
    dt = 0.7 * m.delta_t(0.5)
    series = [a + dt * b + dt**2 / 2 * c + dt**3 / 6 * d for a, b, c, d in zip(q0, L1, L2, L3)]
    m.set_state(*q0)
    m.step(0.0, dt)
    got = m.get_state()
    for g, s in zip(got, series):
        assert np.max(np.abs(g - s)) < 1e-12 * np.abs(s).max()


def test_delta_t_closed_form():
    desc, T, m = make(16, 6, periodic=True)
    assert m.delta_t(0.5) == 0.5 * 0.125 / 36      # 16x16 on [-1,1]^2, N = 6: 1/576
    n, tf = m.solve(0.5, 0.01)
    assert tf == 0.01 and n == 6


@pytest.mark.parametrize("N,tol", [(4, 2e-5), (6, 5e-7), (8, 1e-7)])
def test_periodic_pulse_exact_solution(N, tol):
    """The 16x16 periodic geometry ( on [-1,1]^2, CFL 0.5) with the periodised pulse, against
    the Fourier-series exact solution; spectral convergence in N."""
    desc, T, m = make(16, N, periodic=True)
    base = [(0.1, -0.2, 0.2, 1.0)]
    m.initial_conditions(exact.periodic_images(*base[0]))
    t_end = 0.25
    _, tf = m.solve(0.5, t_end)
    assert tf == t_end
    X, Y = packed_nodes(desc, T, m.N)
    pe, ue, ve = exact.periodic_pulse(X, Y, tf, base, n_modes=24)
    p, u, v = m.get_state()
    assert max(np.abs(p - pe).max(), np.abs(u - ue).max(), np.abs(v - ve).max()) < tol


def test_periodised_pulses_initial_condition():
    """The periodised-pulse initial condition (each pulse and its 8 nearest images, period 2) equals
    the Fourier series of the periodised pulses at t = 0 and the 5 x 5 image sum given as plain
    pulses; at N = 12 the solution then follows the exact one to 5e-8."""
    pulses = np.array([[0.1, -0.2, 0.2, 1.0], [-0.6, 0.7, 0.15, 0.5]])
    desc, T, m = make(8, 12, periodic=True)
    m.initial_conditions(pulses, period=(2.0, 2.0))
    X, Y = packed_nodes(desc, T, m.N)
    pe, ue, ve = exact.periodic_pulse(X, Y, 0.0, pulses, n_modes=30)
    p, u, v = m.get_state()
    assert np.abs(p - pe).max() < 1e-13 and not u.any() and not v.any()
    _, _, m2 = make(8, 12, periodic=True)
    m2.initial_conditions(np.concatenate([exact.periodic_images(*q) for q in pulses]))
    assert np.abs(m2.get_state()[0] - p).max() < 1e-15
    _, tf = m.solve(0.5, 0.1)
    pe, ue, ve = exact.periodic_pulse(X, Y, tf, pulses, n_modes=30)
    p, u, v = m.get_state()
    assert max(np.abs(p - pe).max(), np.abs(u - ue).max(), np.abs(v - ve).max()) < 5e-8


@pytest.mark.parametrize("c", [1.05, 0.8])
def test_wave_speed_delta_t_closed_form(c):
    """Wave speed c: the CFL step is the c = 1 step divided by c (Solver2D_t.cpp:19-29 with the
    constant c of constants.h:15 made a runtime value)."""
    desc, T, m = make(16, 6, periodic=True)
    m.set_wave_speed(c)
    assert m.delta_t(0.5) == 0.5 * 0.125 / (36 * 1.0 * c)


@pytest.mark.parametrize("c", [1.05, 0.8])
def test_wave_speed_operator_scales(c):
    """The speed-c operator is c times the c = 1 operator on any state (jittered walls, so the
    wall ghosts are included)."""
    desc, T, m = make(8, 5, jitter=0.2, periodic=False, bc=nd.BC_WALL)
    rng = np.random.default_rng(31)
    n = m.n
    q = [rng.normal(size=n) for _ in range(3)]
    m.set_state(*q)
    G1 = m.compute_rhs(0.0)
    m.set_wave_speed(c)
    Gc = m.compute_rhs(0.0)
    for a, b in zip(Gc, G1):
        assert np.max(np.abs(a - c * b)) < 1e-13 * np.abs(b).max()


@pytest.mark.parametrize("c", [1.05, 0.8])
def test_wave_speed_periodic_pulse_exact_solution(c):
    """16 x 16 periodic mesh at wave speed c against the Fourier-series solution whose modes turn
    at c|k|; spectral convergence over N = 4, 6, 8 with the c = 1 bounds."""
    errs = []
    for N, tol in ((4, 2e-5), (6, 5e-7), (8, 1e-7)):
        desc, T, m = make(16, N, periodic=True)
        m.set_wave_speed(c)
        base = [(0.1, -0.2, 0.2, 1.0)]
        m.initial_conditions(exact.periodic_images(*base[0]))
        _, tf = m.solve(0.5, 0.25)
        assert tf == 0.25
        X, Y = packed_nodes(desc, T, m.N)
        pe, ue, ve = exact.periodic_pulse(X, Y, tf, base, n_modes=24, c=c)
        p, u, v = m.get_state()
        err = max(np.abs(p - pe).max(), np.abs(u - ue).max(), np.abs(v - ve).max())
        assert err < tol, (c, N, err)
        # the c = 1 solution at the same time is far away: the speed is really c
        pw, _, _ = exact.periodic_pulse(X, Y, tf, base, n_modes=24, c=1.0)
        assert np.abs(p - pw).max() > 100 * tol
        errs.append(err)
    assert errs[2] < errs[1] < errs[0]


@pytest.mark.parametrize("c", [1.05, 0.8])
def test_wave_speed_plane_wave_inflow_exact_solution(c):
    """The reference plane wave g moving at speed c, inflow ghosts at speed c on all sides."""
    errs = []
    for N in (4, 7):
        desc, T, m = make(8, N, jitter=0.1, periodic=False, bc=nd.BC_INFLOW)
        m.set_wave_speed(c)
        m.initial_conditions(None)
        _, tf = m.solve(0.5, 0.2)
        X, Y = packed_nodes(desc, T, m.N)
        pe, ue, ve = exact.plane_wave(X, Y, tf, c=c)
        p, u, v = m.get_state()
        errs.append(max(np.abs(p - pe).max(), np.abs(u - ue).max(), np.abs(v - ve).max()))
    assert errs[1] < 3e-4 and errs[1] < errs[0] / 30, errs


def test_plane_wave_inflow_exact_solution():
    """The reference's own exact solution g(x, t) with inflow ghosts on all four sides."""
    errs = []
    for N in (4, 7):
        desc, T, m = make(8, N, jitter=0.1, periodic=False, bc=nd.BC_INFLOW)
        m.initial_conditions(None)
        _, tf = m.solve(0.5, 0.2)
        X, Y = packed_nodes(desc, T, m.N)
        pe, ue, ve = exact.plane_wave(X, Y, tf)
        p, u, v = m.get_state()
        errs.append(max(np.abs(p - pe).max(), np.abs(u - ue).max(), np.abs(v - ve).max()))
    assert errs[1] < 2e-4 and errs[1] < errs[0] / 40, errs   # measured: 7.2e-3 -> 9.0e-5


def test_energy_non_increasing_periodic():
    desc, T, m = make(8, 6, jitter=0.2, periodic=True)
    rng = np.random.default_rng(9)
    m.initial_conditions(exact.periodic_images(0.0, 0.1, 0.25, 1.0))
    W = quadrature_weights(desc, T, m.N)
    E = []
    t = 0.0
    for _ in range(6):
        p, u, v = m.get_state()
        E.append(0.5 * np.sum(W * (p * p + u * u + v * v)))
        t = m.run_steps(0.5, 5, t)
    assert all(E[k + 1] <= E[k] * (1 + 1e-13) for k in range(len(E) - 1)), E
    assert E[-1] < E[0]


def test_estimator_synthetic_decay(oracle_tables):
    """A field with modal coefficients a_k ~ C exp(-sigma k) along xi gives back sigma."""
    N, sigma, C = 8, 1.3, 0.7
    nf = N + 1
    x = oracle_tables.get(O.O_NODES, N)
    V = np.array([np.polynomial.legendre.legval(x, np.eye(nf)[k]) for k in range(nf)]).T  # (node, mode)
    a = C * np.exp(-sigma * np.arange(nf))
    f1 = V @ a
    p = np.outer(f1, np.ones(nf))          # varies along xi (first index) only
    out = O.element_estimate(oracle_tables, N, p.ravel())
    assert abs(out[1] - sigma) < 1e-10
    zero = O.element_estimate(oracle_tables, N, np.zeros(nf * nf))
    assert np.isnan(zero[1]) and zero[2] == 0.0   # log(0): sigma NaN, no refinement (F6)


# ---------------------------------------------------------------------------------------------
# Hanging (2:1) faces: the h-non-conforming mortar (A12/A14 with offset/scale and two faces per
# element side, Mesh2D_t.cpp:3803-4006)

def hanging_mesh(res, N, mixed, jitter, seed):
    import meshes
    desc = nd.square_mesh(res, x_periodic=True, y_periodic=True, jitter=jitter, seed=seed)
    rng = np.random.default_rng(seed)
    inner = meshes.interior_elements(desc)
    chosen = rng.choice(inner, size=len(inner) // 3, replace=False)
    if mixed:
        desc.quad_N = rng.choice([N, N + 2, N + 4], size=res * res).astype(np.int32)
    fine = meshes.split_elements(desc, chosen)
    T = O.Tables(16, 5)
    return fine, T, O.Mesh(T, fine, N)


def test_hanging_faces_are_built():
    desc, T, m = hanging_mesh(8, 4, False, 0.0, 1)
    # every split element's unsplit neighbour has sides covered by two small faces; the face count
    # is that of the conforming fine mesh minus one face per hanging pair
    K = desc.quads.shape[0]
    assert m.K == K
    edges = {}
    for q in desc.quads:
        for s in range(4):
            edges[(int(q[s]), int(q[(s + 1) % 4]))] = True
    unmatched = sum(1 for (a, b) in edges if (b, a) not in edges)
    n_bars = desc.bars.shape[0]
    # unmatched directed edges: boundary sides (n_bars) + 3 per hanging side (big + two small)
    n_hanging = (unmatched - n_bars) // 3
    assert n_hanging > 0
    assert m.n_faces == (4 * K - unmatched) // 2 + n_bars + 2 * n_hanging


@pytest.mark.parametrize("mixed", [False, True])
def test_hanging_free_stream_and_polynomials(mixed):
    """Constant states give G = 0 and continuous polynomial fields of degree <= N give the exact
    -div f on every element whose faces are interior: the interpolation to the half faces and the
    L2 projection back are exact for these degrees."""
    desc, T, m = hanging_mesh(8, 4, mixed, 0.0, 3)
    n = m.n
    m.set_state(np.full(n, 0.75), np.zeros(n), np.zeros(n))
    assert max(np.abs(g).max() for g in m.compute_rhs(0.0)) < 1e-11
    Ns = m.N
    X, Y = packed_nodes(desc, T, Ns)
    p = X**2 + 3 * X * Y - Y**3 + 0.5
    u = Y**2 - X * Y + 2 * X**3
    v = X * Y + X**2 * Y
    m.set_state(p, u, v)
    Gp, Gu, Gv = m.compute_rhs(0.0)
    ex = (-((-Y + 6 * X**2) + (X + X**2)), -(2 * X + 3 * Y), -(3 * X - 3 * Y**2))
    mask = interior_mask(desc, Ns, 8)
    for got, ref in zip((Gp, Gu, Gv), ex):
        assert np.max(np.abs(got[mask] - ref[mask])) < 1e-9


@pytest.mark.parametrize("mixed,jitter", [(False, 0.2), (True, 0.1)])
def test_hanging_conservation(mixed, jitter):
    desc, T, m = hanging_mesh(8, 4, mixed, jitter, 5)
    rng = np.random.default_rng(9)
    n = m.n
    m.set_state(rng.normal(size=n), rng.normal(size=n), rng.normal(size=n))
    Gp, Gu, Gv = m.compute_rhs(0.0)
    W = quadrature_weights(desc, T, m.N)
    assert abs(np.sum(W * Gp)) < 1e-12 * np.sum(W * np.abs(Gp))


def test_openmp_threads_give_the_same_bits():
    """The oracle's stage loops run under OpenMP (elements, faces and ghosts are written by one
    thread each): 1 and 4 threads give the same state bit for bit on a mixed-order mesh with
    hanging faces and every boundary kind."""
    import meshes
    base = nd.square_mesh(8, bc=(nd.BC_WALL, nd.BC_INFLOW, nd.BC_OUTFLOW, nd.BC_SYMMETRY), jitter=0.1, seed=2)
    desc = meshes.split_elements(base, [e for e in meshes.interior_elements(base)][::5])
    desc.quad_N = np.random.default_rng(5).choice([3, 4, 6], size=len(desc.quads)).astype(np.int32)
    out = []
    for threads in (1, 4):
        assert O.set_threads(threads) == threads
        m = O.Mesh(O.Tables(8, 5), desc, 3)
        m.initial_conditions(None)
        m.run_steps(0.5, 5)
        out.append(np.concatenate(m.get_state()))
    O.set_threads(1)
    assert np.array_equal(out[0], out[1])
