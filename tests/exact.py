"""Exact solutions used to pin the time stepper (test helpers, no reference counterpart).

periodic_pulse: the periodic Fourier-series solution on [-1,1]^2 for an initial pressure that is
a periodised radial Gaussian with u = v = 0. Each mode exp(i k.x), k = pi n, evolves as
p_k cos(|k| t), u_k = -i k_x/|k| p_k sin(|k| t) (and v likewise), for the system
p_t + u_x + v_y = 0, u_t + p_x = 0, v_t + p_y = 0 (wave speed 1). With wave speed c
(p_t + c div(u, v) = 0, (u, v)_t + c grad p = 0) every mode turns at c|k|, so the solution at
time t is the c = 1 solution at time c t.

plane_wave: the reference's own exact solution g(x, t) (analytical_solution.tpp:5-11), moving at
speed c along its direction k (c = 1 in the reference).
"""
import numpy as np

LN2 = np.log(2.0)


def periodic_images(x0, y0, width, amp, reach=2):
    """Pulse list (x0, y0, w, a) of the periodised Gaussian: images at offsets 2*(i, j)."""
    out = []
    for i in range(-reach, reach + 1):
        for j in range(-reach, reach + 1):
            out.append((x0 + 2.0 * i, y0 + 2.0 * j, width, amp))
    return np.array(out)


def periodic_pulse(x, y, t, pulses, n_modes=40, c=1.0):
    """p, u, v at points (x, y) and time t for a sum of periodised Gaussian pulses on [-1,1]^2,
    wave speed c."""
    t = c * t
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = np.arange(-n_modes, n_modes + 1)
    kx, ky = np.meshgrid(np.pi * n, np.pi * n, indexing="ij")
    kx, ky = kx.ravel(), ky.ravel()
    km = np.hypot(kx, ky)
    p = np.zeros_like(x)
    u = np.zeros_like(x)
    v = np.zeros_like(x)
    for (x0, y0, w, a) in pulses:
        alpha = LN2 / w**2
        c = a * np.pi / (4.0 * alpha) * np.exp(-(km**2) / (4.0 * alpha))
        keep = c > 1e-300
        kxk, kyk, kmk, ck = kx[keep], ky[keep], km[keep], c[keep]
        cos_t = np.cos(kmk * t)
        with np.errstate(invalid="ignore", divide="ignore"):
            sin_over = np.where(kmk > 0, np.sin(kmk * t) / np.where(kmk > 0, kmk, 1.0), 0.0)
        theta = np.outer(x.ravel() - x0, kxk) + np.outer(y.ravel() - y0, kyk)
        p += (np.cos(theta) @ (ck * cos_t)).reshape(x.shape)
        s = np.sin(theta)
        u += (s @ (ck * kxk * sin_over)).reshape(x.shape)
        v += (s @ (ck * kyk * sin_over)).reshape(x.shape)
    return p, u, v


def plane_wave(x, y, t, c=1.0):
    kx = ky = 0.707106781186548
    d = 0.120112240878645
    a = kx * (x + 0.2) + ky * (y + 0.2) - c * t
    p = np.exp(-a * a / (d * d))
    return p, p * kx, p * ky


def node_coordinates(desc, nodes_1d):
    """Physical coordinates of every element node, packed like the fields (reference order)."""
    quads = np.asarray(desc.quads)
    P = np.asarray(desc.xy)[quads]          # (K, 4, 2)
    xi = np.asarray(nodes_1d)
    XI, ETA = np.meshgrid(xi, xi, indexing="ij")
    XI, ETA = XI.ravel(), ETA.ravel()
    w0 = (1 - XI) * (1 - ETA)
    w1 = (XI + 1) * (1 - ETA)
    w2 = (XI + 1) * (ETA + 1)
    w3 = (1 - XI) * (ETA + 1)
    X = (P[:, 0, 0:1] * w0 + P[:, 1, 0:1] * w1 + P[:, 2, 0:1] * w2 + P[:, 3, 0:1] * w3) / 4
    Y = (P[:, 0, 1:2] * w0 + P[:, 1, 1:2] * w1 + P[:, 2, 1:2] * w2 + P[:, 3, 1:2] * w3) / 4
    return X.ravel(), Y.ravel()
