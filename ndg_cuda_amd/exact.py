"""Exact solution of the periodic multi-pulse problem at element nodes, and the quadrature L2 error
of a DG state against it (host utilities for accuracy measurements; no reference counterpart).

The system is p_t + c div(u, v) = 0, (u, v)_t + c grad p = 0 on the periodic square [-1, 1]^2 with
an initial pressure that is a sum of periodised radial Gaussians a exp(-ln2 |x - x0|^2 / w^2) and
u = v = 0. With alpha = ln2 / w^2 the Fourier coefficient of one periodised pulse on wave vector
k = pi (m, n) is a pi / (4 alpha) exp(-|k|^2 / (4 alpha)) exp(-i k.x0); each mode evolves as
p_k cos(c|k|t), (u, v)_k = -i k/|k| p_k sin(c|k|t). At nodes that lie on a tensor grid (elements
of an axis-aligned mesh) the mode sum factorises: with Ex[i, m] = exp(i pi m x_i) the pressure on
the grid is Re(Ex M Ey^T), two matrix products per field instead of a sum over every node and mode.
"""
import numpy as np

LN2 = np.log(2.0)


def _modes_needed(pulses, eps=1e-18):
    """Largest |m| whose coefficient can exceed eps times the pulse's amplitude factor."""
    w_min = float(np.min(np.asarray(pulses, dtype=np.float64)[:, 2]))
    alpha = LN2 / w_min**2
    k_max = np.sqrt(4.0 * alpha * -np.log(eps))
    return int(np.ceil(k_max / np.pi)) + 1


def node_coordinates(xy, quads, nodes_1d):
    """Physical coordinates (K, (N+1)^2) of the nodes of the given elements, packed like the
    fields (node i*(N+1)+j at reference coordinates (xi_i, eta_j), bilinear map of the 4 vertices
    in local order), and the Jacobian x_xi y_eta - x_eta y_xi at each node."""
    P = np.asarray(xy, dtype=np.float64)[np.asarray(quads)]          # (K, 4, 2)
    x = np.asarray(nodes_1d, dtype=np.float64)
    XI, ETA = np.meshgrid(x, x, indexing="ij")
    XI, ETA = XI.ravel(), ETA.ravel()
    w = np.stack([(1 - XI) * (1 - ETA), (XI + 1) * (1 - ETA), (XI + 1) * (ETA + 1), (1 - XI) * (ETA + 1)]) / 4
    X = P[:, :, 0] @ w
    Y = P[:, :, 1] @ w
    dxi = np.stack([-(1 - ETA), (1 - ETA), (ETA + 1), -(ETA + 1)]) / 4
    deta = np.stack([-(1 - XI), -(XI + 1), (XI + 1), (1 - XI)]) / 4
    x_xi, y_xi = P[:, :, 0] @ dxi, P[:, :, 1] @ dxi
    x_eta, y_eta = P[:, :, 0] @ deta, P[:, :, 1] @ deta
    return X, Y, x_xi * y_eta - x_eta * y_xi


def _mode_matrices(pulses, t, c, n_modes):
    """M_p, M_u, M_v (complex, (2n+1)^2) so that p = Re(Ex M_p Ey^T) etc. on a tensor grid."""
    m = np.arange(-n_modes, n_modes + 1)
    kx, ky = np.meshgrid(np.pi * m, np.pi * m, indexing="ij")
    km = np.hypot(kx, ky)
    ct = np.cos(km * c * t)
    with np.errstate(invalid="ignore", divide="ignore"):
        st_k = np.where(km > 0, np.sin(km * c * t) / np.where(km > 0, km, 1.0), 0.0)
    A = np.zeros(kx.shape, dtype=np.complex128)
    for (x0, y0, w, a) in np.asarray(pulses, dtype=np.float64):
        alpha = LN2 / w**2
        A += a * np.pi / (4.0 * alpha) * np.exp(-(km**2) / (4.0 * alpha)) * np.exp(-1j * (kx * x0 + ky * y0))
    # sin(k.(x - x0)) = Re(-i exp(i k.(x - x0)))
    return A * ct, -1j * A * kx * st_k, -1j * A * ky * st_k


def periodic_pulses(X, Y, pulses, t, c=1.0, n_modes=None):
    """p, u, v of the periodic multi-pulse solution at points (X, Y) (any shape) at time t. The
    points should take few distinct x and y values (an axis-aligned mesh): the cost is that of
    the tensor grid of the distinct values."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    n_modes = _modes_needed(pulses) if n_modes is None else int(n_modes)
    ux, ix = np.unique(X.ravel(), return_inverse=True)
    uy, iy = np.unique(Y.ravel(), return_inverse=True)
    m = np.arange(-n_modes, n_modes + 1)
    Ex = np.exp(1j * np.pi * np.outer(ux, m))
    Ey = np.exp(1j * np.pi * np.outer(uy, m))
    out = []
    for M in _mode_matrices(pulses, t, c, n_modes):
        grid = ((Ex @ M) @ Ey.T).real
        out.append(grid[ix, iy].reshape(X.shape))
    return tuple(out)


def mesh_nodes(mesh_desc, orders, nodes_of):
    """Node coordinates and Jacobians of every element of a mesh of mixed orders, packed per
    field like Mesh2D_t.get_solution (elements in order, (N_e+1)^2 values each), and the per-node
    quadrature weights w_i w_j. nodes_of(N) and weights are taken from the NDG_t tables:
    nodes_of(N) -> (nodes, weights)."""
    xy, quads = np.asarray(mesh_desc.xy), np.asarray(mesh_desc.quads)
    orders = np.asarray(orders)
    K = quads.shape[0]
    sizes = (orders.astype(np.int64) + 1) ** 2
    off = np.concatenate([[0], np.cumsum(sizes)])
    X = np.empty(off[-1])
    Y = np.empty(off[-1])
    J = np.empty(off[-1])
    W = np.empty(off[-1])
    for n in np.unique(orders):
        sel = np.nonzero(orders == n)[0]
        x, w = nodes_of(int(n))
        Xn, Yn, Jn = node_coordinates(xy, quads[sel], x)
        idx = (off[sel][:, None] + np.arange((n + 1) ** 2)[None, :]).ravel()
        X[idx], Y[idx], J[idx] = Xn.ravel(), Yn.ravel(), Jn.ravel()
        W[idx] = np.tile(np.outer(w, w).ravel(), sel.size)
    assert off[-1] == sizes.sum() and K == orders.size
    return X, Y, J * W


def l2_error(state, exact, weights):
    """Relative L2 error over the three fields with the Gauss quadrature of each element:
    sqrt(sum w J |q - q_ex|^2) / sqrt(sum w J |q_ex|^2)."""
    num = sum(float(np.sum(weights * (np.asarray(a) - b) ** 2)) for a, b in zip(state, exact))
    den = sum(float(np.sum(weights * b**2)) for b in exact)
    return float(np.sqrt(num / den))
