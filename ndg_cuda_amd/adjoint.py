"""Discrete-adjoint form over the C ABI: receivers, the misfit gradient and a torch.autograd wrapper.

No reference counterpart (the reference has no adjoint). The misfit is
J = 1/2 sum_n sum_r (p(x_r, t_n) - d_r(t_n))^2 over the step ends t_n of a run with a frozen step
dt; its gradient with respect to the initial fields (p0, u0, v0) and the wave speed c comes from the
transposed RK3 stages (ndg_adjoint_gradient), with checkpoints every `checkpoint_interval` steps.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import check, lib, ptr


class Receivers:
    """Pressure receivers given as (element, xi, eta) with reference coordinates in [-1, 1]^2."""

    def __init__(self, element, xi_eta):
        self.element = _abi.as_i64(np.asarray(element).reshape(-1))
        self.xi_eta = _abi.as_f64(np.asarray(xi_eta, dtype=np.float64).reshape(-1, 2))
        if self.element.shape[0] != self.xi_eta.shape[0]:
            raise ValueError("Receivers: one (xi, eta) pair per element index")
        self._c = _abi.Receivers(self.element.shape[0], ptr(self.element), ptr(self.xi_eta))

    @property
    def n(self):
        return int(self.element.shape[0])

    @classmethod
    def seeded(cls, n_elements, n, seed):
        """n receivers at seeded elements and reference coordinates inside (-0.9, 0.9)^2."""
        rng = np.random.default_rng(seed)
        return cls(rng.integers(0, n_elements, n), rng.uniform(-0.9, 0.9, (n, 2)))


def set_wave_speed(mesh, c):
    check(lib().ndg_mesh_set_wave_speed(mesh.handle, float(c)))


def record_receivers(mesh, receivers, dt, n_steps):
    """Advances the mesh n_steps steps of dt; returns p at the receivers after each step,
    shape (n_steps, n_receivers)."""
    out = np.empty((int(n_steps), receivers.n), dtype=np.float64)
    check(lib().ndg_record_receivers(mesh.handle, ctypes.byref(receivers._c), float(dt), int(n_steps), ptr(out)))
    return out


def record_receivers_noalloc(mesh, receivers, dt, n_steps):
    """record_receivers through the allocation-free entry point (the receivers and records live in
    the library's fixed 4 MiB per-GPU scratch array; larger requests raise)."""
    out = np.empty((int(n_steps), receivers.n), dtype=np.float64)
    check(lib().ndg_record_receivers_noalloc(mesh.handle, ctypes.byref(receivers._c), float(dt), int(n_steps),
                                             ptr(out)))
    return out


def set_c_derivative(mesh, mode):
    """'recompute' (default: checkpoints, recompute, <G^T mu~, q_s> / c in the transposed stages) or
    'stage_records' (dJ/dc from the RK stage right-hand sides recorded at the receivers in the
    forward pass, valid because the operator is linear and time-invariant; no checkpoints, no
    recompute)."""
    code = {"recompute": _abi.ADJOINT_C_RECOMPUTE, "stage_records": _abi.ADJOINT_C_STAGE_RECORDS}[mode]
    check(lib().ndg_adjoint_set_c_derivative(mesh.handle, code))


def set_store_budget(mesh, nbytes):
    """Device memory adjoint_gradient may spend on stage states beyond one segment: the forward pass
    keeps the last 1 + nbytes // (segment bytes) segments, which are then not recomputed. 0 (the
    default) recomputes every segment but the last. The gradient is the same bit for bit."""
    check(lib().ndg_adjoint_set_store_budget(mesh.handle, int(nbytes)))


def adjoint_gradient(mesh, receivers, dt, n_steps, checkpoint_interval, data=None):
    """(J, dJ/dp0, dJ/du0, dJ/dv0, dJ/dc) from the mesh's current state, which is left unchanged."""
    n = mesh.n_dofs_per_field_
    g = [np.empty(n, dtype=np.float64) for _ in range(3)]
    J, dc = ctypes.c_double(), ctypes.c_double()
    d = None if data is None else _abi.as_f64(np.asarray(data, dtype=np.float64).reshape(int(n_steps), receivers.n))
    check(lib().ndg_adjoint_gradient(mesh.handle, ctypes.byref(receivers._c), float(dt), int(n_steps),
                                     int(checkpoint_interval), ptr(d), ctypes.byref(J), *(ptr(a) for a in g),
                                     ctypes.byref(dc)))
    return J.value, g[0], g[1], g[2], dc.value


def apply_rhs_transpose(mesh, mp, mu, mv):
    """G^T applied to packed adjoint fields (for dot-product tests)."""
    ins = [_abi.as_f64(a) for a in (mp, mu, mv)]
    out = [np.empty(mesh.n_dofs_per_field_, dtype=np.float64) for _ in range(3)]
    check(lib().ndg_apply_rhs_transpose(mesh.handle, *(ptr(a) for a in ins), *(ptr(a) for a in out)))
    return tuple(out)


def misfit_function():
    """torch.autograd.Function computing J(p0, u0, v0, c) on a mesh, with the adjoint gradient as
    its backward (the forward call runs the adjoint too and keeps the gradient)."""
    import torch

    class ReceiverMisfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, p0, u0, v0, c, mesh, receivers, dt, n_steps, checkpoint_interval, data):
            fields = [t.detach().to("cpu", torch.float64).numpy() for t in (p0, u0, v0)]
            mesh.set_solution(*fields)
            set_wave_speed(mesh, float(c))
            J, gp, gu, gv, gc = adjoint_gradient(mesh, receivers, dt, n_steps, checkpoint_interval, data)
            ctx.save_for_backward(*(torch.from_numpy(a).to(p0.device, p0.dtype) for a in (gp, gu, gv)),
                                  torch.tensor(gc, dtype=c.dtype, device=c.device))
            return torch.tensor(J, dtype=p0.dtype, device=p0.device)

        @staticmethod
        def backward(ctx, grad):
            gp, gu, gv, gc = ctx.saved_tensors
            return grad * gp, grad * gu, grad * gv, grad * gc, None, None, None, None, None, None

    return ReceiverMisfit
