"""Calibration harness: gradients of a pile-shape loss with respect to the contact material.

The differentiated run is the FP64 step of the library (include/xdem_calib.h): a pile of
spheres released onto a floor, stepped for a fixed number of steps. The loss compares the
final radial height profile of the pile with a target profile. Its gradient with respect to
theta = (sliding friction, restitution, log Young's modulus) comes from the library's
discrete adjoint (checkpoints every `checkpoint_every` steps, recomputation per segment,
reverse sweep of the contact and integration kernels); `PileLoss` exposes the loss as a
torch.autograd.Function so that a torch optimiser can drive the calibration.

Radial height profile: bins b = 0..nbins-1 of the distance rho from the pile's axis, centres
c_b = (b + 1/2) drho, drho = rho_max / nbins; each sphere weighs w_b = max(0, 1 - |rho - c_b| /
drho) in the two bins around it (linear interpolation); h_b = sum_i w_b(rho_i) (z_i + r_i) /
(sum_i w_b(rho_i) + eps), the weighted mean height of the sphere tops. Loss = mean_b
(h_b - target_b)^2.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .scene import PARITY_QUIRKS, SceneData


class Calibration:
    """One scene's FP64 run and its adjoint (xdem_calib_*). theta[2] is log E; the library's
    modulus scale is s = log E - log E0, E0 the scene's sphere modulus."""

    def __init__(self, scene: SceneData, max_steps: int, checkpoint_every: int = 100, capacity: int = 12,
                 quirks: dict | None = None, device: int = 0):
        self.scene = scene
        self.L = _lib.lib()
        cfg = _lib.CalibConfig()
        cfg.capacity = capacity
        cfg.checkpoint_every = checkpoint_every
        cfg.max_steps = max_steps
        cfg.device = device
        q = dict(PARITY_QUIRKS)
        q.update(quirks or {})
        cfg.quirks = _lib.Quirks(q["gpu_force_law"], q["velocity_not_updated"], q["stale_ft"], q["flat_inertia"])
        self.log_e0 = float(np.log(scene.cp_pp.Ei))
        self._h = C.c_void_p()
        cs, keep = scene.to_c()
        _lib.check(self.L.xdem_calib_create(C.byref(cs), C.byref(cfg), C.byref(self._h)), "xdem_calib_create")
        del keep
        self.np = scene.np

    def theta_of(self, friction: float, restitution: float, youngs: float) -> np.ndarray:
        return np.array([friction, restitution, np.log(youngs)], dtype=np.float64)

    def _lib_theta(self, theta) -> np.ndarray:
        t = np.asarray(theta, dtype=np.float64).reshape(3).copy()
        t[2] = t[2] - self.log_e0
        return t

    def forward(self, theta, nsteps: int, start: int = 0) -> np.ndarray:
        """Runs nsteps from the scene's initial state (start = 0), or continues the last run from
        its step `start` (a multiple of checkpoint_every) at theta; returns the final (np, 4)
        positions. The next backward covers the steps from start on."""
        t = self._lib_theta(theta)
        pos = np.empty((self.np, 4), dtype=np.float64)
        if start:
            _lib.check(self.L.xdem_calib_forward_from(self._h, t.ctypes.data, start, nsteps, pos.ctypes.data),
                       "xdem_calib_forward_from")
        else:
            _lib.check(self.L.xdem_calib_forward(self._h, t.ctypes.data, nsteps, pos.ctypes.data),
                       "xdem_calib_forward")
        return pos

    def state(self):
        n = self.np
        pos, vel, ep, ev = (np.empty((n, 4)), np.empty((n, 3)), np.empty((n, 4)), np.empty((n, 4)))
        _lib.check(self.L.xdem_calib_state(self._h, pos.ctypes.data, vel.ctypes.data, ep.ctypes.data, ev.ctypes.data),
                   "xdem_calib_state")
        return pos, vel, ep, ev

    def backward(self, dL_dpos) -> np.ndarray:
        """d loss / d theta for the loss gradient dL_dpos (np, 3) at the final positions of the
        last forward run."""
        g = np.ascontiguousarray(dL_dpos, dtype=np.float64).reshape(self.np, 3)
        out = np.zeros(3, dtype=np.float64)
        _lib.check(self.L.xdem_calib_backward(self._h, g.ctypes.data, out.ctypes.data), "xdem_calib_backward")
        return out   # d/ds = d/dlog E

    def error_word(self, clear: bool = True) -> int:
        w = C.c_uint()
        _lib.check(self.L.xdem_calib_error_word(self._h, C.byref(w), 1 if clear else 0), "xdem_calib_error_word")
        return w.value

    def close(self) -> None:
        if self._h:
            self.L.xdem_calib_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def radial_profile(pos, center, nbins: int = 64, rho_max: float = 0.06, eps: float = 1e-12, grad: bool = False):
    """Heights h (nbins,) of the radial profile; with grad, also dh/d(x, y, z) as (nbins, np, 3)."""
    p = np.asarray(pos, dtype=np.float64)
    dx, dy = p[:, 0] - center[0], p[:, 1] - center[1]
    rho = np.sqrt(dx * dx + dy * dy)
    top = p[:, 2] + p[:, 3]
    drho = rho_max / nbins
    c = (np.arange(nbins) + 0.5) * drho
    d = rho[None, :] - c[:, None]                      # (nbins, np)
    w = np.maximum(0.0, 1.0 - np.abs(d) / drho)
    W = w.sum(axis=1) + eps
    h = (w * top[None, :]).sum(axis=1) / W
    if not grad:
        return h
    inside = np.abs(d) < drho
    dw = np.where(inside, -np.sign(d) / drho, 0.0)       # dw / drho
    with np.errstate(invalid="ignore", divide="ignore"):
        ux = np.where(rho > 0, dx / rho, 0.0)
        uy = np.where(rho > 0, dy / rho, 0.0)
    g_rho = dw * (top[None, :] - h[:, None]) / W[:, None]
    J = np.empty((nbins, p.shape[0], 3))
    J[:, :, 0] = g_rho * ux[None, :]
    J[:, :, 1] = g_rho * uy[None, :]
    J[:, :, 2] = w / W[:, None]
    return h, J


def profile_loss(pos, target, center, nbins: int = 64, rho_max: float = 0.06):
    """(loss, dloss/dpos (np, 3)) of the profile MSE against target."""
    h, J = radial_profile(pos, center, nbins, rho_max, grad=True)
    r = h - np.asarray(target, dtype=np.float64)
    loss = float(np.mean(r * r))
    g = np.einsum("b,bij->ij", 2.0 * r / nbins, J)
    return loss, g


def pile_center(scene: SceneData):
    """The pile's axis: the mean of the initial sphere centres in x and y."""
    return (float(scene.pos[:, 0].mean()), float(scene.pos[:, 1].mean()))


try:
    import torch

    class PileLoss(torch.autograd.Function):
        """loss = PileLoss.apply(theta, calibration, nsteps, target, center[, nbins, rho_max, start]);
        theta a float64 tensor (friction, restitution, log E). The forward runs the steps (from the
        last run's step `start` when given); the backward runs the adjoint over them."""

        @staticmethod
        def forward(ctx, theta, calib, nsteps, target, center, nbins=64, rho_max=0.06, start=0):
            th = theta.detach().cpu().numpy().astype(np.float64)
            pos = calib.forward(th, nsteps, start)
            loss, g = profile_loss(pos, target, center, nbins, rho_max)
            ctx.calib = calib
            ctx.dpos = g
            return torch.tensor(loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx, grad_out):
            gt = ctx.calib.backward(ctx.dpos) * float(grad_out)
            return torch.from_numpy(gt), None, None, None, None, None, None, None
except ImportError:   # torch is optional for the numpy interface
    PileLoss = None
