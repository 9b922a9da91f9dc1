"""Host-side pins of the calibration adjoint (no GPU):

- xdem_calib_pair_eval (the generic pair contact on doubles and on dual numbers) returns the
  oracle's force for the same pair (ref_pair_force of the FP64 C restatement, fed the overlap,
  relative velocity and normal computed as the reference's p2p kernel computes them), bit for
  bit, for both damping laws and both models;
- its dual-number Jacobian (23 inputs: both positions, velocities and angular velocities, the
  history entry, friction, restitution and the modulus scale) matches central differences of
  the same function;
- the radial-profile loss gradient of the harness matches central differences.
"""
import ctypes as C

import numpy as np
import pytest

import xdem_ref
from xdynamics_dll_amd import _lib, contact_property
from xdynamics_dll_amd.calib import profile_loss, radial_profile


def quirks(gpu_law, stale=0):
    return _lib.Quirks(gpu_law, 0, stale, 1)


def pair_eval(cp, model, q, dt, ir, jr, im, jm, x):
    out = np.zeros(12)
    jac = np.zeros(12 * 23)
    hit = _lib.lib().xdem_calib_pair_eval(C.byref(cp), model, C.byref(q), dt, ir, jr, im, jm,
                                         np.ascontiguousarray(x).ctypes.data, out.ctypes.data, jac.ctypes.data)
    return hit, out, jac.reshape(12, 23)


def sample(rng, ir, jr):
    xi = rng.uniform(-1e-3, 1e-3, 3)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    xj = xi + (ir + jr) * (1.0 - rng.uniform(0.001, 0.01)) * u
    x = np.concatenate([xi, xj, rng.normal(0, 0.05, 3), rng.normal(0, 0.05, 3), rng.normal(0, 20, 3),
                        rng.normal(0, 20, 3), [rng.uniform(-1e-6, 1e-6), rng.normal(0, 0.01)], [0.4, 0.35, 0.0]])
    return x


def cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


@pytest.mark.parametrize("model,gpu_law", [(1, 0), (1, 1), (0, 0)])
def test_pair_eval_matches_oracle_force(model, gpu_law):
    rng = np.random.default_rng(7 + model + 2 * gpu_law)
    ref = xdem_ref.lib()
    for trial in range(50):
        ir, jr = rng.uniform(1e-3, 1.5e-3, 2)
        im, jm = 2500.0 * 4.0 / 3.0 * np.pi * ir ** 3, 2500.0 * 4.0 / 3.0 * np.pi * jr ** 3
        cp = contact_property(2e7, 2e7, 0.3, 0.3, 0.35, 0.45, 0.05, 0.0, 0.1)
        x = sample(rng, ir, jr)
        x[20], x[21] = cp.fric, cp.rest
        q = quirks(gpu_law)
        hit, out, _ = pair_eval(cp, model, q, 2e-5, ir, jr, im, jm, x)
        assert hit == 1
        # the p2p kernel's geometry (impl:613-733) in the same operations
        xi, xj, vi, vj, wi, wj = (x[0:3], x[3:6], x[6:9], x[9:12], x[12:15], x[15:18])
        rp = np.array([xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]])
        dist = np.sqrt(rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2])
        cdist = (ir + jr) - dist
        unit = rp / dist
        cpt = xi + ir * unit
        dcpr, dcpr_j = cpt - xi, cpt - xj
        rv = (vj + cross(wj, dcpr_j)) - (vi + cross(wi, dcpr))
        p = _lib.DemParameters()
        p.contact_model = model
        p.dt = 2e-5
        sd = np.array([x[18], x[19]])
        ft = np.zeros(3)
        fn = np.zeros(3)
        ref.ref_pair_force(C.byref(p), C.byref(q), C.byref(cp), ir, jr, im, jm, cdist, rv.ctypes.data,
                           unit.ctypes.data, sd.ctypes.data, ft.ctypes.data, fn.ctypes.data)
        F = fn + ft
        assert np.array_equal(out[0:3], F), (trial, out[0:3], F)
        assert np.array_equal(out[10:12], sd), trial
        M = cross(dcpr, F)
        assert np.array_equal(out[3:6], M), trial


@pytest.mark.parametrize("model,gpu_law", [(1, 0), (1, 1), (0, 0)])
def test_pair_eval_jacobian_matches_central_differences(model, gpu_law):
    rng = np.random.default_rng(11 + model + 2 * gpu_law)
    q = quirks(gpu_law)
    for trial in range(20):
        ir, jr = rng.uniform(1e-3, 1.5e-3, 2)
        im, jm = 2500.0 * 4.0 / 3.0 * np.pi * ir ** 3, 2500.0 * 4.0 / 3.0 * np.pi * jr ** 3
        cp = contact_property(2e7, 2e7, 0.3, 0.3, 0.35, 0.45, 0.05, 0.0, 0.1)
        x = sample(rng, ir, jr)
        x[22] = rng.uniform(-0.3, 0.3)
        _, out, jac = pair_eval(cp, model, q, 2e-5, ir, jr, im, jm, x)
        scale = np.abs(x).copy()
        scale[0:6] = 1e-3
        scale[scale == 0] = 1.0
        for k in range(23):
            h = 1e-6 * scale[k]
            xp, xm = x.copy(), x.copy()
            xp[k] += h
            xm[k] -= h
            _, op, _ = pair_eval(cp, model, q, 2e-5, ir, jr, im, jm, xp)
            _, om, _ = pair_eval(cp, model, q, 2e-5, ir, jr, im, jm, xm)
            fd = (op - om) / (2 * h)
            # rows of the same kind share a scale (the components of one vector)
            for r0, r1 in ((0, 3), (3, 6), (6, 9), (9, 10), (10, 11), (11, 12)):
                tol = 1e-5 * max(np.abs(jac[r0:r1, :]).max(), np.abs(fd[r0:r1]).max(), 1e-300)
                assert np.abs(jac[r0:r1, k] - fd[r0:r1]).max() <= tol, (trial, k, r0, jac[r0:r1, k], fd[r0:r1])


def test_pair_eval_reports_no_contact():
    cp = contact_property(2e7, 2e7, 0.3, 0.3, 0.35, 0.45, 0.05)
    x = np.zeros(23)
    x[3] = 3e-3
    x[20], x[21] = 0.45, 0.35
    hit, out, jac = pair_eval(cp, 1, quirks(0), 2e-5, 1e-3, 1e-3, 1e-5, 1e-5, x)
    assert hit == 0 and not out.any() and not jac.any()


def test_profile_loss_gradient_matches_central_differences():
    rng = np.random.default_rng(3)
    n = 400
    pos = np.empty((n, 4))
    pos[:, 0] = 0.2 + rng.normal(0, 0.02, n)
    pos[:, 1] = 0.2 + rng.normal(0, 0.02, n)
    pos[:, 2] = rng.uniform(0, 0.02, n)
    pos[:, 3] = rng.uniform(1e-3, 1.5e-3, n)
    target = radial_profile(pos + [0.001, 0, 0.002, 0], (0.2, 0.2))
    loss, g = profile_loss(pos, target, (0.2, 0.2))
    assert loss > 0
    for i in rng.choice(n, 10, replace=False):
        for a in range(3):
            h = 1e-7
            pp, pm = pos.copy(), pos.copy()
            pp[i, a] += h
            pm[i, a] -= h
            fd = (profile_loss(pp, target, (0.2, 0.2))[0] - profile_loss(pm, target, (0.2, 0.2))[0]) / (2 * h)
            assert abs(fd - g[i, a]) <= 1e-5 * np.abs(g).max() + 1e-12, (i, a, fd, g[i, a])
