"""The host HHT-alpha integrator of the driven wheel (include/xdem_mbd.h), checked against
closed forms (no GPU): a force switched on at t = 0 gives the discrete Newmark first step and
then the exact quadratic path, for three values of alpha (which fix beta and gamma); the
constraints hold, the drive turns the wheel at omega, the y multiplier balances the applied y
force; the Newton correction converges in a few iterations."""
import ctypes as C
import math

import numpy as np
import pytest

from xdynamics_dll_amd import _lib


def make(alpha=-0.1, dbi=None, load=(0.0, 0.0, -200.0), omega=2.0, dt=1e-5, gravity=(0.0, 0.0, -9.81)):
    cfg = _lib.HhtWheelConfig()
    cfg.mass = 40.0
    cfg.inertia[:] = (0.258, 0.45, 0.258)
    d = np.zeros(15) if dbi is None else np.asarray(dbi, dtype=np.float64)
    if dbi is None:
        d[1:4] = (0.5, 0.2, 0.3)
        d[7] = 1.0
        d[13] = 0.5 * omega   # e_dot for a turn about y at omega: (0, 0, omega / 2, 0)
    cfg.dbi[:] = list(d)
    cfg.gravity[:] = gravity
    cfg.load[:] = load
    cfg.omega = omega
    cfg.dt = dt
    cfg.alpha = alpha
    h = C.c_void_p()
    _lib.check(_lib.host().xdem_hht_wheel_create(C.byref(cfg), C.byref(h)), "xdem_hht_wheel_create")
    return h, d


def step(h, wrench):
    out = np.zeros(15)
    it = C.c_int()
    w = np.ascontiguousarray(wrench, dtype=np.float64)
    _lib.check(_lib.host().xdem_hht_wheel_step(h, w.ctypes.data, out.ctypes.data, C.byref(it)), "xdem_hht_wheel_step")
    return out, it.value


def state(h):
    d, qdd, lam, t = np.zeros(15), np.zeros(7), np.zeros(5), C.c_double()
    _lib.check(_lib.host().xdem_hht_wheel_state(h, d.ctypes.data, qdd.ctypes.data, lam.ctypes.data, C.byref(t)),
               "xdem_hht_wheel_state")
    return d, qdd, lam, t.value


@pytest.mark.parametrize("alpha", [0.0, -0.1, -0.3])
def test_constant_force_gives_the_exact_quadratic_path(alpha):
    """The Newmark family with gamma = 1/2 - alpha, beta = (1 - alpha)^2 / 4 reproduces a
    constant acceleration exactly, whatever alpha. The initial solve knows gravity and the
    load (a0); the contact force switched on with the first step makes the acceleration a1 from
    then on, so step 1 is v1 = v0 + h((1 - gamma) a0 + gamma a1), x1 = x0 + h v0 +
    h^2 ((1/2 - beta) a0 + beta a1), and every later step the exact quadratic from (x1, v1)."""
    wrench = np.array([30.0, 5.0, 150.0, 0.0, 0.0, 0.0])
    h, d0 = make(alpha=alpha)
    n, dt = 400, 1e-5
    beta, gamma = (1 - alpha) ** 2 / 4, 0.5 - alpha
    for _ in range(n):
        d, it = step(h, wrench)
        assert it <= 5
    a0 = np.array([0.0, 0.0, -9.81 - 200.0 / 40.0])
    a1 = a0 + np.array([30.0, 0.0, 150.0]) / 40.0   # y is held: its force goes to the constraint
    v1 = d0[4:7] + dt * ((1 - gamma) * a0 + gamma * a1)
    x1 = d0[1:4] + dt * d0[4:7] + dt * dt * ((0.5 - beta) * a0 + beta * a1)
    t = (n - 1) * dt
    assert np.allclose(d[1:4], x1 + v1 * t + 0.5 * a1 * t * t, rtol=0, atol=1e-13)
    assert np.allclose(d[4:7], v1 + a1 * t, rtol=0, atol=1e-11)
    _, qdd, _, _ = state(h)
    assert np.allclose(qdd[:3], a1, rtol=1e-9, atol=1e-9)
    _lib.host().xdem_hht_wheel_destroy(h)


def test_drive_and_constraints():
    """The drive turns the wheel about y at omega (e = (cos(w t / 2), 0, sin(w t / 2), 0)), the
    y axis and y position stay, |e| = 1, and the multiplier of the y constraint carries the
    applied y force."""
    h, d0 = make(alpha=-0.1, omega=2.0)
    wrench = np.array([0.0, 7.5, 0.0, 0.0, 0.0, 0.0])
    for k in range(1, 2001):
        d, _ = step(h, wrench)
    t = 2000 * 1e-5
    e = d[7:11]
    assert np.allclose(e, [math.cos(t), 0.0, math.sin(t), 0.0], atol=1e-12)
    assert abs(np.dot(e, e) - 1.0) < 1e-13
    assert abs(d[2] - d0[2]) < 1e-14
    _, _, lam, tt = state(h)
    assert abs(tt - t) < 1e-15
    assert abs(lam[0] - 7.5) < 1e-8   # Q_y - lambda_y = m y_dd = 0
    _lib.host().xdem_hht_wheel_destroy(h)


def test_newton_converges_and_refuses_bad_settings():
    """The correction converges in a few Newton iterations from the Newmark prediction (the
    reference caps it at 100, xIntegratorHHT.cpp:131-139); alpha outside [-1/3, 0] is refused."""
    h, _ = make(alpha=-0.3, omega=5.0)
    its = [step(h, [10.0, 0.0, 80.0, 0.5, 0.2, -0.3])[1] for _ in range(200)]
    assert max(its) <= 4
    _lib.host().xdem_hht_wheel_destroy(h)
    cfg = _lib.HhtWheelConfig()
    cfg.mass, cfg.dt, cfg.alpha = 40.0, 1e-5, -0.5
    cfg.inertia[:] = (1.0, 1.0, 1.0)
    cfg.dbi[7] = 1.0
    out = C.c_void_p()
    assert _lib.host().xdem_hht_wheel_create(C.byref(cfg), C.byref(out)) == _lib.STATUS_INVALID
