"""Host multibody side of the DEM <-> MBD coupling (include/xdem_mbd.h).

HhtWheel is the HHT-alpha index-3 integrator of a driven wheel (xIntegratorHHT::OneStepSimulation,
source/xdynamics_simulation/xIntegratorHHT.cpp:101-183) for one rigid body; cosim_wheel_steps
runs the simulator's coupled loop (xDynamicsSimulator::xRunSimulationThread,
source/xdynamics_simulation/xDynamicSimulator.cpp:347-375) in native code over a step-layer
context with the wheel as its mesh body.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class HhtWheel:
    def __init__(self, mass: float, inertia, dbi, omega: float, dt: float, alpha: float = -0.1,
                 gravity=(0.0, 0.0, -9.81), load=(0.0, 0.0, 0.0), tolerance: float = 0.0, max_iterations: int = 0):
        cfg = _lib.HhtWheelConfig()
        cfg.mass = mass
        cfg.inertia[:] = list(inertia)
        cfg.dbi[:] = list(np.asarray(dbi, dtype=np.float64))
        cfg.gravity[:] = list(gravity)
        cfg.load[:] = list(load)
        cfg.omega = omega
        cfg.dt = dt
        cfg.alpha = alpha
        cfg.tolerance = tolerance
        cfg.max_iterations = max_iterations
        self._h = C.c_void_p()
        _lib.check(_lib.host().xdem_hht_wheel_create(C.byref(cfg), C.byref(self._h)), "xdem_hht_wheel_create")

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def step(self, wrench) -> tuple[np.ndarray, int]:
        """One step under the contact wrench of the DEM step just taken; returns the new pose
        (MM..ED3) and the Newton iterations."""
        out = np.zeros(15)
        it = C.c_int()
        w = np.ascontiguousarray(wrench, dtype=np.float64)
        _lib.check(_lib.host().xdem_hht_wheel_step(self._h, w.ctypes.data, out.ctypes.data, C.byref(it)),
                   "xdem_hht_wheel_step")
        return out, it.value

    def state(self):
        """(pose MM..ED3, accelerations (7), multipliers (5), time)."""
        d, qdd, lam, t = np.zeros(15), np.zeros(7), np.zeros(5), C.c_double()
        _lib.check(_lib.host().xdem_hht_wheel_state(self._h, d.ctypes.data, qdd.ctypes.data, lam.ctypes.data,
                                                    C.byref(t)), "xdem_hht_wheel_state")
        return d, qdd, lam, t.value

    def close(self) -> None:
        if self._h:
            _lib.host().xdem_hht_wheel_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cosim_wheel_steps(ctx, wheel: HhtWheel, mesh: int, nsteps: int) -> np.ndarray:
    """nsteps of the coupled loop in libxdem.so; returns (nsteps, 8): x, z, vx, vz of the body,
    Fx, Fz, My of the contact wrench, Newton iterations. The wheel must have been created by
    the host library that libxdem.so also contains (both load the same integrator code)."""
    track = np.zeros((nsteps, 8))
    _lib.check(ctx.L.xdem_cosim_wheel_steps(ctx._h, wheel.handle, mesh, nsteps, track.ctypes.data),
               "xdem_cosim_wheel_steps")
    return track
