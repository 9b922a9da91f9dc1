"""GPU checks of the calibration adjoint (include/xdem_calib.h, xdynamics_dll_amd/calib.py):

- the calibration's forward run is the FP64 step: final positions, velocities, Euler parameters
  and rates bit-identical with the oracle after 2,500 steps of a 600-sphere pile falling onto
  its floor;
- the adjoint's gradient of the radial-profile loss with respect to (friction, restitution,
  log E) agrees with central differences of the forward run (step 1e-7) within 1e-3 of the
  gradient's size, over 300 and 500 steps (3 and 5 checkpoint segments of 100 steps), with
  the CPU force law (xContact.cpp:210-264). Over longer runs the loss is only piecewise smooth
  in the parameters (a contact that starts a step earlier or later changes the damping impulse
  by a finite amount), and central differences then swing with the step size while the adjoint
  gives the derivative of the piece the run is on. The kernel's law (impl:450) adds a
  tangential term vs sqrt(St m) that does not vanish with the slip velocity, along the slip
  direction e / |e|: a nearly stuck contact then flips a finite force with the sign of a tiny
  velocity, and no derivative of such a run is meaningful;
- a run continued from a checkpoint at new parameters (the finite-horizon gradient) against
  central differences of the continued runs;
- the same gradient through PileLoss, the torch.autograd.Function of the harness.
"""
import numpy as np
import pytest

import xdem_ref
from xdynamics_dll_amd import SceneData
from xdynamics_dll_amd.calib import Calibration, PileLoss, pile_center, profile_loss, radial_profile

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def pile(n=600, seed=1005):
    return SceneData.build(5, seed, np_override=n)


def theta0(scene):
    return np.array([scene.cp_pp.fric, scene.cp_pp.rest, np.log(scene.cp_pp.Ei)])


# the kernel's damping law (impl:432-452); the CPU formula (the default) has a damping
# coefficient beta 2 sqrt(m kn) that exceeds 2 m / dt for the lighter spheres of a dense pile
# at dt = 2e-5 s, where such a run diverges after about 1,500 steps
KERNEL_LAW = dict(gpu_force_law=1)


@pytest.mark.parametrize("n,steps,quirks", [(2000, 2500, KERNEL_LAW), (600, 2000, None)])
def test_forward_is_the_fp64_step(n, steps, quirks):
    scene = pile(n)
    print(f"[calib] forward parity np={scene.np} steps={steps} quirks={quirks}", flush=True)
    cal = Calibration(scene, max_steps=steps, checkpoint_every=100, quirks=quirks)
    cal.forward(theta0(scene), steps)
    pos, vel, ep, ev = cal.state()
    ref = xdem_ref.RefSim(scene, quirks=quirks)
    ref.step(steps)
    for name, g, r in (("pos", pos, ref.pos), ("vel", vel, ref.vel), ("ep", ep, ref.ep), ("ev", ev, ref.ev)):
        assert np.isfinite(r).all()
        assert np.array_equal(g, r), (name, float(np.abs(g - r).max()))
    if n > 1000:   # the 600-sphere pile is still falling: a few contacts
        assert (ref.pp_count > 0).sum() > n // 10 and (ref.pl_count > 0).sum() > n // 10
    assert cal.error_word() == 0
    cal.close()


def loss_at(cal, theta, steps, target, center):
    pos = cal.forward(theta, steps)
    return profile_loss(pos, target, center)


@pytest.mark.parametrize("steps", [300, 500])
def test_adjoint_gradient_matches_central_differences(steps):
    scene = pile()
    center = pile_center(scene)
    cal = Calibration(scene, max_steps=steps, checkpoint_every=100)
    # target: the profile of the same pile at other parameters
    t_target = np.array([0.3, 0.5, np.log(3e7)])
    target = radial_profile(cal.forward(t_target, steps), center)
    th = theta0(scene)
    loss, g = loss_at(cal, th, steps, target, center)
    grad = cal.backward(g)
    print(f"[calib] steps={steps} loss={loss:.6e} adjoint={grad}", flush=True)
    assert np.isfinite(grad).all() and loss > 0
    for h in (1e-7,):
        fd = np.zeros(3)
        for k in range(3):
            tp, tm = th.copy(), th.copy()
            tp[k] += h
            tm[k] -= h
            fd[k] = (loss_at(cal, tp, steps, target, center)[0] - loss_at(cal, tm, steps, target, center)[0]) / (2 * h)
        print(f"[calib] steps={steps} h={h:g} central differences={fd}", flush=True)
        scale = np.abs(fd).max()
        assert np.abs(grad - fd).max() <= 1e-3 * scale, (h, grad, fd)
    assert cal.error_word() == 0
    cal.close()


def test_finite_horizon_gradient_matches_central_differences():
    """A run continued from step 300 of a 500-step run at new parameters: the adjoint over steps
    300..499 against central differences of the continued runs."""
    scene = pile()
    center = pile_center(scene)
    steps, start = 500, 300
    cal = Calibration(scene, max_steps=steps, checkpoint_every=100)
    target = radial_profile(cal.forward(np.array([0.3, 0.5, np.log(3e7)]), steps), center)
    th0 = theta0(scene)
    cal.forward(th0, steps)
    th = th0 + np.array([0.02, -0.03, 0.05])
    h = 1e-7
    fd = np.zeros(3)
    for k in range(3):
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (profile_loss(cal.forward(tp, steps, start), target, center)[0] -
                 profile_loss(cal.forward(tm, steps, start), target, center)[0]) / (2 * h)
    pos = cal.forward(th, steps, start)
    loss, g = profile_loss(pos, target, center)
    grad = cal.backward(g)
    print(f"[calib] continued from {start}: adjoint={grad} central differences={fd}", flush=True)
    assert np.abs(grad - fd).max() <= 1e-3 * np.abs(fd).max(), (grad, fd)
    # the continued run is the full run when the parameters do not change
    p_full = cal.forward(th0, steps)
    assert np.array_equal(cal.forward(th0, steps, start), p_full)
    cal.close()


def test_torch_function_gives_the_adjoint_gradient():
    scene = pile()
    center = pile_center(scene)
    steps = 300
    cal = Calibration(scene, max_steps=steps, checkpoint_every=100)
    target = radial_profile(cal.forward(np.array([0.3, 0.5, np.log(3e7)]), steps), center)
    theta = torch.tensor(theta0(scene), dtype=torch.float64, requires_grad=True)
    loss = PileLoss.apply(theta, cal, steps, target, center)
    loss.backward()
    pos = cal.forward(theta0(scene), steps)
    l2, g = profile_loss(pos, target, center)
    assert float(loss.detach()) == l2
    # the adjoint sums theta's cotangent with atomics: equal up to the order of the additions
    assert np.allclose(theta.grad.numpy(), cal.backward(g), rtol=1e-9, atol=0.0)
    cal.close()
