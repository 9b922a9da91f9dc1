"""The DEM <-> multibody loop on the GPU: the grousered wheel (a 4,096-triangle mesh body driven
at omega about y, free in x and z) over a 20,000-sphere bed, advanced by the host HHT-alpha
integrator under the contact wrench the step layer reduces every step
(xDynamicSimulator.cpp:347-375).

- the native loop (xdem_cosim_wheel_steps) equals the same calls made one by one from Python,
  bit for bit;
- with step graphs enabled, the loop's single steps replay one-step HIP graphs (one per buffer
  parity) and the pose is handed to the device as a kernel argument without a host wait: the
  run equals the eager one bit for bit;
- against the oracle driving the same integrator: the wrench agrees to the reduction's rounding
  every step, so the wheel's pose and the spheres' state agree to 1e-12 m after 30 steps.
"""
import numpy as np
import pytest

import xdem_ref
from mesh_scenes import wheel_over_bed
from xdynamics_dll_amd import DemContext
from xdynamics_dll_amd.mbd import HhtWheel, cosim_wheel_steps

pytestmark = pytest.mark.gpu

OMEGA = 2.0
INERTIA = (0.258, 0.45, 0.258)


def setup(n=20000):
    scene, tris, dbi, cp = wheel_over_bed(n)
    dbi = dbi.copy()
    dbi[13] = 0.5 * OMEGA   # e_dot of a turn about y at omega from e = (1, 0, 0, 0)
    return scene, tris, dbi, cp


def wheel(scene, dbi):
    return HhtWheel(40.0, INERTIA, dbi, OMEGA, scene.dt, alpha=-0.1, gravity=scene.gravity)


def test_native_loop_equals_python_loop():
    scene, tris, dbi, cp = setup()
    print(f"[cosim] np={scene.np} ntri={len(tris)}", flush=True)
    a = DemContext(scene, precision="fp64", fuse_velocity=True, meshes=[(0, tris, dbi, cp)])
    wa = wheel(scene, dbi)
    track = cosim_wheel_steps(a, wa, 0, 30)
    b = DemContext(scene, precision="fp64", fuse_velocity=True, meshes=[(0, tris, dbi, cp)])
    wb = wheel(scene, dbi)
    for k in range(30):
        b.step(1)
        w = b.mesh_wrench(0)
        d, it = wb.step(w)
        b.set_mesh_pose(0, d)
        assert np.array_equal(track[k], [d[1], d[3], d[4], d[6], w[0], w[2], w[4], it]), k
    for f in ("pos", "vel", "ep", "ev"):
        assert np.array_equal(a.read(f), b.read(f)), f
    assert np.array_equal(wa.state()[0], wb.state()[0])
    assert np.abs(track[:, 5]).max() > 0          # the bed carries the wheel
    assert a.error_word() == 0
    a.close()
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_one_step_graphs_equal_eager_steps(precision):
    scene, tris, dbi, cp = setup()
    runs = []
    for sg in (0, 50):
        ctx = DemContext(scene, precision=precision, fuse_velocity=True, step_graph=sg, meshes=[(0, tris, dbi, cp)])
        w = wheel(scene, dbi)
        track = cosim_wheel_steps(ctx, w, 0, 31)
        runs.append((track, {f: ctx.read(f) for f in ("pos", "vel", "ep", "ev", "acc", "ea")}, w.state()[0],
                     ctx.history("p2p"), ctx.history("mesh")))
        assert ctx.error_word() == 0
        ctx.close()
    (ta, sa, pa, ha, ma), (tb, sb, pb, hb, mb) = runs
    assert np.array_equal(ta, tb)
    for f in sa:
        assert np.array_equal(sa[f], sb[f]), f
    assert np.array_equal(pa, pb)
    for x, y in zip(ha + ma, hb + mb):
        assert np.array_equal(x, y)


def test_coupled_loop_matches_oracle():
    scene, tris, dbi, cp = setup()
    gpu = DemContext(scene, precision="fp64", fuse_velocity=True, meshes=[(0, tris, dbi, cp)])
    ref = xdem_ref.RefSim(scene)
    ref.add_mesh(tris, dbi, cp)
    wg, wr = wheel(scene, dbi), wheel(scene, dbi)
    for k in range(30):
        gpu.step(1)
        ref.step(1)
        w_g, w_r = gpu.mesh_wrench(0), ref.mesh["wrench"].copy()
        scale = max(np.abs(w_r).max(), 1e-300)
        assert np.abs(w_g - w_r).max() <= 1e-11 * scale, (k, w_g, w_r)
        dg, _ = wg.step(w_g)
        dr, _ = wr.step(w_r)
        gpu.set_mesh_pose(0, dg)
        ref.set_mesh_pose(dr)
    assert np.abs(dg - dr).max() <= 1e-12
    assert np.abs(gpu.read("pos") - ref.pos).max() <= 1e-12
    v = np.abs(ref.vel).max()
    assert np.abs(gpu.read("vel") - ref.vel).max() <= 1e-9 * v
    gpu.close()
