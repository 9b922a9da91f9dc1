"""GPU parity of mesh bodies in the step layer (xdem_ctx with xdem_ctx_add_mesh) against the
oracle (FP64 C restatement of impl:1930-2109 driven like the reference's step).

- a bumpy terrain body under falling spheres (face, edge and vertex contacts, cohesion, both
  damping laws): FP64 states, triangle histories (counts, ids, (ds, s_dot)) bit-identical,
  body wrench within the reduction's rounding;
- the terrain moved to a new pose before every step for 200 steps;
- mixed precision: forces within 1e-4 of the oracle's, contact sets exact;
- the 4,096-triangle grousered wheel over a 20,000-sphere bed, and over the full 500,000-sphere
  bed with property checks.
"""
import numpy as np
import pytest

import xdem_ref
from mesh_scenes import mesh_scene, wheel_over_bed
from xdynamics_dll_amd import DemContext

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "acc", "ep", "ev", "ea")


def assert_state_equal(gpu, ref, what):
    for f in STATE:
        g, r = gpu.read(f), getattr(ref, f)
        assert np.isfinite(r).all(), (what, f, "oracle")
        assert np.isfinite(g).all(), (what, f, "gpu")
        assert np.array_equal(g, r), (what, f, float(np.abs(g - r).max()))


def assert_mesh_history_equal(gpu, ref, what):
    c, ids, tsd = gpu.history("mesh")
    assert np.array_equal(c, ref.ms_count), what
    for i in np.nonzero(c)[0]:
        assert np.array_equal(ids[i, : c[i]], ref.ms_id[i, : c[i]]), (what, i)
        assert np.array_equal(tsd[i, : c[i]], ref.ms_tsd[i, : c[i]]), (what, i)


def terrain_pair(coh=0.0, quirks=None, precision="fp64", fuse=True):
    # 200 spheres on a 40 mm terrain of 800 triangles, overlapping the terrain and each other by
    # at most 0.5% of r, dt = 1e-6 s: a step test follows the spheres through many steps
    scene, tris, _, dbi, cp = mesh_scene(n=200, coh=coh, max_overlap=0.005, nx=20, separated=True, dt=1e-6)
    gpu = DemContext(scene, precision=precision, fuse_velocity=fuse, quirks=quirks, meshes=[(0, tris, dbi, cp)])
    ref = xdem_ref.RefSim(scene, quirks=quirks)
    ref.add_mesh(tris, dbi, cp)
    return scene, tris, dbi, cp, gpu, ref


@pytest.mark.parametrize("coh,gpu_law,fuse", [(0.0, 0, True), (0.05, 0, True), (0.0, 1, True), (0.0, 0, False)])
def test_terrain_mesh_steps_match_oracle(coh, gpu_law, fuse):
    scene, tris, dbi, _, gpu, ref = terrain_pair(coh=coh, quirks=dict(gpu_force_law=gpu_law), fuse=fuse)
    print(f"[mesh step] np={scene.np} ntri={len(tris)} coh={coh} gpu_law={gpu_law} fuse={fuse}", flush=True)
    done = 0
    for target in (1, 5, 30):
        gpu.step(target - done)
        ref.step(target - done)
        done = target
        assert_state_equal(gpu, ref, f"step {target}")
        assert_mesh_history_equal(gpu, ref, f"step {target}")
        w, r = gpu.mesh_wrench(0), ref.mesh["wrench"]
        assert np.allclose(w, r, rtol=1e-11, atol=1e-12 * max(np.abs(r).max(), 1e-300)), (target, w, r)
    assert ref.ms_count.sum() > scene.np // 4
    assert gpu.error_word() == 0
    gpu.close()


def terrain_pose(dbi0, k, dt):
    t = k * dt
    w = 2 * np.pi * 400.0
    d = dbi0.copy()
    d[1] += 1e-4 * np.sin(w * t)
    d[3] += 1.5e-4 * np.sin(w * t)
    d[4] = 1e-4 * w * np.cos(w * t)
    d[6] = 1.5e-4 * w * np.cos(w * t)
    th = 0.001 * np.sin(w * t)
    e0 = dbi0[7:11]
    q = np.array([np.cos(th / 2), np.sin(th / 2), 0.0, 0.0])   # extra turn about x
    # quaternion product q * e0
    a0, a1, a2, a3 = q
    b0, b1, b2, b3 = e0
    e = np.array([a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3, a0 * b1 + a1 * b0 + a2 * b3 - a3 * b2,
                  a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1, a0 * b3 + a1 * b2 - a2 * b1 + a3 * b0])
    d[7:11] = e / np.linalg.norm(e)
    return d   # the rates of dbi0 are kept (the contact law reads them; they need not match the motion)


def test_moving_terrain_matches_oracle():
    """The terrain body gets a new pose (translation and rotation) before every step for 200
    steps; the library places the triangles and proxies from their body-frame coordinates as
    the oracle does (impl:2544-2601), and the lists stay complete through the body's motion."""
    scene, tris, dbi, _, gpu, ref = terrain_pair()
    for k in range(1, 201):
        d = terrain_pose(dbi, k, scene.dt)
        gpu.set_mesh_pose(0, d)
        ref.set_mesh_pose(d)
        gpu.step(1)
        ref.step(1)
        if k % 50 == 0:
            print(f"[moving mesh] step {k}", flush=True)
            assert_state_equal(gpu, ref, f"step {k}")
            assert_mesh_history_equal(gpu, ref, f"step {k}")
    assert gpu.error_word() == 0
    gpu.close()


def test_terrain_mixed_precision_forces():
    """Mixed precision at the FP64 state of step 10: per-sphere force error normalised by the
    sum of the sphere's contact-force magnitudes <= 1e-4; triangle contact sets exact."""
    scene, tris, dbi, cp, gpu64, ref = terrain_pair()
    gpu64.step(10)
    state = {f: gpu64.read(f) for f in STATE}
    gpu64.close()
    _, _, _, _, mixed, _ = terrain_pair(precision="mixed")
    ref = xdem_ref.RefSim(scene)
    ref.add_mesh(tris, dbi, cp)
    for f in STATE:
        mixed.write(f, state[f])
        ref.write(f, state[f])
    mixed.collide()
    ref.collide()
    fm, fr = mixed.read("force"), ref.force
    scale = np.linalg.norm(fr, axis=1) + 1e-300
    has = np.linalg.norm(fr, axis=1) > 0
    err = (np.linalg.norm(fm - fr, axis=1)[has] / scale[has]).max()
    assert err <= 1e-4, err
    c, ids, _ = mixed.history("mesh")
    assert np.array_equal(c, ref.ms_count)
    for i in np.nonzero(c)[0]:
        assert np.array_equal(np.sort(ids[i, : c[i]]), np.sort(ref.ms_id[i, : c[i]])), i
    mixed.close()


def test_wheel_over_20k_bed_matches_oracle():
    scene, tris, dbi, cp = wheel_over_bed(20000)
    print(f"[wheel 20k] np={scene.np} ntri={len(tris)}", flush=True)
    # the default context and one whose pairs go through the pair pass on every step
    gpus = [DemContext(scene, precision="fp64", fuse_velocity=True, meshes=[(0, tris, dbi, cp)], merge_pairs_below=m)
            for m in (0, 1)]
    ref = xdem_ref.RefSim(scene)
    ref.add_mesh(tris, dbi, cp)
    done = 0
    for target in (1, 10, 40):
        for gpu in gpus:
            gpu.step(target - done)
        ref.step(target - done)
        done = target
        for m, gpu in enumerate(gpus):
            assert_state_equal(gpu, ref, f"merge {m} step {target}")
            assert_mesh_history_equal(gpu, ref, f"merge {m} step {target}")
    assert ref.ms_count.sum() > 0
    for gpu in gpus:
        info = gpu.info()
        assert info["ntriangle"] == 4096 and info["mesh_spheres"] > 0
        assert gpu.error_word() == 0
        gpu.close()


def test_wheel_over_full_bed_properties():
    scene, tris, dbi, cp = wheel_over_bed(500000)
    print(f"[wheel 500k] np={scene.np}", flush=True)
    gpu = DemContext(scene, precision="mixed", p2p_capacity=16, fuse_velocity=True, meshes=[(0, tris, dbi, cp)])
    gpu.step(1)
    # first step: the bed spheres are at rest and the wheel is still, so every contact force is
    # the elastic push of an overlap, which lifts the wheel
    w = gpu.mesh_wrench(0)
    assert np.isfinite(w).all() and w[2] > 0, w
    gpu.step(99)
    gpu.synchronize()
    assert gpu.error_word() == 0
    assert np.isfinite(gpu.read("pos")).all()
    info = gpu.info()
    assert info["mesh_spheres"] > 0
    c, _, _ = gpu.history("mesh")
    assert c.sum() > 0
    assert np.isfinite(gpu.mesh_wrench(0)).all()
    gpu.close()
