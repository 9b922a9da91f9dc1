"""GPU parity on the benchmark workloads and the boundary's callers.

- the 10,000-sphere box lattice (seeded generator) in FP64 against the oracle through the
  first contacts with the floor and each other: bit-identical states;
- members of the batched sweep advanced together in one context, each against its own oracle
  run, and the full 256-scene batch with property checks;
- the DHS force model (contact_model 0) in FP64 and mixed precision;
- a wall moved through 1,000 poses (xdem_ctx_set_body_pose) against the oracle;
- a compiled C++ caller of xIntegratorVV::OneStepSimulation and the Python mirror of the
  facade, both bit-identical to DemContext.
"""
import os
import subprocess

import numpy as np
import pytest

import xdem_ref
from conftest import ROOT
from scenes import compressed_lattice
from xdynamics_dll_amd import DemContext, SceneData, xIntegratorVV

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "acc", "ep", "ev", "ea")


def assert_state_equal(gpu, ref, what, scene=0):
    for f in STATE:
        g = gpu.read(f, scene)
        r = getattr(ref, f)
        assert np.array_equal(g, r), (what, f, float(np.abs(g - r).max()))


def test_box_lattice_10k_fp64_matches_oracle():
    """The 20 x 20 x 25 lattice (seed 1001) falls onto the floor from step ~1,500; through
    step 2,200 (~700 sphere and 300 wall contacts) the FP64 step is bit-identical to the
    oracle: positions, velocities, accelerations, Euler parameters and their rates."""
    scene = SceneData.build(1, 1001)
    # the same run with the pairs walked by the pair pass on every step (merge_pairs_below 1)
    gpus = [DemContext(scene, precision="fp64", fuse_velocity=True, merge_pairs_below=m) for m in (0, 1)]
    ref = xdem_ref.RefSim(scene)
    done = 0
    for target in (1, 10, 100, 1500, 1800, 2200):
        print(f"[cfg1 lattice] np={scene.np} step {target}", flush=True)
        for gpu in gpus:
            gpu.step(target - done)
        ref.step(target - done)
        done = target
        for m, gpu in enumerate(gpus):
            assert_state_equal(gpu, ref, f"merge {m} step {target}")
    for gpu in gpus:
        c, _, _ = gpu.history("p2p")
        assert np.array_equal(c, ref.pp_count) and ref.pp_count.sum() > 100
        cw, _, _ = gpu.history("plane")
        assert np.array_equal(cw, ref.pl_count) and ref.pl_count.sum() > 100
        assert gpu.error_word() == 0
        gpu.close()


def test_batch_members_in_one_context_match_oracle():
    """Eight members of the batched sweep (seed 1004, variants 0..7; 2,000-64,000 spheres, own
    materials, walls and time steps) advanced together in one context: each scene's FP64 state
    equals its own oracle run after 1, 10 and 400 steps."""
    scenes = [SceneData.build(4, 1004, variant=k) for k in range(8)]
    print(f"[cfg4 members] np={[s.np for s in scenes]}", flush=True)
    # the default context and one whose pairs go through the pair pass on every step
    gpus = [DemContext(scenes, precision="fp64", fuse_velocity=True, merge_pairs_below=m) for m in (0, 1)]
    refs = [xdem_ref.RefSim(s) for s in scenes]
    done = 0
    for target in (1, 10, 400):
        for gpu in gpus:
            gpu.step(target - done)
        for k, r in enumerate(refs):
            r.step(target - done)
            for m, gpu in enumerate(gpus):
                assert_state_equal(gpu, r, f"merge {m} scene {k} step {target}", scene=k)
        done = target
    contacts = sum(int(r.pp_count.sum() + r.pl_count.sum()) for r in refs)
    print(f"[cfg4 members] contacts at step 400: {contacts}", flush=True)
    assert contacts > 0
    for gpu in gpus:
        assert gpu.error_word() == 0
        gpu.close()


def test_batch_256_scenes_full_size():
    """The whole batched sweep (256 scenes, ~4.4M spheres) in one context, mixed precision,
    200 steps: the state stays finite, no error bit is set, every scene keeps its spheres
    inside its box, and the list builds run."""
    scenes = [SceneData.build(4, 1004, variant=k) for k in range(256)]
    n = sum(s.np for s in scenes)
    print(f"[cfg4 batch] 256 scenes, {n} spheres", flush=True)
    gpu = DemContext(scenes, precision="mixed", p2p_capacity=16, fuse_velocity=True, step_graph=50)
    gpu.step(200)
    gpu.synchronize()
    assert gpu.error_word() == 0
    info = gpu.info()
    assert info["np"] == n and info["list_builds"] >= 1
    for k in (0, 17, 128, 255):
        p = gpu.read("pos", k)
        side = scenes[k].planes[2].xw.x   # the wall at x = side
        assert np.isfinite(p).all()
        assert (p[:, :3] > -p[:, 3:4]).all() and (p[:, :3] < side + p[:, 3:4]).all(), k
    gpu.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dhs_model_matches_oracle(precision):
    """contact_model 0 (DHS: linear spring k_s = k_n s_ratio, damping v_s = beta s_ratio,
    impl:454-474) on the compressed lattice with walls: FP64 forces, histories and states
    bit-identical to the oracle; mixed forces within 1e-4 (normalised per sphere by the sum
    of its pair-force magnitudes), pair sets exact."""
    scene = compressed_lattice(contact_model=0, sratio=0.8)
    gpu = DemContext(scene, precision=precision)
    ref = xdem_ref.RefSim(scene)
    for step in range(3):
        gpu.collide()
        mag = np.zeros(scene.np)
        ref.collide(pair_magnitude=mag)
        f = gpu.read("force")
        if precision == "fp64":
            assert np.array_equal(f, ref.force), step
            assert np.array_equal(gpu.read("moment"), ref.moment), step
        else:
            norm = mag + np.linalg.norm(ref.force, axis=1)
            has = norm > 0
            err = (np.linalg.norm(f - ref.force, axis=1)[has] / norm[has]).max()
            assert err <= 1e-4, (step, err)
        c, ids, _ = gpu.history("p2p")
        assert np.array_equal(c, ref.pp_count)
        for i in range(scene.np):
            assert np.array_equal(np.sort(ids[i, : c[i]]), np.sort(ref.pp_id[i, : ref.pp_count[i]])), i
        assert ref.pp_count.sum() > 0
        gpu.finish()
        ref.finish()
        if precision == "fp64":
            # the next step starts from the oracle's state in both paths
            pass
        for fld in STATE:
            gpu.write(fld, getattr(ref, fld))
    assert gpu.error_word() == 0
    gpu.close()


def floor_pose(dbi0, k, dt):
    """Pose k of the floor: it oscillates in x and z and rocks about y (MM..ED3, with the
    velocity and the Euler-parameter rate of that motion)."""
    t = k * dt
    w = 2 * np.pi * 25.0
    d = dbi0.copy()
    d[1] += 2e-4 * np.sin(w * t)
    d[3] += 1e-4 * np.sin(w * t)
    d[4] = 2e-4 * w * np.cos(w * t)
    d[6] = 1e-4 * w * np.cos(w * t)
    th = 0.02 * np.sin(w * t)
    thd = 0.02 * w * np.cos(w * t)
    d[7:11] = (np.cos(th / 2), 0.0, np.sin(th / 2), 0.0)
    d[11:15] = (-0.5 * thd * np.sin(th / 2), 0.0, 0.5 * thd * np.cos(th / 2), 0.0)
    return d


@pytest.mark.parametrize("fuse", [False, True])
def test_moving_wall_through_1000_poses_matches_oracle(fuse):
    """The floor of the compressed lattice (a dynamic body, 2 kg) is moved to a new pose before
    every step for 1,000 steps, in translation and rotation. The library rebuilds the corners
    from the local corners fixed at the scene's pose (xParticlePlaneContact::update), as the
    oracle does; states and the floor's wrench stay bit-identical to the oracle's."""
    scene = compressed_lattice()
    info, dbi = scene.planes[0], scene.plane_dbi[0].copy()
    dbi[0] = 2.0
    scene.plane_dbi[0] = dbi
    gpu = DemContext(scene, precision="fp64", fuse_velocity=fuse)
    ref = xdem_ref.RefSim(scene)
    for k in range(1, 1001):
        d = floor_pose(dbi, k, scene.dt)
        gpu.set_body_pose(0, 0, d)
        ref.set_plane_pose(0, d)
        gpu.step(1)
        ref.step(1)
        if k % 100 == 0:
            print(f"[moving floor] fuse={fuse} step {k}", flush=True)
            assert_state_equal(gpu, ref, f"step {k}")
            # the wrench is a sum over spheres whose order (unspecified in the reference's
            # thrust::reduce) differs from the oracle's left-to-right sum
            w, r = gpu.wrench()[0], ref.wrench[0]
            assert np.allclose(w, r, rtol=1e-11, atol=1e-12 * np.abs(r).max()), (k, w, r)
    assert ref.pl_count.sum() > 0
    assert gpu.error_word() == 0
    gpu.close()


def test_cpp_facade_caller_matches_context(tmp_path):
    """tests/facade_caller.cpp drives xIntegratorVV::OneStepSimulation from C++ (built by
    __graft_entry__.build()); its positions and velocities after 300 steps of a 2,000-sphere
    dam-break column equal DemContext's with the same default configuration, bit for bit."""
    exe = os.path.join(ROOT, "tests", "_build", "facade_caller")
    assert os.path.exists(exe), "run __graft_entry__.build() first"
    out = tmp_path / "state.bin"
    r = subprocess.run([exe, "2", "1002", "2000", "300", str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = out.read_bytes()
    n, err = np.frombuffer(raw[:8], dtype=np.uint32)
    vals = np.frombuffer(raw[8:], dtype=np.float64)
    pos, vel = vals[: 4 * n].reshape(n, 4), vals[4 * n:].reshape(n, 3)
    assert err == 0
    scene = SceneData.build(2, 1002, np_override=2000)
    ctx = DemContext(scene, precision="fp64")
    ctx.step(300)
    assert np.array_equal(ctx.read("pos"), pos)
    assert np.array_equal(ctx.read("vel"), vel)
    ctx.close()


def test_python_facade_mirror_matches_context():
    """The Python mirror of the facade (xIntegratorVV.OneStepSimulation, SaveStepResult,
    set_dem_data) steps exactly like DemContext."""
    scene = compressed_lattice()
    vv = xIntegratorVV()
    vv.Initialize(scene)
    ctx = DemContext(scene, precision="fp64")
    for k in range(20):
        vv.OneStepSimulation(k * scene.dt, k)
    ctx.step(20)
    res = vv.SaveStepResult(20)
    for f in STATE:
        assert np.array_equal(res[f], ctx.read(f)), f
    vv.set_dem_data(None, res["pos"], res["vel"], None, None, None, None)
    vv.OneStepSimulation(0.0, 21)
    ctx.step(1)
    assert np.array_equal(vv.SaveStepResult(21)["pos"], ctx.read("pos"))
    assert vv.num_particles() == scene.np
    ctx.close()
