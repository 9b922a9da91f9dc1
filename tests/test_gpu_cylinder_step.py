"""GPU parity of cylinder bodies in the step layer (xdem_ctx_add_cylinder) against the oracle
(the FP64 C restatement of cylinder_contact_force_kernel, impl:1327-1460, driven like the
reference's step, cylinders after the walls):

- spheres inside a spinning drum (inner side wall, closed lids from inside; no trigonometry on
  that path): FP64 states and cylinder histories bit-identical over 40 steps, with the adaptive
  neighbour lists and with lists rebuilt every step, a dynamic drum's wrench within the
  reduction's rounding;
- the drum turned about its axis and shaken along it before every step (pose updates that keep
  or wear out the lists);
- spheres around a rod (outer side wall, lid planes, rims: atan, sin and cos from the device and
  the host libraries), with and without a lid ring and unequal radii: states within 1e-11 of
  their scale, histories' counts and ids exact;
- two scenes with four cylinders added out of scene order, one scene without a cylinder of id
  0 (the slots are never restarted and overflow, as the reference's);
- mixed precision: forces within 1e-4 of the oracle's, cylinder contact sets exact;
- a part file with the cylinder history, restored into a new context, continuing bit for bit.
"""
import numpy as np
import pytest

import xdem_ref
import xdem_part_ref
from cylinder_scenes import Y_TO_Z, drum, quat_mul, rate_of, rod
from xdynamics_dll_amd import DemContext, _lib, make_cylinder

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "acc", "ep", "ev", "ea")


def assert_state(gpu, ref, what, scene=0, rtol=0.0):
    for f in STATE:
        g, r = gpu.read(f, scene), getattr(ref, f)
        assert np.isfinite(r).all(), (what, f, "oracle")
        assert np.isfinite(g).all(), (what, f, "gpu")
        if rtol == 0.0:
            assert np.array_equal(g, r), (what, f, float(np.abs(g - r).max()))
        else:
            scale = max(float(np.abs(r).max()), 1e-300)
            assert np.abs(g - r).max() <= rtol * scale, (what, f, float(np.abs(g - r).max()), scale)


def assert_cyl_history(gpu, ref, what, scene=0, tsd_rtol=0.0):
    c, ids, tsd = gpu.history("cylinder", scene)
    assert np.array_equal(c, ref.cy_count), (what, int(np.abs(c.astype(int) - ref.cy_count).max()))
    scale = max(float(np.abs(ref.cy_tsd).max()), 1e-300)
    for i in np.nonzero(c)[0]:
        n = min(int(c[i]), 3)
        assert np.array_equal(ids[i, :n], ref.cy_id[i, :n]), (what, i)
        if tsd_rtol == 0.0:
            assert np.array_equal(tsd[i, :n], ref.cy_tsd[i, :n]), (what, i)
        else:
            assert np.abs(tsd[i, :n] - ref.cy_tsd[i, :n]).max() <= tsd_rtol * scale, (what, i)


def drum_pair(skin=0.0, precision="fp64", mass=0.0, **kw):
    scene, info, dbi, cp = drum(mass=mass, **kw)
    gpu = DemContext(scene, precision=precision, fuse_velocity=True, skin=skin, cylinders=[(0, info, dbi, cp)])
    ref = xdem_ref.RefSim(scene)
    ref.add_cylinder(info, dbi, cp)
    return scene, info, dbi, cp, gpu, ref


@pytest.mark.parametrize("skin,coh,mass", [(0.0, 0.0, 0.0), (-1.0, 0.0, 0.0), (0.0, 0.05, 4.0)])
def test_drum_steps_match_oracle(skin, coh, mass):
    scene, info, dbi, cp, gpu, ref = drum_pair(skin=skin, coh=coh, mass=mass)
    print(f"[cylinder step] drum np={scene.np} skin={skin} coh={coh} mass={mass}", flush=True)
    done = 0
    for target in (1, 5, 40):
        gpu.step(target - done)
        ref.step(target - done)
        done = target
        assert_state(gpu, ref, f"step {target}")
        assert_cyl_history(gpu, ref, f"step {target}")
        w, r = gpu.cylinder_wrench(0), ref.cylinders[0]["wrench"]
        if mass > 0.0:
            assert np.abs(r).max() > 0.0
            assert np.allclose(w, r, rtol=1e-11, atol=1e-12 * np.abs(r).max()), (target, w, r)
        else:
            assert not w.any()
    # contacts with the side wall and with both lids (ids 0 and 1000)
    ids = set(ref.cy_id[ref.cy_count > 0, 0].tolist())
    assert {0, 1000} <= ids, ids
    assert ref.overflow["cylinder"] == 0
    assert gpu.error_word() == 0
    gpu.close()


def test_moving_drum_matches_oracle():
    """Before every step the drum turns about its axis (the lists stay valid) and, every fifth
    step, shifts along it by 50 um (the lists wear out)."""
    scene, info, dbi0, cp, gpu, ref = drum_pair()
    w = 20.0
    print(f"[cylinder step] moving drum np={scene.np}", flush=True)
    for k in range(60):
        th = w * k * scene.dt
        e = quat_mul(np.array([np.cos(th / 2), 0.0, 0.0, np.sin(th / 2)]), Y_TO_Z)   # a turn about z first
        d = dbi0.copy()
        d[7:11] = e
        d[11:15] = rate_of(e, [0.0, 0.0, w])
        d[3] += 5e-5 * ((k // 5) % 2)
        gpu.set_cylinder_pose(0, d)
        ref.set_cylinder_pose(0, d)
        gpu.step(1)
        ref.step(1)
        if k % 10 == 9:
            assert_state(gpu, ref, f"step {k}")
            assert_cyl_history(gpu, ref, f"step {k}")
    # the shifts wore the lists out: builds beyond the first (12 shifts of 50 um against an
    # adaptive skin of up to 0.4 of the smallest radius)
    assert gpu.info()["list_builds"] >= 3
    assert gpu.error_word() == 0
    gpu.close()


@pytest.mark.parametrize("r_top,thickness", [(None, 0.0), (0.0055, 0.0), (0.0055, 0.002)])
def test_rod_steps_match_oracle(r_top, thickness):
    scene, info, dbi, cp = rod(r_top=r_top, thickness=thickness)
    print(f"[cylinder step] rod np={scene.np} r_top={r_top} thickness={thickness}", flush=True)
    gpu = DemContext(scene, fuse_velocity=True, cylinders=[(0, info, dbi, cp)])
    ref = xdem_ref.RefSim(scene)
    ref.add_cylinder(info, dbi, cp)
    done = 0
    for target in (1, 5, 30):
        gpu.step(target - done)
        ref.step(target - done)
        done = target
        assert_state(gpu, ref, f"step {target}", rtol=1e-11)
        assert_cyl_history(gpu, ref, f"step {target}", tsd_rtol=1e-9)
    assert (ref.cy_count > 0).sum() > scene.np // 2
    assert gpu.error_word() == 0
    gpu.close()


def test_two_scenes_cylinders_out_of_order():
    """Scene 0: a drum (id 0) with an inner rod (id 1) appending to its slots; scene 1: a drum
    with id 1 and a ring with id 2 and no id 0, so its slots are never restarted and overflow;
    the cylinders are added as (1, drum), (0, drum), (1, ring), (0, rod)."""
    s0, d0, dbi0, cp = drum(n=300, seed=41)
    s1, d1, dbi1, _ = drum(n=300, seed=42, cyl_id=1, center=(0.3, 0.08, 0.05))
    rod0, rdbi0 = make_cylinder([0.1, 0.08 - 0.008, 0.05], [0.1, 0.08 + 0.008, 0.05], 0.004, 0.004, cyl_id=1)
    rdbi0[7:11] = Y_TO_Z
    # a ring just wider than the drum: its lids are the drum's (ids 1001 and 1002 both stored)
    ring1, gdbi1 = make_cylinder([0.3, 0.08 - 0.008, 0.05], [0.3, 0.08 + 0.008, 0.05], 0.0202, 0.0202, cyl_id=2)
    gdbi1[7:11] = Y_TO_Z
    # keep scene 0's spheres clear of the rod
    keep = np.hypot(s0.pos[:, 0] - 0.1, s0.pos[:, 1] - 0.08) > 0.004 + s0.pos[:, 3] * 1.001
    s0 = type(s0).from_positions(s0.pos[keep], 2500.0, cp, dt=s0.dt, vel=s0.vel[keep], gravity=s0.gravity)
    cyls = [(1, d1, dbi1, cp), (0, d0, dbi0, cp), (1, ring1, gdbi1, cp), (0, rod0, rdbi0, cp)]
    print(f"[cylinder step] two scenes np={s0.np}+{s1.np}", flush=True)
    gpu = DemContext([s0, s1], fuse_velocity=True, cylinders=cyls)
    refs = [xdem_ref.RefSim(s0), xdem_ref.RefSim(s1)]
    refs[0].add_cylinder(d0, dbi0, cp)
    refs[0].add_cylinder(rod0, rdbi0, cp)
    refs[1].add_cylinder(d1, dbi1, cp)
    refs[1].add_cylinder(ring1, gdbi1, cp)
    done = 0
    for target in (1, 4, 20):
        gpu.step(target - done)
        for r in refs:
            r.step(target - done)
        done = target
        for si, r in enumerate(refs):
            assert_state(gpu, r, f"scene {si} step {target}", scene=si)
            assert_cyl_history(gpu, r, f"scene {si} step {target}", scene=si)
    assert refs[1].overflow["cylinder"] > 0 and refs[0].overflow["cylinder"] == 0
    assert gpu.error_word() & _lib.ERR_CYL_OVERFLOW
    gpu.close()


def test_drum_mixed_precision_forces():
    """Mixed precision at the FP64 state of step 10: per-sphere force error over the oracle's
    force <= 1e-4; cylinder contact sets exact."""
    scene, info, dbi, cp, gpu64, _ = drum_pair()
    gpu64.step(10)
    state = {f: gpu64.read(f) for f in STATE}
    gpu64.close()
    _, _, _, _, mixed, ref = drum_pair(precision="mixed")
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
    c, ids, _ = mixed.history("cylinder")
    assert np.array_equal(c, ref.cy_count)
    for i in np.nonzero(c)[0]:
        assert np.array_equal(np.sort(ids[i, : c[i]]), np.sort(ref.cy_id[i, : c[i]])), i
    mixed.close()


def test_part_file_with_cylinder_history(tmp_path):
    scene, info, dbi, cp, gpu, _ = drum_pair()
    gpu.step(20)
    path = str(tmp_path / "part.bin")
    gpu.save_part(path, 2e-5)
    with open(path, "rb") as f:
        buf = f.read()
    parts = xdem_part_ref.parse(buf, {"p2p": gpu.capacity, "cylinder": 3})
    c, ids, tsd = gpu.history("cylinder")
    assert np.array_equal(parts["cylinder_count"], c)
    assert np.array_equal(parts["cylinder_id"], ids) and np.array_equal(parts["cylinder_tsd"], tsd)
    other = DemContext(scene, fuse_velocity=True, cylinders=[(0, info, dbi, cp)])
    assert other.load_part(path, history=True) == 2e-5
    gpu.step(10)
    other.step(10)
    for f in STATE:
        assert np.array_equal(gpu.read(f), other.read(f)), f
    c2, ids2, _ = other.history("cylinder")
    assert np.array_equal(c2, gpu.history("cylinder")[0])
    gpu.close()
    other.close()
