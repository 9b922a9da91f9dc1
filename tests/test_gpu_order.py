"""The stored sphere order (xdem_ctx_config.keep_order): a context stores each scene's spheres
in the bucket order of their initial positions, so that a caller's order (a file's, a
generator's, a shuffle) does not scatter the neighbours a kernel gathers; every call that
takes or returns per-sphere data uses the caller's order.

- a shuffled pack stepped with the stored order and with the caller's order: states and
  histories equal bit for bit (the visiting order breaks its ties by the caller's ids), FP64
  and mixed, one scene and two;
- the shuffled pack against the oracle run in the same (shuffled) ids;
- writes and reads of fields and histories are each other's inverse;
- xdem_ctx_sphere_order is a permutation, the identity with keep_order;
- xdem_ctx_reorder between steps (plain packs in both step modes, the coupled wheel loop with
  its triangle and wrench slots) leaves the run bit-identical to an uninterrupted one.
"""
import dataclasses

import numpy as np
import pytest

import xdem_ref
from scenes import compressed_lattice, random_pack
from xdynamics_dll_amd import DemContext

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "acc", "ep", "ev", "ea")


def shuffled(scene, seed):
    perm = np.random.default_rng(seed).permutation(scene.np)
    return dataclasses.replace(scene, pos=scene.pos[perm].copy(), mass=scene.mass[perm].copy(),
                               inertia=scene.inertia[perm].copy(),
                               vel=None if scene.vel is None else scene.vel[perm].copy())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_stored_order_equals_callers_order(precision):
    scene = shuffled(random_pack(n=3000), 11)
    print(f"[order] np={scene.np} precision={precision}", flush=True)
    a = DemContext(scene, precision=precision, fuse_velocity=True)
    b = DemContext(scene, precision=precision, fuse_velocity=True, keep_order=True)
    oa, ob = a.sphere_order(), b.sphere_order()
    assert np.array_equal(np.sort(oa), np.arange(scene.np)) and not np.array_equal(oa, np.arange(scene.np))
    assert np.array_equal(ob, np.arange(scene.np))
    for n in (1, 7, 30):
        a.step(n)
        b.step(n)
        for f in STATE:
            assert np.array_equal(a.read(f), b.read(f)), (precision, n, f)
        for kind in ("p2p", "plane"):
            for x, y in zip(a.history(kind), b.history(kind)):
                assert np.array_equal(x, y), (precision, n, kind)
    assert a.error_word() == 0 and b.error_word() == 0
    a.close()
    b.close()


def test_stored_order_two_scenes():
    scenes = [shuffled(compressed_lattice(), 5), shuffled(random_pack(n=2000, seed=4), 6)]
    a = DemContext(scenes, precision="fp64", fuse_velocity=True, merge_pairs_below=1)
    b = DemContext(scenes, precision="fp64", fuse_velocity=True, keep_order=True)
    a.step(12)
    b.step(12)
    for s in range(2):
        for f in STATE:
            assert np.array_equal(a.read(f, s), b.read(f, s)), (s, f)
        for x, y in zip(a.history("p2p", s), b.history("p2p", s)):
            assert np.array_equal(x, y), s
    a.close()
    b.close()


def test_shuffled_pack_matches_oracle():
    scene = shuffled(random_pack(n=3000), 12)
    gpu = DemContext(scene, precision="fp64", fuse_velocity=True)
    ref = xdem_ref.RefSim(scene)
    for n in (1, 10, 40):
        gpu.step(n)
        ref.step(n)
        for f in ("pos", "vel", "ep", "ev"):
            g, r = gpu.read(f), getattr(ref, f)
            assert float(np.abs(g - r).max() / np.abs(r).max()) <= 1e-10, (n, f)
        c, ids, _ = gpu.history("p2p")
        assert np.array_equal(c, ref.pp_count)
        for i in range(scene.np):
            assert np.array_equal(ids[i, : c[i]], ref.pp_id[i, : c[i]]), (n, i)
    assert ref.pp_count.sum() > 0
    assert gpu.error_word() == 0
    gpu.close()


def test_writes_and_reads_are_inverse():
    scene = shuffled(random_pack(n=2000, seed=9), 13)
    ctx = DemContext(scene, precision="fp64", fuse_velocity=True, p2p_capacity=12)
    rng = np.random.default_rng(2)
    v = rng.normal(size=(scene.np, 3))
    ctx.write("vel", v)
    assert np.array_equal(ctx.read("vel"), v)
    cap = 12
    count = rng.integers(0, 4, scene.np).astype(np.uint32)
    ids = np.zeros((scene.np, cap), dtype=np.uint32)
    tsd = np.zeros((scene.np, cap, 2))
    for i in range(scene.np):
        ids[i, : count[i]] = rng.choice(scene.np, count[i], replace=False)
        tsd[i, : count[i]] = rng.normal(size=(count[i], 2))
    ctx.write_history("p2p", count, ids, tsd)
    c2, i2, t2 = ctx.history("p2p")
    assert np.array_equal(c2, count)
    for i in range(scene.np):
        assert np.array_equal(i2[i, : count[i]], ids[i, : count[i]]), i
        assert np.array_equal(t2[i, : count[i]], tsd[i, : count[i]]), i
    ctx.close()


@pytest.mark.parametrize("precision,fuse", [("fp64", True), ("fp64", False), ("mixed", True)])
def test_reorder_mid_run_is_bit_identical(precision, fuse):
    """xdem_ctx_reorder between steps gathers every carried per-sphere array into the order of
    the current positions and rebuilds the lists: the run goes on bit for bit."""
    scene = shuffled(random_pack(n=3000), 14)
    a = DemContext(scene, precision=precision, fuse_velocity=fuse, keep_order=True)
    b = DemContext(scene, precision=precision, fuse_velocity=fuse, keep_order=True)
    a.step(13)
    a.reorder()
    assert not np.array_equal(a.sphere_order(), np.arange(scene.np))
    a.step(20)
    a.reorder()
    a.step(7)
    b.step(40)
    for f in STATE:
        assert np.array_equal(a.read(f), b.read(f)), (precision, fuse, f)
    for kind in ("p2p", "plane"):
        for x, y in zip(a.history(kind), b.history(kind)):
            assert np.array_equal(x, y), (precision, fuse, kind)
    assert a.error_word() == 0 and b.error_word() == 0
    a.close()
    b.close()


def test_reorder_in_the_coupled_wheel_loop():
    """The wheel over a 20,000-sphere bed (triangle slots, the body's wrench slots) re-sorted
    mid-run: the coupled track and the bed's state equal an uninterrupted run's."""
    from mesh_scenes import wheel_over_bed
    from xdynamics_dll_amd.mbd import HhtWheel, cosim_wheel_steps

    scene, tris, dbi, cp = wheel_over_bed(20000)
    dbi = dbi.copy()
    dbi[13] = 1.0
    runs = []
    for cut in (True, False):
        ctx = DemContext(scene, precision="fp64", fuse_velocity=True, step_graph=50, meshes=[(0, tris, dbi, cp)])
        w = HhtWheel(40.0, (0.258, 0.45, 0.258), dbi, 2.0, scene.dt, alpha=-0.1, gravity=scene.gravity)
        if cut:
            t1 = cosim_wheel_steps(ctx, w, 0, 12)
            ctx.reorder()
            t2 = cosim_wheel_steps(ctx, w, 0, 12)
            track = np.concatenate([t1, t2])
        else:
            track = cosim_wheel_steps(ctx, w, 0, 24)
        runs.append((track, {f: ctx.read(f) for f in STATE}, ctx.history("mesh"), ctx.history("p2p")))
        assert ctx.error_word() == 0
        ctx.close()
    (ta, sa, ma, pa), (tb, sb, mb, pb) = runs
    assert np.array_equal(ta, tb)
    assert np.abs(ta[:, 5]).max() > 0
    for f in STATE:
        assert np.array_equal(sa[f], sb[f]), f
    for x, y in zip(ma + pa, mb + pb):
        assert np.array_equal(x, y)


def test_reorder_with_cylinders_and_two_scenes():
    """A drum (cylinder slots) next to a shuffled pack in one context, re-sorted mid-run: both
    scenes go on bit for bit."""
    from cylinder_scenes import drum

    dscene, info, dbi, cp = drum()
    pack = shuffled(random_pack(n=2000, seed=8), 15)
    runs = []
    for cut in (True, False):
        ctx = DemContext([dscene, pack], precision="fp64", fuse_velocity=True, cylinders=[(0, info, dbi, cp)])
        if cut:
            ctx.step(9)
            ctx.reorder()
            ctx.step(21)
        else:
            ctx.step(30)
        runs.append(({(f, s): ctx.read(f, s) for f in STATE for s in (0, 1)},
                     ctx.history("cylinder", 0), ctx.history("p2p", 0), ctx.history("p2p", 1)))
        assert ctx.error_word() == 0
        ctx.close()
    (sa, *ha), (sb, *hb) = runs
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    for x, y in zip(ha, hb):
        for u, v in zip(x, y):
            assert np.array_equal(u, v)
    assert runs[0][1][0].sum() > 0   # the drum's slots are in use
