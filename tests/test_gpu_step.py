"""GPU parity of the step layer (xdem_ctx) against the oracle (FP64 C restatement).

FP64 path: forces, moments, rolling sums, histories and states equal the oracle's to
1e-12 relative (the two compute the same operations in the same order; the bound only
absorbs libm differences in log/pow); pair sets are compared bit-exactly.
Mixed path: per-sphere net-force error normalised by the sum of that sphere's pair-force
magnitudes (FP64) <= 1e-4, pair sets bit-exact at the same input state.

Sphere pairs are walked by the contact kernel itself when the full pass is small and by the
pair pass (k_pairs, lanes packed over the pairs of 64 spheres) when it holds at least
merge_pairs_below spheres; these scenes are below the default threshold, so the tests run
with merge_pairs_below = 1 (every step through the pair pass) as well as with the default.
"""
import numpy as np
import pytest

import xdem_ref
from scenes import block_scene, compressed_lattice, random_pack, sphere_drop_scene
from xdynamics_dll_amd import DemContext

pytestmark = pytest.mark.gpu

SCENES = {
    "lattice": lambda: compressed_lattice(),
    "block": block_scene,
    "drop": sphere_drop_scene,
    "cohesive_pack": lambda: random_pack(n=3000),
}


def rel(a, b):
    s = np.abs(b).max()
    return 0.0 if s == 0 else float(np.abs(a - b).max() / s)


def pair_dict(count, ids, tsd, i):
    return {int(ids[i, s]): tuple(tsd[i, s]) for s in range(count[i])}


def compare_histories(gpu, ref, tol=1e-12):
    c, ids, tsd = gpu.history("p2p")
    assert np.array_equal(c, ref.pp_count)
    for i in range(ref.np):
        g = pair_dict(c, ids, tsd, i)
        r = pair_dict(ref.pp_count, ref.pp_id, ref.pp_tsd, i)
        assert g.keys() == r.keys(), i
        for k in g:
            assert np.allclose(g[k], r[k], rtol=tol, atol=1e-300), (i, k, g[k], r[k])
    c, ids, tsd = gpu.history("plane")
    assert np.array_equal(c, ref.pl_count)
    assert np.array_equal(ids, ref.pl_id)
    assert np.allclose(tsd, ref.pl_tsd, rtol=tol, atol=1e-300)


# merge_pairs_below: 0 = the library's threshold (these scenes: the contact kernel walks the
# pairs), 1 = the pair pass on every step
REGIMES = [0, 1]


@pytest.mark.parametrize("merge", REGIMES)
@pytest.mark.parametrize("name", list(SCENES))
def test_fp64_collide_matches_oracle(name, merge):
    scene = SCENES[name]()
    gpu = DemContext(scene, precision="fp64", merge_pairs_below=merge)
    ref = xdem_ref.RefSim(scene)
    print(f"[fp64 collide] scene={name} np={scene.np} merge={merge}", flush=True)
    for step in range(3):
        gpu.collide()
        ref.collide()
        assert rel(gpu.read("pos"), ref.pos) == 0.0
        for f in ("force", "moment", "tmax", "rres"):
            e = rel(gpu.read(f), getattr(ref, f))
            assert e <= 1e-12, (name, step, f, e)
        compare_histories(gpu, ref)
        gpu.finish()
        ref.finish()
        for f in ("vel", "acc", "ev", "ea", "ep"):
            e = rel(gpu.read(f), getattr(ref, f))
            assert e <= 1e-12, (name, step, f, e)
    assert gpu.error_word() == 0
    gpu.close()


@pytest.mark.parametrize("merge", REGIMES)
@pytest.mark.parametrize("fuse", [False, True])
@pytest.mark.parametrize("name", ["lattice", "block"])
def test_fp64_fused_steps_match_oracle(name, fuse, merge):
    scene = SCENES[name]()
    gpu = DemContext(scene, precision="fp64", fuse_velocity=fuse, merge_pairs_below=merge)
    ref = xdem_ref.RefSim(scene)
    for n in (1, 9, 40):
        gpu.step(n)
        ref.step(n)
        for f in ("pos", "vel", "ep", "ev"):
            e = rel(gpu.read(f), getattr(ref, f))
            assert e <= 1e-10, (name, n, f, e)
    assert gpu.error_word() == 0


@pytest.mark.parametrize("merge", REGIMES)
@pytest.mark.parametrize("name", ["lattice", "block", "cohesive_pack"])
def test_mixed_force_error_and_pairs(name, merge):
    scene = SCENES[name]()
    gpu = DemContext(scene, precision="mixed", merge_pairs_below=merge)
    ref = xdem_ref.RefSim(scene)
    mag = np.zeros(scene.np)
    gpu.collide()
    ref.collide(pair_magnitude=mag)
    c, ids, _ = gpu.history("p2p")
    for i in range(scene.np):
        assert np.array_equal(np.sort(ids[i, : c[i]]), np.sort(ref.pp_id[i, : ref.pp_count[i]])), i
    f = gpu.read("force")
    has = mag > 0
    # wall forces are part of the net force: normalise by sphere-pair + own magnitude
    norm = mag + np.linalg.norm(ref.force, axis=1)
    err = np.linalg.norm(f - ref.force, axis=1)[has] / norm[has]
    print(f"[mixed] {name}: max normalised force error {err.max():.3e}")
    assert err.max() <= 1e-4


@pytest.mark.parametrize("merge", REGIMES)
@pytest.mark.parametrize("precision,fuse", [("fp64", False), ("fp64", True), ("mixed", True)])
def test_multi_scene_equals_single_scenes(precision, fuse, merge):
    a, b = compressed_lattice(), block_scene()
    both = DemContext([a, b], precision=precision, fuse_velocity=fuse, merge_pairs_below=merge)
    sa = DemContext(a, precision=precision, fuse_velocity=fuse)
    sb = DemContext(b, precision=precision, fuse_velocity=fuse)
    both.step(5)
    sa.step(5)
    sb.step(5)
    for f in ("pos", "vel", "ev"):
        assert np.array_equal(both.read(f, 0), sa.read(f)), f
        assert np.array_equal(both.read(f, 1), sb.read(f)), f


@pytest.mark.parametrize("fuse", [False, True])
def test_graph_replay_equals_eager(fuse):
    scene = compressed_lattice()
    eager = DemContext(scene, precision="fp64", fuse_velocity=fuse)
    graph = DemContext(scene, precision="fp64", fuse_velocity=fuse)
    graph.graph_capture(2)
    for k in range(5):
        graph.graph_launch(1)
        eager.step(2)
        graph.synchronize()
        for f in ("pos", "vel", "ev"):
            assert np.array_equal(graph.read(f), eager.read(f)), (k, f)


@pytest.mark.parametrize("merge", REGIMES)
def test_history_overflow_is_flagged_not_written_past_capacity(merge):
    scene = compressed_lattice()
    gpu = DemContext(scene, precision="fp64", p2p_capacity=3, merge_pairs_below=merge)
    ref = xdem_ref.RefSim(scene, capacity=3)
    gpu.collide()
    ref.collide()
    assert gpu.error_word() & 0x1
    assert ref.overflow["p2p"] > 0
    assert rel(gpu.read("force"), ref.force) <= 1e-12
    c, ids, _ = gpu.history("p2p")
    assert c.max() == 3
    assert np.array_equal(c, ref.pp_count)


@pytest.mark.parametrize("q", [dict(gpu_force_law=1), dict(velocity_not_updated=1), dict(stale_ft=0),
                               dict(flat_inertia=0)])
@pytest.mark.parametrize("merge", REGIMES)
def test_quirk_profiles_match_oracle(q, merge):
    scene = compressed_lattice()
    gpu = DemContext(scene, precision="fp64", quirks=q, merge_pairs_below=merge)
    ref = xdem_ref.RefSim(scene, quirks=q)
    gpu.step(4)
    ref.step(4)
    for f in ("pos", "vel", "ev"):
        assert rel(gpu.read(f), getattr(ref, f)) <= 1e-10, (q, f)


def test_dynamic_wall_wrench_matches_oracle():
    scene = compressed_lattice()
    scene.plane_dbi[0] = scene.plane_dbi[0].copy()
    scene.plane_dbi[0][0] = 1.0   # floor with mass: a dynamic body
    gpu = DemContext(scene, precision="fp64")
    ref = xdem_ref.RefSim(scene)
    gpu.step(2)
    ref.step(2)
    w = gpu.wrench()
    assert np.allclose(w[0], ref.wrench[0], rtol=1e-12, atol=1e-18)
    assert np.all(w[1:] == 0.0)


@pytest.mark.parametrize("fuse", [False, True])
def test_state_written_between_steps_matches_oracle(fuse):
    """Writing the state after steps discards the drift the last step precomputed (and the
    bucket counts it added): the following steps must follow the new state."""
    scene = compressed_lattice()
    gpu = DemContext(scene, precision="fp64", fuse_velocity=fuse)
    ref = xdem_ref.RefSim(scene)
    gpu.step(3)
    ref.step(3)
    rng = np.random.default_rng(11)
    vel = ref.vel + rng.normal(0, 0.01, size=ref.vel.shape)
    pos = ref.pos.copy()
    pos[:, :3] += rng.normal(0, 2e-5, size=(scene.np, 3))
    for name, val in (("vel", vel), ("pos", pos)):
        gpu.write(name, val)
        ref.write(name, val)
    gpu.step(4)
    ref.step(4)
    for f in ("pos", "vel", "ev"):
        assert rel(gpu.read(f), getattr(ref, f)) <= 1e-10, f
    assert gpu.error_word() == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pair_pass_and_contact_kernel_walks_are_bit_identical(precision):
    """The same scene stepped with the pairs walked by the contact kernel and by the pair pass:
    states and histories equal bit for bit (both add the pairs in rank order with the same
    operations)."""
    scene = random_pack(n=3000)
    a = DemContext(scene, precision=precision, fuse_velocity=True)
    b = DemContext(scene, precision=precision, fuse_velocity=True, merge_pairs_below=1)
    for n in (1, 7, 30):
        a.step(n)
        b.step(n)
        for f in ("pos", "vel", "ep", "ev", "acc", "ea"):
            assert np.array_equal(a.read(f), b.read(f)), (precision, n, f)
        for kind in ("p2p", "plane"):
            for x, y in zip(a.history(kind), b.history(kind)):
                assert np.array_equal(x, y), (precision, n, kind)
    assert a.error_word() == 0 and b.error_word() == 0
    a.close()
    b.close()
