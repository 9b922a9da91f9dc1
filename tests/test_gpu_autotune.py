"""The launch-shape tuner (xdem_autotune) and the claim it rests on: the full-pass block
count, the bucket budget and the adaptive skin's cap change the speed of a step, never its
results (bit for bit)."""
import numpy as np
import pytest

from scenes import compressed_lattice, random_pack
from xdynamics_dll_amd import DemContext, autotune

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "acc", "ep", "ev", "ea")


@pytest.mark.parametrize("fb,bps,cap", [(64, 2, 0.0), (256, 8, 0.4), (2048, 4, 0.8)])
def test_launch_shapes_do_not_change_results(fb, bps, cap):
    scene = random_pack(n=3000)
    ref = DemContext(scene, precision="fp64", fuse_velocity=True)
    ref.step(30)
    ctx = DemContext(scene, precision="fp64", fuse_velocity=True, full_blocks=fb, buckets_per_sphere=bps,
                     skin_cap=cap)
    ctx.step(30)
    for f in STATE:
        assert np.array_equal(ctx.read(f), ref.read(f)), (fb, bps, f)
    c0, i0, t0 = ref.history("p2p")
    c1, i1, t1 = ctx.history("p2p")
    assert np.array_equal(c0, c1)
    for i in np.nonzero(c0)[0]:
        assert np.array_equal(i0[i, : c0[i]], i1[i, : c1[i]]) and np.array_equal(t0[i, : c0[i]], t1[i, : c1[i]])
    assert ctx.error_word() == 0


def test_autotune_picks_a_candidate_and_caches_it(tmp_path):
    scenes = [compressed_lattice(), random_pack(n=2000)]
    cache = str(tmp_path / "tune.txt")
    r = autotune(scenes, precision="mixed", warmup=5, steps=10, cache_path=cache)
    print(f"[autotune] {r['full_blocks']} blocks, {r['buckets_per_sphere']} buckets/sphere, "
          f"{len(r['trials'])} trials", flush=True)
    assert not r["from_cache"]
    assert r["full_blocks"] in (256, 512, 1024, 2048) and r["buckets_per_sphere"] in (2, 4, 8)
    assert r["skin_cap"] in (0.2, 0.4, 0.8)
    assert 4 <= len(r["trials"]) <= 8
    assert r["ms_per_step"] == min(t["ms_per_step"] for t in r["trials"]) > 0
    again = autotune(scenes, precision="mixed", warmup=5, steps=10, cache_path=cache)
    assert again["from_cache"] and again["full_blocks"] == r["full_blocks"]
    assert again["buckets_per_sphere"] == r["buckets_per_sphere"] and again["key"] == r["key"]
    assert again["skin_cap"] == r["skin_cap"]
    # another kind of scene has another key
    other = autotune([compressed_lattice(n_side=4)], precision="fp64", warmup=2, steps=4, cache_path=cache)
    assert other["key"] != r["key"] and not other["from_cache"]
