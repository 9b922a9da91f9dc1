"""Neighbour lists with a skin (grid and lists rebuilt only after a sphere has moved half the
skin) against lists rebuilt every step and against the oracle.

The lists only select candidates; the pairs that pass the reference's predicate are visited
in the reference's order (cell offset, then id), re-sorted when a sphere has changed cells
since the build. So a context with a skin must give bit-identical states and histories
(including the order of the history slots, which is the visiting order) to one without,
and the slot order must equal the oracle's. The scene is a compressed polydisperse pack
with fast random velocities, so spheres change cells and contacts form and break between
list builds.
"""
import numpy as np
import pytest

import xdem_ref
from scenes import random_pack
from xdynamics_dll_amd import DemContext, SceneData, box_walls, contact_property

pytestmark = pytest.mark.gpu

FIELDS = ("pos", "vel", "acc", "ep", "ev", "ea")


def shaken_pack(n=2000, speed=1.0, seed=21):
    s = random_pack(n=n, seed=seed)
    rng = np.random.default_rng(seed)
    vel = rng.normal(0.0, speed, size=(s.np, 3))
    return SceneData.from_positions(s.pos, 2600.0, s.cp_pp, walls=list(zip(s.planes, s.plane_dbi)),
                                    wall_cp=s.plane_cp[0], dt=5e-6, vel=vel)


def run(scene, skin, precision, fuse, steps):
    ctx = DemContext(scene, precision=precision, fuse_velocity=fuse, skin=skin)
    ctx.step(steps)
    out = {f: ctx.read(f) for f in FIELDS}
    out["p2p"] = ctx.history("p2p")
    out["plane"] = ctx.history("plane")
    out["builds"] = ctx.info()["list_builds"]
    out["err"] = ctx.error_word()
    return out


@pytest.mark.parametrize("precision,fuse,skin", [("fp64", True, 1.0e-3), ("fp64", False, 1.0e-3),
                                                 ("mixed", True, 1.0e-3), ("fp64", True, 0.0)])
def test_skin_lists_are_bit_identical_to_rebuilding_every_step(precision, fuse, skin):
    """A fixed 1 mm skin, and the default adaptive skin (its scale changes between builds)."""
    scene = shaken_pack()
    steps = 120
    a = run(scene, skin, precision, fuse, steps)
    b = run(scene, -1.0, precision, fuse, steps)
    print("list builds with skin", a["builds"], "without", b["builds"], flush=True)
    assert a["err"] == 0 and b["err"] == 0
    assert b["builds"] == steps
    assert 1 < a["builds"] < (steps // 4 if skin > 0 else steps)
    for f in FIELDS:
        assert np.array_equal(a[f], b[f]), f
    for h in ("p2p", "plane"):
        for x, y in zip(a[h], b[h]):
            assert np.array_equal(x, y), h


@pytest.mark.parametrize("merge", [0, 1])
def test_skin_history_slot_order_matches_oracle(merge):
    scene = shaken_pack(n=1500, seed=5)
    gpu = DemContext(scene, precision="fp64", fuse_velocity=True, skin=1.0e-3, merge_pairs_below=merge)
    ref = xdem_ref.RefSim(scene)
    for n in (7, 33):
        gpu.step(n)
        ref.step(n)
        c, ids, tsd = gpu.history("p2p")
        assert np.array_equal(c, ref.pp_count)
        for i in range(scene.np):
            assert np.array_equal(ids[i, :c[i]], ref.pp_id[i, :c[i]]), i
        for f in ("pos", "vel", "ep", "ev"):
            g, r = gpu.read(f), getattr(ref, f)
            assert float(np.abs(g - r).max() / np.abs(r).max()) <= 1e-10, (n, f)
    assert gpu.info()["list_builds"] < 40
    assert gpu.error_word() == 0


def bouncing_spheres(n_side=12, seed=8):
    """Spheres falling onto the floor with sideways velocity: no spin in flight, spin from
    the friction of the impact, free spinning flight after the bounce."""
    rng = np.random.default_rng(seed)
    g = (np.arange(n_side) + 0.5) * 3.0e-3
    x, y = np.meshgrid(g, g, indexing="ij")
    r = 1.0e-3
    z = r + rng.uniform(5e-6, 4e-4, size=x.size)
    pos = np.stack([x.ravel(), y.ravel(), z, np.full(x.size, r)], axis=1)
    vel = np.stack([rng.normal(0, 0.3, x.size), rng.normal(0, 0.3, x.size), -rng.uniform(0.3, 1.0, x.size)], axis=1)
    cp = contact_property(1e7, 1e7, 0.3, 0.3, 0.6, 0.5, 0.02, 0.0)
    w = box_walls(n_side * 3.0e-3, n_side * 3.0e-3, 0.03)
    return SceneData.from_positions(pos, 2500.0, cp, walls=w, dt=1e-5, vel=vel)


def test_spin_stable_spheres_through_impact_match_oracle():
    """Rotation writes are skipped for spheres whose rotation state does not change; the
    spheres here go from no spin (skipped) through a frictional impact (full pass) to
    spinning flight (written every step) and must follow the oracle throughout."""
    scene = bouncing_spheres()
    gpu = DemContext(scene, precision="fp64", fuse_velocity=True)
    ref = xdem_ref.RefSim(scene)
    spun = False
    for n in (20, 60, 220):
        gpu.step(n)
        ref.step(n)
        for f in ("pos", "vel", "ep", "ev", "ea"):
            g, r = gpu.read(f), getattr(ref, f)
            s = np.abs(r).max()
            assert s == 0 or float(np.abs(g - r).max() / s) <= 1e-10, (n, f)
        spun = spun or bool(np.abs(ref.ev).max() > 0)
    assert spun
    assert gpu.error_word() == 0


def test_step_graph_replay_is_bit_identical_to_eager_launches():
    """With step_graph = 32, xdem_ctx_step replays a 32-step HIP graph for long runs; the same
    steps launched one by one (step_graph = 0) must give the same bits, also across a state
    write (which forces eager steps until the drift and the buffer parities fit the graph
    again)."""
    scene = shaken_pack(n=1500, seed=9)
    out = []
    for sg in (32, 0):
        ctx = DemContext(scene, precision="fp64", fuse_velocity=True, step_graph=sg)
        ctx.step(70)
        vel = ctx.read("vel")
        vel[:, 2] += 0.01
        ctx.write("vel", vel)
        ctx.step(65)
        out.append({f: ctx.read(f) for f in FIELDS} | {"p2p": ctx.history("p2p")})
        assert ctx.error_word() == 0
    for f in FIELDS:
        assert np.array_equal(out[0][f], out[1][f]), f
    for x, y in zip(out[0]["p2p"], out[1]["p2p"]):
        assert np.array_equal(x, y)


def test_bucket_table_size_does_not_change_results():
    """The bucket grid only selects candidates: a table of 16M buckets (4096 scan tiles,
    scanned in two levels) gives the same bits as the default one (one level) over steps
    that all rebuild the lists."""
    s = random_pack(n=3000, seed=4)
    vel = np.random.default_rng(4).normal(0.0, 1.0, size=(s.np, 3))
    # a 2 x 1 x 1 m box around the 4 x 4 x 8 cm pack: the grid covers the box
    scene = SceneData.from_positions(s.pos, 2600.0, s.cp_pp, walls=box_walls(2.0, 1.0, 1.0), wall_cp=s.plane_cp[0],
                                     dt=5e-6, vel=vel)
    out = []
    for bps in (0, 8192):
        ctx = DemContext(scene, precision="fp64", fuse_velocity=True, buckets_per_sphere=bps, skin=-1.0)
        ctx.step(60)
        info = ctx.info()
        out.append({f: ctx.read(f) for f in FIELDS} | {"p2p": ctx.history("p2p"), "info": info})
        assert ctx.error_word() == 0
    print("buckets", out[0]["info"]["nbucket"], out[1]["info"]["nbucket"],
          "builds", out[0]["info"]["list_builds"], out[1]["info"]["list_builds"], flush=True)
    assert out[0]["info"]["nbucket"] <= 2048 * 4096 < out[1]["info"]["nbucket"]
    assert out[1]["info"]["list_builds"] > 2
    for f in FIELDS:
        assert np.array_equal(out[0][f], out[1][f]), f
    for x, y in zip(out[0]["p2p"], out[1]["p2p"]):
        assert np.array_equal(x, y)


def test_contexts_stepping_concurrently_match_sequential_runs():
    """Three contexts whose steps are enqueued back to back (their rebuild kernels, each
    waiting between its phases, run concurrently on the device) give the same bits as one
    context alone, with no time-out in the error word."""
    scene = shaken_pack(n=1500, seed=12)
    alone = DemContext(scene, precision="fp64", fuse_velocity=True, skin=-1.0)
    alone.step(30)
    ref = {f: alone.read(f) for f in FIELDS}
    assert alone.error_word() == 0
    ctxs = [DemContext(scene, precision=p, fuse_velocity=True, skin=-1.0) for p in ("fp64", "fp64", "mixed")]
    for _ in range(30):
        for c in ctxs:
            c.step(1)
    for c in ctxs:
        c.synchronize()
        assert c.error_word() == 0
        assert c.info()["list_builds"] == 30
    for c in ctxs[:2]:
        for f in FIELDS:
            assert np.array_equal(c.read(f), ref[f]), f


def test_rebuild_wait_limit_is_a_warning_and_the_build_still_completes():
    """With a wait limit of one try, the blocks of the list rebuild that find no work unit left
    give up at once instead of waiting for the phase (XDEM_ERR_GRID_SYNC, a warning). The build
    still completes with the blocks that stay (the block that finishes a phase's last unit
    always passes its own wait), so the run is bit-identical to the default limit: positions,
    velocities and the sphere history after every step, with a list build on every step."""
    from xdynamics_dll_amd._lib import ERR_BUILD_INCOMPLETE, ERR_GRID_SYNC

    scene = shaken_pack(n=1500, seed=13)
    bad = DemContext(scene, precision="fp64", fuse_velocity=True, skin=-1.0, rebuild_wait_limit=1)
    good = DemContext(scene, precision="fp64", fuse_velocity=True, skin=-1.0)
    warned = 0
    for step in range(6):
        bad.step(1)
        good.step(1)
        w = bad.error_word()
        warned |= w
        assert (w & ~ERR_GRID_SYNC) == 0 and not (w & ERR_BUILD_INCOMPLETE), (step, hex(w))
        assert good.error_word() == 0
        for f in ("pos", "vel", "ev"):
            assert np.array_equal(bad.read(f), good.read(f)), (step, f)
        cb, ib, tb = bad.history("p2p")
        cg, ig, tg = good.history("p2p")
        assert np.array_equal(cb, cg), step
        for i in np.nonzero(cg)[0]:
            assert np.array_equal(ib[i, : cg[i]], ig[i, : cg[i]]), (step, i)
            assert np.array_equal(tb[i, : cg[i]], tg[i, : cg[i]]), (step, i)
    print(f"[wait limit 1] warning seen: {bool(warned & ERR_GRID_SYNC)}", flush=True)
    bad.close()
    good.close()
