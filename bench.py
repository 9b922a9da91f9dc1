"""Benchmark of the DEM time step on one MI355X: the 1,000,000-sphere dam-break column.

Workload (seeded): 1,000,000 spheres, r ~ U[1, 2] mm, rho = 2600, E = 5e7 Pa, nu = 0.25,
e = 0.4, mu = 0.5, mu_r = 0.05, JKR cohesion 0.05 J/m^2 between spheres, packed by
sequential insertion in a 0.3 x 0.3 x 0.6 m column inside a 1.2 x 0.3 x 0.7 m five-wall box
(the gate is absent from t = 0), dt = 5e-6 s. The timed region is K full steps of the mixed
precision path (FP32 contact arithmetic; FP64 positions, Verlet state, tangential history
and force sums), all state resident in HBM. The same run times the FP64 path, measures the
mixed path's force error against the FP64 path at the same state, the contact kernel's
duration with HIP events on its stream, and the CPU restatement on the host cores.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0
SEED = 1002


def step_bytes(cbar: float) -> float:
    """Algorithmic HBM bytes per sphere-step of the whole step, FP64 state in the reference
    layout: state read 208 + state write 176 + grid 36 + history 2*(8 + 20*cbar)."""
    return 436.0 + 40.0 * cbar


def merged_collide_kernel_bytes(spheres: float, wall_contacts: float, pairs: float) -> float:
    """k_collide walking the pairs too: its own bytes without the pair sums' read (480 - 80),
    plus the pair walk's pass mask 8 and old and new counts 4 + 4 per sphere and 44 per
    passing pair (the own state is read once)."""
    return (480.0 - 80.0 + 16.0) * spheres + 40.0 * wall_contacts + 44.0 * pairs


# Algorithmic bytes per sphere of k_lite_move (the translation of every sphere outside the
# step's full pass): the translational part of the per-sphere state traffic, read pos 32,
# vel 24, acc 24, mass 8 and write pos 32, vel 24, acc 24.
LITE_MOVE_BYTES = 168.0


def pairs_kernel_bytes(spheres: float, pairs: float) -> float:
    """Algorithmic bytes of one k_pairs launch: per full-pass sphere its own pos 32, vel 24,
    mass 8, ep 32, ev 32, list flags 4, pass mask 8, old count 4, new count 4 and the ten pair
    sums 80 (= 228); per listed pair that passes, its list entry 4, the old history slot read
    (id 4, tsd 16) and the new one written (id 4, tsd 16) (= 44). A neighbour's state is some
    other sphere's own state: it is counted once, as that sphere's (re-reads hit the caches
    or are waste)."""
    return 228.0 * spheres + 44.0 * pairs


def collide_kernel_bytes(spheres: float, wall_contacts: float) -> float:
    """Algorithmic bytes of one fused k_collide launch: per full-pass sphere its state read
    (pos 32, vel 24, acc 24, ep 32, ev 32, ea 32, mass 8, inertia 24 = 208), the pair sums of
    k_pairs 80, the list flags 4 and path length 4, the state written (pos 32, vel 24, acc 24,
    ep 32, ev 32, ea 32 = 176) and the wall-slot count read and written 8 (= 480); per wall
    contact its slot read and written (id 4 + tsd 16, twice = 40)."""
    return 480.0 * spheres + 40.0 * wall_contacts


# The library walks the pairs inside k_collide (k_pairs returns at once) when the step's full
# pass holds fewer spheres than this.
MERGE_PAIRS_BELOW = 65536


def kernel_rooflines(tm, n, info, n_pp, n_wall, fuse):
    """Per-step time and achieved HBM bandwidth of the step's kernels from the event-timed
    steps: k_lite_move (spheres outside the full pass), k_pairs and k_collide (the full-pass
    spheres, contact_spheres of the last step). On a small full pass k_collide also walks the
    pairs (their bytes count for it) and k_pairs only reads the queue counters."""
    steps = max(tm["steps"], 1)
    full = info["contact_spheres"]
    merged = full < MERGE_PAIRS_BELOW
    out = {}
    if fuse and tm["lite_move_ms"] > 0:
        ms = tm["lite_move_ms"] / steps
        b = LITE_MOVE_BYTES * (n - full)
        out["k_lite_move"] = {"ms": ms, "bytes": b, "units": n - full, "bytes_per_unit": LITE_MOVE_BYTES}
    if tm.get("pair_pass_ms", 0) > 0 and not merged:
        ms = tm["pair_pass_ms"] / steps
        b = pairs_kernel_bytes(full, n_pp)
        out["k_pairs"] = {"ms": ms, "bytes": b, "units": full, "bytes_per_unit": b / max(full, 1)}
    if fuse and tm.get("collide_ms", 0) > 0:
        ms = tm["collide_ms"] / steps
        b = merged_collide_kernel_bytes(full, n_wall, n_pp) if merged else collide_kernel_bytes(full, n_wall)
        out["k_collide"] = {"ms": ms, "bytes": b, "units": full, "bytes_per_unit": b / max(full, 1),
                            "walks_pairs": merged}
    for k in out.values():
        k["gbps"] = k["bytes"] / (k["ms"] * 1e-3) / 1e9
        k["frac"] = k["gbps"] / HBM_PEAK_GBPS
    return out


def dominant_roofline(kr, traffic=None):
    """The roofline object of the kernel with the longest time per step."""
    name = max(kr, key=lambda k: kr[k]["ms"])
    k = kr[name]
    return {"bound": "hbm", "achieved": k["gbps"], "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": k["frac"],
            "traffic": traffic.get(name) if traffic else None, "kernel": name, "kernel_ms": k["ms"],
            "bytes_per_unit": k["bytes_per_unit"], "units": k["units"]}


def run_gpu(scene, precision, steps, warmup, capacity, fuse=False, buckets_per_sphere=0, skin=0.0, graph=0,
            full_blocks=0, merge_below=0, skin_cap=0.0):
    import torch
    from xdynamics_dll_amd import DemContext

    ctx = DemContext(scene, precision=precision, p2p_capacity=capacity, fuse_velocity=fuse,
                     buckets_per_sphere=buckets_per_sphere, skin=skin, full_blocks=full_blocks,
                     merge_pairs_below=merge_below, skin_cap=skin_cap)
    ctx.step(warmup)
    ctx.synchronize()
    torch.cuda.synchronize()
    # a graph of at most the timed steps (an even count): short timed runs replay one graph
    # too instead of launching their steps one by one
    graph = min(graph, steps - steps % 2) if graph else 0
    if graph >= 2:
        ctx.graph_capture(graph)   # graph steps recorded (not run) once; replays do the timed steps
        ctx.synchronize()
    else:
        graph = 0
    info0 = ctx.info()
    t0 = time.perf_counter()
    if graph:
        ctx.graph_launch(steps // graph)
        ctx.step(steps % graph)
    else:
        ctx.step(steps)
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    n_pp, n_wall = ctx.contact_count()
    info = ctx.info()
    info["timed_list_builds"] = info["list_builds"] - info0["list_builds"]
    # the phase breakdown comes from a separate run of further steps with HIP events around
    # the phases (the events add stream-ordering work of their own, so the timed run has none)
    ctx.set_timing(True)
    ctx.step(min(steps, 200))
    ctx.synchronize()
    ctx.set_timing(False)
    tm = ctx.timing()
    err = ctx.error_word()
    ctx.close()
    return {"seconds": t1 - t0, "timing": tm, "n_pp": n_pp, "n_wall": n_wall, "err": err, "info": info}


def force_error(scene, capacity, nsteps=40):
    """Mixed and FP64 GPU contact forces against the FP64 oracle at one state: an FP64
    context is advanced nsteps, its state is copied into a fresh mixed context, a fresh FP64
    context and the oracle, and each runs one contact pass. Per sphere, the mixed error is
    |F_mixed - F_oracle| / (sum of that sphere's |F_pair| + |F_oracle|); the FP64 GPU error
    is max |F_fp64 - F_oracle| / max |F_oracle|. Pair sets are compared exactly."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import xdem_ref
    from xdynamics_dll_amd import DemContext

    fields = ("pos", "vel", "acc", "ep", "ev", "ea")
    a = DemContext(scene, precision="fp64", p2p_capacity=capacity)
    a.step(nsteps)
    state = {f: a.read(f) for f in fields}
    a.close()
    b = DemContext(scene, precision="mixed", p2p_capacity=capacity)
    c = DemContext(scene, precision="fp64", p2p_capacity=capacity)
    r = xdem_ref.RefSim(scene, capacity=capacity)
    for f in fields:
        b.write(f, state[f])
        c.write(f, state[f])
        r.write(f, state[f])
    b.collide()
    c.collide()
    mag = np.zeros(scene.np)
    r.collide(pair_magnitude=mag)
    fr = r.force
    fb, fc = b.read("force"), c.read("force")
    norm = mag + np.linalg.norm(fr, axis=1)
    has = norm > 0
    mixed_err = float((np.linalg.norm(fb - fr, axis=1)[has] / norm[has]).max()) if has.any() else 0.0
    fp64_err = float(np.abs(fc - fr).max() / max(np.abs(fr).max(), 1e-300))

    def same_pairs(ctx):
        cnt, ids, _ = ctx.history("p2p")
        if not np.array_equal(cnt, r.pp_count):
            return False
        return all(np.array_equal(np.sort(ids[i, : cnt[i]]), np.sort(r.pp_id[i, : r.pp_count[i]]))
                   for i in np.nonzero(cnt)[0])

    out = {"mixed_force_err": mixed_err, "fp64_force_err": fp64_err, "pairs_exact_mixed": same_pairs(b),
           "pairs_exact_fp64": same_pairs(c), "contacts_checked": int(r.pp_count.sum()), "state_step": nsteps,
           "error_words": [b.error_word(), c.error_word()]}
    b.close()
    c.close()
    return out


def host_cpu():
    """The host's CPU model and the cores this process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, usable, os.cpu_count() or 1


def cpu_baseline(scene, capacity, budget_s=12.0, state=None, mesh=None, label=None):
    """The oracle (FP64 C restatement, OpenMP over spheres) stepping the same scene on the host:
    first on every core this process may use, then on one thread. state (field -> array) and
    mesh ((triangles, pose, ContactProperty)) set it to a given step of the GPU run."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import xdem_ref

    model, usable, total = host_cpu()

    def timed(threads, budget):
        xdem_ref.set_threads(threads)
        ref = xdem_ref.RefSim(scene, capacity=capacity)
        if mesh is not None:
            ref.add_mesh(*mesh)
        for f, v in (state or {}).items():
            ref.write(f, v)
        t0 = time.perf_counter()
        ref.step(1)
        first = time.perf_counter() - t0
        nsteps = max(1, min(20, int(budget / max(first, 1e-3))))
        t0 = time.perf_counter()
        ref.step(nsteps)
        return scene.np * nsteps / (time.perf_counter() - t0), nsteps

    v_all, n_all = timed(usable, budget_s)
    v_one, n_one = timed(1, 0.5 * budget_s)
    what = label or f"the same {scene.np:,}-sphere scene"
    return {"value": v_all, "unit": "particle-steps/s", "cores": usable, "kind": "port",
            "sample": f"{n_all} steps ({usable} threads) and {n_one} steps (1 thread) of {what} after 1 untimed "
                      "step each, FP64 C restatement, OpenMP over spheres",
            "value_1_thread": v_one, "cpu_model": model, "host_cpus": total}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_lite_move_dam_break_1m.json")


def measured_traffic(np_, fused):
    """HBM bytes per dispatch of k_lite_move from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the ratio the same counter shows
    on a plain device copy of known size), or None when no summary matches this workload."""
    if not fused:
        return None
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("k_lite_move")
    if not k or t.get("np") != np_:
        return None
    return (k["fetch_kib"] * t["copy_scale_fetch"] + k["write_kib"] * t["copy_scale_write"]) * 1024.0


WHEEL_TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_wheel_500k.json")


def wheel_traffic(np_, kernel="k_pairs"):
    """HBM bytes per dispatch of a full-pass kernel of the wheel workload from the committed
    rocprofv3 --pmc summary (scaled as in measured_traffic), or None when the summary is of
    another bed size."""
    try:
        with open(WHEEL_TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    k = t.get("kernels", {}).get(kernel)
    if not k or t.get("np") != np_:
        return None
    return (k["fetch_kib"] * t["copy_scale_fetch"] + k["write_kib"] * t["copy_scale_write"]) * 1024.0


def copy_probe(mib):
    """A plain device-to-device copy of mib MiB, repeated 5 times: the calibration kernel
    for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 5
    print(json.dumps({"copy_probe_mib": mib, "bytes_read": n * 4, "bytes_written": n * 4,
                      "ms": dt * 1e3, "gbps": 2 * n * 4 / dt / 1e9}), flush=True)


def batch_force_error(scenes, capacity, nsteps=40):
    """Per-scene force error of the batch at one state: an FP64 context of all scenes is
    advanced nsteps, its state copied into a fresh mixed context, a fresh FP64 context and one
    oracle per scene, and each runs one contact pass. Per sphere the error is |F - F_oracle| /
    (sum of the sphere's |F_pair| + |F_oracle|); per scene its maximum over the spheres with a
    force. Pair sets are compared exactly."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import xdem_ref
    from xdynamics_dll_amd import DemContext

    model, usable, _ = host_cpu()
    xdem_ref.set_threads(usable)
    fields = ("pos", "vel", "acc", "ep", "ev", "ea")
    a = DemContext(scenes, precision="fp64", p2p_capacity=capacity)
    a.step(nsteps)
    states = [{f: a.read(f, k) for f in fields} for k in range(len(scenes))]
    a.close()
    b = DemContext(scenes, precision="mixed", p2p_capacity=capacity)
    c = DemContext(scenes, precision="fp64", p2p_capacity=capacity)
    for k in range(len(scenes)):
        for f in fields:
            b.write(f, states[k][f], k)
            c.write(f, states[k][f], k)
    b.collide()
    c.collide()
    mixed_err, fp64_err, pairs_exact, contacts = [], [], True, 0
    for k, sc in enumerate(scenes):
        r = xdem_ref.RefSim(sc, capacity=capacity)
        for f in fields:
            r.write(f, states[k][f])
        mag = np.zeros(sc.np)
        r.collide(pair_magnitude=mag)
        norm = mag + np.linalg.norm(r.force, axis=1)
        has = norm > 0
        for ctx, out in ((b, mixed_err), (c, fp64_err)):
            f = ctx.read("force", k)
            out.append(float((np.linalg.norm(f - r.force, axis=1)[has] / norm[has]).max()) if has.any() else 0.0)
            cnt, ids, _ = ctx.history("p2p", k)
            if not np.array_equal(cnt, r.pp_count):
                pairs_exact = False
            else:
                for i in np.nonzero(cnt)[0]:
                    if not np.array_equal(np.sort(ids[i, : cnt[i]]), np.sort(r.pp_id[i, : r.pp_count[i]])):
                        pairs_exact = False
                        break
        contacts += int(r.pp_count.sum())
    errs = [b.error_word(), c.error_word()]
    b.close()
    c.close()
    me, fe = np.array(mixed_err), np.array(fp64_err)
    return {"mixed_force_err_max": float(me.max()), "mixed_force_err_median": float(np.median(me)),
            "worst_scene": int(me.argmax()), "fp64_force_err_max": float(fe.max()), "pairs_exact": pairs_exact,
            "contacts_checked": contacts, "state_step": nsteps, "error_words": errs,
            "mixed_force_err_per_scene": [float(x) for x in me]}


def batch_main(args):
    """The 256-scene batched sweep (seed 1004): independent box-settling scenes of
    floor(exp(U[ln 2000, ln 64000])) spheres each, r ~ U[0.5, 2] mm, per-scene random
    friction, restitution and modulus, advanced together in one context (one launch per
    kernel for all scenes). Mixed precision, fused velocity update. The full-pass blocks and
    the bucket budget come from the autotuner (xdem_autotune) unless given. The default run is
    1,000 steps: scenes 245, 158 and 212 (restitution 0.10-0.15) go non-finite after ~1,750 and
    ~2,000 steps, the oracle's run of scene 245 as well. The CPU force law's damping
    coefficient beta 2 sqrt(m* kn) (xContact.cpp:244, kn = 4/3 E* sqrt(R*) with no overlap
    factor) gives c dt / m* ~ 3.7 for two of the smallest spheres of scene 245, above the
    explicit update's limit of 2."""
    import tempfile

    import torch
    from xdynamics_dll_amd import SceneData, autotune

    t_build = time.perf_counter()
    scenes = [SceneData.build(4, 1004, variant=k) for k in range(args.nscene)]
    t_build = time.perf_counter() - t_build
    n = sum(s.np for s in scenes)
    fuse = not args.split
    tuned = None
    skin_cap = args.skin_cap
    if not args.full_blocks and not args.buckets_per_sphere and not args.skin_cap:
        t0 = time.perf_counter()
        cache = os.path.join(tempfile.gettempdir(), "xdem_autotune_bench.txt")
        # trials over the timed run's own steps: the skin's cap (list builds against listed
        # pairs) pays only once the settling scenes move fast, after several hundred steps
        tuned = autotune(scenes, precision="mixed", fuse_velocity=fuse, p2p_capacity=args.capacity,
                         warmup=args.warmup, steps=args.steps, cache_path=cache)
        tuned["seconds"] = time.perf_counter() - t0
        args.full_blocks, args.buckets_per_sphere = tuned["full_blocks"], tuned["buckets_per_sphere"]
        skin_cap = tuned["skin_cap"]
        print(f"[batch] autotuned: {tuned['full_blocks']} full-pass blocks, {tuned['buckets_per_sphere']} buckets "
              f"per sphere, skin cap {skin_cap:g} ({tuned['ms_per_step']:.4f} ms/step, {len(tuned['trials'])} trials)",
              file=sys.stderr, flush=True)
    mixed = run_gpu(scenes, "mixed", args.steps, args.warmup, args.capacity, fuse, args.buckets_per_sphere, args.skin,
                    args.graph, args.full_blocks, skin_cap=skin_cap)
    value = n * args.steps / mixed["seconds"]
    f64 = run_gpu(scenes, "fp64", args.steps, args.warmup, args.capacity, fuse,
                  args.buckets_per_sphere, args.skin, args.graph, args.full_blocks, skin_cap=skin_cap)
    f64_value = n * args.steps / f64["seconds"]
    tm = mixed["timing"]
    cbar = (mixed["n_pp"] + mixed["n_wall"]) / n
    out = {
        "metric": "particle-steps/s, batched independent scenes",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": mixed["seconds"] / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32-contact/fp64-state", "data": "synthetic",
        "config": {"workload": f"batched sweep: {args.nscene} box-settling scenes, seed 1004, "
                               "N_s = floor(exp(U[ln 2000, ln 64000])), r ~ U[0.5, 2] mm, "
                               "dt = min(1e-5 s, 0.2 Rayleigh time) per scene",
                   "np_total": n, "nscene": args.nscene, "nbucket": mixed["info"]["nbucket"],
                   "parallelism": "single GPU, all scenes in one context"},
        "roofline": dominant_roofline(kernel_rooflines(tm, n, mixed["info"], mixed["n_pp"], mixed["n_wall"], fuse)),
        "kernels": kernel_rooflines(tm, n, mixed["info"], mixed["n_pp"], mixed["n_wall"], fuse),
        "mean_contacts_per_sphere": cbar,
        "contact_pass_ms_per_step": tm["contact_ms"] / max(tm["steps"], 1),
        "rebuild_kernel_ms_per_step": tm["grid_ms"] / max(tm["steps"], 1),
        "touch_ms_per_step": tm["touch_ms"] / max(tm["steps"], 1),
        "list_builds_timed": mixed["info"]["timed_list_builds"],
        "list_spheres": mixed["info"]["list_spheres"],
        "contact_spheres": mixed["info"]["contact_spheres"],
        "error_word": mixed["err"], "scene_build_s": t_build, "fp64_value": f64_value,
        "mixed_over_fp64": value / f64_value,
        "autotune": tuned, "full_blocks": args.full_blocks, "buckets_per_sphere": args.buckets_per_sphere,
        "skin_cap": skin_cap,
    }
    if not args.no_fp64:
        out["force_check"] = batch_force_error(scenes, args.capacity)
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import xdem_ref

        model, usable, total = host_cpu()

        def sweep(threads, budget):
            xdem_ref.set_threads(threads)
            done, t = 0, 0.0
            for sc in scenes:   # whole scenes, 5 steps each, until the budget of CPU time
                ref = xdem_ref.RefSim(sc, capacity=args.capacity)
                t0 = time.perf_counter()
                ref.step(5)
                t += time.perf_counter() - t0
                done += sc.np * 5
                if t > budget:
                    break
            return done / t, done

        v_all, d_all = sweep(usable, 10.0)
        v_one, d_one = sweep(1, 5.0)
        out["cpu_baseline"] = {"value": v_all, "unit": "particle-steps/s", "cores": usable, "kind": "port",
                               "sample": f"the first scenes of the batch, 5 steps each ({d_all} particle-steps on "
                                         f"{usable} threads, {d_one} on 1), FP64 C restatement, OpenMP over spheres",
                               "value_1_thread": v_one, "cpu_model": model, "host_cpus": total}
    print(json.dumps(out), flush=True)


def wheel_main(args):
    """The wheel on a granular bed (seed 1003): 500,000 spheres r ~ U[2, 3] mm inserted loose
    into a 1.0 x 0.4 m bin and settled under gravity (untimed), then a grousered wheel (4,096
    triangles, R = 0.15 m, width 0.1 m, 40 kg) set on the bed with its lowest point 1 mm in it
    and driven at 2 rad/s about y by the host HHT integrator (alpha = -0.1) under a 200 N load,
    sphere-triangle friction 0.6, dt = 1e-5 s, mixed precision. Timed: the coupled steps (DEM
    step with the wheel's wrench reduced on the device, the HHT step on the host, the new pose
    back), run by the native loop xdem_cosim_wheel_steps."""
    import torch
    from xdynamics_dll_amd import DemContext, SceneData, contact_property, wheel_mesh
    from xdynamics_dll_amd.mbd import HhtWheel, cosim_wheel_steps

    t_build = time.perf_counter()
    scene = SceneData.build(3, 1003, np_override=args.np)
    tris = wheel_mesh()
    n = scene.np
    L, W = scene.planes[2].xw.x, scene.planes[4].xw.y
    xc, yc = 0.5 * L, 0.5 * W
    R = float(np.sqrt((tris.reshape(-1, 3)[:, [0, 2]] ** 2).sum(1)).max())   # rim or grouser tip
    park = np.zeros(15)
    park[0], park[7] = 40.0, 1.0
    park[1:4] = (xc, yc, 1.0)          # above the bin while the bed settles
    cp_wheel = contact_property(2e11, 5e7, 0.3, 0.25, 0.4, 0.6, 0.05)
    quirks = dict(gpu_force_law=1) if args.force_law == "kernel" else None
    ctx = DemContext(scene, precision="mixed", p2p_capacity=args.capacity, fuse_velocity=True, step_graph=50,
                     meshes=[(0, tris, park, cp_wheel)], full_blocks=args.full_blocks, quirks=quirks,
                     skin_cap=args.skin_cap)
    t_build = time.perf_counter() - t_build
    t0 = time.perf_counter()
    done = 0
    while done < args.settle:      # progress every 5,000 steps
        k = min(5000, args.settle - done)
        ctx.step(k)
        ctx.synchronize()
        done += k
        print(f"[wheel] settled {done} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    t_settle = time.perf_counter() - t0
    # the settled bed's spheres re-sorted by their current buckets (the stored order was that
    # of the loose column)
    ctx.reorder()
    pos = ctx.read("pos")
    vel = ctx.read("vel")
    under = (np.abs(pos[:, 0] - xc) < 0.08) & (np.abs(pos[:, 1] - yc) < 0.06)
    top = float((pos[under, 2] + pos[under, 3]).max())
    speed = np.linalg.norm(vel, axis=1)
    bed = {"top_m": top, "max_speed_after_settle": float(speed.max()), "mean_speed_after_settle": float(speed.mean()),
           "kinetic_energy_after_settle_J": float(0.5 * (scene.mass * speed ** 2).sum())}
    omega = 2.0
    start = park.copy()
    start[3] = top + R - 1e-3
    start[13] = 0.5 * omega            # e_dot of a turn about y at omega from e = (1, 0, 0, 0)
    ctx.set_mesh_pose(0, start)
    inertia = (40.0 * (3 * 0.15 ** 2 + 0.1 ** 2) / 12.0, 0.5 * 40.0 * 0.15 ** 2, 40.0 * (3 * 0.15 ** 2 + 0.1 ** 2) / 12.0)
    hht = HhtWheel(40.0, inertia, start, omega, scene.dt, alpha=-0.1, gravity=scene.gravity, load=(0.0, 0.0, -200.0))
    warm = cosim_wheel_steps(ctx, hht, 0, args.warmup)
    ctx.synchronize()
    info0 = ctx.info()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    track = cosim_wheel_steps(ctx, hht, 0, args.steps)
    ctx.synchronize()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    info = ctx.info()
    n_pp, n_wall = ctx.contact_count()
    n_mesh = int(ctx.history("mesh")[0].sum())
    pose = hht.state()[0]
    state = {f: ctx.read(f) for f in ("pos", "vel", "acc", "ep", "ev", "ea")}
    # phase breakdown: 200 further steps launched one by one with HIP events (wheel held at
    # its last pose)
    ctx.set_timing(True)
    ctx.step(200)
    ctx.synchronize()
    ctx.set_timing(False)
    tm = ctx.timing()
    err = ctx.error_word()
    ctx.close()
    value = n * args.steps / t
    kr = kernel_rooflines(tm, n, info, n_pp, n_wall, True)
    track = np.concatenate([warm, track])
    out = {
        "metric": "particle-steps/s, wheel on a granular bed (DEM + host HHT coupling)",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32-contact/fp64-state", "data": "synthetic",
        "config": {"workload": f"wheel on a granular bed: {n:,} spheres r~U[2,3] mm settled {args.settle} steps in a "
                               f"{L:.2f}x{W:.2f} m bin, 4,096-triangle grousered wheel R=0.15 m, 40 kg, "
                               "2 rad/s about y, 200 N load, HHT alpha=-0.1, mu=0.6, dt=1e-5 s, mixed",
                   "np": n, "ntriangle": len(tris), "parallelism": "single GPU + host HHT",
                   "force_law": args.force_law},
        "roofline": dominant_roofline(kr, {"k_pairs": wheel_traffic(n, "k_pairs"),
                                           "k_collide": wheel_traffic(n, "k_collide")}
                                      if args.force_law == "cpu" else None),
        "kernels": kr,
        "mean_contacts_per_sphere": (n_pp + n_wall + n_mesh) / n,
        "sphere_contacts": n_pp, "wall_contacts": n_wall, "mesh_contacts": n_mesh,
        "contact_spheres": info["contact_spheres"], "mesh_spheres": info["mesh_spheres"],
        "contact_pass_ms_per_step": tm["contact_ms"] / max(tm["steps"], 1),
        "mesh_pass_ms_per_step": tm["mesh_pass_ms"] / max(tm["steps"], 1),
        "rebuild_kernel_ms_per_step": tm["grid_ms"] / max(tm["steps"], 1),
        "touch_ms_per_step": tm["touch_ms"] / max(tm["steps"], 1),
        "list_builds_timed": info["list_builds"] - info0["list_builds"],
        "bed": bed,
        "wheel": {"x_travel_m": float(track[-1, 0] - start[1]), "sinkage_m": float(start[3] - track[-1, 1]),
                  "mean_Fz_N_last_1000": float(track[-1000:, 5].mean()),
                  "mean_Fx_N_last_1000": float(track[-1000:, 4].mean()),
                  "max_newton_iterations": int(track[:, 7].max())},
        "error_word": err, "scene_build_s": t_build, "settle_s": t_settle,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, args.capacity, state=state, mesh=(tris, pose, cp_wheel),
                                           label=f"the {n:,}-sphere bed with the wheel at the end of the timed run")
    print(json.dumps(out), flush=True)


def lattice_main(args):
    """The lattice box (seed 1001): 10,000 spheres r = 1 mm on a 2.2 mm cubic lattice with
    +-0.05 mm jitter, at rest in a 0.1 x 0.1 x 0.3 m five-wall box, E = 1e7 Pa, e = 0.3,
    mu = 0.3, mu_r = 0.01, gravity -z, dt = 1e-5 s, FP64. Timed: the steps after warmup (the
    lattice falls and collapses onto the floor)."""
    import torch
    from xdynamics_dll_amd import SceneData

    scene = SceneData.build(1, 1001, np_override=args.np)
    n = scene.np
    fuse = not args.split
    r = run_gpu(scene, "fp64", args.steps, args.warmup, args.capacity, fuse, args.buckets_per_sphere, args.skin,
                args.graph, args.full_blocks, args.merge_below, skin_cap=args.skin_cap)
    value = n * args.steps / r["seconds"]
    tm = r["timing"]
    kr = kernel_rooflines(tm, n, r["info"], r["n_pp"], r["n_wall"], fuse)
    out = {
        "metric": "particle-steps/s, lattice box, FP64",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["seconds"] / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"lattice box: {n:,} spheres r=1 mm, 2.2 mm lattice +-0.05 mm jitter, "
                               "0.1x0.1x0.3 m 5-wall box, dt=1e-5 s, FP64", "np": n, "seed": 1001,
                   "parallelism": "single GPU"},
        "roofline": dominant_roofline(kr),
        "kernels": kr,
        "mean_contacts_per_sphere": (r["n_pp"] + r["n_wall"]) / n,
        "contact_spheres": r["info"]["contact_spheres"],
        "contact_pass_ms_per_step": tm["contact_ms"] / max(tm["steps"], 1),
        "rebuild_kernel_ms_per_step": tm["grid_ms"] / max(tm["steps"], 1),
        "touch_ms_per_step": tm["touch_ms"] / max(tm["steps"], 1),
        "list_builds_timed": r["info"]["timed_list_builds"],
        "error_word": r["err"],
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, args.capacity)
    print(json.dumps(out), flush=True)


def calib_main(args):
    """The calibration workload (seed 1005): 20,000 spheres r ~ U[1, 1.5] mm released from a
    30 mm cylinder region onto a rough floor, FP64 steps over 0.08 s (16,000 steps at dt = 5e-6 s:
    at the nominal 2e-5 s the CPU force law's damping, beta 2 sqrt(m kn), exceeds 2 m / dt for
    the lighter spheres and the run diverges; the kernel's law is stable there but has no
    meaningful derivative, its tangential force keeping a finite size along the direction of a
    vanishing slip velocity). Target: the
    64-bin radial height profile of a hidden run (pile seed 2005) at friction 0.45, restitution
    0.35, E = 2e7 Pa. Timed: one loss and gradient at the initial guess (friction 0.3,
    restitution 0.5, E = 3e7 Pa): the forward run with checkpoints every 100 steps, then the
    adjoint (recomputation per segment and the reverse sweep). Checked against central
    differences with step 1e-4 (and 1e-7) on (friction, restitution, log E)."""
    import torch
    from xdynamics_dll_amd import SceneData
    from xdynamics_dll_amd.calib import Calibration, pile_center, profile_loss, radial_profile

    steps = args.calib_steps
    hidden = SceneData.build(5, 2005, np_override=args.np)
    scene = SceneData.build(5, 1005, np_override=args.np)
    n = scene.np
    every = 100
    t0 = time.perf_counter()
    hidden.dt = scene.dt = args.calib_dt
    hcal = Calibration(hidden, max_steps=steps, checkpoint_every=every)
    target = radial_profile(hcal.forward(hcal.theta_of(0.45, 0.35, 2e7), steps), pile_center(hidden))
    hcal.close()
    t_target = time.perf_counter() - t0
    center = pile_center(scene)
    cal = Calibration(scene, max_steps=steps, checkpoint_every=every)
    theta = cal.theta_of(0.3, 0.5, 3e7)
    cal.forward(theta, 10)   # warm-up (first launches)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pos = cal.forward(theta, steps)
    t_fwd = time.perf_counter() - t0
    loss, g = profile_loss(pos, target, center)
    t0 = time.perf_counter()
    grad = cal.backward(g)
    t_bwd = time.perf_counter() - t0
    print(f"[calib] loss {loss:.6e} gradient {grad} forward {t_fwd:.2f} s adjoint {t_bwd:.2f} s", file=sys.stderr,
          flush=True)
    fds = {}
    for h in (1e-4, 1e-7):
        fd = np.zeros(3)
        for k in range(3):
            tp, tm = theta.copy(), theta.copy()
            tp[k] += h
            tm[k] -= h
            fd[k] = (profile_loss(cal.forward(tp, steps), target, center)[0] -
                     profile_loss(cal.forward(tm, steps), target, center)[0]) / (2 * h)
        fds[h] = fd
        print(f"[calib] central differences h={h:g}: {fd}", file=sys.stderr, flush=True)
    # the finite-horizon gradient: the run continued from step steps - horizon at theta, the
    # adjoint over those steps, central differences of the continued runs
    horizon = args.calib_horizon
    start = steps - horizon
    cal.forward(theta, steps)
    fd_h = np.zeros(3)
    hh = 1e-7
    for k in range(3):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += hh
        tm[k] -= hh
        fd_h[k] = (profile_loss(cal.forward(tp, steps, start), target, center)[0] -
                   profile_loss(cal.forward(tm, steps, start), target, center)[0]) / (2 * hh)
    loss_h, g_h = profile_loss(cal.forward(theta, steps, start), target, center)
    t0 = time.perf_counter()
    grad_h = cal.backward(g_h)
    t_bwd_h = time.perf_counter() - t0
    print(f"[calib] last {horizon} steps: adjoint {grad_h} central differences {fd_h}", file=sys.stderr, flush=True)
    # where the whole-run gradient stays meaningful: runs of 500, 1,000 and 2,000 steps from the
    # start (targets from the hidden pile at the same times), adjoint against central differences
    study = []
    for hs in (500, 1000, 2000):
        hc = Calibration(hidden, max_steps=hs, checkpoint_every=every)
        tgt = radial_profile(hc.forward(hc.theta_of(0.45, 0.35, 2e7), hs), pile_center(hidden))
        hc.close()
        loss_s, g_s = profile_loss(cal.forward(theta, hs), tgt, center)
        grad_s = cal.backward(g_s)
        fd_s = np.zeros(3)
        for k in range(3):
            tp, tm = theta.copy(), theta.copy()
            tp[k] += hh
            tm[k] -= hh
            fd_s[k] = (profile_loss(cal.forward(tp, hs), tgt, center)[0] -
                       profile_loss(cal.forward(tm, hs), tgt, center)[0]) / (2 * hh)
        study.append({"steps": hs, "loss": loss_s, "adjoint": grad_s.tolist(), "central_differences_1e-07": fd_s.tolist(),
                      "max_difference_over_max_fd": float(np.abs(grad_s - fd_s).max() / np.abs(fd_s).max())})
        print(f"[calib] {hs} steps from the start: adjoint {grad_s} central differences {fd_s}", file=sys.stderr,
              flush=True)
    err = cal.error_word()
    cal.close()
    value = n * steps / (t_fwd + t_bwd)
    out = {
        "metric": "particle-steps/s of one loss + gradient (forward with checkpoints + adjoint), pile calibration",
        "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": steps, "warmup": 10,
        "ms_per_step": (t_fwd + t_bwd) / steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"pile calibration: {n:,} spheres r~U[1,1.5] mm from a 30 mm cylinder region onto a "
                               f"floor, {steps} steps dt={args.calib_dt:g} s, checkpoints every {every} steps, loss = "
                               "MSE of the 64-bin radial height profile vs a hidden run (seed 2005, mu=0.45, e=0.35, "
                               "E=2e7), the CPU force law of xContact.cpp",
                   "np": n, "seed": 1005, "parallelism": "single GPU"},
        "forward_s": t_fwd, "adjoint_s": t_bwd, "target_run_s": t_target,
        "theta": {"friction": theta[0], "restitution": theta[1], "log_E": theta[2]},
        "loss": loss, "gradient": grad.tolist(),
        "central_differences": {f"{h:g}": fd.tolist() for h, fd in fds.items()},
        "relative_difference": {f"{h:g}": (np.abs(fd - grad) / np.maximum(np.abs(fd), 1e-300)).tolist()
                                for h, fd in fds.items()},
        "horizon_study": study,
        "finite_horizon": {"steps": horizon, "start": start, "gradient": grad_h.tolist(),
                           "central_differences_1e-07": fd_h.tolist(), "adjoint_s": t_bwd_h,
                           "relative_difference": (np.abs(fd_h - grad_h) / np.abs(fd_h).max()).tolist()},
        "error_word": err,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 2,000; the batch: 1,000, as three of its scenes diverge under the "
                         "CPU force law after ~1,750 steps)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--np", type=int, default=0, help="sphere count override (default 1,000,000)")
    ap.add_argument("--capacity", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--split", action="store_true",
                    help="rolling + velocity update in a kernel of their own (default: inside the contact pass)")
    ap.add_argument("--graph", type=int, default=50,
                    help="replay the timed steps from a HIP graph of this many steps (even; 0: eager launches)")
    ap.add_argument("--skin-cap", type=float, default=0.0,
                    help="the adaptive skin's cap as a fraction of the smallest radius (0: the library's 0.4; "
                         "the batch: autotuned)")
    ap.add_argument("--skin", type=float, default=0.0,
                    help="neighbour-list skin in m (0: adaptive, <0: rebuild every step)")
    ap.add_argument("--buckets-per-sphere", type=int, default=0,
                    help="bucket budget of the grid per sphere (0 = library default)")
    ap.add_argument("--workload", choices=["dam_break", "batch", "wheel", "lattice", "calib"], default="dam_break",
                    help="dam_break: the 1,000,000-sphere column (default); batch: the batched scene sweep; "
                         "wheel: the wheel on a 500,000-sphere bed with HHT coupling; lattice: the 10,000-sphere "
                         "lattice box in FP64; calib: loss and adjoint gradient of the 20,000-sphere pile")
    ap.add_argument("--settle", type=int, default=20000, help="untimed settling steps of the wheel's bed")
    ap.add_argument("--force-law", choices=["cpu", "kernel"], default="cpu",
                    help="wheel: the damping law of the sphere contacts, the CPU path's (xContact.cpp:238-264, "
                         "the parity profile) or the GPU kernel's (impl:432-474)")
    ap.add_argument("--calib-steps", type=int, default=16000, help="steps of the calibration workload")
    ap.add_argument("--calib-dt", type=float, default=5e-6, help="time step of the calibration workload")
    ap.add_argument("--calib-horizon", type=int, default=100,
                    help="steps of the calibration's finite-horizon gradient (a multiple of 100)")
    ap.add_argument("--full-blocks", type=int, default=0,
                    help="blocks of the full-pass kernels (0 = library default)")
    ap.add_argument("--merge-below", type=int, default=0,
                    help="full passes smaller than this walk their pairs in the contact kernel (0 = library default)")
    ap.add_argument("--nscene", type=int, default=256, help="scenes of the batch workload")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 1000 if args.workload == "batch" else 2000
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU layout (one spatial slab "
              f"of spheres per GPU with halo exchange) is not built", file=sys.stderr)
        sys.exit(2)

    import torch
    from xdynamics_dll_amd import SceneData

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.workload == "batch":
        batch_main(args)
        return
    if args.workload == "wheel":
        wheel_main(args)
        return
    if args.workload == "calib":
        calib_main(args)
        return
    if args.workload == "lattice":
        lattice_main(args)
        return
    t_build = time.perf_counter()
    scene = SceneData.build(2, SEED, np_override=args.np)
    t_build = time.perf_counter() - t_build
    n = scene.np

    fuse = not args.split
    mixed = run_gpu(scene, "mixed", args.steps, args.warmup, args.capacity, fuse, args.buckets_per_sphere, args.skin,
                    args.graph, args.full_blocks, args.merge_below, skin_cap=args.skin_cap)
    value = n * args.steps / mixed["seconds"]
    print(f"[dam] mixed leg: {mixed['seconds'] / args.steps * 1e3:.4f} ms per step", file=sys.stderr, flush=True)
    cbar = (mixed["n_pp"] + mixed["n_wall"]) / n
    tm = mixed["timing"]
    step_gbps = step_bytes(cbar) * n * args.steps / mixed["seconds"] / 1e9
    kr = kernel_rooflines(tm, n, mixed["info"], mixed["n_pp"], mixed["n_wall"], fuse)

    out = {
        "metric": "particle-steps/s @1M spheres, mixed-FP32 vs FP64 GPU (<=1e-4 force err), GB/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-contact/fp64-state",
        "data": "synthetic",
        "config": {"workload": "dam-break column, 1,000,000 polydisperse spheres r~U[1,2] mm, dt=5e-6 s, "
                               "5 walls, mixed FP32 contact arithmetic (FP64 positions/history/sums)",
                   "np": n, "seed": SEED, "p2p_capacity": args.capacity, "parallelism": "single GPU",
                   "nbucket": mixed["info"]["nbucket"]},
        "roofline": dominant_roofline(kr, {"k_lite_move": measured_traffic(n, fuse)}),
        "kernels": kr,
        "step_gbps": step_gbps,
        "step_roofline_frac": step_gbps / HBM_PEAK_GBPS,
        "mean_contacts_per_sphere": cbar,
        "contact_pass_ms_per_step": tm["contact_ms"] / max(tm["steps"], 1),
        "rebuild_kernel_ms_per_step": tm["grid_ms"] / max(tm["steps"], 1),
        "touch_ms_per_step": tm["touch_ms"] / max(tm["steps"], 1),
        "finish_ms_per_step": tm["finish_ms"] / max(tm["steps"], 1),
        "list_builds_timed": mixed["info"]["timed_list_builds"],
        "list_spheres": mixed["info"]["list_spheres"],
        "contact_spheres": mixed["info"]["contact_spheres"],
        "error_word": mixed["err"],
        "scene_build_s": t_build,
    }
    if not args.no_fp64:
        f64 = run_gpu(scene, "fp64", args.steps, args.warmup, args.capacity,
                      fuse, args.buckets_per_sphere, args.skin, args.graph, args.full_blocks, args.merge_below)
        f64_value = n * args.steps / f64["seconds"]
        print(f"[dam] FP64 leg: {f64['seconds'] / args.steps * 1e3:.4f} ms per step", file=sys.stderr, flush=True)
        out["fp64_value"] = f64_value
        out["mixed_over_fp64"] = value / f64_value
        out["force_check"] = force_error(scene, args.capacity)
        print("[dam] force check done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        cb = cpu_baseline(scene, args.capacity)
        out["cpu_baseline"] = cb
        if not args.no_fp64:
            out["fp64_gpu_over_cpu"] = out["fp64_value"] / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
