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


def contact_kernel_bytes(cbar: float, fused: bool) -> float:
    """Algorithmic bytes per sphere of the contact pass.
    Fused (k_lite_move + k_lite_turn for the spheres outside this step's full pass, k_collide
    for the others: contacts, rolling, velocity update and the next step's drift): own state
    read pos 32, vel 24, acc 24, ep 32, ev 32, ea 32, mass 8, inertia 24 (= 208); state write
    pos 32, vel 24, acc 24, ep 32, ev 32, ea 32 (= 176); list flags 2*4, path length 2*4;
    sphere and wall history: counts 2*4, entries 2*20*cbar.
    Split (k_collide_lite + k_collide, contacts only): read pos 32, vel 24, ep 32, ev 32,
    mass 8 (= 128); write force, moment, rolling torque sum and magnitude 24+24+24+8 (= 80);
    history 16 + 40*cbar."""
    return (408.0 if fused else 224.0) + 40.0 * cbar


# Algorithmic bytes per sphere of k_lite_move (the translation of every sphere outside the
# step's full pass): the translational part of the per-sphere state traffic, read pos 32,
# vel 24, acc 24, mass 8 and write pos 32, vel 24, acc 24.
LITE_MOVE_BYTES = 168.0


def roofline(tm, n, cbar, fuse, info, traffic):
    """Roofline object of the dominant kernel. Fused mode: k_lite_move, timed with HIP events
    recorded on the context's stream right before and after it; its units are the spheres
    outside the full pass (n - contact_spheres of the last step). Split mode: the contact
    pass (both streams, fork to join)."""
    steps = max(tm["steps"], 1)
    if fuse:
        k_ms = tm["lite_move_ms"] / steps
        units = n - info["contact_spheres"]
        achieved = LITE_MOVE_BYTES * units / (k_ms * 1e-3) / 1e9
        return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                "kernel": "k_lite_move (translation of the spheres outside the full contact pass)",
                "kernel_ms": k_ms, "bytes_per_unit": LITE_MOVE_BYTES, "units": units}
    k_ms = tm["contact_ms"] / steps
    achieved = contact_kernel_bytes(cbar, False) * n / (k_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
            "kernel": "contact pass: k_collide_lite, k_collide on a second stream (contacts)",
            "kernel_ms": k_ms, "bytes_per_unit": contact_kernel_bytes(cbar, False), "units": n}


def run_gpu(scene, precision, steps, warmup, capacity, fuse=False, buckets_per_sphere=0, skin=0.0, graph=0):
    import torch
    from xdynamics_dll_amd import DemContext

    ctx = DemContext(scene, precision=precision, p2p_capacity=capacity, fuse_velocity=fuse,
                     buckets_per_sphere=buckets_per_sphere, skin=skin)
    ctx.step(warmup)
    ctx.synchronize()
    torch.cuda.synchronize()
    if graph:
        ctx.graph_capture(graph)   # graph steps recorded once; replays do the timed steps
        ctx.synchronize()
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


def cpu_baseline(scene, capacity, budget_s=12.0):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import xdem_ref

    threads = min(16, os.cpu_count() or 1)
    xdem_ref.set_threads(threads)
    ref = xdem_ref.RefSim(scene, capacity=capacity)
    t0 = time.perf_counter()
    ref.step(1)
    first = time.perf_counter() - t0
    nsteps = max(1, min(20, int(budget_s / max(first, 1e-3))))
    t0 = time.perf_counter()
    ref.step(nsteps)
    t = time.perf_counter() - t0
    return {"value": scene.np * nsteps / t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps} steps of the same 1,000,000-sphere scene after 1 untimed step, "
                      f"FP64 C restatement, OpenMP over spheres"}


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


def batch_main(args):
    """The 256-scene batched sweep (seed 1004): independent box-settling scenes of
    floor(exp(U[ln 2000, ln 64000])) spheres each, r ~ U[0.5, 2] mm, per-scene random
    friction, restitution and modulus, advanced together in one context (one launch per
    kernel for all scenes). Mixed precision, fused velocity update."""
    import torch
    from xdynamics_dll_amd import SceneData

    t_build = time.perf_counter()
    scenes = [SceneData.build(4, 1004, variant=k) for k in range(args.nscene)]
    t_build = time.perf_counter() - t_build
    n = sum(s.np for s in scenes)
    fuse = not args.split
    mixed = run_gpu(scenes, "mixed", args.steps, args.warmup, args.capacity, fuse, args.buckets_per_sphere, args.skin, args.graph)
    value = n * args.steps / mixed["seconds"]
    f64 = run_gpu(scenes, "fp64", args.steps, args.warmup, args.capacity, fuse,
                  args.buckets_per_sphere, args.skin, args.graph)
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
        "roofline": roofline(tm, n, cbar, fuse, mixed["info"], None),
        "mean_contacts_per_sphere": cbar,
        "contact_pass_ms_per_step": tm["contact_ms"] / max(tm["steps"], 1),
        "rebuild_kernel_ms_per_step": tm["grid_ms"] / max(tm["steps"], 1),
        "touch_ms_per_step": tm["touch_ms"] / max(tm["steps"], 1),
        "list_builds_timed": mixed["info"]["timed_list_builds"],
        "list_spheres": mixed["info"]["list_spheres"],
        "contact_spheres": mixed["info"]["contact_spheres"],
        "error_word": mixed["err"], "scene_build_s": t_build, "fp64_value": f64_value,
        "mixed_over_fp64": value / f64_value,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import xdem_ref

        threads = min(16, os.cpu_count() or 1)
        xdem_ref.set_threads(threads)
        done, t = 0, 0.0
        for sc in scenes:   # whole scenes, 5 steps each, until about 10 s of CPU work
            ref = xdem_ref.RefSim(sc, capacity=args.capacity)
            t0 = time.perf_counter()
            ref.step(5)
            t += time.perf_counter() - t0
            done += sc.np * 5
            if t > 10.0:
                break
        out["cpu_baseline"] = {"value": done / t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
                               "sample": f"the first scenes of the batch, 5 steps each ({done} particle-steps), "
                                         "FP64 C restatement, OpenMP over spheres"}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--np", type=int, default=0, help="sphere count override (default 1,000,000)")
    ap.add_argument("--capacity", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--split", action="store_true",
                    help="rolling + velocity update in a kernel of their own (default: inside the contact pass)")
    ap.add_argument("--graph", type=int, default=50,
                    help="replay the timed steps from a HIP graph of this many steps (even; 0: eager launches)")
    ap.add_argument("--skin", type=float, default=0.0,
                    help="neighbour-list skin in m (0: adaptive, <0: rebuild every step)")
    ap.add_argument("--buckets-per-sphere", type=int, default=0,
                    help="bucket budget of the grid per sphere (0 = library default)")
    ap.add_argument("--workload", choices=["dam_break", "batch"], default="dam_break",
                    help="dam_break: the 1,000,000-sphere column (default); batch: the batched scene sweep")
    ap.add_argument("--nscene", type=int, default=256, help="scenes of the batch workload")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the multi-GPU layout is described "
              f"in DESIGN.md and not built", file=sys.stderr)
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
    t_build = time.perf_counter()
    scene = SceneData.build(2, SEED, np_override=args.np)
    t_build = time.perf_counter() - t_build
    n = scene.np

    fuse = not args.split
    mixed = run_gpu(scene, "mixed", args.steps, args.warmup, args.capacity, fuse, args.buckets_per_sphere, args.skin, args.graph)
    value = n * args.steps / mixed["seconds"]
    cbar = (mixed["n_pp"] + mixed["n_wall"]) / n
    tm = mixed["timing"]
    step_gbps = step_bytes(cbar) * n * args.steps / mixed["seconds"] / 1e9

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
        "roofline": roofline(tm, n, cbar, fuse, mixed["info"], measured_traffic(n, fuse)),
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
                      fuse, args.buckets_per_sphere, args.skin, args.graph)
        f64_value = n * args.steps / f64["seconds"]
        out["fp64_value"] = f64_value
        out["mixed_over_fp64"] = value / f64_value
        out["force_check"] = force_error(scene, args.capacity)
    if not args.no_cpu:
        cb = cpu_baseline(scene, args.capacity)
        out["cpu_baseline"] = cb
        if not args.no_fp64:
            out["fp64_gpu_over_cpu"] = out["fp64_value"] / cb["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
