"""Python view of the step layer (xdem_ctx) and of the reference's DEM entry points.

`DemContext` owns one device context with one or many scenes. `xIntegratorVV` mirrors the
method set of the reference's xDiscreteElementMethodSimulation / xIntegratorVV
(include/xdynamics_simulation/xDiscreteElementMethodSimulation.h:20-37,
xIngegratorVV.h:13-14) on top of it. All arithmetic runs in the gfx950 library.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from .scene import PARITY_QUIRKS, SceneData


_HISTORY_KINDS = {"p2p": _lib.HISTORY_P2P, "plane": _lib.HISTORY_PLANE, "cylinder": _lib.HISTORY_CYLINDER,
                  "mesh": _lib.HISTORY_MESH}


class DemContext:
    def __init__(self, scenes, precision: str = "fp64", p2p_capacity: int = 12, quirks: dict | None = None,
                 grid_dim_max: int = 512, device: int = 0, fuse_velocity: bool = False,
                 buckets_per_sphere: int = 0, skin: float = 0.0, step_graph: int = 0,
                 rebuild_wait_limit: int = 0, meshes=(), full_blocks: int = 0, cylinders=(),
                 merge_pairs_below: int = 0, keep_order: bool = False, skin_cap: float = 0.0):
        if isinstance(scenes, SceneData):
            scenes = [scenes]
        self.scenes = list(scenes)
        self.L = _lib.lib()
        cfg = _lib.CtxConfig()
        cfg.precision = {"fp64": _lib.PRECISION_FP64, "mixed": _lib.PRECISION_MIXED}[precision]
        cfg.p2p_capacity = p2p_capacity
        q = dict(PARITY_QUIRKS)
        q.update(quirks or {})
        cfg.quirks = _lib.Quirks(q["gpu_force_law"], q["velocity_not_updated"], q["stale_ft"], q["flat_inertia"])
        cfg.grid_dim_max = grid_dim_max
        cfg.fuse_velocity = 1 if fuse_velocity else 0
        cfg.device = device
        cfg.buckets_per_sphere = buckets_per_sphere
        cfg.skin = skin   # 0: adaptive (up to skin_cap of the smallest radius); < 0: lists rebuilt every step
        cfg.step_graph = step_graph   # > 0: step() replays HIP graphs of that many steps; 0: eager launches
        cfg.rebuild_wait_limit = rebuild_wait_limit   # 0: ~1 s before a rebuild phase gives up
        cfg.full_blocks = full_blocks   # 0: 1024 blocks for the full-pass kernels
        cfg.merge_pairs_below = merge_pairs_below   # 0: 65536; 1: the pair pass always runs
        cfg.keep_order = 1 if keep_order else 0   # 0: spheres stored in bucket order (API calls keep the caller's)
        cfg.skin_cap = skin_cap   # 0: the adaptive skin goes up to 0.4 of the smallest radius
        self.precision = precision
        self.capacity = p2p_capacity
        self._h = C.c_void_p()
        _lib.check(self.L.xdem_ctx_create(C.byref(cfg), C.byref(self._h)), "xdem_ctx_create")
        for s in self.scenes:
            cs, keep = s.to_c()
            _lib.check(self.L.xdem_ctx_add_scene(self._h, C.byref(cs), None), "xdem_ctx_add_scene")
            del keep
        # mesh bodies: (scene index, vertices (T, 3, 3) in the body frame, pose (15), ContactProperty)
        self.meshes = []
        for scene, verts, dbi, cp in meshes:
            v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).reshape(-1, 9))
            m = _lib.Mesh()
            m.ntriangle = v.shape[0]
            m.vertices = v.ctypes.data_as(C.POINTER(C.c_double))
            m.dbi[:] = [float(x) for x in dbi]
            m.cp = cp
            idx = C.c_uint()
            _lib.check(self.L.xdem_ctx_add_mesh(self._h, scene, C.byref(m), C.byref(idx)), "xdem_ctx_add_mesh")
            self.meshes.append((scene, v.shape[0]))
        # cylinder bodies: (scene index, CylinderInfo, pose (15), ContactProperty), evaluated
        # after the scene's walls in this order
        self.cylinders = []
        for scene, info, dbi, cp in cylinders:
            d = np.ascontiguousarray(dbi, dtype=np.float64)
            if d.size != 15:
                raise ValueError("a cylinder pose has 15 values (MM..ED3)")
            idx = C.c_uint()
            _lib.check(self.L.xdem_ctx_add_cylinder(self._h, scene, C.byref(info), d.ctypes.data, C.byref(cp),
                                                    C.byref(idx)), "xdem_ctx_add_cylinder")
            self.cylinders.append(scene)
        _lib.check(self.L.xdem_ctx_commit(self._h), "xdem_ctx_commit")

    # ------------------------------------------------------------------ stepping
    def step(self, n: int = 1) -> None:
        _lib.check(self.L.xdem_ctx_step(self._h, n), "xdem_ctx_step")

    def collide(self) -> None:
        _lib.check(self.L.xdem_ctx_collide(self._h), "xdem_ctx_collide")

    def finish(self) -> None:
        _lib.check(self.L.xdem_ctx_finish(self._h), "xdem_ctx_finish")

    def synchronize(self) -> None:
        _lib.check(self.L.xdem_ctx_synchronize(self._h), "xdem_ctx_synchronize")

    def graph_capture(self, nsteps: int) -> None:
        _lib.check(self.L.xdem_ctx_graph_capture(self._h, nsteps), "xdem_ctx_graph_capture")

    def graph_launch(self, n: int) -> None:
        _lib.check(self.L.xdem_ctx_graph_launch(self._h, n), "xdem_ctx_graph_launch")

    # ------------------------------------------------------------------ state
    def read(self, name: str, scene: int = 0) -> np.ndarray:
        f, w = _lib.FIELD[name]
        n = self.scenes[scene].np
        out = np.empty(n * w, dtype=np.float64)
        _lib.check(self.L.xdem_ctx_read(self._h, f, scene, out.ctypes.data), "xdem_ctx_read")
        return out.reshape(n, w) if w > 1 else out

    def write(self, name: str, values: np.ndarray, scene: int = 0) -> None:
        f, w = _lib.FIELD[name]
        a = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if a.size != self.scenes[scene].np * w:
            raise ValueError(f"{name}: expected {self.scenes[scene].np * w} values, got {a.size}")
        _lib.check(self.L.xdem_ctx_write(self._h, f, scene, a.ctypes.data), "xdem_ctx_write")

    def device_ptr(self, name: str) -> int:
        p = C.c_void_p()
        _lib.check(self.L.xdem_ctx_device_ptr(self._h, _lib.FIELD[name][0], C.byref(p)), "xdem_ctx_device_ptr")
        return p.value

    def reorder(self) -> None:
        """Re-sorts the stored sphere order by the current positions (xdem_ctx_reorder)."""
        _lib.check(self.L.xdem_ctx_reorder(self._h), "xdem_ctx_reorder")

    def sphere_order(self, scene: int = 0) -> np.ndarray:
        """order[j]: the caller's index of the sphere stored j-th (the rows of device_ptr)."""
        out = np.empty(self.scenes[scene].np, dtype=np.uint32)
        _lib.check(self.L.xdem_ctx_sphere_order(self._h, scene, out.ctypes.data_as(C.POINTER(C.c_uint))),
                   "xdem_ctx_sphere_order")
        return out

    def history(self, kind: str = "p2p", scene: int = 0):
        n = self.scenes[scene].np
        k = _HISTORY_KINDS[kind]
        cap = {"p2p": self.capacity, "plane": 3, "cylinder": 3, "mesh": 5}[kind]
        count = np.empty(n, dtype=np.uint32)
        ids = np.empty(n * cap, dtype=np.uint32)
        tsd = np.empty(n * cap * 2, dtype=np.float64)
        _lib.check(self.L.xdem_ctx_read_history(self._h, k, scene, count.ctypes.data, ids.ctypes.data,
                                                tsd.ctypes.data), "xdem_ctx_read_history")
        return count, ids.reshape(n, cap), tsd.reshape(n, cap, 2)

    def write_history(self, kind: str, count, ids, tsd, scene: int = 0) -> None:
        """Inverse of history(): count (np,), ids (np, cap), tsd (np, cap, 2)."""
        k = _HISTORY_KINDS[kind]
        c = np.ascontiguousarray(count, dtype=np.uint32)
        i = np.ascontiguousarray(ids, dtype=np.uint32)
        t = np.ascontiguousarray(tsd, dtype=np.float64)
        _lib.check(self.L.xdem_ctx_write_history(self._h, k, scene, c.ctypes.data, i.ctypes.data, t.ctypes.data),
                   "xdem_ctx_write_history")

    def save_part(self, path: str, ct: float, scene: int = 0) -> None:
        """Part file of one scene in the reference's step-result layout (xResultManager.cpp:828-925)."""
        _lib.check(self.L.xdem_ctx_save_part(self._h, scene, os.fsencode(path), ct), "xdem_ctx_save_part")

    def load_part(self, path: str, scene: int = 0, history: bool = False, positions_only: bool = False) -> float:
        """Restores the state (and with history=True the contact histories) from a part file;
        returns its time (xDynamicSimulator.cpp:392-492). positions_only: the positions alone
        (xParticleManager::SetCurrentParticlesFromPartResult)."""
        ct = C.c_double()
        flags = (_lib.PART_HISTORY if history else 0) | (_lib.PART_POSITIONS if positions_only else 0)
        _lib.check(self.L.xdem_ctx_load_part(self._h, scene, os.fsencode(path), flags, C.byref(ct)),
                   "xdem_ctx_load_part")
        return ct.value

    def error_word(self, clear: bool = True) -> int:
        w = C.c_uint()
        _lib.check(self.L.xdem_ctx_error_word(self._h, C.byref(w), 1 if clear else 0), "xdem_ctx_error_word")
        return w.value

    def wrench(self, scene: int = 0) -> np.ndarray:
        nw = len(self.scenes[scene].planes)
        out = np.zeros(6 * max(nw, 1))
        _lib.check(self.L.xdem_ctx_body_wrench(self._h, scene, out.ctypes.data), "xdem_ctx_body_wrench")
        return out[: 6 * nw].reshape(nw, 6)

    def set_body_pose(self, scene: int, wall: int, dbi) -> None:
        a = np.ascontiguousarray(dbi, dtype=np.float64)
        _lib.check(self.L.xdem_ctx_set_body_pose(self._h, scene, wall, a.ctypes.data), "xdem_ctx_set_body_pose")

    def set_mesh_pose(self, mesh: int, dbi) -> None:
        a = np.ascontiguousarray(dbi, dtype=np.float64)
        _lib.check(self.L.xdem_ctx_set_mesh_pose(self._h, mesh, a.ctypes.data), "xdem_ctx_set_mesh_pose")

    def set_cylinder_pose(self, cylinder: int, dbi) -> None:
        a = np.ascontiguousarray(dbi, dtype=np.float64)
        _lib.check(self.L.xdem_ctx_set_cylinder_pose(self._h, cylinder, a.ctypes.data), "xdem_ctx_set_cylinder_pose")

    def cylinder_wrench(self, cylinder: int = 0) -> np.ndarray:
        out = np.zeros(6, dtype=np.float64)
        _lib.check(self.L.xdem_ctx_cylinder_wrench(self._h, cylinder, out.ctypes.data), "xdem_ctx_cylinder_wrench")
        return out

    def mesh_wrench(self, mesh: int = 0) -> np.ndarray:
        out = np.zeros(6)
        _lib.check(self.L.xdem_ctx_mesh_wrench(self._h, mesh, out.ctypes.data), "xdem_ctx_mesh_wrench")
        return out

    def contact_count(self):
        a, b = C.c_ulonglong(), C.c_ulonglong()
        _lib.check(self.L.xdem_ctx_contact_count(self._h, C.byref(a), C.byref(b)), "xdem_ctx_contact_count")
        return a.value, b.value

    def info(self) -> dict:
        i = _lib.CtxInfo()
        _lib.check(self.L.xdem_ctx_info_get(self._h, C.byref(i)), "xdem_ctx_info_get")
        return {n: getattr(i, n) for n, _ in i._fields_}

    def set_timing(self, on: bool) -> None:
        _lib.check(self.L.xdem_ctx_set_timing(self._h, 1 if on else 0), "xdem_ctx_set_timing")

    def timing(self):
        out = (C.c_double * 9)()
        _lib.check(self.L.xdem_ctx_timing(self._h, out), "xdem_ctx_timing")
        return {"contact_ms": out[0], "grid_ms": out[1], "steps": int(out[2]), "finish_ms": out[3],
                "touch_ms": out[4], "lite_move_ms": out[5], "collide_ms": out[6], "mesh_pass_ms": out[7],
                "pair_pass_ms": out[8]}

    def close(self) -> None:
        if self._h:
            self.L.xdem_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(scenes, precision: str = "mixed", fuse_velocity: bool = True, p2p_capacity: int = 12,
             warmup: int = 20, steps: int = 40, cache_path: str | None = None, step_graph: int = 0) -> dict:
    """Launch shapes of the step layer for these scenes (xdem_autotune): full-pass blocks,
    buckets per sphere and the adaptive skin's cap, chosen by timing short runs, cached per
    scene signature in cache_path. Returns the choice, whether it came from the cache, and
    every trial's time."""
    if isinstance(scenes, SceneData):
        scenes = [scenes]
    L = _lib.lib()
    cfg = _lib.CtxConfig()
    cfg.precision = {"fp64": _lib.PRECISION_FP64, "mixed": _lib.PRECISION_MIXED}[precision]
    cfg.p2p_capacity = p2p_capacity
    q = dict(PARITY_QUIRKS)
    cfg.quirks = _lib.Quirks(q["gpu_force_law"], q["velocity_not_updated"], q["stale_ft"], q["flat_inertia"])
    cfg.fuse_velocity = 1 if fuse_velocity else 0
    cfg.step_graph = step_graph
    keeps = [sc.to_c() for sc in scenes]
    arr = (_lib.Scene * len(scenes))(*(k[0] for k in keeps))
    res = _lib.AutotuneResult()
    path = None if cache_path is None else os.fsencode(cache_path)
    _lib.check(L.xdem_autotune(arr, len(scenes), C.byref(cfg), warmup, steps, path, C.byref(res)), "xdem_autotune")
    del keeps
    return {"full_blocks": res.full_blocks, "buckets_per_sphere": res.buckets_per_sphere, "skin_cap": res.skin_cap,
            "ms_per_step": res.ms_per_step, "from_cache": bool(res.from_cache), "key": res.key.decode(),
            "trials": [{"full_blocks": res.trial_full_blocks[k], "buckets_per_sphere": res.trial_buckets_per_sphere[k],
                        "skin_cap": res.trial_skin_cap[k], "ms_per_step": res.trial_ms_per_step[k]}
                       for k in range(res.ntrial)]}


class xIntegratorVV:
    """Method set of xDiscreteElementMethodSimulation + xIntegratorVV over DemContext."""

    def __init__(self):
        self.ctx: DemContext | None = None
        self._scenes: list[SceneData] = []

    def Initialize(self, scenes, precision: str = "fp64", **kw) -> int:
        kw.setdefault("keep_order", True)   # Position()/Velocity() are device pointers in the caller's order
        self.ctx = DemContext(scenes, precision=precision, **kw)
        self._scenes = self.ctx.scenes
        return 0

    def Initialized(self) -> bool:
        return self.ctx is not None

    def OneStepSimulation(self, ct: float, cstep: int) -> int:
        self.ctx.step(1)
        return 0

    def SaveStepResult(self, pt: int) -> dict:
        return {n: self.ctx.read(n) for n in ("pos", "vel", "acc", "ep", "ev", "ea")}

    def CriticalTimeStep(self, min_rad: float) -> float:
        s = self._scenes[0]
        return _lib.host().xdem_critical_time_step(min_rad, s.density, s.cp_pp.Ei, s.cp_pp.pri)

    def updateObjectFromMBD(self, poses: dict | None = None) -> None:
        for (scene, wall), dbi in (poses or {}).items():
            self.ctx.set_body_pose(scene, wall, dbi)

    def Position(self) -> int:
        return self.ctx.device_ptr("pos")

    def Velocity(self) -> int:
        return self.ctx.device_ptr("vel")

    def num_particles(self) -> int:
        return sum(s.np for s in self._scenes)

    def set_from_part_result(self, path: str) -> float:
        """xDynamicsSimulator::set_from_part_result: the state from a part file, contact
        histories kept as they are (the reference's contact-manager restart is empty)."""
        return self.ctx.load_part(path)

    def set_dem_data(self, cpos, pos, vel, acc, ep, ev, ea) -> int:
        for name, a in (("pos", pos), ("vel", vel), ("acc", acc), ("ep", ep), ("ev", ev), ("ea", ea)):
            if a is not None:
                self.ctx.write(name, a)
        return 0
