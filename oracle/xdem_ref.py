"""ORACLE (test infrastructure only): Python driver of the FP64 C restatement in xdem_ref.c.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this module, as the
checker or as the timed CPU baseline; the product never imports it.

RefSim holds one scene in the reference's AoS arrays and runs the stages of
xIntegratorVV::OneStepSimulation (source/xdynamics_simulation/xIntegratorVV.cpp:26-64) with
the contacts of xContactManager::deviceCollision (xContactManager.cpp:458-472): spheres,
then walls by id, then rolling resistance, then the velocity update.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libxdem_ref.so")

PARITY_QUIRKS = dict(gpu_force_law=0, velocity_not_updated=0, stale_ft=1, flat_inertia=1)


class RefQuirks(C.Structure):
    _fields_ = [("gpu_force_law", C.c_int), ("velocity_not_updated", C.c_int), ("stale_ft", C.c_int),
                ("flat_inertia", C.c_int)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P, U, D, I = C.c_void_p, C.c_uint, C.c_double, C.c_int
        sig = {
            "ref_set_threads": ([I], None),
            "ref_get_threads": ([], I),
            "ref_vv_update_position": ([P] * 7 + [U], None),
            "ref_vv_update_velocity": ([P] * 11 + [U], None),
            "ref_calculate_hash_and_index": ([P] * 4 + [U, U], None),
            "ref_reorder_data_and_find_cell_start": ([P] * 5 + [U, U], None),
            "ref_calculate_p2p": ([P] * 18 + [U, U], U),
            "ref_plane_contact_force": ([P] * 23 + [U, U], U),
            "ref_particle_polygon_collision": ([P] * 26 + [U, U, U], U),
            "ref_decide_rolling_friction_moment": ([P] * 8 + [U], None),
            "ref_update_mesh_object_data": ([P] * 5 + [U], None),
            "ref_reduction": ([P, U], D),
            "ref_closest_pt_point_triangle": ([P, P, P], I),
            "ref_plane_detection": ([P, P, D, P], D),
            "ref_calculate_uceom": ([P] * 5, None),
            "ref_to_angular_velocity": ([P] * 3, None),
            "ref_pair_force": ([P, P, P, D, D, D, D, D, P, P, P, P, P], None),
            "ref_cohesion_limit_depth": ([D, D, P], D),
            "ref_cohesion_force": ([D, D, P, D], D),
            "ref_fit_sphere_to_triangle": ([P, P, P, D, P], None),
            "ref_set_pair_magnitude_output": ([P], None),
            "ref_plane_local_points": ([P, P, P], None),
            "ref_plane_update": ([P, P, U, P], None),
            "ref_mesh_define": ([P, P, U, P, P], None),
            "ref_cylinder_detection": ([P, P, P, D, P, P], D),
            "ref_cylinder_inner_lid": ([P, P, P, D, P, P], D),
            "ref_cylinder_contact_force": ([P] * 23 + [U], U),
        }
        for name, (a, r) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = a
            fn.restype = r
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().ref_set_threads(int(n))


def quirks_struct(q: dict | None = None) -> RefQuirks:
    d = dict(PARITY_QUIRKS)
    d.update(q or {})
    return RefQuirks(d["gpu_force_law"], d["velocity_not_updated"], d["stale_ft"], d["flat_inertia"])


def _p(a):
    return a.ctypes.data


class RefTriangle(C.Structure):
    _fields_ = [("id", C.c_int), ("tid", C.c_uint), ("P", C.c_double * 3), ("Q", C.c_double * 3),
                ("R", C.c_double * 3)]


class RefSim:
    """One scene on the CPU restatement. `scene` is an xdynamics_dll_amd.SceneData or any
    object with the same attributes (pos, vel, mass, inertia, cp_pp, planes, plane_dbi,
    plane_cp, contact_model, dt, gravity, world_origin, cell_size)."""

    def __init__(self, scene, quirks: dict | None = None, capacity: int = 12, grid=(128, 128, 128)):
        L = lib()
        self.L = L
        n = scene.pos.shape[0]
        self.np = n
        self.cap = capacity
        self.q = quirks_struct(quirks)
        self.prm = scene.dem_parameters(grid=grid)
        self.ncell = grid[0] * grid[1] * grid[2]
        self.pos = np.ascontiguousarray(scene.pos, dtype=np.float64).copy()
        self.vel = np.zeros((n, 3)) if scene.vel is None else np.ascontiguousarray(scene.vel, dtype=np.float64).copy()
        self.acc = np.zeros((n, 3))
        self.acc[:, 1] = scene.gravity[1]      # acc = (0, g.y, 0), xDiscreteElementMethodSimulation.cpp:74-80
        self.ep = np.zeros((n, 4))
        self.ep[:, 0] = 1.0
        self.ev = np.zeros((n, 4))
        self.ea = np.zeros((n, 4))
        self.force = np.zeros((n, 3))
        self.moment = np.zeros((n, 3))
        self.mass = np.ascontiguousarray(scene.mass, dtype=np.float64).copy()
        self.inertia = np.ascontiguousarray(scene.inertia, dtype=np.float64).copy()
        self.tmax = np.zeros((n, 3))
        self.rres = np.zeros(n)
        self.cp_pp = scene.cp_pp
        self.pp_count = np.zeros(n, dtype=np.uint32)
        self.pp_id = np.zeros((n, capacity), dtype=np.uint32)
        self.pp_tsd = np.zeros((n, capacity, 2))
        self.planes = list(scene.planes)
        self.plane_dbi = [np.ascontiguousarray(d, dtype=np.float64) for d in scene.plane_dbi]
        self.plane_cp = list(scene.plane_cp)
        self.pl_count = np.zeros(n, dtype=np.uint32)
        self.pl_id = np.zeros((n, 3), dtype=np.uint32)
        self.pl_tsd = np.zeros((n, 3, 2))
        self.hash = np.zeros(n, dtype=np.uint32)
        self.index = np.zeros(n, dtype=np.uint32)
        self.sorted_index = np.zeros(n, dtype=np.uint32)
        self.cstart = np.zeros(self.ncell, dtype=np.uint32)
        self.cend = np.zeros(self.ncell, dtype=np.uint32)
        self.bf = [np.zeros(n) for _ in range(6)]
        self.wrench = np.zeros((len(self.planes), 6))
        self.overflow = {"p2p": 0, "plane": 0, "cylinder": 0, "mesh": 0}
        self.mesh = None
        self.cylinders = []
        self.cy_count = np.zeros(n, dtype=np.uint32)
        self.cy_id = np.zeros((n, 3), dtype=np.uint32)
        self.cy_tsd = np.zeros((n, 3, 2))

    def add_cylinder(self, info, dbi, cp) -> None:
        """A cylinder body (xParticleCylinderContact, evaluated by cylinder_contact_force_kernel,
        impl:1327-1460): info a ref_cylinder_info-layout structure, pose dbi (15 doubles),
        material cp. Cylinders are evaluated after the walls, in the order added, on 3 slots
        per sphere."""
        self.cylinders.append(dict(info=info, dbi=np.ascontiguousarray(dbi, dtype=np.float64).copy(), cp=cp,
                                   wrench=np.zeros(6)))

    def set_cylinder_pose(self, k: int, dbi) -> None:
        self.cylinders[k]["dbi"] = np.ascontiguousarray(dbi, dtype=np.float64).copy()

    def add_mesh(self, vertices, dbi, cp) -> None:
        """One rigid mesh body (xParticleMeshObjectContact::define,
        xParticleMeshObjectContact.cpp:77-167): vertices (T, 3, 3) in the body frame, pose dbi
        (15 doubles), material cp. Proxy spheres are fitted at this pose and kept in the body
        frame; they are hashed with the spheres (entries np .. np + T - 1) and the cell grows to
        twice the largest proxy radius (xDiscreteElementMethodSimulation.cpp:91-100). The
        triangle history has the reference's 5 slots per sphere."""
        v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float64).reshape(-1, 9))
        T = v.shape[0]
        d = np.ascontiguousarray(dbi, dtype=np.float64).copy()
        dlocal = np.zeros((T, 3))
        rad = np.zeros(T)
        self.L.ref_mesh_define(_p(d), _p(v), T, _p(dlocal), _p(rad))
        tri = (RefTriangle * T)()
        for t in range(T):
            tri[t].id = t
            tri[t].tid = t
        sphere = np.zeros((T, 4))
        sphere[:, 3] = rad
        self.mesh = dict(T=T, vlist=v, dlocal=dlocal, dbi=d, cp=cp, tri=tri, sphere=sphere, wrench=np.zeros(6))
        self.prm.cell_size = max(self.prm.cell_size, 2.0 * float(rad.max()))
        n = self.np
        self.hash = np.zeros(n + T, dtype=np.uint32)
        self.index = np.zeros(n + T, dtype=np.uint32)
        self.sorted_index = np.zeros(n + T, dtype=np.uint32)
        self.ms_count = np.zeros(n, dtype=np.uint32)
        self.ms_id = np.zeros((n, 5), dtype=np.uint32)
        self.ms_tsd = np.zeros((n, 5, 2))
        self.set_mesh_pose(d)

    def set_mesh_pose(self, dbi) -> None:
        """Places the mesh at pose dbi (xParticleMeshObjectContact::update, impl:2544-2601)."""
        m = self.mesh
        m["dbi"] = np.ascontiguousarray(dbi, dtype=np.float64).copy()
        self.L.ref_update_mesh_object_data(_p(m["dbi"]), _p(m["vlist"]), _p(m["sphere"]), _p(m["dlocal"]),
                                           C.addressof(m["tri"]), m["T"])

    def triangles(self) -> np.ndarray:
        """Global triangles (T, 3, 3) at the current pose."""
        m = self.mesh
        return np.frombuffer(bytes(m["tri"]), dtype=np.float64).reshape(m["T"], 10)[:, 1:].reshape(m["T"], 3, 3)

    def set_plane_pose(self, k: int, dbi) -> None:
        """Moves wall k to pose dbi (15 doubles) from its local corners at the first pose
        (xParticlePlaneContact::update)."""
        if not hasattr(self, "_plane_local"):
            self._plane_local = []
            for j, pl in enumerate(self.planes):
                loc = np.zeros(12)
                self.L.ref_plane_local_points(C.byref(pl), _p(self.plane_dbi[j]), _p(loc))
                self._plane_local.append(loc)
        d = np.ascontiguousarray(dbi, dtype=np.float64).copy()
        pl = type(self.planes[k])()
        self.L.ref_plane_update(_p(self._plane_local[k]), _p(d), self.planes[k].id, C.byref(pl))
        self.planes[k] = pl
        self.plane_dbi[k] = d

    def write(self, name: str, values) -> None:
        getattr(self, name)[...] = np.asarray(values, dtype=np.float64).reshape(getattr(self, name).shape)

    # -- stages of one step
    def integrate_position(self):
        self.L.ref_vv_update_position(C.byref(self.prm), _p(self.pos), _p(self.ep), _p(self.vel), _p(self.ev),
                                      _p(self.acc), _p(self.ea), self.np)

    def grid(self):
        self.L.ref_calculate_hash_and_index(C.byref(self.prm), _p(self.hash), _p(self.index), _p(self.pos), 0, self.np)
        n = self.np
        if self.mesh is not None:
            # the proxy spheres follow the particles in the hash (impl:290-315, sid = np)
            self.L.ref_calculate_hash_and_index(C.byref(self.prm), _p(self.hash), _p(self.index),
                                                _p(self.mesh["sphere"]), self.np, self.mesh["T"])
            n += self.mesh["T"]
        self.L.ref_reorder_data_and_find_cell_start(_p(self.hash), _p(self.index), _p(self.cstart), _p(self.cend),
                                                    _p(self.sorted_index), n, self.ncell)

    def contacts(self, pair_magnitude: np.ndarray | None = None):
        """All contacts; pair_magnitude (np,) receives sum |F_pair| per sphere when given."""
        L, q = self.L, C.byref(self.q)
        L.ref_set_pair_magnitude_output(None if pair_magnitude is None else _p(pair_magnitude))
        self.overflow["p2p"] += L.ref_calculate_p2p(
            C.byref(self.prm), q, _p(self.pos), _p(self.ep), _p(self.vel), _p(self.ev), _p(self.force),
            _p(self.moment), _p(self.mass), _p(self.tmax), _p(self.rres), _p(self.pp_count), _p(self.pp_id),
            _p(self.pp_tsd), _p(self.sorted_index), _p(self.cstart), _p(self.cend), C.byref(self.cp_pp),
            self.np, self.cap)
        L.ref_set_pair_magnitude_output(None)
        for k, pl in enumerate(self.planes):
            self.overflow["plane"] += L.ref_plane_contact_force(
                C.byref(self.prm), q, C.byref(pl), _p(self.plane_dbi[k]), C.byref(self.plane_cp[k]),
                *(_p(b) for b in self.bf), _p(self.pos), _p(self.ep), _p(self.vel), _p(self.ev), _p(self.force),
                _p(self.moment), _p(self.mass), _p(self.tmax), _p(self.rres), _p(self.pl_count), _p(self.pl_id),
                _p(self.pl_tsd), self.np, 3)
            for c in range(6):
                self.wrench[k, c] = L.ref_reduction(_p(self.bf[c]), self.np)
        for cy in self.cylinders:
            self.overflow["cylinder"] += L.ref_cylinder_contact_force(
                C.byref(self.prm), q, C.byref(cy["info"]), _p(cy["dbi"]), C.byref(cy["cp"]),
                *(_p(b) for b in self.bf), _p(self.pos), _p(self.ep), _p(self.vel), _p(self.ev), _p(self.force),
                _p(self.moment), _p(self.mass), _p(self.tmax), _p(self.rres), _p(self.cy_count), _p(self.cy_id),
                _p(self.cy_tsd), self.np)
            for c in range(6):
                cy["wrench"][c] = L.ref_reduction(_p(self.bf[c]), self.np)
        if self.mesh is not None:
            m = self.mesh
            self.overflow["mesh"] += L.ref_particle_polygon_collision(
                C.byref(self.prm), q, C.addressof(m["tri"]), _p(m["dbi"]), *(_p(b) for b in self.bf), _p(self.pos),
                _p(self.ep), _p(self.vel), _p(self.ev), _p(self.force), _p(self.moment), _p(self.mass), _p(self.tmax),
                _p(self.rres), _p(self.ms_count), _p(self.ms_id), _p(self.ms_tsd), _p(self.sorted_index),
                _p(self.cstart), _p(self.cend), C.byref(m["cp"]), self.np, m["T"], 5)
            for c in range(6):
                m["wrench"][c] = L.ref_reduction(_p(self.bf[c]), self.np)

    def rolling(self):
        self.L.ref_decide_rolling_friction_moment(C.byref(self.prm), C.byref(self.q), _p(self.tmax), _p(self.rres),
                                                  _p(self.inertia), _p(self.ep), _p(self.ev), _p(self.moment),
                                                  self.np)

    def integrate_velocity(self):
        self.L.ref_vv_update_velocity(C.byref(self.prm), C.byref(self.q), _p(self.vel), _p(self.acc), _p(self.ep),
                                      _p(self.ev), _p(self.ea), _p(self.force), _p(self.moment), _p(self.mass),
                                      _p(self.inertia), self.np)

    def collide(self, pair_magnitude: np.ndarray | None = None):
        """Position update, grid and all contacts (force/moment/tmax/rres left set)."""
        self.integrate_position()
        self.grid()
        self.contacts(pair_magnitude)

    def finish(self):
        self.rolling()
        self.integrate_velocity()

    def step(self, n: int = 1):
        for _ in range(n):
            self.collide()
            self.finish()

    def pair_sets(self):
        """Per-sphere sorted stored sphere-contact ids."""
        return [np.sort(self.pp_id[i, : self.pp_count[i]]) for i in range(self.np)]
