"""Scenes for the DEM step: spheres, walls, materials and step settings in host memory.

A scene is what the reference assembles from its Excel model before
xDiscreteElementMethodSimulation::Initialize (xDiscreteElementMethodSimulation.cpp:35-246):
positions with radii, masses and inertias (xParticleManager.cpp:1111-1131), walls as
device_plane_info plus a 15-double body pose (xParticlePlaneContact.cpp:96-131), the
contact properties of each contact object (xContact.cpp:304-322), gravity, time step and
the neighbour grid (origin (-1,-1,-1), cell = 2 x largest radius,
xDiscreteElementMethodSimulation.cpp:94-102).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib

PARITY_QUIRKS = dict(gpu_force_law=0, velocity_not_updated=0, stale_ft=1, flat_inertia=1)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def contact_property(E_i: float, E_j: float, nu_i: float, nu_j: float, rest: float, fric: float,
                     rfric: float, coh: float = 0.0, sratio: float = 0.0) -> _lib.ContactProperty:
    cp = _lib.ContactProperty()
    _lib.host().xdem_contact_property_make(E_i, E_j, nu_i, nu_j, rest, fric, rfric, coh, sratio, C.byref(cp))
    return cp


def make_plane(p0, p1, p2, p3, wall_id: int, mass: float = 0.0):
    """Wall from four corners (reference xPlaneObject::define + xParticlePlaneContact::update)."""
    info = _lib.PlaneInfo()
    dbi = np.zeros(15, dtype=np.float64)
    pts = [np.ascontiguousarray(np.asarray(p, dtype=np.float64)) for p in (p0, p1, p2, p3)]
    rc = _lib.host().xdem_make_plane(*(q.ctypes.data for q in pts), wall_id, mass, C.byref(info), dbi.ctypes.data)
    _lib.check(rc, "xdem_make_plane")
    return info, dbi


def box_walls(L: float, W: float, H: float, origin=(0.0, 0.0, 0.0)):
    """Floor and four side walls of [0,L]x[0,W]x[0,H] (+origin), normals inward, ids 0..4."""
    o = np.asarray(origin, dtype=np.float64)
    c = [
        [(0, 0, 0), (L, 0, 0), (L, W, 0), (0, W, 0)],
        [(0, 0, 0), (0, W, 0), (0, W, H), (0, 0, H)],
        [(L, 0, 0), (L, 0, H), (L, W, H), (L, W, 0)],
        [(0, 0, 0), (0, 0, H), (L, 0, H), (L, 0, 0)],
        [(0, W, 0), (L, W, 0), (L, W, H), (0, W, H)],
    ]
    out = []
    for k, q in enumerate(c):
        out.append(make_plane(*(o + np.asarray(p, dtype=np.float64) for p in q), k))
    return out


@dataclass
class SceneData:
    pos: np.ndarray                      # (np, 4) x, y, z, r
    mass: np.ndarray                     # (np,)
    inertia: np.ndarray                  # (np, 3)
    cp_pp: _lib.ContactProperty
    vel: np.ndarray | None = None        # (np, 3)
    planes: list = field(default_factory=list)      # PlaneInfo
    plane_dbi: list = field(default_factory=list)   # (15,) arrays
    plane_cp: list = field(default_factory=list)    # ContactProperty
    contact_model: int = 1
    dt: float = 1e-5
    nsteps: int = 1
    gravity: tuple = (0.0, 0.0, -9.81)
    world_origin: tuple = (-1.0, -1.0, -1.0)
    cell_size: float = 0.0
    density: float = 2500.0

    @property
    def np(self) -> int:
        return int(self.pos.shape[0])

    @classmethod
    def build(cls, workload: int, seed: int, np_override: int = 0, variant: int = 0) -> "SceneData":
        """Seeded generator of the benchmark workloads (see xdem_host.h)."""
        s = _lib.Scene()
        rc = _lib.host().xdem_scene_build(workload, seed, np_override, variant, C.byref(s))
        _lib.check(rc, "xdem_scene_build")
        try:
            n = s.np
            pos = np.ctypeslib.as_array(s.pos, shape=(n * 4,)).copy().reshape(n, 4)
            vel = np.ctypeslib.as_array(s.vel, shape=(n * 3,)).copy().reshape(n, 3)
            mass = np.ctypeslib.as_array(s.mass, shape=(n,)).copy()
            inertia = np.ctypeslib.as_array(s.inertia, shape=(n * 3,)).copy().reshape(n, 3)
            planes, dbis, cps = [], [], []
            for k in range(s.nplane):
                pi = _lib.PlaneInfo()
                C.memmove(C.byref(pi), C.byref(s.planes[k]), C.sizeof(pi))
                planes.append(pi)
                dbis.append(np.array([s.plane_dbi[15 * k + q] for q in range(15)], dtype=np.float64))
                cp = _lib.ContactProperty()
                C.memmove(C.byref(cp), C.byref(s.plane_cp[k]), C.sizeof(cp))
                cps.append(cp)
            cp_pp = _lib.ContactProperty()
            C.memmove(C.byref(cp_pp), C.byref(s.cp_pp), C.sizeof(cp_pp))
            return cls(pos=pos, vel=vel, mass=mass, inertia=inertia, cp_pp=cp_pp, planes=planes, plane_dbi=dbis,
                       plane_cp=cps, contact_model=s.contact_model, dt=s.dt, nsteps=s.nsteps,
                       gravity=tuple(s.gravity), world_origin=tuple(s.world_origin), cell_size=s.cell_size,
                       density=s.density)
        finally:
            _lib.host().xdem_scene_free(C.byref(s))

    @classmethod
    def from_positions(cls, pos: np.ndarray, density: float, cp_pp, walls=(), wall_cp=None, dt: float = 1e-5,
                       gravity=(0.0, 0.0, -9.81), vel: np.ndarray | None = None, contact_model: int = 1):
        """Scene from given spheres (x, y, z, r) with reference mass/inertia and grid rules."""
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 4)
        n = pos.shape[0]
        mass = np.zeros(n)
        inertia = np.zeros((n, 3))
        _lib.host().xdem_mass_inertia(pos.ctypes.data, n, density, mass.ctypes.data, inertia.ctypes.data)
        planes = [w[0] for w in walls]
        dbis = [np.asarray(w[1], dtype=np.float64) for w in walls]
        cps = [wall_cp if wall_cp is not None else cp_pp for _ in walls]
        return cls(pos=pos, vel=None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(n, 3),
                   mass=mass, inertia=inertia, cp_pp=cp_pp, planes=planes, plane_dbi=dbis, plane_cp=cps,
                   contact_model=contact_model, dt=dt, gravity=tuple(gravity),
                   cell_size=2.0 * float(pos[:, 3].max()), density=density)

    def to_c(self):
        """ctypes Scene pointing into this object's arrays (keep the returned holder alive)."""
        keep = {}
        pos = np.ascontiguousarray(self.pos, dtype=np.float64).ravel()
        vel = np.zeros(3 * self.np) if self.vel is None else np.ascontiguousarray(self.vel, dtype=np.float64).ravel()
        mass = np.ascontiguousarray(self.mass, dtype=np.float64)
        inertia = np.ascontiguousarray(self.inertia, dtype=np.float64).ravel()
        keep.update(pos=pos, vel=vel, mass=mass, inertia=inertia)
        s = _lib.Scene()
        s.np = self.np
        s.pos, s.vel, s.mass, s.inertia = _ptr(pos), _ptr(vel), _ptr(mass), _ptr(inertia)
        nw = len(self.planes)
        s.nplane = nw
        if nw:
            pl = (_lib.PlaneInfo * nw)(*self.planes)
            dbi = np.ascontiguousarray(np.concatenate([np.asarray(d, dtype=np.float64) for d in self.plane_dbi]))
            cps = (_lib.ContactProperty * nw)(*self.plane_cp)
            keep.update(pl=pl, dbi=dbi, cps=cps)
            s.planes = C.cast(pl, C.POINTER(_lib.PlaneInfo))
            s.plane_dbi = _ptr(dbi)
            s.plane_cp = C.cast(cps, C.POINTER(_lib.ContactProperty))
        s.cp_pp = self.cp_pp
        s.contact_model = self.contact_model
        s.dt = self.dt
        s.nsteps = self.nsteps
        s.gravity[:] = self.gravity
        s.world_origin[:] = self.world_origin
        s.cell_size = self.cell_size if self.cell_size > 0 else 2.0 * float(self.pos[:, 3].max())
        s.density = self.density
        s.min_radius = float(self.pos[:, 3].min())
        s.max_radius = float(self.pos[:, 3].max())
        keep["scene"] = s
        return s, keep

    def dem_parameters(self, grid=(128, 128, 128), ncell=None) -> _lib.DemParameters:
        """device_dem_parameters as xDiscreteElementMethodSimulation::Initialize fills it
        (xDiscreteElementMethodSimulation.cpp:179-210)."""
        p = _lib.DemParameters()
        p.contact_model = self.contact_model
        p.np = self.np
        p.nplane = len(self.planes)
        p.grid_size = _lib.uint3(*grid)
        p.ncell = ncell if ncell is not None else grid[0] * grid[1] * grid[2]
        p.dt = self.dt
        p.half2dt = 0.5 * self.dt * self.dt
        p.cell_size = self.cell_size if self.cell_size > 0 else 2.0 * float(self.pos[:, 3].max())
        p.gravity = _lib.double3(*self.gravity)
        p.world_origin = _lib.double3(*self.world_origin)
        return p


def make_cylinder(p0, p1, r_top: float, r_bottom: float, length: float | None = None, thickness: float = 0.0,
                  empty: int = 0, cyl_id: int = 0, mass: float = 0.0):
    """Cylinder body from its end points (xCylinderObject::SetupDataFromStructure,
    xCylinderObject.cpp:77-109, then xParticleCylinderContact::define, xParticleCylinderContact.cpp:68-80):
    the body position is the midpoint, the pose's Euler parameters the identity, the ends kept
    relative to the position, len_rr = (length, r_bottom, r_top). The rim region of the contact
    takes the body's local y axis as the cylinder axis, so a cylinder meant to have rims is
    given along y. Returns (CylinderInfo, dbi)."""
    a = np.asarray(p0, dtype=np.float64)
    b = np.asarray(p1, dtype=np.float64)
    pos = 0.5 * (a + b)
    info = _lib.CylinderInfo()
    info.id = cyl_id
    info.empty_part = empty
    info.thickness = thickness
    info.len_rr = _lib.double3(float(np.linalg.norm(b - a)) if length is None else length, r_bottom, r_top)
    info.pbase = _lib.double3(*(a - pos))
    info.ptop = _lib.double3(*(b - pos))
    dbi = np.zeros(15)
    dbi[0] = mass
    dbi[1:4] = pos
    dbi[7] = 1.0
    return info, dbi


def _generate(fn, data, *extra, seed: int = 1):
    H = _lib.host()
    n = C.c_uint()
    state = C.c_uint(seed)
    _lib.check(getattr(H, fn)(C.byref(data), C.byref(state), *extra, None, C.byref(n)), fn)
    pos = np.zeros((n.value, 4))
    _lib.check(getattr(H, fn)(C.byref(data), C.byref(state), *extra, pos.ctypes.data, C.byref(n)), fn)
    return pos[: n.value], state.value


def generate_cube(dx, dy, dz, lx, ly, lz, minr, maxr, seed: int = 1):
    """CreateCubeParticle's spheres (xParticleManager.cpp:646-712), (n, 4), and the generator
    state after the draws (seed: the C library's rand() state, 1 = never seeded)."""
    return _generate("xdem_create_cube_particles", _lib.CubeParticleData(dx, dy, dz, lx, ly, lz, minr, maxr), seed=seed)


def generate_plane(dx, layers, dz, direction, corner, minr, maxr, seed: int = 1, reference_quirks: bool = False):
    """CreatePlaneParticle's spheres (xParticleManager.cpp:714-792)."""
    d = _lib.PlaneParticleData(dx, layers, dz, *direction, *corner, minr, maxr)
    return _generate("xdem_create_plane_particles", d, 1 if reference_quirks else 0, seed=seed)


def generate_circle(diameter, center, direction, minr, maxr, seed: int = 1):
    """CreateCircleParticle's spheres (xParticleManager.cpp:799-880)."""
    d = _lib.CircleParticleData(diameter, *center, *direction, minr, maxr)
    return _generate("xdem_create_circle_particles", d, seed=seed)


def wheel_mesh(radius: float = 0.15, width: float = 0.1, n_theta: int = 170, n_y: int = 8, n_grouser: int = 12,
               grouser_height: float = 0.015, grouser_thickness: float = 0.006, n_ring: int = 188) -> np.ndarray:
    """Triangles (T, 3, 3) of a grousered wheel in its body frame, axis along y (xdem_wheel_mesh):
    4,096 triangles with the defaults."""
    n = C.c_uint()
    H = _lib.host()
    args = (radius, width, n_theta, n_y, n_grouser, grouser_height, grouser_thickness, n_ring)
    _lib.check(H.xdem_wheel_mesh(*args, None, C.byref(n)), "xdem_wheel_mesh")
    v = np.zeros(9 * n.value)
    _lib.check(H.xdem_wheel_mesh(*args, v.ctypes.data, C.byref(n)), "xdem_wheel_mesh")
    return v.reshape(n.value, 3, 3)
