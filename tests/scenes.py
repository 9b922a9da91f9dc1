"""Test scenes shared by the CPU and GPU tests (host data only)."""
from __future__ import annotations

import numpy as np

from conftest import load_fixture
from xdynamics_dll_amd import SceneData, box_walls, contact_property, make_plane


def compressed_lattice(n_side=6, pitch=1.98e-3, seed=7, coh=0.0, walls=True, rfric=0.01, contact_model=1,
                       sratio=0.0):
    """n_side^3 spheres (r = 1 mm) closer than a diameter: every lattice neighbour overlaps,
    the bottom layer overlaps the floor; random velocities drive the tangential law."""
    rng = np.random.default_rng(seed)
    g = np.arange(n_side) * pitch + 1.0e-3 - 1e-6
    x, y, z = np.meshgrid(g + 2e-3, g + 2e-3, g, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel(), np.full(x.size, 1.0e-3)], axis=1)
    vel = rng.uniform(-0.05, 0.05, size=(pos.shape[0], 3))
    cp = contact_property(1e7, 1e7, 0.3, 0.3, 0.3, 0.3, rfric, coh, sratio)
    w = box_walls(0.004 + n_side * pitch, 0.004 + n_side * pitch, 0.03) if walls else ()
    return SceneData.from_positions(pos, 2500.0, cp, walls=w, dt=1e-5, vel=vel, contact_model=contact_model)


def sphere_drop_scene(coh=0.0):
    """Reference fixture resource/sphere_drop_particles.bin: 31,329 spheres r = 0.5 mm in a
    settled bed, y up; a floor under the lowest layer, gravity -y."""
    pos, vel = load_fixture("sphere_drop_particles")
    r = pos[:, 3].max()
    yf = pos[:, 1].min() - r + 2e-6
    x0, x1 = pos[:, 0].min() - 2 * r, pos[:, 0].max() + 2 * r
    z0, z1 = pos[:, 2].min() - 2 * r, pos[:, 2].max() + 2 * r
    floor = make_plane((x0, yf, z0), (x0, yf, z1), (x1, yf, z1), (x1, yf, z0), 0)
    cp = contact_property(1e7, 1e7, 0.3, 0.3, 0.5, 0.4, 0.02, coh)
    return SceneData.from_positions(pos, 2500.0, cp, walls=[floor], dt=1e-5, gravity=(0.0, -9.81, 0.0), vel=vel)


def block_scene():
    """Reference fixture resource/particleposition.bin: 8,400 spheres r = 1 mm, z up, in a
    60 mm box (floor and four sides)."""
    pos, vel = load_fixture("particleposition")
    walls = box_walls(0.06, 0.06, 0.06, origin=(-0.03, -0.03, 0.0))
    cp = contact_property(2e7, 2e7, 0.25, 0.25, 0.4, 0.5, 0.05)
    return SceneData.from_positions(pos, 2600.0, cp, walls=walls, dt=1e-5, vel=vel)


def random_pack(n=3000, seed=3, overlap=0.02, rmin=1e-3, rmax=2e-3, coh=0.05):
    """Polydisperse random pack with small overlaps (pack shrunk after insertion), for the
    cohesion branch and mixed precision."""
    s = SceneData.build(2, seed, np_override=n)
    pos = s.pos.copy()
    c = pos[:, :3].mean(0)
    pos[:, :3] = c + (pos[:, :3] - c) * (1.0 - overlap)
    cp = contact_property(5e7, 5e7, 0.25, 0.25, 0.4, 0.5, 0.05, coh)
    wall_cp = contact_property(5e7, 5e7, 0.25, 0.25, 0.4, 0.5, 0.05, 0.0)
    rng = np.random.default_rng(seed)
    vel = rng.normal(0, 0.02, size=(pos.shape[0], 3))
    sc = SceneData.from_positions(pos, 2600.0, cp, walls=list(zip(s.planes, s.plane_dbi)), wall_cp=wall_cp,
                                  dt=5e-6, vel=vel)
    return sc
