"""Scenes with cylinder bodies for the step-layer tests: spheres inside a horizontal drum (the
inner side wall and the closed lids from inside) and around a horizontal rod (the outer side
wall, the lid planes beyond the ends and the rims), placed with overlaps of at most a small
fraction of a radius, so that a test can follow them through many steps."""
import numpy as np

from xdynamics_dll_amd import SceneData, contact_property, make_cylinder

# Euler parameters turning the body's y axis (the cylinder axis) onto the global z axis (a
# quarter turn about x) and onto the global x axis (a quarter turn about z, clockwise)
Y_TO_Z = np.array([np.sqrt(0.5), np.sqrt(0.5), 0.0, 0.0])
Y_TO_X = np.array([np.sqrt(0.5), 0.0, 0.0, -np.sqrt(0.5)])


def rot(e):
    e0, e1, e2, e3 = e
    return 2.0 * np.array([[e0 * e0 + e1 * e1 - 0.5, e1 * e2 - e0 * e3, e1 * e3 + e0 * e2],
                           [e1 * e2 + e0 * e3, e0 * e0 + e2 * e2 - 0.5, e2 * e3 - e0 * e1],
                           [e1 * e3 - e0 * e2, e2 * e3 + e0 * e1, e0 * e0 + e3 * e3 - 0.5]])


def rate_of(e, w):
    """Euler parameter rates of a body at e turning at angular velocity w (global):
    0.5 G(e)^T w, G = [-e1 e0 -e3 e2; -e2 e3 e0 -e1; -e3 -e2 e1 e0]."""
    e0, e1, e2, e3 = e
    G = np.array([[-e1, e0, -e3, e2], [-e2, e3, e0, -e1], [-e3, -e2, e1, e0]])
    return 0.5 * G.T @ np.asarray(w, dtype=np.float64)


def quat_mul(a, b):
    a0, a1, a2, a3 = a
    b0, b1, b2, b3 = b
    return np.array([a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3, a0 * b1 + a1 * b0 + a2 * b3 - a3 * b2,
                     a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1, a0 * b3 + a1 * b2 - a2 * b1 + a3 * b0])


def _place(candidate, n, rng, max_overlap, tries=200000):
    """n spheres from candidate(rng, k) -> (centre, r), rejecting any that overlaps an accepted
    one by more than max_overlap of the smaller radius."""
    cs, rs = np.zeros((n, 3)), np.zeros(n)
    k = 0
    for _ in range(tries):
        if k == n:
            break
        c, r = candidate(rng, k)
        if k:
            d = np.linalg.norm(cs[:k] - c, axis=1)
            if np.any(d < rs[:k] + r - max_overlap * np.minimum(rs[:k], r)):
                continue
        cs[k], rs[k] = c, r
        k += 1
    assert k == n, f"placed {k} of {n}"
    return cs, rs


def drum(n=600, seed=21, R=0.02, length=0.016, spin=20.0, coh=0.0, mass=0.0, dt=1e-6, empty=0, cyl_id=0,
         center=(0.1, 0.08, 0.05), max_overlap=0.005):
    """Spheres (r ~ U[0.8, 1.2] mm) inside a closed drum with its axis along z, turning at `spin`
    rad/s about it: a third touching the side wall, a third at the lids, a third inside.
    Returns (SceneData, CylinderInfo, dbi, ContactProperty)."""
    C = np.asarray(center, dtype=np.float64)
    A = rot(Y_TO_Z)

    def cand(rng, k):
        r = rng.uniform(0.8e-3, 1.2e-3)
        phi = rng.uniform(-np.pi, np.pi)
        q = np.array([np.sin(phi), 0.0, np.cos(phi)])
        d = rng.uniform(0.0, max_overlap) * r
        kind = k % 3
        if kind == 0:     # at the side wall
            loc = (R - r + d) * q + [0.0, rng.uniform(-length / 2 + r, length / 2 - r), 0.0]
        elif kind == 1:   # at a lid
            s = rng.choice([-1.0, 1.0])
            loc = rng.uniform(0.0, R - 1.05 * r) * q + [0.0, s * (length / 2 - r + d), 0.0]
        else:             # inside, clear of the walls
            loc = rng.uniform(0.0, R - 1.2 * r) * q + [0.0, rng.uniform(-length / 2 + 1.2 * r, length / 2 - 1.2 * r), 0.0]
        return C + A @ loc, r

    rng = np.random.default_rng(seed)
    cs, rs = _place(cand, n, rng, max_overlap)
    pos = np.concatenate([cs, rs[:, None]], axis=1)
    vel = rng.normal(0.0, 0.02, (n, 3))
    cp = contact_property(1e8, 2e9, 0.3, 0.25, 0.5, 0.4, 0.02, coh)
    scene = SceneData.from_positions(pos, 2500.0, cp, dt=dt, vel=vel, gravity=(0.0, -9.81, 0.0))
    info, dbi = make_cylinder(C - [0.0, length / 2, 0.0], C + [0.0, length / 2, 0.0], R, R, mass=mass,
                              empty=empty, cyl_id=cyl_id)
    dbi[7:11] = Y_TO_Z
    dbi[11:15] = rate_of(Y_TO_Z, [0.0, 0.0, spin])
    return scene, info, dbi, cp


def rod(n=200, seed=31, R=0.006, length=0.03, r_top=None, thickness=0.0, coh=0.0, mass=0.0, dt=1e-6,
        center=(0.05, 0.04, 0.03), max_overlap=0.005):
    """Spheres (r ~ U[0.8, 1.2] mm) around a rod with its axis along x: 24 beyond the ends
    within the rim radius (the lid planes), 24 at the rims, the rest outside the side wall.
    r_top (default R) is the rim radius, R the side wall's. Returns (SceneData, CylinderInfo,
    dbi, ContactProperty)."""
    C = np.asarray(center, dtype=np.float64)
    A = rot(Y_TO_X)
    rr = R if r_top is None else r_top

    def cand(rng, k):
        r = rng.uniform(0.8e-3, 1.2e-3)
        phi = rng.uniform(-np.pi, np.pi)
        q = np.array([np.sin(phi), 0.0, np.cos(phi)])
        d = rng.uniform(-0.5, 1.0) * max_overlap * r
        kind = 1 if k < 24 else (2 if k < 48 else 0)
        s = rng.choice([-1.0, 1.0])
        if kind == 0:     # outside the side wall
            loc = (R + r - d) * q + [0.0, rng.uniform(-length / 2, length / 2), 0.0]
        elif kind == 1:   # beyond an end, within the rim radius
            loc = rng.uniform(0.0, 0.95) * rr * q + [0.0, s * (length / 2 + r - d), 0.0]
        else:             # at a rim
            a = rng.uniform(0.05, np.pi / 2 - 0.05)
            loc = (rr + (r - d) * np.cos(a)) * q + [0.0, s * (length / 2 + (r - d) * np.sin(a)), 0.0]
        return C + A @ loc, r

    rng = np.random.default_rng(seed)
    cs, rs = _place(cand, n, rng, max_overlap)
    pos = np.concatenate([cs, rs[:, None]], axis=1)
    vel = rng.normal(0.0, 0.02, (n, 3))
    cp = contact_property(1e8, 2e9, 0.3, 0.25, 0.5, 0.4, 0.02, coh)
    scene = SceneData.from_positions(pos, 2500.0, cp, dt=dt, vel=vel, gravity=(0.0, -9.81, 0.0))
    info, dbi = make_cylinder(C - [0.0, length / 2, 0.0], C + [0.0, length / 2, 0.0], rr, R, mass=mass,
                              thickness=thickness)
    dbi[7:11] = Y_TO_X
    return scene, info, dbi, cp
