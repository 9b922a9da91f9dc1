"""Triangle-mesh test scenes (host data only): a bumpy terrain of small triangles carried by a
moving, rotated rigid body, and spheres placed in face, edge and vertex contact with it."""
from __future__ import annotations

import ctypes as C

import numpy as np

import xdem_ref
from xdynamics_dll_amd import SceneData, _lib, contact_property, wheel_mesh


def terrain(nx=10, h=2.0e-3, amp=6.0e-4, seed=3):
    """(ntri, 3, 3) local vertices of an nx-by-nx grid of squares of side h, each split into
    two triangles, with vertex heights uniform in [-amp, amp] (ridges and peaks give edge and
    vertex contacts)."""
    rng = np.random.default_rng(seed)
    z = rng.uniform(-amp, amp, size=(nx + 1, nx + 1))
    v = lambda i, j: np.array([i * h, j * h, z[i, j]])
    tris = []
    for i in range(nx):
        for j in range(nx):
            a, b, c, d = v(i, j), v(i + 1, j), v(i + 1, j + 1), v(i, j + 1)
            tris.append([a, b, c])
            tris.append([a, c, d])
    return np.array(tris)


def fit_proxies(tris):
    """Proxy sphere of every triangle in the body frame (xUtilityFunction.cpp:379-415 with
    the factor 0.8 of xParticleMeshObjectContact.cpp:120): (ntri, 4) centre + radius."""
    L = xdem_ref.lib()
    out = np.zeros((len(tris), 4))
    for t, (P, Q, R) in enumerate(tris):
        o = np.zeros(4)
        L.ref_fit_sphere_to_triangle(*(np.ascontiguousarray(x).ctypes.data for x in (P, Q, R)), 0.8, o.ctypes.data)
        out[t] = o
    return out


def body_pose(seed=4):
    """MM..ED3 pose of the mesh body: 10 kg, offset, a small rotation, translating and
    spinning (e . e_dot = 0)."""
    rng = np.random.default_rng(seed)
    e = np.array([1.0, 0.03, -0.02, 0.05])
    e /= np.linalg.norm(e)
    ed = rng.normal(size=4) * 0.05
    ed -= np.dot(ed, e) * e
    dbi = np.zeros(15)
    dbi[0] = 10.0
    dbi[1:4] = (0.004, -0.003, 0.002)
    dbi[4:7] = (0.01, -0.005, -0.02)
    dbi[7:11] = e
    dbi[11:15] = ed
    return dbi


def rot(e):
    e0, e1, e2, e3 = e
    return 2.0 * np.array([[e0 * e0 + e1 * e1 - 0.5, e1 * e2 - e0 * e3, e1 * e3 + e0 * e2],
                           [e1 * e2 + e0 * e3, e0 * e0 + e2 * e2 - 0.5, e2 * e3 - e0 * e1],
                           [e1 * e3 - e0 * e2, e2 * e3 + e0 * e1, e0 * e0 + e3 * e3 - 0.5]])


def closest_distance(tris_g, p):
    """Distance from p to the mesh by brute force over all triangles (projection when it
    falls inside a triangle, else the nearest clamped edge point)."""
    a, b, c = tris_g[:, 0], tris_g[:, 1], tris_g[:, 2]
    ab, ac = b - a, c - a
    n = np.cross(ab, ac)
    n /= np.linalg.norm(n, axis=1)[:, None]
    dn = ((p - a) * n).sum(1)
    aq = (p - dn[:, None] * n) - a
    d00, d01, d11 = (ab * ab).sum(1), (ab * ac).sum(1), (ac * ac).sum(1)
    d20, d21 = (aq * ab).sum(1), (aq * ac).sum(1)
    den = d00 * d11 - d01 * d01
    v, w = (d11 * d20 - d01 * d21) / den, (d00 * d21 - d01 * d20) / den
    inside = (v >= 0) & (w >= 0) & (v + w <= 1)
    best = np.abs(dn[inside]).min() if inside.any() else np.inf

    def seg(s_, e_):
        t = np.clip(((p - s_) * (e_ - s_)).sum(1) / ((e_ - s_) ** 2).sum(1), 0.0, 1.0)
        return np.linalg.norm(p - (s_ + t[:, None] * (e_ - s_)), axis=1).min()

    return min(best, seg(a, b), seg(b, c), seg(c, a))


def mesh_scene(n=400, seed=5, coh=0.0, max_overlap=0.1, nx=10, separated=False, dt=1e-5):
    """Spheres r ~ U[0.8, 1.2] mm over an nx-by-nx terrain: 60% overlap it by up to max_overlap
    of r, 15% sit within 0.2 um outside it (the JKR separation band when cohesive), the rest just
    out of reach. With separated, a sphere overlaps the spheres placed before it by at most
    max_overlap of the smaller radius (a scene to follow through many steps; without it the
    spheres may overlap each other deeply, which only a one-step test can use). Returns
    (SceneData, local triangles, proxies, dbi, contact property). The reference's CPU Hertz law
    damps with 2 beta sqrt(m kn) dv (xContact.cpp:244), about 1.6 N s/m for these spheres: an
    explicit step of 1e-5 s is unstable for it (c dt / m > 2), so a scene followed through many
    steps takes dt = 1e-6."""
    rng = np.random.default_rng(seed)
    tris = terrain(nx=nx)
    prox = fit_proxies(tris)
    dbi = body_pose()
    A = rot(dbi[7:11])
    tris_g = tris @ A.T + dbi[1:4]
    lo, hi = tris_g.reshape(-1, 3).min(0), tris_g.reshape(-1, 3).max(0)
    pos = []
    tries = 0
    while len(pos) < n:
        tries += 1
        if tries > 100 * n:
            raise RuntimeError(f"placed {len(pos)} of {n} spheres")
        r = rng.uniform(0.8e-3, 1.2e-3)
        nrm = A[:, 2]
        # start above the terrain and move along the body normal until the distance to the
        # mesh is the target
        p = np.array([rng.uniform(lo[0] + 1e-3, hi[0] - 1e-3), rng.uniform(lo[1] + 1e-3, hi[1] - 1e-3),
                      hi[2]]) + 3e-3 * nrm
        u = rng.random()
        if u < 0.6:
            target = r * rng.uniform(1.0 - max_overlap, 1.0)   # overlapping
        elif u < 0.75:
            target = r + rng.uniform(0.0, 2e-7)          # inside the JKR separation band
        else:
            target = r * rng.uniform(1.0, 1.1)           # just out of reach
        for _ in range(30):
            d = closest_distance(tris_g, p)
            if abs(d - target) < 1e-9:
                break
            p = p + (target - d) * nrm
        if abs(closest_distance(tris_g, p) - target) >= 1e-7:
            continue
        if separated and pos:
            q = np.asarray(pos)
            d = np.linalg.norm(q[:, :3] - p, axis=1)
            if (d < q[:, 3] + r - max_overlap * np.minimum(q[:, 3], r)).any():
                continue
        pos.append([*p, r])
    pos = np.array(pos)
    cp = contact_property(4e7, 2e8, 0.28, 0.3, 0.45, 0.55, 0.03, coh)
    scene = SceneData.from_positions(pos, 2500.0, cp, walls=(), dt=dt,
                                     vel=rng.normal(0, 0.03, size=(n, 3)))
    return scene, tris, prox, dbi, cp


def triangle_infos(n):
    arr = (_lib.TriangleInfo * n)()
    for t in range(n):
        arr[t].id = t
        arr[t].tid = 0
    return arr


def wheel_over_bed(n, seed=1003, sink=1.5e-3):
    """The grousered wheel (4,096 triangles, R = 0.15 m, width 0.1 m, 40 kg) over the loose bed
    generator of n spheres, its lowest point `sink` below the highest sphere top under it."""
    scene = SceneData.build(3, seed, np_override=n)
    tris = wheel_mesh()
    assert len(tris) == 4096
    p = scene.pos
    L = scene.planes[2].xw.x
    W = scene.planes[4].xw.y
    xc, yc = 0.5 * L, 0.5 * W
    under = (np.abs(p[:, 0] - xc) < 0.06) & (np.abs(p[:, 1] - yc) < 0.05)
    top = (p[under, 2] + p[under, 3]).max()
    low = -tris.reshape(-1, 3)[:, 2].min()   # the lowest point below the axis (rim or grouser)
    dbi = np.zeros(15)
    dbi[0] = 40.0
    dbi[1:4] = (xc, yc, top + low - sink)
    dbi[7] = 1.0
    cp = contact_property(2e11, 5e7, 0.3, 0.25, 0.4, 0.6, 0.05)
    return scene, tris, dbi, cp
