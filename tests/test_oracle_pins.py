"""Pins of the oracle (the FP64 C restatement under oracle/) by checks that do not restate its
formulas: brute force, the reference's own particle files, physical identities the force
laws must satisfy, independent geometry, and conservation laws. Each check is chosen so that
a plausible slip in the restatement (a dropped term, a swapped material index, a wrong sign
or factor) fails it. CPU only.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest
from scipy.optimize import minimize_scalar
from scipy.spatial import cKDTree

import xdem_ref
from scenes import block_scene, compressed_lattice, sphere_drop_scene
from xdynamics_dll_amd import SceneData, _lib, contact_property, make_plane

L = xdem_ref.lib()


def f64(x):
    return np.ascontiguousarray(x, dtype=np.float64)


def ptr(a):
    return a.ctypes.data


def effective_modulus(Ei, Ej, nui, nuj):
    return 1.0 / ((1.0 - nui ** 2) / Ei + (1.0 - nuj ** 2) / Ej)


def effective_shear_modulus(Ei, Ej, nui, nuj):
    """Mindlin: 1/G* = (2 - nu_i)/G_i + (2 - nu_j)/G_j with G = E / (2 (1 + nu))."""
    Gi, Gj = Ei / (2 * (1 + nui)), Ej / (2 * (1 + nuj))
    return 1.0 / ((2 - nui) / Gi + (2 - nuj) / Gj)


def params(model=1, dt=1e-5):
    s = compressed_lattice(n_side=2, walls=False)
    p = s.dem_parameters()
    p.contact_model = model
    p.dt = dt
    return p


def quirks(gpu_law=0):
    return xdem_ref.quirks_struct(dict(gpu_force_law=gpu_law))


def pair_force(p, q, cp, ri, rj, mi, mj, cdist, dv, unit, sd=(0.0, 0.0), ft=(0.0, 0.0, 0.0)):
    dv, unit = f64(dv), f64(unit)
    sd, ft, fn = f64(sd).copy(), f64(ft).copy(), np.zeros(3)
    L.ref_pair_force(C.byref(p), C.byref(q), C.byref(cp), ri, rj, mi, mj, cdist, ptr(dv), ptr(unit), ptr(sd),
                     ptr(ft), ptr(fn))
    return fn, ft, sd


# A material pair with different moduli and Poisson ratios on the two sides, so that a
# swapped index in the effective moduli changes the result.
EI, EJ, NUI, NUJ, REST, FRIC = 3.0e7, 8.0e7, 0.21, 0.34, 0.35, 0.45
RI, RJ, MI, MJ = 1.3e-3, 0.8e-3, 2.1e-5, 0.6e-5
RSTAR, MSTAR = RI * RJ / (RI + RJ), MI * MJ / (MI + MJ)
ESTAR = effective_modulus(EI, EJ, NUI, NUJ)
GSTAR = effective_shear_modulus(EI, EJ, NUI, NUJ)
UNIT = np.array([0.6, 0.0, 0.8])
TANG = np.array([0.0, 1.0, 0.0])


def cp_pair(coh=0.0, sratio=0.0, fric=FRIC):
    return contact_property(EI, EJ, NUI, NUJ, REST, fric, 0.0, coh, sratio)


# ------------------------------------------------------------------ grid, sort, pair sets
def lattice_compressed():
    """The 10,000-sphere jittered lattice (2.2 mm pitch, r = 1 mm, +-0.05 mm jitter) pressed
    to 0.91 of its pitch, so that most but not all lattice neighbours overlap."""
    s = SceneData.build(1, 1001)
    pos = s.pos.copy()
    c = pos[:, :3].min(0)
    pos[:, :3] = c + (pos[:, :3] - c) * 0.91
    return SceneData.from_positions(pos, 2500.0, s.cp_pp, walls=(), dt=s.dt)


SCENES = {"sphere_drop": sphere_drop_scene, "block": block_scene, "lattice_cfg1": lattice_compressed}


@pytest.mark.parametrize("name", list(SCENES))
def test_hash_sort_and_cell_bounds(name):
    """impl:54-72, 290-360: 128^3 hash, stable sort by (hash, id), first/last+1 per cell."""
    s = SCENES[name]()
    ref = xdem_ref.RefSim(s)
    ref.grid()
    n = s.np
    g = np.floor((s.pos[:, :3] - np.asarray(s.world_origin)) / s.cell_size).astype(np.int64)
    h = (((g[:, 2] & 127) * 128 + (g[:, 1] & 127)) * 128 + (g[:, 0] & 127)).astype(np.uint32)
    order = np.argsort(h, kind="stable")
    assert np.array_equal(ref.hash, h[order])
    assert np.array_equal(ref.index, order.astype(np.uint32))
    assert np.array_equal(ref.sorted_index, order.astype(np.uint32))
    counts = np.bincount(h, minlength=ref.ncell)
    occupied = np.nonzero(counts)[0]
    first = np.searchsorted(h[order], occupied, side="left")
    assert np.array_equal(ref.cstart[occupied], first)
    assert np.array_equal(ref.cend[occupied], first + counts[occupied])
    assert np.all(ref.cstart[counts == 0] == 0xFFFFFFFF)
    assert n == counts.sum()


@pytest.mark.parametrize("name", list(SCENES))
def test_pair_sets_equal_brute_force(name):
    """impl:613-733 stores every pair with r_i + r_j - |p_j - p_i| > 0 (no cohesion here);
    the brute force uses a k-d tree for candidates and the exact predicate."""
    s = SCENES[name]()
    ref = xdem_ref.RefSim(s, capacity=16)
    ref.collide()
    assert ref.overflow["p2p"] == 0
    pos = ref.pos
    tree = cKDTree(pos[:, :3])
    cand = tree.query_pairs(2.0 * pos[:, 3].max() * (1 + 1e-9), output_type="ndarray")
    i, j = cand[:, 0], cand[:, 1]
    d = pos[j, :3] - pos[i, :3]
    dist = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
    cdist = (pos[i, 3] + pos[j, 3]) - dist
    sure = cdist > 1e-15
    unsure = np.abs(cdist) <= 1e-15
    expect = [set() for _ in range(s.np)]
    for a, b in zip(i[sure], j[sure]):
        expect[a].add(int(b))
        expect[b].add(int(a))
    skip = set(map(tuple, cand[unsure]))
    got = [set(map(int, ref.pp_id[k, : ref.pp_count[k]])) for k in range(s.np)]
    for k in range(s.np):
        extra = {m for m in got[k] ^ expect[k] if (min(k, m), max(k, m)) not in skip}
        assert not extra, (k, extra)
    assert sum(len(e) for e in expect) > s.np // 2   # the fixture really is a contact network


# ------------------------------------------------------------------------ force laws
@pytest.mark.parametrize("model", [1, 0])
def test_hertz_normal_force(model):
    """F_n = -(4/3) E* sqrt(R*) delta^1.5 along the unit vector, E* from both materials."""
    p, q = params(model), quirks()
    for delta in (1e-7, 3e-6, 2e-5):
        fn, _, _ = pair_force(p, q, cp_pair(sratio=0.3), RI, RJ, MI, MJ, delta, [0, 0, 0], UNIT)
        expect = -(4.0 / 3.0) * ESTAR * math.sqrt(RSTAR) * delta ** 1.5 * UNIT
        assert np.allclose(fn, expect, rtol=1e-13, atol=0), (delta, fn, expect)


@pytest.mark.parametrize("gpu_law", [0, 1])
def test_normal_damping_is_the_restitution_damping_ratio(gpu_law):
    """The damping coefficient beta must be the damping ratio whose linear spring-dashpot
    restitution is e: e = exp(-pi beta / sqrt(1 - beta^2)). The CPU law scales it by the
    critical damping 2 sqrt(m* k_n) (xContact.cpp:244), the kernels' law does not (impl:438)."""
    p, q = params(1), quirks(gpu_law)
    delta, v = 4e-6, 0.37
    kn = (4.0 / 3.0) * ESTAR * math.sqrt(RSTAR)
    f0, _, _ = pair_force(p, q, cp_pair(), RI, RJ, MI, MJ, delta, [0, 0, 0], UNIT)
    f1, _, _ = pair_force(p, q, cp_pair(), RI, RJ, MI, MJ, delta, v * UNIT, UNIT)
    fdn = float(np.dot(f1 - f0, UNIT))
    beta = fdn / v if gpu_law else fdn / (2.0 * math.sqrt(MSTAR * kn) * v)
    assert 0 < beta < 1
    assert math.isclose(math.exp(-math.pi * beta / math.sqrt(1 - beta * beta)), REST, rel_tol=1e-9)


def test_tangential_history_under_constant_slip_hertz_mindlin():
    """With the slip rate s' constant from rest, ds_n = dt (2n - 1) s' and s'_prev = s'
    (trapezoidal history, impl:443-447), and the CPU law's tangential force is
    St ds + vs sqrt(St m*) s' with St = 8 G* sqrt(R* delta), vs = 2 sqrt(5/6) beta."""
    dt, sdot, delta = 1e-5, 0.02, 5e-6
    p, q = params(1, dt), quirks()
    cp = cp_pair(fric=1e6)                    # no Coulomb cap
    lne = math.log(REST)
    beta = -lne / math.sqrt(lne * lne + math.pi * math.pi)
    St = 8.0 * GSTAR * math.sqrt(RSTAR * delta)
    vs = 2.0 * math.sqrt(5.0 / 6.0) * beta
    sd = (0.0, 0.0)
    for n in range(1, 8):
        _, ft, sd = pair_force(p, q, cp, RI, RJ, MI, MJ, delta, sdot * TANG, UNIT, sd=sd)
        assert math.isclose(sd[0], dt * (2 * n - 1) * sdot, rel_tol=1e-12)
        assert sd[1] == sdot
        expect = St * sd[0] + vs * math.sqrt(St * MSTAR) * sdot
        assert np.allclose(ft, expect * TANG, rtol=1e-12, atol=0)


def test_tangential_force_dhs_model():
    """DHS (model 0): k_s = k_n sratio, c_s = beta sratio; first step ds = dt s'."""
    dt, sdot, delta, ratio = 1e-5, 0.03, 5e-6, 0.7
    p, q = params(0, dt), quirks()
    kn = (4.0 / 3.0) * ESTAR * math.sqrt(RSTAR)
    lne = math.log(REST)
    beta = -lne / math.sqrt(lne * lne + math.pi * math.pi)
    _, ft, sd = pair_force(p, q, cp_pair(sratio=ratio, fric=1e6), RI, RJ, MI, MJ, delta, sdot * TANG, UNIT)
    expect = kn * ratio * (dt * sdot) + beta * ratio * sdot
    assert np.allclose(ft, expect * TANG, rtol=1e-12, atol=0)


def test_coulomb_cap():
    """Once the spring force exceeds mu |F_n|, |F_t| = mu |F_n| along the slip."""
    p, q = params(1), quirks()
    fn, ft, _ = pair_force(p, q, cp_pair(), RI, RJ, MI, MJ, 5e-6, 0.1 * TANG, UNIT, sd=(1e-3, 0.1))
    assert math.isclose(np.linalg.norm(ft), FRIC * np.linalg.norm(fn), rel_tol=1e-14)
    assert np.allclose(ft / np.linalg.norm(ft), TANG, rtol=0, atol=1e-15)


def test_stale_tangential_force_when_slip_direction_undefined():
    """impl:443-451: with |e| = 0 the tangential force keeps the value carried in."""
    p, q = params(1), quirks()
    carried = (1e-4, -2e-4, 3e-4)
    _, ft, sd = pair_force(p, q, cp_pair(), RI, RJ, MI, MJ, 5e-6, 0.2 * UNIT, UNIT, sd=(1e-6, 0.5), ft=carried)
    assert np.array_equal(ft, carried)
    assert np.array_equal(sd, [1e-6, 0.5])


# ---------------------------------------------------------------------------- cohesion
def test_jkr_force_satisfies_the_jkr_load_relation():
    """JKR: an external load F and the contact radius a satisfy
    F = 4 E* a^3 / (3 R*) - sqrt(8 pi w E* a^3). The oracle returns the adhesive term
    cf = -sqrt(8 pi w E* a^3) for a given F; recover a^3 from it and check the relation."""
    w = 0.05
    cp = cp_pair(coh=w)
    c0 = 3.0 * math.pi * w * RSTAR
    for F in (-0.5 * c0, -0.3 * c0, 0.0, 0.7 * c0, 5 * c0, 40 * c0):
        cf = L.ref_cohesion_force(RI, RJ, C.byref(cp), F)
        a3 = cf * cf / (8.0 * math.pi * w * ESTAR)
        assert cf < 0
        assert math.isclose(4.0 * ESTAR * a3 / (3.0 * RSTAR) + cf, F, rel_tol=1e-9, abs_tol=1e-12 * c0), F


def test_jkr_pull_off_force():
    """Below the pull-off load -1.5 pi w R* (textbook JKR) the contact radius stays at its
    pull-off value, so the load relation returns exactly the pull-off force."""
    w = 0.05
    cp = cp_pair(coh=w)
    F_po = -1.5 * math.pi * w * RSTAR
    for F in (F_po, 2 * F_po, 10 * F_po):
        cf = L.ref_cohesion_force(RI, RJ, C.byref(cp), F)
        a3 = cf * cf / (8.0 * math.pi * w * ESTAR)
        assert math.isclose(4.0 * ESTAR * a3 / (3.0 * RSTAR) + cf, F_po, rel_tol=1e-9)


def test_cohesion_separation_depth_is_the_fixed_grip_instability():
    """The JKR approach delta(a) = a^2/R* - sqrt(2 pi w a / E*) has its minimum (separation
    under fixed grips) at -3/4 (pi^2 w^2 R* / E*^2)^(1/3); found here by minimisation."""
    for w in (0.01, 0.05, 0.3):
        cp = cp_pair(coh=w)
        got = L.ref_cohesion_limit_depth(RI, RJ, C.byref(cp))
        f = lambda a: a * a / RSTAR - math.sqrt(2 * math.pi * w * a / ESTAR)
        res = minimize_scalar(f, bounds=(1e-12, 1e-3), method="bounded", options={"xatol": 1e-16})
        assert math.isclose(got, res.fun, rel_tol=1e-7), (w, got, res.fun)


# ------------------------------------------------------------------------ wall geometry
def _plane(Lx=0.1, Wy=0.05):
    info, dbi = make_plane((0, 0, 0), (Lx, 0, 0), (Lx, Wy, 0), (0, Wy, 0), 0)
    return info, Lx, Wy


def _detect(info, x, r):
    u = np.zeros(3)
    c = L.ref_plane_detection(C.byref(info), ptr(f64(x)), r, ptr(u))
    return c, u


def test_plane_face_edge_and_corner_distances():
    """impl:735-815 against analytic distances: face (one-sided normal -uw, Q7), edge (unit
    from the sphere towards the edge line), corner (unit from the corner to the sphere, as
    written in the reference), and no contact outside the reach."""
    info, Lx, Wy = _plane()
    r = 1e-3
    # face: 0.4 mm above the face
    c, u = _detect(info, (0.03, 0.02, 4e-4), r)
    assert math.isclose(c, r - 4e-4, rel_tol=1e-12) and np.allclose(u, [0, 0, -1], atol=1e-15)
    # edge y = 0: the centre 0.3 mm beyond the edge and 0.5 mm above the plane
    x = np.array([0.05, -3e-4, 5e-4])
    c, u = _detect(info, x, r)
    d = math.hypot(3e-4, 5e-4)
    assert math.isclose(c, r - d, rel_tol=1e-12)
    assert np.allclose(u, np.array([0, 3e-4, -5e-4]) / d, atol=1e-13)
    # edge y = Wy
    c, u = _detect(info, (0.05, Wy + 2e-4, 6e-4), r)
    assert math.isclose(c, r - math.hypot(2e-4, 6e-4), rel_tol=1e-12)
    # edge x = Lx
    c, u = _detect(info, (Lx + 1e-4, 0.02, 7e-4), r)
    assert math.isclose(c, r - math.hypot(1e-4, 7e-4), rel_tol=1e-12)
    # corner (Lx, Wy, 0)
    x = np.array([Lx + 2e-4, Wy + 3e-4, 4e-4])
    c, u = _detect(info, x, r)
    X = x - np.array([Lx, Wy, 0])
    assert math.isclose(c, r - np.linalg.norm(X), rel_tol=1e-12)
    assert np.allclose(u, X / np.linalg.norm(X), atol=1e-13)
    # out of reach beyond the corner, and beyond the edge x = 0 (the unreachable branch)
    assert _detect(info, (Lx + 9e-4, Wy + 9e-4, 9e-4), r)[0] == 0.0
    assert _detect(info, (-2e-4, 0.02, 3e-4), r)[0] == 0.0


# ------------------------------------------------------------------- triangle geometry
def closest_point_independent(a, b, c, p):
    """Projection onto the triangle's plane if it lies inside, else the nearest of the
    three clamped edge projections (not Ericson's region tests)."""
    n = np.cross(b - a, c - a)
    n = n / np.linalg.norm(n)
    q = p - np.dot(p - a, n) * n
    M = np.stack([b - a, c - a], axis=1)
    lam, *_ = np.linalg.lstsq(M, q - a, rcond=None)
    if lam[0] >= 0 and lam[1] >= 0 and lam.sum() <= 1:
        return q
    best, bd = None, np.inf
    for s, e in ((a, b), (b, c), (c, a)):
        t = np.clip(np.dot(p - s, e - s) / np.dot(e - s, e - s), 0.0, 1.0)
        x = s + t * (e - s)
        d = np.linalg.norm(p - x)
        if d < bd:
            best, bd = x, d
    return best


def test_closest_point_on_triangle():
    rng = np.random.default_rng(11)
    kinds = set()
    for _ in range(3000):
        a, b, c = rng.normal(size=(3, 3))
        p = rng.normal(size=3) * 2.0
        t = _lib.TriangleInfo()
        for name, v in zip("PQR", (a, b, c)):
            setattr(t, name, _lib.double3(*v))
        out = np.zeros(3)
        ct = L.ref_closest_pt_point_triangle(C.byref(t), ptr(f64(p)), ptr(out))
        kinds.add(ct)
        exp = closest_point_independent(a, b, c, p)
        assert np.allclose(out, exp, rtol=0, atol=1e-9), (a, b, c, p, out, exp)
    assert kinds == {0, 1, 2}


def test_triangle_proxy_sphere_passes_through_the_vertices():
    """xUtilityFunction.cpp:379-415: the proxy sphere's centre lies on the normal through
    the circumcentre, so all three vertices are on the sphere, and its offset from the
    circumcentre is ft times the radius (to the iteration's tolerance 1e-4)."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        P, Q, R = rng.normal(size=(3, 3))
        out = np.zeros(4)
        L.ref_fit_sphere_to_triangle(ptr(f64(P)), ptr(f64(Q)), ptr(f64(R)), 0.8, ptr(out))
        c, r = out[:3], out[3]
        for v in (P, Q, R):
            assert math.isclose(np.linalg.norm(v - c), r, rel_tol=1e-8)
        A = np.stack([Q - P, R - P])
        rhs = 0.5 * np.array([np.dot(Q, Q) - np.dot(P, P), np.dot(R, R) - np.dot(P, P)])
        n = np.cross(Q - P, R - P)
        cc = np.linalg.solve(np.vstack([A, n]), np.append(rhs, np.dot(n, P)))   # circumcentre
        off = c - cc
        assert np.linalg.norm(np.cross(off, n)) <= 1e-7 * np.linalg.norm(off) * np.linalg.norm(n)
        assert abs(np.linalg.norm(off) / r - 0.8) < 2e-4


# ------------------------------------------------------------------- rigid-body rotation
def G(e):
    return np.array([[-e[1], e[0], e[3], -e[2]], [-e[2], -e[3], e[0], e[1]], [-e[3], e[2], -e[1], e[0]]])


def test_uceom_satisfies_euler_equations():
    """impl:74-133: the 4x4 solve must give Euler's equations in the body frame,
    J w' + w x (J w) = n with w = 2 G(e) e', w' = 2 G(e) e'', and the unit-norm constraint
    e . e'' = -e' . e'."""
    rng = np.random.default_rng(2)
    for _ in range(200):
        e = rng.normal(size=4)
        e /= np.linalg.norm(e)
        d = rng.normal(size=4)
        d -= np.dot(d, e) * e
        J = rng.uniform(0.5, 2.0, size=3) * 1e-9
        n = rng.normal(size=3) * 1e-9
        out = np.zeros(4)
        L.ref_calculate_uceom(ptr(f64(J)), ptr(f64(e)), ptr(f64(d)), ptr(f64(n)), ptr(out))
        w = 2 * G(e) @ d
        wdot = 2 * G(e) @ out
        lhs = J * wdot + np.cross(w, J * w)
        assert np.allclose(lhs, n, rtol=1e-9, atol=1e-9 * np.abs(n).max())
        assert math.isclose(np.dot(e, out), -np.dot(d, d), rel_tol=1e-9, abs_tol=1e-12)


def test_angular_velocity_is_the_global_rate():
    """impl:34-42: w = 2 E(e) e' is the global angular velocity: A(e) (2 G(e) e') with
    A = E G^T."""
    rng = np.random.default_rng(4)
    for _ in range(100):
        e = rng.normal(size=4)
        e /= np.linalg.norm(e)
        d = rng.normal(size=4)
        d -= np.dot(d, e) * e
        E = np.array([[-e[1], e[0], -e[3], e[2]], [-e[2], e[3], e[0], -e[1]], [-e[3], -e[2], e[1], e[0]]])
        A = E @ G(e).T
        out = np.zeros(3)
        L.ref_to_angular_velocity(ptr(f64(e)), ptr(f64(d)), ptr(out))
        assert np.allclose(out, A @ (2 * G(e) @ d), rtol=1e-12, atol=1e-14)
        assert np.allclose(A @ A.T, np.eye(3), atol=1e-14)


def test_rolling_resistance_clamp():
    """impl:2420-2454 with the reference's flat inertia off: the added moment is along
    T = J dt w - Tmax, of magnitude min(Mr, |T|); sums are cleared after use."""
    rng = np.random.default_rng(8)
    n = 64
    p = params(1, 1e-5)
    q = xdem_ref.quirks_struct(dict(flat_inertia=0))
    ep = rng.normal(size=(n, 4))
    ep /= np.linalg.norm(ep, axis=1)[:, None]
    ev = rng.normal(size=(n, 4)) * 10
    ev -= (ev * ep).sum(1)[:, None] * ep
    inertia = np.repeat(rng.uniform(1e-10, 1e-9, size=(n, 1)), 3, axis=1)
    tmax = rng.normal(size=(n, 3)) * 1e-14
    rres = rng.uniform(0, 3e-13, size=n)
    moment = np.zeros((n, 3))
    w = np.array([2 * (np.array([[-e[1], e[0], -e[3], e[2]], [-e[2], e[3], e[0], -e[1]],
                                 [-e[3], -e[2], e[1], e[0]]]) @ d) for e, d in zip(ep, ev)])
    T = inertia[:, :1] * p.dt * w - tmax
    r0 = rres.copy()
    L.ref_decide_rolling_friction_moment(C.byref(p), C.byref(q), ptr(tmax), ptr(rres), ptr(f64(inertia)),
                                         ptr(ep), ptr(ev), ptr(moment), n)
    mag = np.linalg.norm(moment, axis=1)
    assert np.allclose(mag, np.minimum(r0, np.linalg.norm(T, axis=1)), rtol=1e-12)
    assert np.allclose(np.cross(moment, T), 0.0, atol=1e-12 * np.abs(T).max() * np.abs(moment).max())
    assert (np.einsum("ij,ij->i", moment, T) > 0).all()
    assert np.all(tmax == 0) and np.all(rres == 0)
    assert (r0 < np.linalg.norm(T, axis=1)).any() and (r0 > np.linalg.norm(T, axis=1)).any()


# ------------------------------------------------------------------ Verlet, conservation
@pytest.mark.parametrize("gdir", ["y", "z"])
def test_free_fall(gdir):
    """One sphere, no walls. With gravity along y the initial acceleration (0, g.y, 0)
    (xDiscreteElementMethodSimulation.cpp:74-80) is g, so x_n = x0 + g (n dt)^2 / 2; along z
    it starts at 0, so x_n = x0 + g dt^2 n (n - 1) / 2."""
    g = (0.0, -9.81, 0.0) if gdir == "y" else (0.0, 0.0, -9.81)
    pos = np.array([[0.1, 0.2, 0.3, 1e-3]])
    cp = contact_property(1e7, 1e7, 0.3, 0.3, 0.5, 0.3, 0.0)
    dt = 1e-4
    s = SceneData.from_positions(pos, 2500.0, cp, walls=(), dt=dt, gravity=g)
    ref = xdem_ref.RefSim(s)
    n = 50
    ref.step(n)
    k = n * n if gdir == "y" else n * (n - 1)
    expect = pos[0, :3] + 0.5 * np.asarray(g) * dt * dt * k
    assert np.allclose(ref.pos[0, :3], expect, rtol=1e-13, atol=1e-15)
    assert ref.pos[0, 3] == pos[0, 3]


@pytest.mark.parametrize("coh", [0.0, 0.05])
def test_newton_third_law(coh):
    """Without rotation, every pair force enters its two spheres with opposite signs, so
    the contact forces of an isolated cluster sum to zero (to rounding)."""
    s = compressed_lattice(n_side=6, walls=False, coh=coh)
    ref = xdem_ref.RefSim(s)
    ref.collide()
    assert ref.pp_count.sum() > 0
    total = ref.force.sum(0)
    assert np.abs(total).max() <= 1e-13 * np.abs(ref.force).sum()
    assert np.abs(ref.force).max() > 0


def test_wall_forces_balance_the_body_wrench():
    """The body force of each wall is minus the sum of the wall's forces on the spheres."""
    s = compressed_lattice()
    ref = xdem_ref.RefSim(s)
    bare = xdem_ref.RefSim(s)
    bare.planes, bare.plane_dbi, bare.plane_cp = [], [], []
    ref.collide()
    bare.collide()
    wall_force = (ref.force - bare.force).sum(0)
    body = ref.wrench[:, :3].sum(0)
    assert np.abs(wall_force).max() > 0
    assert np.allclose(body, -wall_force, rtol=1e-12, atol=1e-12 * np.abs(wall_force).max())


def test_reduction_is_an_fp64_sum():
    rng = np.random.default_rng(9)
    x = rng.normal(size=100003) * 10.0 ** rng.uniform(-5, 5, size=100003)
    got = L.ref_reduction(ptr(x), x.size)
    assert abs(got - math.fsum(x)) <= x.size * np.finfo(float).eps * np.abs(x).sum()


def test_mesh_face_contact_is_hertz_and_balances_the_body():
    """A sphere pressed on the face of one large triangle of a body at rest: the force is the
    Hertz normal force with R* = r (the wall side has no radius), along the face normal, and
    the body receives the opposite force (impl:1748-1835)."""
    from mesh_scenes import triangle_infos

    r, delta = 1e-3, 4e-6
    pos = np.array([[0.01, 0.01, r - delta, r]])
    cp = contact_property(EI, EJ, NUI, NUJ, REST, FRIC, 0.0)
    s = SceneData.from_positions(pos, 2500.0, cp, walls=(), dt=1e-5)
    prm = s.dem_parameters()
    q = xdem_ref.quirks_struct({})
    tri = triangle_infos(1)
    for name, v in zip("PQR", ((-0.05, -0.05, 0.0), (0.1, -0.05, 0.0), (0.01, 0.1, 0.0))):
        setattr(tri[0], name, _lib.double3(*v))
    dbi = np.zeros(15); dbi[0] = 5.0; dbi[7] = 1.0
    ncell = 128 ** 3
    sph = np.array([[0.0, 0.0, 0.0, 0.02]])
    h = np.zeros(2, np.uint32); idx = np.zeros(2, np.uint32)
    L.ref_calculate_hash_and_index(C.byref(prm), ptr(h), ptr(idx), ptr(pos), 0, 1)
    # the proxy sphere binned in the sphere's own cell, so the scan finds the triangle
    sph[0, :3] = pos[0, :3]
    L.ref_calculate_hash_and_index(C.byref(prm), ptr(h), ptr(idx), ptr(sph), 1, 1)
    cs = np.zeros(ncell, np.uint32); ce = np.zeros(ncell, np.uint32); srt = np.zeros(2, np.uint32)
    L.ref_reorder_data_and_find_cell_start(ptr(h), ptr(idx), ptr(cs), ptr(ce), ptr(srt), 2, ncell)
    F = np.zeros((1, 3)); M = np.zeros((1, 3)); T = np.zeros((1, 3)); R = np.zeros(1)
    cnt = np.zeros(1, np.uint32); ids = np.zeros((1, 5), np.uint32); tsd = np.zeros((1, 5, 2))
    bf = [np.zeros(1) for _ in range(6)]
    z4, z3 = np.zeros((1, 4)), np.zeros((1, 3))
    ep = np.array([[1.0, 0.0, 0.0, 0.0]])
    L.ref_particle_polygon_collision(C.byref(prm), C.byref(q), C.addressof(tri), ptr(dbi), *(ptr(b) for b in bf),
                                     ptr(pos), ptr(ep), ptr(z3), ptr(z4), ptr(F), ptr(M), ptr(s.mass), ptr(T), ptr(R),
                                     ptr(cnt), ptr(ids), ptr(tsd), ptr(srt), ptr(cs), ptr(ce), C.byref(cp), 1, 1, 5)
    kn = (4.0 / 3.0) * ESTAR * math.sqrt(r)
    assert cnt[0] == 1 and ids[0, 0] == 0
    fz = kn * delta ** 1.5
    # x and y only carry the rounding of the closest point (barycentric), ~1e-14 relative
    assert np.allclose(F[0], [0.0, 0.0, fz], rtol=1e-12, atol=1e-12 * fz)
    assert np.allclose([bf[0][0], bf[1][0], bf[2][0]], -F[0], rtol=1e-12, atol=1e-12 * fz)


# ------------------------------------------------------------------ mesh contact selection
def _mesh_collide(tris, sphere, vel=(0.0, 0.0, 0.0), ev=(0.0, 0.0, 0.0, 0.0), body=None, cp=None):
    """Runs the oracle's sphere-triangle pass (impl:1930-2109) for one sphere against the
    given global triangles (T, 3, 3). Every proxy centre is put at the sphere's centre, so the
    27-cell scan visits the triangles in index order. Returns (force, moment, count, ids,
    body force/moment)."""
    from mesh_scenes import triangle_infos

    tris = np.asarray(tris, dtype=np.float64)
    nt = len(tris)
    pos = np.array([sphere], dtype=np.float64)
    cp = cp if cp is not None else contact_property(EI, EJ, NUI, NUJ, REST, FRIC, 0.0)
    s = SceneData.from_positions(pos, 2500.0, cp, walls=(), dt=1e-5)
    prm = s.dem_parameters()
    q = xdem_ref.quirks_struct({})
    tri = triangle_infos(nt)
    for t in range(nt):
        for name, v in zip("PQR", tris[t]):
            setattr(tri[t], name, _lib.double3(*v))
    dbi = np.zeros(15) if body is None else np.asarray(body, dtype=np.float64).copy()
    if body is None:
        dbi[0] = 5.0
        dbi[7] = 1.0
    ncell = 128 ** 3
    sph = np.tile(np.r_[pos[0, :3], 0.0], (nt, 1))
    h = np.zeros(1 + nt, np.uint32); idx = np.zeros(1 + nt, np.uint32)
    L.ref_calculate_hash_and_index(C.byref(prm), ptr(h), ptr(idx), ptr(pos), 0, 1)
    L.ref_calculate_hash_and_index(C.byref(prm), ptr(h), ptr(idx), ptr(sph), 1, nt)
    cs = np.zeros(ncell, np.uint32); ce = np.zeros(ncell, np.uint32); srt = np.zeros(1 + nt, np.uint32)
    L.ref_reorder_data_and_find_cell_start(ptr(h), ptr(idx), ptr(cs), ptr(ce), ptr(srt), 1 + nt, ncell)
    F = np.zeros((1, 3)); M = np.zeros((1, 3)); T = np.zeros((1, 3)); R = np.zeros(1)
    cnt = np.zeros(1, np.uint32); ids = np.zeros((1, 5), np.uint32); tsd = np.zeros((1, 5, 2))
    bf = [np.zeros(1) for _ in range(6)]
    v3 = np.array([vel], dtype=np.float64)
    ep = np.array([[1.0, 0.0, 0.0, 0.0]])
    evv = np.array([ev], dtype=np.float64)
    L.ref_particle_polygon_collision(C.byref(prm), C.byref(q), C.addressof(tri), ptr(dbi), *(ptr(b) for b in bf),
                                     ptr(pos), ptr(ep), ptr(v3), ptr(evv), ptr(F), ptr(M), ptr(s.mass), ptr(T), ptr(R),
                                     ptr(cnt), ptr(ids), ptr(tsd), ptr(srt), ptr(cs), ptr(ce), C.byref(cp), 1, nt, 5)
    return F[0], M[0], int(cnt[0]), ids[0, : cnt[0]].copy(), np.array([b[0] for b in bf])


R_T, KN_T = 1e-3, (4.0 / 3.0) * ESTAR * math.sqrt(1e-3)   # sphere radius and its Hertz stiffness on a mesh


def test_ridge_gives_one_edge_contact():
    """Two triangles meeting in a ridge along x (a roof falling away at 45 degrees on both
    sides). A sphere straight above the ridge is closest to the shared edge in both triangles
    (the edge region, not the face), so each yields an edge candidate at the same point; the
    second equals the first within 1e-9 and is dropped (checkOverlab, impl:1729-1746,
    2019-2030): one contact, the Hertz force k_n delta^1.5 along +z, delta = r - h."""
    s, o = 4e-3, np.array([0.013, -0.007, 0.002])
    ridge_a, ridge_b = o, o + (s, 0.0, 0.0)
    tris = [[ridge_a, ridge_b, o + (s / 2, -s, -s)], [ridge_b, ridge_a, o + (s / 2, s, -s)]]
    h = R_T - 3e-6
    F, M, n, ids, _ = _mesh_collide(tris, (o[0] + s / 2, o[1], o[2] + h, R_T))
    assert n == 1 and ids[0] == 0
    fz = KN_T * (R_T - h) ** 1.5
    assert np.allclose(F, [0.0, 0.0, fz], rtol=1e-12, atol=1e-12 * fz)
    assert np.abs(M).max() <= 1e-12 * fz * R_T   # a normal force through the centre has no moment


def test_apex_of_a_fan_of_six_gives_one_vertex_contact():
    """A convex pyramid of six triangles sharing the apex: a sphere straight above the apex is
    in the apex's vertex region of every triangle, so six vertex candidates at the same point;
    all but the first are dropped as duplicates. One contact along +z with delta = r - h."""
    for apex in ((0.011, 0.023, 0.005), (-0.02, 0.004, 0.031)):
        F, n, ids, fz = _apex_fan(np.array(apex))
        assert n == 1 and ids[0] == 0
        assert np.allclose(F, [0.0, 0.0, fz], rtol=1e-12, atol=1e-12 * fz)


def _apex_fan(apex):
    ring = [apex + np.array([3e-3 * math.cos(k * math.pi / 3), 3e-3 * math.sin(k * math.pi / 3), -2e-3])
            for k in range(6)]
    tris = [[apex, ring[k], ring[(k + 1) % 6]] for k in range(6)]
    h = R_T - 2e-6
    F, _, n, ids, _ = _mesh_collide(tris, (apex[0], apex[1], apex[2] + h, R_T))
    return F, n, ids, KN_T * (R_T - h) ** 1.5


def test_vertex_and_edge_contacts_at_the_origin_are_dropped():
    """Reference quirk of the duplicate test (impl:1983-1984, 2015-2026): the previous edge and
    vertex contact points start at (0, 0, 0), so the first edge or vertex candidate within
    1e-9 m of the world origin counts as a duplicate and is dropped. The apex fan with its apex
    at the origin has no contact at all; face contacts there are kept."""
    F, n, _, _ = _apex_fan(np.zeros(3))
    assert n == 0 and not F.any()


def test_face_contact_hides_the_neighbours_edge_contact():
    """Two coplanar triangles sharing the diagonal of a square. A sphere whose projection lies
    inside the first triangle, 0.1 mm from the shared edge, touches the first in its face and
    the second at the shared edge (distance sqrt(h^2 + d^2)). Edges apply only when no face
    does (impl:2037-2100): one contact, the face one, k_n (r - h)^1.5 along the face normal."""
    s, o = 4e-3, np.array([0.01, 0.02, 0.003])
    a, b, c, d = (o + q for q in ((0.0, 0.0, 0.0), (s, 0.0, 0.0), (s, s, 0.0), (0.0, s, 0.0)))
    tris = [[a, b, d], [b, c, d]]
    dist = 2e-5
    off = dist / math.sqrt(2.0)
    p = o + (s / 2 - off, s / 2 - off, 0.0)   # inside the first, 20 um from the diagonal b-d
    h = R_T - 4e-6
    assert math.hypot(h, dist) < R_T          # the neighbour's edge is within reach too
    F, M, n, ids, _ = _mesh_collide(tris, (p[0], p[1], p[2] + h, R_T))
    assert n == 1 and ids[0] == 0
    fz = KN_T * (R_T - h) ** 1.5
    assert np.allclose(F, [0.0, 0.0, fz], rtol=1e-12, atol=1e-12 * fz)


def test_projection_exactly_on_a_shared_edge_is_a_face_of_both_triangles():
    """Reference quirk of closestPtPointTriangle (impl:1655-1705): the face test runs last and
    accepts va, vb, vc >= 0, so a projection exactly on an edge (one barycentric weight exactly
    0) is a face contact, overriding the edge region. With coordinates exact in binary the
    weight is exactly 0 for both triangles sharing the edge (va = d3 d6 - d5 d4 = L^4/4 - L^4/4
    for the first, vb = 0 for the second), so both apply as faces: twice the Hertz force."""
    Lq = 2.0 ** -8
    tris = [[(0.0, 0.0, 0.0), (Lq, 0.0, 0.0), (0.0, Lq, 0.0)], [(Lq, 0.0, 0.0), (Lq, Lq, 0.0), (0.0, Lq, 0.0)]]
    h = R_T - 2.0 ** -18
    F, _, n, ids, _ = _mesh_collide(tris, (Lq / 2, Lq / 2, h, R_T))
    assert n == 2 and list(ids) == [0, 1]
    fz = KN_T * (R_T - h) ** 1.5
    assert np.allclose(F, [0.0, 0.0, 2.0 * fz], rtol=1e-12, atol=1e-12 * fz)


def test_mesh_pose_update_is_the_rigid_transform():
    """ref_update_mesh_object_data (impl:2544-2601): every vertex and proxy centre is the body
    position plus A(e) times its local coordinates, A the rotation of the Euler parameters
    (checked against scipy's quaternion rotation, computed independently)."""
    from scipy.spatial.transform import Rotation

    from mesh_scenes import body_pose, fit_proxies, terrain, triangle_infos

    tris = terrain(nx=4)
    nt = len(tris)
    prox = fit_proxies(tris)
    dbi = body_pose(seed=8)
    e = dbi[7:11]
    Rm = Rotation.from_quat([e[1], e[2], e[3], e[0]]).as_matrix()   # scipy: (x, y, z, w)
    vlist = np.ascontiguousarray(tris.reshape(nt, 9))
    dlocal = np.ascontiguousarray(prox[:, :3])
    sph = np.zeros((nt, 4)); sph[:, 3] = prox[:, 3]
    tri = triangle_infos(nt)
    L.ref_update_mesh_object_data(ptr(dbi), ptr(vlist), ptr(sph), ptr(dlocal), C.addressof(tri), nt)
    got = np.frombuffer(bytes(tri), dtype=np.float64).reshape(nt, 10)[:, 1:].reshape(nt, 3, 3)
    expect = tris @ Rm.T + dbi[1:4]
    assert np.allclose(got, expect, rtol=0, atol=1e-15)
    assert np.allclose(sph[:, :3], prox[:, :3] @ Rm.T + dbi[1:4], rtol=0, atol=1e-15)
    assert np.array_equal(sph[:, 3], prox[:, 3])


# ------------------------------------------------------------------ moment assembly
def test_sphere_pair_moments_are_lever_arm_cross_force():
    """One pair with sliding and spin: the moment on each sphere is (its radius times the unit
    vector towards the other) cross its contact force, so M_i / r_i = M_j / r_j = u x F_i
    (impl:700-703). The forces come from the oracle; the lever arms from the geometry."""
    ri, rj = 1.2e-3, 0.8e-3
    d = ri + rj - 5e-6
    u = np.array([0.6, 0.0, 0.8])
    pos = np.array([[0.01, 0.01, 0.01, ri], [0.01 + d * u[0], 0.01, 0.01 + d * u[2], rj]])
    cp = contact_property(EI, EJ, NUI, NUJ, REST, FRIC, 0.0)
    vel = np.array([[0.02, -0.03, 0.0], [-0.01, 0.02, 0.01]])
    s = SceneData.from_positions(pos, 2500.0, cp, walls=(), dt=1e-5, gravity=(0.0, 0.0, 0.0), vel=vel)
    ref = xdem_ref.RefSim(s)
    ref.ev[:] = [[0.0, 3.0, -2.0, 1.0], [0.0, -1.0, 2.0, 0.5]]
    ref.grid()
    ref.contacts()
    Fi, Fj = ref.force
    assert np.linalg.norm(np.cross(Fi, u)) > 1e-3 * np.linalg.norm(Fi)   # a tangential part
    assert np.allclose(ref.moment[0], np.cross(ri * u, Fi), rtol=1e-12, atol=1e-15)
    assert np.allclose(ref.moment[1], np.cross(-rj * u, Fj), rtol=1e-12, atol=1e-15)


def test_wall_and_mesh_body_moments_are_minus_lever_arm_cross_force():
    """A sphere sliding on a wall and on a triangle: the sphere's moment is (r u) x F with u
    the unit vector from the centre to the contact point, and the body's moment is
    -(contact point - body position) x F (impl:1040-1050, 1820-1830)."""
    r, delta = 1e-3, 3e-6
    cp = contact_property(EI, EJ, NUI, NUJ, REST, FRIC, 0.0)
    # wall: the floor z = 0 of a 0.1 m square, body position at its centre
    floor = make_plane((0, 0, 0), (0.1, 0, 0), (0.1, 0.1, 0), (0, 0.1, 0), 0, mass=3.0)
    pos = np.array([[0.03, 0.07, r - delta, r]])
    s = SceneData.from_positions(pos, 2500.0, cp, walls=[floor], dt=1e-5, gravity=(0.0, 0.0, 0.0),
                                 vel=np.array([[0.05, -0.02, 0.0]]))
    ref = xdem_ref.RefSim(s)
    ref.grid()
    ref.contacts()
    F = ref.force[0]
    cpt = pos[0, :3] + r * np.array([0.0, 0.0, -1.0])
    xb = floor[1][1:4]
    assert abs(F[0]) > 1e-6 * abs(F[2])
    assert np.allclose(ref.moment[0], np.cross(cpt - pos[0, :3], F), rtol=1e-12, atol=1e-15)
    assert np.allclose(ref.wrench[0, 3:], -np.cross(cpt - xb, F), rtol=1e-12, atol=1e-15)
    assert np.allclose(ref.wrench[0, :3], -F, rtol=1e-12, atol=1e-15)
    # triangle face, body position away from the contact
    tris = [[(-0.05, -0.05, 0.0), (0.1, -0.05, 0.0), (0.01, 0.1, 0.0)]]
    body = np.zeros(15); body[0] = 5.0; body[1:4] = (0.02, -0.01, 0.005); body[7] = 1.0
    sp = (0.01, 0.01, r - delta, r)
    F, M, n, _, bfm = _mesh_collide(tris, sp, vel=(0.04, 0.03, 0.0), body=body)
    assert n == 1 and np.hypot(F[0], F[1]) > 1e-6 * abs(F[2])
    cpt = np.array([sp[0], sp[1], 0.0])
    assert np.allclose(M, np.cross(cpt - np.array(sp[:3]), F), rtol=1e-12, atol=1e-15)
    assert np.allclose(bfm[:3], -F, rtol=1e-12, atol=1e-15)
    assert np.allclose(bfm[3:], -np.cross(cpt - body[1:4], F), rtol=1e-12, atol=1e-15)


# ------------------------------------------------------------------ free rotation
@pytest.mark.parametrize("axis", [(0.0, 0.0, 1.0), (0.48, -0.6, 0.64)])
def test_free_spin_of_a_sphere(axis):
    """A sphere spinning freely (isotropic inertia, no force, no contact) against the exact
    rotation e(t) = (cos(w t / 2), sin(w t / 2) n), omega = w n constant. The position update
    normalises e every step (impl:135-147), so |e| = 1 to rounding; the rates e_dot are not
    projected back, so e . e_dot drifts by O(dt) and the scheme converges at first order:
    halving dt halves the errors of e and of omega = 2 L(e) e_dot (checked over a 2 rad turn)."""
    w = 40.0
    n = np.asarray(axis) / np.linalg.norm(axis)

    def run(dt, steps):
        pos = np.array([[0.0, 0.0, 0.0, 1e-3]])
        cp = contact_property(EI, EJ, NUI, NUJ, REST, FRIC, 0.0)
        s = SceneData.from_positions(pos, 2500.0, cp, walls=(), dt=dt, gravity=(0.0, 0.0, 0.0))
        ref = xdem_ref.RefSim(s)
        ref.ev[0] = np.r_[0.0, 0.5 * w * n]   # e = (1, 0, 0, 0): omega = 2 (e_dot_1, e_dot_2, e_dot_3)
        ref.step(steps)
        e = ref.ep[0].copy()
        ed = ref.ev[0].copy()
        om = np.zeros(3)
        L.ref_to_angular_velocity(ptr(e), ptr(ed), ptr(om))
        exact = np.r_[math.cos(0.5 * w * dt * steps), math.sin(0.5 * w * dt * steps) * n]
        return e, np.abs(e - exact).max(), np.linalg.norm(om - w * n) / w

    e1, err1, om1 = run(1e-4, 500)
    e2, err2, om2 = run(5e-5, 1000)
    for e in (e1, e2):
        assert abs(np.linalg.norm(e) - 1.0) <= 1e-14
    assert err1 < 5e-3 and om1 < 1e-2
    assert 1.8 < err1 / err2 < 2.2, (err1, err2)
    assert 1.8 < om1 / om2 < 2.2, (om1, om2)


def test_library_and_oracle_fit_the_same_proxy_spheres():
    """The step layer fits its proxy spheres with the library's FitSphereToTriangle and the
    oracle with its own restatement (xUtilityFunction.cpp:379-415): they must agree bit for bit,
    as the largest proxy radius sets the reference's cell (and so the visiting order)."""
    rng = np.random.default_rng(17)
    H = _lib.host()
    for _ in range(2000):
        P, Q, R = (np.ascontiguousarray(rng.normal(size=3) * 10.0 ** rng.uniform(-4, 0)) for _ in range(3))
        a, b = np.zeros(4), np.zeros(4)
        H.xdem_fit_sphere_to_triangle(P.ctypes.data, Q.ctypes.data, R.ctypes.data, 0.8, a.ctypes.data)
        L.ref_fit_sphere_to_triangle(ptr(P), ptr(Q), ptr(R), 0.8, ptr(b))
        assert np.array_equal(a, b) or (np.isnan(a).all() and np.isnan(b).all())


def test_wheel_mesh_geometry():
    """The grousered wheel: 4,096 triangles; rim vertices at R = 0.15 m, grouser tops at
    R + 15 mm, side rings inside R; the width spans y in [-0.05, 0.05]."""
    from xdynamics_dll_amd import wheel_mesh

    t = wheel_mesh()
    assert t.shape == (4096, 3, 3)
    v = t.reshape(-1, 3)
    r = np.hypot(v[:, 0], v[:, 2])
    assert abs(r.max() - 0.165) < 1e-12 and abs(r.min() - 0.135) < 1e-12
    assert abs(v[:, 1].min() + 0.05) < 1e-15 and abs(v[:, 1].max() - 0.05) < 1e-15
    rim = t[: 2 * 170 * 8].reshape(-1, 3)
    assert np.allclose(np.hypot(rim[:, 0], rim[:, 2]), 0.15, rtol=0, atol=1e-15)


def test_oracle_mesh_body_balances_its_spheres():
    """The oracle's step with a mesh body (RefSim.add_mesh): the body's wrench force is minus
    the sum of the triangle forces on the spheres (the sphere forces with and without the mesh)."""
    from mesh_scenes import mesh_scene

    scene, tris, _, dbi, cp = mesh_scene(n=200)
    with_mesh = xdem_ref.RefSim(scene)
    with_mesh.add_mesh(tris, dbi, cp)
    bare = xdem_ref.RefSim(scene)
    with_mesh.collide()
    bare.collide()
    f = (with_mesh.force - bare.force).sum(0)
    assert np.abs(f).max() > 0 and with_mesh.ms_count.sum() > 0
    assert np.allclose(with_mesh.mesh["wrench"][:3], -f, rtol=1e-10, atol=1e-12 * np.abs(f).max())
