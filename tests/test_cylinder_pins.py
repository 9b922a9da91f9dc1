"""Pins of the oracle's sphere-cylinder contact (cylinder_contact_force_kernel, impl:1327-1460,
its detection impl:1189-1325) against closed forms: the overlap and direction of every region
computed from the geometry alone, the rim point as the nearest point of the rim circle in all
four quadrants, the Hertz magnitude of a still contact, the history ids and the slot rules
(a first cylinder restarts the slots, a later one appends; an open lid has no contact)."""
import ctypes as C
import math

import numpy as np
import pytest

import xdem_ref
from xdynamics_dll_amd import _lib, contact_property, make_cylinder

R, L = 0.05, 0.12
CENTER = np.array([0.3, 0.2, 0.1])


def rot(e):
    e0, e1, e2, e3 = e
    return 2.0 * np.array([[e0 * e0 + e1 * e1 - 0.5, e1 * e2 - e0 * e3, e1 * e3 + e0 * e2],
                           [e1 * e2 + e0 * e3, e0 * e0 + e2 * e2 - 0.5, e2 * e3 - e0 * e1],
                           [e1 * e3 - e0 * e2, e2 * e3 + e0 * e1, e0 * e0 + e3 * e3 - 0.5]])


def cylinder(empty=0, thickness=0.0, cyl_id=0, turned=True):
    """A cylinder along the local y axis, turned by a fixed rotation."""
    info, dbi = make_cylinder(CENTER - [0, L / 2, 0], CENTER + [0, L / 2, 0], R, R, thickness=thickness,
                              empty=empty, cyl_id=cyl_id, mass=3.0)
    if turned:
        e = np.array([0.9, 0.2, -0.3, 0.25])
        dbi[7:11] = e / np.linalg.norm(e)
    return info, dbi


def frame(dbi):
    A = rot(dbi[7:11])
    return A, dbi[1:4]


def detect(info, dbi, p, r):
    u = np.zeros(3)
    inner = C.c_int()
    d = xdem_ref.lib().ref_cylinder_detection(C.byref(info), dbi.ctypes.data, np.ascontiguousarray(p).ctypes.data, r,
                                              u.ctypes.data, C.byref(inner))
    return d, u, inner.value


def inner_lid(info, dbi, p, r):
    u, c = np.zeros(3), np.zeros(3)
    d = xdem_ref.lib().ref_cylinder_inner_lid(C.byref(info), dbi.ctypes.data, np.ascontiguousarray(p).ctypes.data, r,
                                              u.ctypes.data, c.ctypes.data)
    return d, u, c


def at_local(dbi, loc):
    A, o = frame(dbi)
    return o + A @ np.asarray(loc, dtype=np.float64)


@pytest.mark.parametrize("phi", [0.3, 2.0, -1.2, 3.0])
def test_side_wall_from_outside_and_inside(phi):
    info, dbi = cylinder()
    A, o = frame(dbi)
    r, d = 0.004, 1.3e-4
    radial = A @ np.array([math.sin(phi), 0.0, math.cos(phi)])
    # outside: centre at R + r - d from the axis, between the ends
    p = at_local(dbi, [0, 0.02, 0]) + (R + r - d) * radial
    cd, u, inner = detect(info, dbi, p, r)
    assert cd == pytest.approx(d, rel=1e-9) and inner == 0
    assert np.allclose(u, -radial, atol=1e-12)          # from the sphere towards the axis
    # inside: centre at R - r + d
    p = at_local(dbi, [0, -0.03, 0]) + (R - r + d) * radial
    cd, u, inner = detect(info, dbi, p, r)
    assert cd == pytest.approx(d, rel=1e-9) and inner == 1
    assert np.allclose(u, radial, atol=1e-12)           # towards the wall, away from the axis


def test_on_the_axis_is_inner_with_zero_overlap():
    info, dbi = cylinder(turned=False)
    cd, u, inner = detect(info, dbi, CENTER + [0, 0.01, 0], 0.004)
    assert cd == 0 and inner == 1 and np.all(u == 0)


@pytest.mark.parametrize("top", [True, False])
def test_lid_from_outside(top):
    info, dbi = cylinder()
    A, _ = frame(dbi)
    r, d = 0.003, 2e-4
    s = 1.0 if top else -1.0
    p = at_local(dbi, [0.4 * R, s * (L / 2 + r - d), -0.2 * R])
    cd, u, inner = detect(info, dbi, p, r)
    assert cd == pytest.approx(d, rel=1e-8) and inner == 0
    assert np.allclose(u, -s * A[:, 1], atol=1e-12)     # from the end towards the body position


@pytest.mark.parametrize("qx,qz", [(1, 1), (-1, 1), (1, -1), (-1, -1), (0, -1), (1, 0)])
@pytest.mark.parametrize("top", [True, False])
def test_rim_point_is_the_nearest_point_of_the_rim_circle(qx, qz, top):
    info, dbi = cylinder()
    A, _ = frame(dbi)
    r = 0.004
    s = 1.0 if top else -1.0
    q = np.array([qx * 0.7, 0.0, qz * 0.9])
    q /= np.linalg.norm(q)
    # centre 1.5 mm beyond the end and outside the radius: nearest rim point at R q in the
    # end plane (local), at distance sqrt(dr^2 + dy^2)
    dr, dy = 0.0022, 0.0015
    loc = (R + dr) * q + np.array([0.0, s * (L / 2 + dy), 0.0])
    p = at_local(dbi, loc)
    rim = at_local(dbi, R * q + np.array([0.0, s * L / 2, 0.0]))
    cd, u, inner = detect(info, dbi, p, r)
    dist = math.hypot(dr, dy)
    assert cd == pytest.approx(r - dist, rel=1e-9) and inner == 0
    assert np.allclose(u, (rim - p) / dist, atol=1e-10)


def test_out_of_reach_is_minus_one():
    info, dbi = cylinder()
    p = at_local(dbi, [R + 0.02, L / 2 + 0.02, 0.0])
    assert detect(info, dbi, p, 0.004)[0] == -1.0


@pytest.mark.parametrize("empty,top,expect_contact", [(0, True, True), (0, False, True), (1, True, False),
                                                      (1, False, True), (2, False, False), (2, True, True)])
def test_inner_lid_and_open_ends(empty, top, expect_contact):
    info, dbi = cylinder(empty=empty)
    A, _ = frame(dbi)
    r, d = 0.004, 1e-4
    s = 1.0 if top else -1.0
    p = at_local(dbi, [0.01, s * (L / 2 - r + d), -0.005])
    cd, u, cpt = inner_lid(info, dbi, p, r)
    if expect_contact:
        assert cd == pytest.approx(d, rel=1e-8)
        assert np.allclose(u, s * A[:, 1], atol=1e-12)
        assert np.allclose(cpt, p + r * u, atol=1e-15)
    else:
        assert cd == 0.0 and np.all(u == 0)




def test_still_side_contact_is_hertz_with_ids_and_slots():
    """A sphere resting against the outer wall: |F| = (4/3) E* sqrt(r) d^1.5 along the normal,
    moment zero, body force opposite; one slot with the cylinder's id. Inside the tube at a
    closed lid: two slots, id then id + 1000."""
    from xdynamics_dll_amd import SceneData
    info, dbi = cylinder()
    A, _ = frame(dbi)
    r, d = 0.004, 1.0e-4
    radial = A @ np.array([0.6, 0.0, 0.8])
    p_out = at_local(dbi, [0, 0.01, 0]) + (R + r - d) * radial
    # inside, at the top lid and the side wall at once
    p_in = at_local(dbi, [0.0, L / 2 - r + d, 0.0]) + (R - r + d) * radial
    pos = np.array([[*p_out, r], [*p_in, r]])
    cp = contact_property(1e8, 2e9, 0.3, 0.25, 0.5, 0.4, 0.02)
    scene = SceneData.from_positions(pos, 2500.0, cp, dt=1e-5)
    ref = xdem_ref.RefSim(scene)
    dbi_d = np.ascontiguousarray(dbi)
    count = np.zeros(2, np.uint32)
    ids = np.zeros((2, 3), np.uint32)
    tsd = np.zeros((2, 3, 2))
    bf = [np.zeros(2) for _ in range(6)]
    ovf = xdem_ref.lib().ref_cylinder_contact_force(
        C.byref(ref.prm), C.byref(ref.q), C.byref(info), dbi_d.ctypes.data, C.byref(cp),
        *(b.ctypes.data for b in bf), ref.pos.ctypes.data, ref.ep.ctypes.data, ref.vel.ctypes.data,
        ref.ev.ctypes.data, ref.force.ctypes.data, ref.moment.ctypes.data, ref.mass.ctypes.data,
        ref.tmax.ctypes.data, ref.rres.ctypes.data, count.ctypes.data, ids.ctypes.data, tsd.ctypes.data, 2)
    assert ovf == 0
    Eeq = 1.0 / ((1 - 0.3 ** 2) / 1e8 + (1 - 0.25 ** 2) / 2e9)
    kn = 4.0 / 3.0 * Eeq * math.sqrt(r)
    f = ref.force[0]
    assert np.linalg.norm(f) == pytest.approx(kn * d ** 1.5, rel=1e-9)
    assert np.allclose(f / np.linalg.norm(f), radial, atol=1e-9)      # pushed away from the axis
    assert np.allclose(ref.moment[0], 0.0, atol=1e-12 * np.linalg.norm(f))
    assert np.allclose([bf[0][0], bf[1][0], bf[2][0]], -f, rtol=1e-12)
    assert count[0] == 1 and ids[0, 0] == 0
    assert count[1] == 2 and list(ids[1, :2]) == [0, 1000]
    # a second cylinder (id 2) appends after the entries already there, up to the 3 slots
    info2, _ = cylinder(cyl_id=2)
    ovf = xdem_ref.lib().ref_cylinder_contact_force(
        C.byref(ref.prm), C.byref(ref.q), C.byref(info2), dbi_d.ctypes.data, C.byref(cp),
        *(b.ctypes.data for b in bf), ref.pos.ctypes.data, ref.ep.ctypes.data, ref.vel.ctypes.data,
        ref.ev.ctypes.data, ref.force.ctypes.data, ref.moment.ctypes.data, ref.mass.ctypes.data,
        ref.tmax.ctypes.data, ref.rres.ctypes.data, count.ctypes.data, ids.ctypes.data, tsd.ctypes.data, 2)
    assert count[0] == 2 and list(ids[0, :2]) == [0, 2]
    assert count[1] == 3 and list(ids[1]) == [0, 1000, 2] and ovf == 1    # its lid entry did not fit
