"""GPU parity of the sphere-cylinder launcher (xdem_cu_cylinder_contact_force, the reference's
cu_cylinder_contact_force) against the oracle: 4,000 spheres around a turned, moving, spinning
cylinder (outside and inside the tube, at the lids from both sides, at the rims), three launches
with the histories carried and the spheres moved between them; with and without cohesion, both
damping laws, and a second cylinder (id 1) appending to the same slots. Forces, moments, rolling
sums and body forces within 1e-12 of the oracle's largest value (moments and rolling sums of the
largest force times the largest radius: atan, sin and cos of the rim region come from the device
and the host math libraries, and at deep overlaps an ulp of the rim point turns the contact
direction); slot counts and ids exact."""
import ctypes as C

import numpy as np
import pytest

import xdem_ref
from xdynamics_dll_amd import SceneData, _lib, contact_property, make_cylinder

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

R, L = 0.05, 0.12
CENTER = np.array([0.3, 0.2, 0.1])


def rot(e):
    e0, e1, e2, e3 = e
    return 2.0 * np.array([[e0 * e0 + e1 * e1 - 0.5, e1 * e2 - e0 * e3, e1 * e3 + e0 * e2],
                           [e1 * e2 + e0 * e3, e0 * e0 + e2 * e2 - 0.5, e2 * e3 - e0 * e1],
                           [e1 * e3 - e0 * e2, e2 * e3 + e0 * e1, e0 * e0 + e3 * e3 - 0.5]])


def dev(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def host(t, dtype=None):
    a = t.cpu().numpy()
    return a.view(np.uint32) if dtype == np.uint32 else a


def ptr(t):
    return C.c_void_p(t.data_ptr())


def call(name, *args):
    rc = getattr(_lib.lib(), name)(*args)
    assert rc == 0, (name, rc)


def spheres(n, seed):
    """Centres in the local frame of the cylinder (axis y) spread over every region, within a
    sphere radius of the surfaces."""
    rng = np.random.default_rng(seed)
    out = []
    r = rng.uniform(0.002, 0.004, n)
    for k in range(n):
        kind = k % 5
        phi = rng.uniform(-np.pi, np.pi)
        q = np.array([np.sin(phi), 0.0, np.cos(phi)])
        d = rng.uniform(-0.6, 0.95) * r[k]
        if kind == 0:      # outside the side wall
            c = (R + r[k] - d) * q + [0, rng.uniform(-L / 2, L / 2), 0]
        elif kind == 1:    # inside the side wall
            c = (R - r[k] + d) * q + [0, rng.uniform(-L / 2, L / 2), 0]
        elif kind == 2:    # beyond a lid, within the radius
            s = rng.choice([-1.0, 1.0])
            c = rng.uniform(0, 0.95) * R * q + [0, s * (L / 2 + r[k] - d), 0]
        elif kind == 3:    # inside at a lid (and maybe the side)
            s = rng.choice([-1.0, 1.0])
            c = rng.uniform(0, 1.0) * R * q + [0, s * (L / 2 - r[k] + d), 0]
        else:              # at a rim
            s = rng.choice([-1.0, 1.0])
            a = rng.uniform(0, np.pi / 2)
            c = (R + (r[k] - d) * np.cos(a)) * q + [0, s * (L / 2 + (r[k] - d) * np.sin(a)), 0]
        out.append(c)
    return np.array(out), r


def setup(n, seed, coh, gpu_law):
    loc, r = spheres(n, seed)
    info, dbi = make_cylinder(CENTER - [0, L / 2, 0], CENTER + [0, L / 2, 0], R, R, mass=3.0)
    e = np.array([0.9, 0.2, -0.3, 0.25])
    e /= np.linalg.norm(e)
    dbi[7:11] = e
    ed = np.array([0.3, -0.5, 0.2, 0.4])
    dbi[11:15] = ed - np.dot(ed, e) * e     # spinning: e . e_dot = 0
    dbi[4:7] = (0.02, -0.01, 0.03)
    A = rot(e)
    pos = np.concatenate([loc @ A.T + CENTER, r[:, None]], axis=1)
    rng = np.random.default_rng(seed + 1)
    vel = rng.normal(0, 0.05, (n, 3))
    cp = contact_property(1e8, 2e9, 0.3, 0.25, 0.5, 0.4, 0.02, coh)
    scene = SceneData.from_positions(pos, 2500.0, cp, dt=1e-5, vel=vel)
    return scene, info, dbi, cp


@pytest.mark.parametrize("coh,gpu_law", [(0.0, 0), (0.0, 1), (0.3, 0)])
def test_cylinder_launcher_matches_oracle(coh, gpu_law):
    n = 4000
    scene, info, dbi, cp = setup(n, 11, coh, gpu_law)
    print(f"[cylinder] n={n} coh={coh} gpu_law={gpu_law}", flush=True)
    ref = xdem_ref.RefSim(scene, quirks=dict(gpu_force_law=gpu_law))
    prm = scene.dem_parameters()
    call("xdem_setDEMSymbolicParameter", C.byref(prm))
    call("xdem_set_quirks", C.byref(_lib.Quirks(gpu_law, 0, 1, 1)))
    info2, _ = make_cylinder(CENTER - [0, L / 2, 0], CENTER + [0, L / 2, 0], R * 0.99, R * 0.99, cyl_id=1)
    d_info, d_info2 = dev(np.frombuffer(bytes(info), np.uint8).copy()), dev(np.frombuffer(bytes(info2), np.uint8).copy())
    d_dbi = dev(dbi)
    d_cp = dev(np.frombuffer(bytes(cp), np.float64).copy())
    count = np.zeros(n, np.uint32)
    ids = np.zeros((n, 3), np.uint32)
    tsd = np.zeros((n, 3, 2))
    g_count, g_ids, g_tsd = dev(count), dev(ids), dev(tsd)
    bf = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(6)]
    call("xdem_set_body_force_buffers", *(ptr(b) for b in bf))
    rbf = [np.zeros(n) for _ in range(6)]
    rng = np.random.default_rng(5)
    for step in range(3):
        f = {k: np.zeros((n, 3)) for k in ("force", "moment", "tmax")}
        rres = np.zeros(n)
        g = {k: dev(v) for k, v in f.items()}
        g_rres = dev(rres)
        g_pos, g_ep, g_vel, g_ev, g_mass = (dev(ref.pos), dev(ref.ep), dev(ref.vel), dev(ref.ev), dev(ref.mass))
        for k, d_i in ((info, d_info), (info2, d_info2)):
            call("xdem_cu_cylinder_contact_force", ptr(d_i), ptr(d_dbi), ptr(d_cp), ptr(g_pos), ptr(g_ep), ptr(g_vel),
                 ptr(g_ev), ptr(g["force"]), ptr(g["moment"]), ptr(g_mass), ptr(g["tmax"]), ptr(g_rres), ptr(g_count),
                 ptr(g_ids), ptr(g_tsd), n)
            xdem_ref.lib().ref_cylinder_contact_force(
                C.byref(ref.prm), C.byref(ref.q), C.byref(k), dbi.ctypes.data, C.byref(cp),
                *(b.ctypes.data for b in rbf), ref.pos.ctypes.data, ref.ep.ctypes.data, ref.vel.ctypes.data,
                ref.ev.ctypes.data, f["force"].ctypes.data, f["moment"].ctypes.data, ref.mass.ctypes.data,
                f["tmax"].ctypes.data, rres.ctypes.data, count.ctypes.data, ids.ctypes.data, tsd.ctypes.data, n)
            torch.cuda.synchronize()
            for j in range(6):
                gb, rb = host(bf[j]), rbf[j]
                assert np.abs(gb - rb).max() <= 1e-12 * max(np.abs(rb).max(), 1e-300), (step, k.id, j)
        # a moment is r u x F with u nearly along F: its scale is |F| r, not its own size
        fscale = max(np.abs(f["force"]).max(), 1e-300)
        scales = {"force": fscale, "moment": fscale * ref.pos[:, 3].max(), "tmax": fscale * ref.pos[:, 3].max()}
        for name in ("force", "moment", "tmax"):
            gv, rv = host(g[name]), f[name]
            assert np.abs(gv - rv).max() <= 1e-12 * scales[name], (step, name)
        assert np.abs(host(g_rres) - rres).max() <= 1e-12 * max(np.abs(rres).max(), 1e-300)
        gc, gi, gt = host(g_count, np.uint32), host(g_ids, np.uint32), host(g_tsd)
        assert np.array_equal(gc, count), step
        for i in np.nonzero(count)[0]:
            assert np.array_equal(gi[i, : count[i]], ids[i, : count[i]]), (step, i)
            assert np.abs(gt[i, : count[i]] - tsd[i, : count[i]]).max() <= 1e-12 * max(np.abs(tsd).max(), 1e-300)
        assert count.sum() > n // 2
        # move the spheres a little for the next launch (the histories carry over)
        ref.pos[:, :3] += rng.normal(0, 2e-5, (n, 3))
    w = C.c_uint()
    call("xdem_get_error_word", C.byref(w), 1)
    # the second cylinder appends: spheres with two entries before it overflow the 3 slots
    assert w.value in (0, _lib.ERR_CYL_OVERFLOW)
