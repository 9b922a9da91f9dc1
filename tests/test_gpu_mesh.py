"""GPU parity of the mesh path of the launcher layer against the oracle: mesh pose update
(impl:2544-2601), proxy-sphere hashing and sort (impl:290-360), sphere-triangle contact with
the reference's face/edge/vertex rules and 5-slot history (impl:1655-2109), body force
buffers. Forces within 1e-12 relative, everything else bit-exact."""
import ctypes as C

import numpy as np
import pytest

import xdem_ref
from mesh_scenes import mesh_scene, triangle_infos
from xdynamics_dll_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
CAP = 5


def dev(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def host(t, u32=False):
    a = t.cpu().numpy()
    return a.view(np.uint32) if u32 else a


def ptr(t):
    return C.c_void_p(t.data_ptr())


def call(name, *args):
    rc = getattr(_lib.lib(), name)(*args)
    assert rc == 0, (name, rc)


def nptr(a):
    return a.ctypes.data


@pytest.mark.parametrize("coh", [0.0, 0.05])
@pytest.mark.parametrize("gpu_law", [0, 1])
def test_mesh_contact_matches_oracle(coh, gpu_law):
    scene, tris, prox, dbi, cp = mesh_scene(coh=coh)
    n, nt = scene.np, len(tris)
    print(f"[mesh] np={n} ntri={nt} coh={coh} gpu_law={gpu_law}", flush=True)
    L = xdem_ref.lib()
    prm = scene.dem_parameters()
    q = xdem_ref.quirks_struct(dict(gpu_force_law=gpu_law))
    ncell = 128 ** 3
    rng = np.random.default_rng(1)
    pos = scene.pos.copy()
    vel = scene.vel.copy()
    ep = np.zeros((n, 4)); ep[:, 0] = 1.0
    ev = rng.normal(size=(n, 4)) * 0.5
    ev -= (ev * ep).sum(1)[:, None] * ep
    mass = scene.mass.copy()
    vlist = np.ascontiguousarray(tris.reshape(nt, 9))
    dlocal = np.ascontiguousarray(prox[:, :3])
    # --- oracle
    r_sph = np.zeros((nt, 4)); r_sph[:, 3] = prox[:, 3]
    r_tri = triangle_infos(nt)
    L.ref_update_mesh_object_data(nptr(dbi), nptr(vlist), nptr(r_sph), nptr(dlocal), C.addressof(r_tri), nt)
    r_hash = np.zeros(n + nt, np.uint32); r_index = np.zeros(n + nt, np.uint32)
    L.ref_calculate_hash_and_index(C.byref(prm), nptr(r_hash), nptr(r_index), nptr(pos), 0, n)
    # proxies go to entries n .. n + nt - 1: the launcher writes hash[sid + t] (impl:290-302)
    L.ref_calculate_hash_and_index(C.byref(prm), nptr(r_hash), nptr(r_index), nptr(r_sph), n, nt)
    r_cs = np.zeros(ncell, np.uint32); r_ce = np.zeros(ncell, np.uint32); r_sorted = np.zeros(n + nt, np.uint32)
    L.ref_reorder_data_and_find_cell_start(nptr(r_hash), nptr(r_index), nptr(r_cs), nptr(r_ce), nptr(r_sorted),
                                          n + nt, ncell)
    r_force = np.zeros((n, 3)); r_moment = np.zeros((n, 3)); r_tmax = np.zeros((n, 3)); r_rres = np.zeros(n)
    r_cnt = np.zeros(n, np.uint32); r_id = np.zeros((n, CAP), np.uint32); r_tsd = np.zeros((n, CAP, 2))
    # a previous step's history for some spheres: ids of nearby triangles with nonzero (ds, s_dot)
    for i in range(0, n, 3):
        r_cnt[i] = 2
        r_id[i, :2] = rng.integers(0, nt, 2)
        r_tsd[i, :2] = rng.normal(size=(2, 2)) * 1e-6
    h_cnt, h_id, h_tsd = r_cnt.copy(), r_id.copy(), r_tsd.copy()
    bf = [np.zeros(n) for _ in range(6)]
    ovf = L.ref_particle_polygon_collision(C.byref(prm), C.byref(q), C.addressof(r_tri), nptr(dbi),
                                           *(nptr(b) for b in bf), nptr(pos), nptr(ep), nptr(vel), nptr(ev),
                                           nptr(r_force), nptr(r_moment), nptr(mass), nptr(r_tmax), nptr(r_rres),
                                           nptr(r_cnt), nptr(r_id), nptr(r_tsd), nptr(r_sorted), nptr(r_cs),
                                           nptr(r_ce), C.byref(cp), n, nt, CAP)
    # --- GPU, through the launchers
    gq = _lib.Quirks(gpu_law, 0, 1, 1)
    gprm = scene.dem_parameters()
    gprm.nsphere = nt   # the launchers' workspace covers the particles and the proxy spheres
    call("xdem_setDEMSymbolicParameter", C.byref(gprm))
    call("xdem_set_quirks", C.byref(gq))
    d_vlist, d_dlocal, d_dbi = dev(vlist), dev(dlocal), dev(dbi)
    g_sph0 = np.zeros((nt, 4)); g_sph0[:, 3] = prox[:, 3]
    d_sph = dev(g_sph0)
    d_tri = dev(np.frombuffer(bytes(triangle_infos(nt)), dtype=np.uint8).copy())
    call("xdem_cu_update_meshObjectData", ptr(d_vlist), ptr(d_sph), ptr(d_dlocal), ptr(d_tri), ptr(d_dbi), nt)
    d_pos, d_vel, d_ep, d_ev, d_mass = dev(pos), dev(vel), dev(ep), dev(ev), dev(mass)
    d_hash = dev(np.zeros(n + nt, np.uint32)); d_index = dev(np.zeros(n + nt, np.uint32))
    call("xdem_cu_calculateHashAndIndex", ptr(d_hash), ptr(d_index), ptr(d_pos), 0, n)
    call("xdem_cu_calculateHashAndIndexForPolygonSphere", ptr(d_hash), ptr(d_index), n, nt, ptr(d_sph))
    d_cs, d_ce = dev(np.zeros(ncell, np.uint32)), dev(np.zeros(ncell, np.uint32))
    d_sorted = dev(np.zeros(n + nt, np.uint32))
    call("xdem_cu_reorderDataAndFindCellStart", ptr(d_hash), ptr(d_index), ptr(d_cs), ptr(d_ce), ptr(d_sorted),
         n + nt, ncell)
    d_force, d_moment = dev(np.zeros((n, 3))), dev(np.zeros((n, 3)))
    d_tmax, d_rres = dev(np.zeros((n, 3))), dev(np.zeros(n))
    d_cnt, d_id, d_tsd = dev(h_cnt), dev(h_id), dev(h_tsd)
    d_cp = dev(np.frombuffer(bytes(cp), dtype=np.float64).copy())
    d_bf = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(6)]
    call("xdem_set_body_force_buffers", *(ptr(b) for b in d_bf))
    call("xdem_cu_particle_polygonObject_collision", ptr(d_tri), ptr(d_dbi), ptr(d_pos), ptr(d_ep), ptr(d_vel),
         ptr(d_ev), ptr(d_force), ptr(d_moment), ptr(d_mass), ptr(d_tmax), ptr(d_rres), ptr(d_cnt), ptr(d_id),
         ptr(d_tsd), ptr(d_sph), ptr(d_sorted), ptr(d_cs), ptr(d_ce), ptr(d_cp), n, nt)
    torch.cuda.synchronize()
    # --- mesh update and grid: bit-exact
    g_tri = np.frombuffer(host(d_tri).tobytes(), dtype=np.float64).reshape(nt, 10)[:, 1:]
    r_tri_arr = np.frombuffer(bytes(r_tri), dtype=np.float64).reshape(nt, 10)[:, 1:]
    assert np.array_equal(g_tri, r_tri_arr)
    assert np.array_equal(host(d_sph), r_sph)
    assert np.array_equal(host(d_hash, True), r_hash)
    assert np.array_equal(host(d_sorted, True), r_sorted)
    assert np.array_equal(host(d_cs, True), r_cs)
    assert np.array_equal(host(d_ce, True), r_ce)
    # --- contacts
    assert r_cnt.sum() > n // 2, "the scene should put most spheres in contact"
    assert np.array_equal(host(d_cnt, True), r_cnt)
    gid = host(d_id, True)
    for i in range(n):
        assert np.array_equal(gid[i, : r_cnt[i]], r_id[i, : r_cnt[i]]), i
    gts = host(d_tsd)
    for i in range(n):
        c = r_cnt[i]
        assert np.allclose(gts[i, :c], r_tsd[i, :c], rtol=1e-12, atol=1e-300), i
    for name, g, r in (("force", d_force, r_force), ("moment", d_moment, r_moment), ("tmax", d_tmax, r_tmax),
                       ("rres", d_rres, r_rres)):
        g = host(g)
        assert np.abs(g - r).max() <= 1e-12 * max(np.abs(r).max(), 1e-300), name
    for c in range(6):
        g = host(d_bf[c])
        assert np.abs(g - bf[c]).max() <= 1e-12 * max(np.abs(bf[c]).max(), 1e-300), c
    w = C.c_uint()
    call("xdem_get_error_word", C.byref(w), 1)
    assert bool(w.value & _lib.ERR_MESH_OVERFLOW) == (ovf > 0)
