"""GPU parity of the launcher layer (the reference's launcher names and array layouts)
against the oracle. Device buffers are torch tensors (plumbing only)."""
import ctypes as C

import numpy as np
import pytest

import xdem_ref
from scenes import block_scene, compressed_lattice, sphere_drop_scene
from xdynamics_dll_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a: np.ndarray):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        return torch.from_numpy(a.view(np.int32)).cuda()
    return torch.from_numpy(a).cuda()


def host(t, dtype=None):
    a = t.cpu().numpy()
    if dtype == np.uint32:
        return a.view(np.uint32)
    return a


def ptr(t):
    return C.c_void_p(t.data_ptr())


def call(name, *args):
    rc = getattr(_lib.lib(), name)(*args)
    assert rc == 0, (name, rc)


def setup_params(scene, grid=(128, 128, 128)):
    prm = scene.dem_parameters(grid=grid)
    call("xdem_setDEMSymbolicParameter", C.byref(prm))
    q = _lib.Quirks(0, 0, 1, 1)
    call("xdem_set_quirks", C.byref(q))
    return prm


class Arrays:
    """Reference-layout device arrays of one scene, mirrored from a RefSim's state."""

    def __init__(self, ref):
        self.pos = dev(ref.pos); self.vel = dev(ref.vel); self.acc = dev(ref.acc)
        self.ep = dev(ref.ep); self.ev = dev(ref.ev); self.ea = dev(ref.ea)
        self.force = dev(ref.force); self.moment = dev(ref.moment)
        self.mass = dev(ref.mass); self.inertia = dev(ref.inertia)
        self.tmax = dev(ref.tmax); self.rres = dev(ref.rres)
        self.pp_count = dev(ref.pp_count); self.pp_id = dev(ref.pp_id); self.pp_tsd = dev(ref.pp_tsd)
        self.pl_count = dev(ref.pl_count); self.pl_id = dev(ref.pl_id); self.pl_tsd = dev(ref.pl_tsd)
        n = ref.np
        self.hash = dev(np.zeros(n, np.uint32)); self.index = dev(np.zeros(n, np.uint32))
        self.sorted = dev(np.zeros(n, np.uint32))
        self.cstart = dev(np.zeros(ref.ncell, np.uint32)); self.cend = dev(np.zeros(ref.ncell, np.uint32))


@pytest.mark.parametrize("make", [compressed_lattice, block_scene, sphere_drop_scene])
def test_launchers_one_step_match_oracle(make):
    scene = make()
    print(f"[launchers] np={scene.np}", flush=True)
    ref = xdem_ref.RefSim(scene)
    d = Arrays(ref)
    n = scene.np
    setup_params(scene)
    cp = dev(np.frombuffer(bytes(scene.cp_pp), dtype=np.float64).copy())
    # --- position update (impl:135-147)
    call("xdem_vv_update_position", ptr(d.pos), ptr(d.ep), ptr(d.vel), ptr(d.ev), ptr(d.acc), ptr(d.ea), n)
    ref.integrate_position()
    torch.cuda.synchronize()
    assert np.array_equal(host(d.pos), ref.pos)
    assert np.array_equal(host(d.ep), ref.ep)
    # --- hash, sort, cell bounds: bit-exact
    call("xdem_cu_calculateHashAndIndex", ptr(d.hash), ptr(d.index), ptr(d.pos), 0, n)
    call("xdem_cu_reorderDataAndFindCellStart", ptr(d.hash), ptr(d.index), ptr(d.cstart), ptr(d.cend), ptr(d.sorted),
         n, ref.ncell)
    ref.grid()
    torch.cuda.synchronize()
    assert np.array_equal(host(d.hash, np.uint32), ref.hash)
    assert np.array_equal(host(d.sorted, np.uint32), ref.sorted_index)
    assert np.array_equal(host(d.cstart, np.uint32), ref.cstart)
    assert np.array_equal(host(d.cend, np.uint32), ref.cend)
    # --- sphere contacts (impl:613-733)
    call("xdem_cu_calculate_p2p", ptr(d.pos), ptr(d.ep), ptr(d.vel), ptr(d.ev), ptr(d.force), ptr(d.moment),
         ptr(d.mass), ptr(d.tmax), ptr(d.rres), ptr(d.pp_count), ptr(d.pp_id), ptr(d.pp_tsd), ptr(d.sorted),
         ptr(d.cstart), ptr(d.cend), ptr(cp), n)
    # --- walls, one launch each (impl:950-1080), body-force buffers registered
    bf = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(6)]
    call("xdem_set_body_force_buffers", *(ptr(b) for b in bf))
    wrench = []
    for k, pl in enumerate(scene.planes):
        pinfo = dev(np.frombuffer(bytes(pl), dtype=np.uint8).copy())
        dbi = dev(np.asarray(scene.plane_dbi[k], dtype=np.float64))
        wcp = dev(np.frombuffer(bytes(scene.plane_cp[k]), dtype=np.float64).copy())
        call("xdem_cu_plane_contact_force", ptr(pinfo), ptr(dbi), ptr(wcp), ptr(d.pos), ptr(d.ep), ptr(d.vel),
             ptr(d.ev), ptr(d.force), ptr(d.moment), ptr(d.mass), ptr(d.tmax), ptr(d.rres), ptr(d.pl_count),
             ptr(d.pl_id), ptr(d.pl_tsd), n)
        out = C.c_double()
        call("xdem_reduction", ptr(bf[2]), n, C.byref(out))
        wrench.append(out.value)
        torch.cuda.synchronize()
    ref.contacts()
    torch.cuda.synchronize()
    for name in ("force", "moment", "tmax", "rres"):
        g, r = host(getattr(d, name)), getattr(ref, name)
        assert np.abs(g - r).max() <= 1e-12 * max(np.abs(r).max(), 1e-300), name
    pc = host(d.pp_count, np.uint32)
    assert np.array_equal(pc, ref.pp_count)
    ids = host(d.pp_id, np.uint32)
    for i in range(n):
        assert np.array_equal(np.sort(ids[i, : pc[i]]), np.sort(ref.pp_id[i, : ref.pp_count[i]])), i
    assert np.array_equal(host(d.pl_count, np.uint32), ref.pl_count)
    for k in range(len(scene.planes)):
        assert abs(wrench[k] - ref.wrench[k, 2]) <= 1e-12 * max(abs(ref.wrench[k, 2]), 1e-300)
    # --- rolling resistance + velocity update
    call("xdem_cu_decide_rolling_friction_moment", ptr(d.tmax), ptr(d.rres), ptr(d.inertia), ptr(d.ep), ptr(d.ev),
         ptr(d.moment), n)
    call("xdem_vv_update_velocity", ptr(d.vel), ptr(d.acc), ptr(d.ep), ptr(d.ev), ptr(d.ea), ptr(d.force),
         ptr(d.moment), ptr(d.mass), ptr(d.inertia), n)
    ref.finish()
    torch.cuda.synchronize()
    for name in ("vel", "acc", "ev", "ea"):
        g, r = host(getattr(d, name)), getattr(ref, name)
        assert np.abs(g - r).max() <= 1e-12 * max(np.abs(r).max(), 1e-300), name
    w = C.c_uint()
    call("xdem_get_error_word", C.byref(w), 1)
    assert w.value == 0


@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 100003])
@pytest.mark.parametrize("ncell", [128 ** 3, 1 << 24])
def test_reorder_sort_is_stable_and_exact(n, ncell):
    rng = np.random.default_rng(n + ncell)
    # clustered keys (many equal) and spread keys
    keys = np.where(rng.random(n) < 0.5, rng.integers(0, 64, n), rng.integers(0, ncell, n)).astype(np.uint32)
    index = np.arange(n, dtype=np.uint32)
    d_hash, d_index = dev(keys), dev(index)
    d_cs, d_ce = dev(np.zeros(ncell, np.uint32)), dev(np.zeros(ncell, np.uint32))
    d_sorted = dev(np.zeros(n, np.uint32))
    call("xdem_reserve_workspace", n)   # set-up: the launchers themselves never allocate
    call("xdem_cu_reorderDataAndFindCellStart", ptr(d_hash), ptr(d_index), ptr(d_cs), ptr(d_ce), ptr(d_sorted), n,
         ncell)
    torch.cuda.synchronize()
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(host(d_hash, np.uint32), keys[order])
    assert np.array_equal(host(d_sorted, np.uint32), index[order])
    cs = np.full(ncell, 0xFFFFFFFF, np.uint32)
    ce = np.zeros(ncell, np.uint32)
    sk = keys[order]
    first = np.r_[True, sk[1:] != sk[:-1]]
    last = np.r_[sk[1:] != sk[:-1], True]
    cs[sk[first]] = np.nonzero(first)[0]
    ce[sk[last]] = np.nonzero(last)[0] + 1
    assert np.array_equal(host(d_cs, np.uint32), cs)
    assert np.array_equal(host(d_ce, np.uint32), ce)


def test_reduction_fp64():
    rng = np.random.default_rng(5)
    for n in (1, 255, 256, 1000, 262145):
        a = rng.normal(size=n)
        out = C.c_double()
        t = dev(a)
        call("xdem_reserve_workspace", 1)
        call("xdem_reduction", ptr(t), n, C.byref(out))
        exact = float(np.sum(a.astype(np.longdouble)))
        assert abs(out.value - exact) <= 1e-13 * np.abs(a).sum()


def test_caller_owned_workspace_and_the_workspace_bound():
    """The sort runs in a caller-owned workspace of exactly xdem_workspace_size(n) bytes and
    gives the same result as in the library's; a launch beyond the workspace is refused with
    XDEM_E_INVALID and XDEM_ERR_WORKSPACE, and leaves its arrays untouched."""
    n, ncell = 5000, 1 << 12
    rng = np.random.default_rng(11)
    keys = rng.integers(0, ncell, n).astype(np.uint32)
    order = np.argsort(keys, kind="stable")
    size = C.c_ulonglong()
    call("xdem_workspace_size", n, C.byref(size))
    ws = torch.empty(size.value + 256, dtype=torch.uint8, device="cuda")
    base = (ws.data_ptr() + 255) // 256 * 256
    L = _lib.lib()
    assert L.xdem_set_workspace(C.c_void_p(base), size.value - 1, n) == _lib.STATUS_INVALID
    call("xdem_set_workspace", C.c_void_p(base), size.value, n)
    try:
        d_hash, d_index = dev(keys), dev(np.arange(n, dtype=np.uint32))
        d_cs, d_ce, d_sorted = dev(np.zeros(ncell, np.uint32)), dev(np.zeros(ncell, np.uint32)), dev(np.zeros(n, np.uint32))
        call("xdem_cu_reorderDataAndFindCellStart", ptr(d_hash), ptr(d_index), ptr(d_cs), ptr(d_ce), ptr(d_sorted),
             n, ncell)
        torch.cuda.synchronize()
        assert np.array_equal(host(d_sorted, np.uint32), order.astype(np.uint32))
        # one element more than the workspace holds: refused, nothing written
        big = dev(np.zeros(n + 1, np.uint32))
        rc = L.xdem_cu_reorderDataAndFindCellStart(ptr(big), ptr(big), ptr(d_cs), ptr(d_ce), ptr(big), n + 1, ncell)
        assert rc == _lib.STATUS_INVALID
        w = C.c_uint()
        call("xdem_get_error_word", C.byref(w), 1)
        assert w.value == _lib.ERR_WORKSPACE
        assert np.array_equal(host(d_sorted, np.uint32), order.astype(np.uint32))
    finally:
        call("xdem_set_workspace", None, 0, 0)
    del ws
