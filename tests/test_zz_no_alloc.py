"""The allocation-free cell sort (xdem_cu_reorderDataAndFindCellStart_noalloc) against the
workspace one, through the C ABI: two calls in a row, the second smaller, bit-identical sorted
keys, sorted index and cell tables; sizes beyond its scratch are refused."""
import ctypes as C

import numpy as np
import pytest

from xdynamics_dll_amd import _lib

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def ptr(t):
    return C.c_void_p(t.data_ptr())


def sort_with(name, keys, ncell):
    n = len(keys)
    d_hash, d_index = dev(keys), dev(np.arange(n, dtype=np.uint32))
    d_cs = dev(np.full(ncell, 7, np.uint32))
    d_ce = dev(np.full(ncell, 7, np.uint32))
    d_sorted = dev(np.zeros(n, np.uint32))
    rc = getattr(_lib.lib(), name)(ptr(d_hash), ptr(d_index), ptr(d_cs), ptr(d_ce), ptr(d_sorted), n, ncell)
    assert rc == 0, (name, rc)
    torch.cuda.synchronize()
    return [t.cpu().numpy().view(np.uint32) for t in (d_hash, d_index, d_cs, d_ce, d_sorted)]


def test_noalloc_sort_matches_the_workspace_sort():
    rng = np.random.default_rng(21)
    _lib.call("xdem_reserve_workspace", 300007)
    for n, ncell in ((300007, 1 << 21), (4099, 128 ** 3)):
        print(f"[noalloc sort] n={n} ncell={ncell}", flush=True)
        keys = np.where(rng.random(n) < 0.3, rng.integers(0, 50, n), rng.integers(0, ncell, n)).astype(np.uint32)
        a = sort_with("xdem_cu_reorderDataAndFindCellStart", keys, ncell)
        b = sort_with("xdem_cu_reorderDataAndFindCellStart_noalloc", keys, ncell)
        for x, y, what in zip(a, b, ("hash", "index", "cstart", "cend", "sorted")):
            assert np.array_equal(x, y), (n, what)
        assert np.array_equal(b[4], np.argsort(keys, kind="stable").astype(np.uint32))


def test_noalloc_sort_refuses_sizes_beyond_its_scratch():
    big = (1 << 20) + 1
    z = torch.zeros(4, dtype=torch.int32, device="cuda")
    rc = _lib.lib().xdem_cu_reorderDataAndFindCellStart_noalloc(ptr(z), ptr(z), ptr(z), ptr(z), ptr(z), big, 4)
    assert rc == _lib.STATUS_INVALID
