"""Part files of the step layer on the GPU (xdem_ctx_save_part / xdem_ctx_load_part).

- the file holds the layout of the reference's step-result file (xResultManager.cpp:828-925):
  the oracle's writer (oracle/xdem_part_ref.py) given the same arrays writes the same bytes;
- a run continued from a part file with its histories equals the run that wrote it, bit for
  bit (sphere walls and a triangle mesh; the new context's lists and skin start afresh);
- without XDEM_PART_HISTORY the histories are kept as they were (the reference's restart);
- files of another scene and histories past their slots are refused and change nothing.
"""
import numpy as np
import pytest

import xdem_part_ref as part
from mesh_scenes import mesh_scene
from scenes import compressed_lattice
from xdynamics_dll_amd import DemContext, XdemError

pytestmark = pytest.mark.gpu

STATE = ("pos", "vel", "acc", "ep", "ev", "ea")


def make(name, fuse=True):
    if name == "lattice":
        scene = compressed_lattice()
        return scene, {"caps": {"p2p": 12, "plane": 3}}, lambda: DemContext(scene, fuse_velocity=fuse)
    scene, tris, _, dbi, cp = mesh_scene(n=200, max_overlap=0.005, nx=20, separated=True, dt=1e-6)
    return scene, {"caps": {"p2p": 12, "mesh": 5}}, lambda: DemContext(scene, fuse_velocity=fuse,
                                                                        meshes=[(0, tris, dbi, cp)])


def histories(ctx, caps):
    return {k: ctx.history(k) for k in caps}


def assert_histories_equal(a, b):
    for k in a:
        (ca, ia, ta), (cb, ib, tb) = a[k], b[k]
        assert np.array_equal(ca, cb), k
        for i in np.nonzero(ca)[0]:
            assert np.array_equal(ia[i, : ca[i]], ib[i, : cb[i]]), (k, i)
            assert np.array_equal(ta[i, : ca[i]], tb[i, : cb[i]]), (k, i)


@pytest.mark.parametrize("name,fuse", [("lattice", True), ("lattice", False), ("terrain", True)])
def test_part_file_layout_and_restart(name, fuse, tmp_path):
    scene, meta, new = make(name, fuse)
    caps = meta["caps"]
    print(f"[part] {name} np={scene.np} fuse={fuse}", flush=True)
    a = new()
    a.step(40)
    path = str(tmp_path / "part0001.bin")
    ct = 40 * scene.dt
    a.save_part(path, ct)
    buf = open(path, "rb").read()
    # layout: the oracle's writer given the arrays read back writes the same bytes
    state = {f: a.read(f) for f in ("mass",) + STATE}
    hist = histories(a, caps)
    ref_path = str(tmp_path / "oracle.bin")
    part.write(ref_path, ct, state, hist)
    assert buf == open(ref_path, "rb").read()
    p = part.parse(buf, caps)
    assert p["ct"] == ct and p["np"] == scene.np and p["nc"] == scene.np
    assert np.array_equal(p["pos"], state["pos"])
    # restart with the histories: continues bit-identically
    a.step(40)
    b = new()
    assert b.load_part(path, history=True) == ct
    b.step(40)
    for f in STATE:
        assert np.array_equal(a.read(f), b.read(f)), f
    assert_histories_equal(histories(a, caps), histories(b, caps))
    assert a.error_word() == 0 and b.error_word() == 0
    a.close()
    b.close()


def test_restart_without_histories_keeps_them(tmp_path):
    scene, meta, new = make("lattice")
    caps = meta["caps"]
    a = new()
    a.step(10)
    path = str(tmp_path / "part.bin")
    a.save_part(path, 1.0)
    b = new()
    b.step(3)
    before = histories(b, caps)
    assert before["p2p"][0].sum() > 0
    assert b.load_part(path) == 1.0
    for f in STATE:
        assert np.array_equal(a.read(f), b.read(f)), f
    assert_histories_equal(before, histories(b, caps))
    a.close()
    b.close()


def test_positions_only_restart(tmp_path):
    """SetCurrentParticlesFromPartResult: the positions from the file, everything else kept."""
    scene, meta, new = make("lattice")
    a = new()
    a.step(20)
    path = str(tmp_path / "part.bin")
    a.save_part(path, 0.5)
    b = new()
    b.step(3)
    before = {f: b.read(f) for f in STATE}
    assert b.load_part(path, positions_only=True) == 0.5
    assert np.array_equal(b.read("pos"), a.read("pos"))
    for f in STATE[1:]:
        assert np.array_equal(b.read(f), before[f]), f
    with pytest.raises(XdemError):
        b.load_part(path, history=True, positions_only=True)
    a.close()
    b.close()


def test_foreign_file_and_overfull_history_are_refused(tmp_path):
    scene, meta, new = make("lattice")
    a = new()
    a.step(5)
    path = str(tmp_path / "part.bin")
    a.save_part(path, 0.0)
    small = compressed_lattice(n_side=4)
    b = DemContext(small)
    pos = b.read("pos")
    with pytest.raises(XdemError):
        b.load_part(path)                       # another scene's sphere count
    trunc = str(tmp_path / "trunc.bin")
    open(trunc, "wb").write(open(path, "rb").read()[:-8])
    with pytest.raises(XdemError):
        a.load_part(trunc)
    assert np.array_equal(b.read("pos"), pos)
    c, ids, tsd = a.history("p2p")
    before = a.history("p2p")
    c2 = c.copy()
    c2[0] = 13                                  # past the 12 slots
    with pytest.raises(XdemError):
        a.write_history("p2p", c2, ids, tsd)
    ids2 = ids.copy()
    j = int(np.nonzero(c)[0][0])
    ids2[j, 0] = scene.np                       # a sphere outside the scene
    with pytest.raises(XdemError):
        a.write_history("p2p", c, ids2, tsd)
    assert_histories_equal({"p2p": before}, {"p2p": a.history("p2p")})
    a.close()
    b.close()
