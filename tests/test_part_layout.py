"""The part-file layout of the oracle (oracle/xdem_part_ref.py) against the reference's own
version-2 part file resource/input/part0104.bin (kept as tests/golden/part0104_layout.npz by
tests/golden/make_fixtures.py, which parsed the whole file with this layout: it fills the
4,085,440 bytes exactly, and the parsed arrays passed the checks listed in the fixture)."""
import os

import numpy as np

import xdem_part_ref as part

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "part0104_layout.npz")
CHECKS = {"radii_5mm", "mass_equal_positive", "ep_unit", "cluster_pos_radius_zero_or_positive", "p2p_count_le_cap",
          "plane_count_le_cap", "p2p_ids_in_range", "tsd_finite"}


def test_layout_fills_the_reference_part_file_with_one_slot_split():
    f = np.load(FIX)
    n, nc, size = int(f["np"]), int(f["nc"]), int(f["size"])
    assert (n, nc) == (10584, 2646)
    assert set(f["checks"].tolist()) == CHECKS
    assert part.file_size(n, nc, {"p2p": 12, "plane": 3}) == size
    # the size fixes the sum of the slot counts (15); the split 12 + 3 is fixed by the arrays
    # parsed with it (wall counts <= 3, sphere ids < np: the fixture's checks)
    fits = [(a, b) for a in range(1, 33) for b in range(0, 9)
            if part.file_size(n, nc, {"p2p": a, **({"plane": b} if b else {})}) == size]
    assert (12, 3) in fits and all(a + b == 15 for a, b in fits)
    # the version-1 layout (no mass block) is 8 nc bytes short of it
    assert part.file_size(n, nc, {"p2p": 12, "plane": 3}, version=1) == size - 8 * nc


def test_header_and_mass_block_of_the_reference_file():
    f = np.load(FIX)
    head = f["head"].tobytes()
    assert np.frombuffer(head[:8], "<f8")[0] == f["ct"]
    assert tuple(np.frombuffer(head[8:16], "<u4")) == (int(f["np"]), int(f["nc"]))
    off = part.offsets(int(f["np"]), int(f["nc"]), {"p2p": 12, "plane": 3})
    assert off["mass"] == (16, 8 * int(f["nc"]))
    assert np.array_equal(np.frombuffer(head[16:80], "<f8"), f["mass"])
    # the spheres are 5 mm in radius and the clusters' Euler parameters are unit quaternions
    assert np.all(f["pos"][:, 3] == 0.005)
    assert np.allclose(np.linalg.norm(f["ep"], axis=1), 1.0, atol=1e-12)


def test_reference_restart_reads_the_mass_block_as_positions():
    """xDynamicSimulator.cpp:413-427 reads pos right after the counts: on a version-2 file the
    first sphere's (x, y, z, r) it gets are the first four cluster masses."""
    f = np.load(FIX)
    ro = part.reference_restart_offsets(int(f["np"]), int(f["nc"]))
    off = part.offsets(int(f["np"]), int(f["nc"]), {"p2p": 12, "plane": 3})
    assert ro["pos"] == 16 and off["pos"][0] == 16 + 8 * int(f["nc"])
    head = f["head"].tobytes()
    misread = np.frombuffer(head[ro["pos"]:ro["pos"] + 32], "<f8")
    assert np.array_equal(misread, f["mass"][:4])
    assert not np.array_equal(misread, f["pos"][0])
    for k in ("vel", "acc", "ep", "ev", "ea"):
        assert off[k][0] - ro[k] == 8 * int(f["nc"])


def test_histories_of_the_reference_file():
    f = np.load(FIX)
    c, ids, tsd = f["p2p_count"], f["p2p_id"], f["p2p_tsd"]
    # every sphere contact is stored by both spheres
    assert int(c.sum()) % 2 == 0 and c.max() <= 12 and f["plane_count"].max() <= 3
    for i in range(ids.shape[0]):
        assert (ids[i, : c[i]] < int(f["np"])).all()
        assert np.isfinite(tsd[i, : c[i]]).all()


def test_write_then_parse_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    n = 37
    state = {"mass": rng.random(n), "pos": rng.random((n, 4)), "vel": rng.random((n, 3)), "acc": rng.random((n, 3)),
             "ep": rng.random((n, 4)), "ev": rng.random((n, 4)), "ea": rng.random((n, 4))}
    hist = {"p2p": (rng.integers(0, 13, n), rng.integers(0, n, (n, 12)), rng.random((n, 12, 2))),
            "mesh": (rng.integers(0, 6, n), rng.integers(0, 99, (n, 5)), rng.random((n, 5, 2)))}
    p = tmp_path / "part.bin"
    part.write(str(p), 0.25, state, hist)
    a = part.parse(p.read_bytes(), {"p2p": 12, "mesh": 5})
    assert a["ct"] == 0.25 and a["np"] == n and a["nc"] == n
    for k, v in state.items():
        assert np.array_equal(a[k], v)
    for k, (cnt, i, t) in hist.items():
        assert np.array_equal(a[f"{k}_count"], cnt) and np.array_equal(a[f"{k}_id"], i)
        assert np.array_equal(a[f"{k}_tsd"], t)
    assert p.stat().st_size == part.file_size(n, n, {"p2p": 12, "mesh": 5})
