"""Writes the golden fixtures under tests/golden/ from the reference's particle-state data.

Run in a container where the reference tree is mounted (it does not exist on GPU hosts):
    python tests/golden/make_fixtures.py /root/reference

Inputs are the reference's own resource files, read as plain data:
  resource/sphere_drop_particles.bin  31,329 spheres, r = 0.5 mm, a settled bed (y up)
  resource/particleposition.bin        8,400 spheres, r = 1 mm, a settled block (z up)
Layout (legacy state): double ct; uint32 N; double4 pos[N] (x, y, z, r); then 6N doubles,
of which the first 3N are the velocities. Only positions, radii and velocities are kept.

  resource/input/part0104.bin          a version-2 part file of a clustered model (10,584
                                       spheres in 2,646 clusters) with sphere (12 slots) and
                                       wall (3 slots) histories: parsed with the oracle's
                                       part-file layout (oracle/xdem_part_ref.py), which must
                                       account for every byte; the checks on the parsed arrays
                                       and a few of their rows are kept (part0104_layout.npz).
"""
from __future__ import annotations

import os
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_state(path: str):
    b = open(path, "rb").read()
    ct = struct.unpack("d", b[:8])[0]
    n = struct.unpack("I", b[8:12])[0]
    pos = np.frombuffer(b[12:12 + 32 * n], dtype=np.float64).reshape(n, 4).copy()
    rest = np.frombuffer(b[12 + 32 * n:], dtype=np.float64)
    vel = rest[: 3 * n].reshape(n, 3).copy() if rest.size >= 3 * n else np.zeros((n, 3))
    return ct, pos, vel


def part_layout_fixture(ref_root: str) -> None:
    sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
    import xdem_part_ref as part

    src = os.path.join(ref_root, "resource", "input", "part0104.bin")
    buf = open(src, "rb").read()
    caps = {"p2p": 12, "plane": 3}
    a = part.parse(buf, caps)   # raises unless the layout fills the file exactly
    n, nc = a["np"], a["nc"]
    checks = {
        "radii_5mm": bool(np.all(a["pos"][:, 3] == 0.005)),
        "mass_equal_positive": bool(np.all(a["mass"] == a["mass"][0]) and a["mass"][0] > 0),
        "ep_unit": bool(np.abs(np.linalg.norm(a["ep"], axis=1) - 1).max() < 1e-9),
        "cluster_pos_radius_zero_or_positive": bool(np.all(a["cluster_pos"][:, 3] >= 0)),
        "p2p_count_le_cap": bool(a["p2p_count"].max() <= 12),
        "plane_count_le_cap": bool(a["plane_count"].max() <= 3),
        "p2p_ids_in_range": bool(all((a["p2p_id"][i, : a["p2p_count"][i]] < n).all() for i in range(n))),
        "tsd_finite": bool(np.isfinite(a["p2p_tsd"]).all() and np.isfinite(a["plane_tsd"]).all()),
    }
    assert all(checks.values()), checks
    out = os.path.join(HERE, "part0104_layout.npz")
    np.savez_compressed(out, size=np.int64(len(buf)), ct=np.float64(a["ct"]), np=np.uint32(n), nc=np.uint32(nc),
                        checks=np.array(sorted(checks)), mass=a["mass"][:8], pos=a["pos"][:8], ep=a["ep"][:8],
                        cluster_pos=a["cluster_pos"][:8], p2p_count=a["p2p_count"], plane_count=a["plane_count"],
                        p2p_id=a["p2p_id"][:64], p2p_tsd=a["p2p_tsd"][:64],
                        head=np.frombuffer(buf[:16 + 8 * 8], dtype=np.uint8).copy())
    print(f"{out}: size={len(buf)} np={n} nc={nc} ct={a['ct']} p2p contacts={int(a['p2p_count'].sum())} "
          f"wall contacts={int(a['plane_count'].sum())}")


def main(ref_root: str) -> None:
    for name in ("sphere_drop_particles", "particleposition"):
        src = os.path.join(ref_root, "resource", name + ".bin")
        ct, pos, vel = read_state(src)
        out = os.path.join(HERE, name + ".npz")
        np.savez_compressed(out, ct=np.float64(ct), pos=pos, vel=vel)
        print(f"{out}: n={pos.shape[0]} ct={ct}")
    part_layout_fixture(ref_root)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
