"""Writes the golden fixtures under tests/golden/ from the reference's particle-state data.

Run in a container where the reference tree is mounted (it does not exist on GPU hosts):
    python tests/golden/make_fixtures.py /root/reference

Inputs are the reference's own resource files, read as plain data:
  resource/sphere_drop_particles.bin  31,329 spheres, r = 0.5 mm, a settled bed (y up)
  resource/particleposition.bin        8,400 spheres, r = 1 mm, a settled block (z up)
Layout (legacy state): double ct; uint32 N; double4 pos[N] (x, y, z, r); then 6N doubles,
of which the first 3N are the velocities. Only positions, radii and velocities are kept.
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


def main(ref_root: str) -> None:
    for name in ("sphere_drop_particles", "particleposition"):
        src = os.path.join(ref_root, "resource", name + ".bin")
        ct, pos, vel = read_state(src)
        out = os.path.join(HERE, name + ".npz")
        np.savez_compressed(out, ct=np.float64(ct), pos=pos, vel=vel)
        print(f"{out}: n={pos.shape[0]} ct={ct}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
