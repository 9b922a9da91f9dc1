"""ORACLE (test infrastructure only): the reference's part-file layout restated in numpy.

Only tests/ use this module, as the checker of the library's part files; the product never
imports it.

Layout of xResultManager::export_step_data_to_file (source/xdynamics_manager/xResultManager.cpp
:828-925, VERSION_NUMBER 2 of include/xdynamics_decl.h:18), native little-endian, no padding:

    double ct
    uint32 nparticles, uint32 nclusters                     (when nparticles > 0)
    double mass[nclusters]                                  (version > 1)
    double pos[4 np], vel[3 nc], acc[3 nc], ep[4 nc], ev[4 nc], ea[4 nc]
    double cluster_pos[4 nc]                                (np != nc)
    multibody block                                         (models with generalised coordinates)
    for each history kind present (sphere, wall, cylinder, triangle), with its slot count cap:
        uint32 count[np], uint32 id[np cap], double tsd[2 np cap]

The restart reader (xDynamicSimulator.cpp:392-492) reads ct, np, nc and then pos directly,
without the version-2 mass block; `reference_restart_offsets` gives the offsets it reads at.
"""
from __future__ import annotations

import numpy as np

HISTORY_KINDS = ("p2p", "plane", "cylinder", "mesh")   # xResultManager.cpp:889-912, in file order


def sections(np_: int, nc: int, caps: dict, version: int = 2, mbd_bytes: int = 0):
    """[(name, dtype, shape)] of a part file in file order; caps maps a history kind of
    HISTORY_KINDS to its slot count (kinds absent from caps are not in the file)."""
    out = [("ct", "<f8", ()), ("np", "<u4", ()), ("nc", "<u4", ())]
    if version > 1:
        out.append(("mass", "<f8", (nc,)))
    out += [("pos", "<f8", (np_, 4)), ("vel", "<f8", (nc, 3)), ("acc", "<f8", (nc, 3)),
            ("ep", "<f8", (nc, 4)), ("ev", "<f8", (nc, 4)), ("ea", "<f8", (nc, 4))]
    if np_ != nc:
        out.append(("cluster_pos", "<f8", (nc, 4)))
    if mbd_bytes:
        out.append(("mbd", "u1", (mbd_bytes,)))
    for k in HISTORY_KINDS:
        if k in caps:
            cap = caps[k]
            out += [(f"{k}_count", "<u4", (np_,)), (f"{k}_id", "<u4", (np_, cap)), (f"{k}_tsd", "<f8", (np_, cap, 2))]
    return out


def _nbytes(dt, shape):
    return int(np.dtype(dt).itemsize * int(np.prod(shape, dtype=np.int64)))


def file_size(np_: int, nc: int, caps: dict, version: int = 2, mbd_bytes: int = 0) -> int:
    return sum(_nbytes(dt, sh) for _, dt, sh in sections(np_, nc, caps, version, mbd_bytes))


def offsets(np_: int, nc: int, caps: dict, version: int = 2, mbd_bytes: int = 0) -> dict:
    """name -> (byte offset, byte length)."""
    out, o = {}, 0
    for name, dt, sh in sections(np_, nc, caps, version, mbd_bytes):
        n = _nbytes(dt, sh)
        out[name] = (o, n)
        o += n
    return out


def parse(buf: bytes, caps: dict, version: int = 2, mbd_bytes: int = 0) -> dict:
    """Arrays of a part file; raises ValueError when the bytes do not fill the layout exactly."""
    head = np.frombuffer(buf[:16], dtype="<u4")
    np_, nc = int(head[2]), int(head[3])
    secs = sections(np_, nc, caps, version, mbd_bytes)
    total = sum(_nbytes(dt, sh) for _, dt, sh in secs)
    if total != len(buf):
        raise ValueError(f"layout of {total} bytes for a file of {len(buf)}")
    out, o = {}, 0
    for name, dt, sh in secs:
        n = _nbytes(dt, sh)
        a = np.frombuffer(buf[o:o + n], dtype=dt).reshape(sh)
        out[name] = a.item() if sh == () else a.copy()
        o += n
    return out


def write(path: str, ct: float, state: dict, histories: dict) -> None:
    """A single-sphere (np = nc) version-2 part file from a state (mass, pos, vel, acc, ep, ev, ea)
    and histories {kind: (count, id, tsd)} in the order of HISTORY_KINDS."""
    n = int(state["pos"].shape[0])
    with open(path, "wb") as f:
        f.write(np.float64(ct).tobytes())
        f.write(np.array([n, n], dtype="<u4").tobytes())
        for k in ("mass", "pos", "vel", "acc", "ep", "ev", "ea"):
            f.write(np.ascontiguousarray(state[k], dtype="<f8").tobytes())
        for k in HISTORY_KINDS:
            if k in histories:
                c, i, t = histories[k]
                f.write(np.ascontiguousarray(c, dtype="<u4").tobytes())
                f.write(np.ascontiguousarray(i, dtype="<u4").tobytes())
                f.write(np.ascontiguousarray(t, dtype="<f8").tobytes())


def reference_restart_offsets(np_: int, nc: int) -> dict:
    """Byte offsets at which xDynamicsSimulator::set_from_part_result (xDynamicSimulator.cpp:
    413-427) reads the state arrays: it goes from the counts straight to pos, so on a version-2
    file every array starts 8 nc bytes early (pos takes the mass block and the first rows)."""
    o = 16
    out = {}
    for name, w, n in (("pos", 4, np_), ("vel", 3, nc), ("acc", 3, nc), ("ep", 4, nc), ("ev", 4, nc), ("ea", 4, nc)):
        out[name] = o
        o += 8 * w * n
    return out
