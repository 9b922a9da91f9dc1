"""The reference's particle generators (xParticleManager.cpp:646-880) restated in
xdem_generators.cpp: the C library generator behind frand() against its published first
outputs, and each layout against the closed forms of its own text (counts, spacing, rings),
non-overlap, the jitter and radius ranges, and the plane generator's quirks."""
import ctypes as C
import math

import numpy as np
import pytest

from xdynamics_dll_amd import _lib, generate_circle, generate_cube, generate_plane


def test_crt_rand_known_answers():
    # the first outputs of the reference platform's rand() from its initial state (1), and
    # after srand(0)
    s = C.c_uint(1)
    H = _lib.host()
    assert [H.xdem_crt_rand(C.byref(s)) for _ in range(10)] == [41, 18467, 6334, 26500, 19169, 15724, 11478, 29358,
                                                               26962, 24464]
    s = C.c_uint(0)
    assert [H.xdem_crt_rand(C.byref(s)) for _ in range(3)] == [38, 7719, 21238]
    s = C.c_uint(1)
    assert H.xdem_crt_frand(C.byref(s)) == 41 / 32767.0


def pair_min_distance(p):
    d = np.linalg.norm(p[:, None, :3] - p[None, :, :3], axis=2)
    np.fill_diagonal(d, np.inf)
    return d.min()


@pytest.mark.parametrize("dims,r", [((0.02, 0.03, 0.025), 1e-3), ((0.05, 0.02, 0.02), 2.5e-3)])
def test_cube_lattice(dims, r):
    dx, dy, dz = dims
    lo = np.array([0.1, -0.2, 0.05])
    pos, state = generate_cube(dx, dy, dz, *lo, r, r)
    nd = [int(d / (2 * r)) - 1 for d in dims]
    assert len(pos) == nd[0] * nd[1] * nd[2]
    space = [(d - 2 * r * n) / (n + 1) for d, n in zip(dims, nd)]
    # sphere (x, y, z) of the lattice: origin l + r + space, step 2r + space, jitter in
    # [0, r space], x fastest
    idx = np.array([(x, y, z) for z in range(nd[2]) for y in range(nd[1]) for x in range(nd[0])])
    base = lo + r + np.array(space) + idx * (2 * r + np.array(space))
    jit = pos[:, :3] - base
    assert (jit >= -1e-15).all() and (jit <= r * np.array(space) + 1e-15).all()
    assert np.all(pos[:, 3] == r)
    assert pair_min_distance(pos) >= 2 * r
    # three draws per sphere, z first: the first sphere's z jitter is the first draw
    s = C.c_uint(1)
    first = _lib.host().xdem_crt_frand(C.byref(s))
    assert pos[0, 2] == pytest.approx(base[0, 2] + r * space[2] * first, abs=1e-18)
    assert state == _state_after(3 * len(pos))


def _state_after(ndraws, seed=1):
    s = C.c_uint(seed)
    for _ in range(ndraws):
        _lib.host().xdem_crt_rand(C.byref(s))
    return s.value


def test_cube_radii_drawn_after_positions():
    pos, state = generate_cube(0.02, 0.02, 0.02, 0, 0, 0, 0.6e-3, 1e-3, seed=7)
    assert (pos[:, 3] >= 0.6e-3).all() and (pos[:, 3] <= 1e-3).all() and len(np.unique(pos[:, 3])) > 10
    assert state == _state_after(4 * len(pos), seed=7)
    same, _ = generate_cube(0.02, 0.02, 0.02, 0, 0, 0, 0.6e-3, 1e-3, seed=7)
    assert np.array_equal(pos, same)


def test_plane_layers_repeat_and_its_equal_radius_quirk():
    r = 1e-3
    pos, _ = generate_plane(0.02, 2.5, 0.012, (1.0, 0.0, 0.0), (0.3, 0.1, 0.2), r, r)
    nx, nz = int(0.02 / (2 * r)) - 1, int(0.012 / (2 * r)) - 1
    assert len(pos) == 3 * nx * nz                      # dy = 2.5 counts three layers
    layer = nx * nz
    assert np.array_equal(pos[:layer], pos[layer:2 * layer])   # at the same place
    sx = (0.02 - 2 * nx * r) / (nx + 1)
    sz = (0.012 - 2 * nz * r) / (nz + 1)
    # u = x: the frame is the identity, the layer spans (y, z) from the corner
    assert np.allclose(pos[0, :3], [0.3, 0.1 + sx + r + sx, 0.2 + r + 2 * sz])
    assert np.allclose(pos[1, 1] - pos[0, 1], 2 * r + sx)
    assert pair_min_distance(pos[:layer]) >= 2 * r
    q, _ = generate_plane(0.02, 2.5, 0.012, (1.0, 0.0, 0.0), (0.3, 0.1, 0.2), r, r, reference_quirks=True)
    assert np.array_equal(q[0], pos[-1]) and not q[1:].any()


def test_circle_rings():
    r, D = 1e-3, 0.05
    c = np.array([0.1, 0.2, 0.3])
    pos, _ = generate_circle(D, c, (1.0, 0.0, 0.0), r, r)
    cr = D / 2
    nr = int(cr / (2 * r)) - 1
    space = (cr - (2 * r * nr + r)) / (nr + 1)
    counts = []
    for i in range(1, nr + 1):
        _r = i * (2 * r + space)
        counts.append(int(2 * math.pi / ((2 * r + space) / _r)))
    assert len(pos) == sum(counts)
    rel = pos[:, :3] - c
    assert np.allclose(rel[:, 0], 0.0)                     # the plane through c normal to u = x
    radii = np.linalg.norm(rel, axis=1)
    k = 0
    for i, n in enumerate(counts, start=1):
        assert np.allclose(radii[k:k + n], i * (2 * r + space))
        k += n
    assert pair_min_distance(pos) >= 2 * r * (1 - 1e-12)
    assert (pos[:, 3] == r).all()


def test_small_boxes_are_refused():
    with pytest.raises(_lib.XdemError):
        generate_cube(0.003, 0.02, 0.02, 0, 0, 0, 1e-3, 1e-3)
