"""ctypes bindings of the C ABI (include/xdem_capi.h, include/xdem_host.h, include/xdem_types.h).

The structures mirror the C structs byte for byte; the C structs in turn are byte-compatible
with the reference's device structs (include/xdynamics_parallel/xParallelDEM_decl.cuh:128-257).
Loading fails loudly when the native library has not been built: there is no Python fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
BUILD_DIR = os.path.join(_HERE, "_build")
LIB_PATH = os.path.join(BUILD_DIR, "libxdem.so")
HOST_LIB_PATH = os.path.join(BUILD_DIR, "libxdem_host.so")

XDEM_OK = 0
STATUS_INVALID = -1
STATUS_NOT_READY = -3
STATUS = {0: "XDEM_OK", -1: "XDEM_E_INVALID", -2: "XDEM_E_HIP", -3: "XDEM_E_NOT_READY",
          -4: "XDEM_E_DEVICE_ERROR", -5: "XDEM_E_UNSUPPORTED"}

ERR_P2P_OVERFLOW = 0x1
ERR_PLANE_OVERFLOW = 0x2
ERR_MESH_OVERFLOW = 0x4
ERR_NONFINITE = 0x8
ERR_MESH_LIST = 0x10
ERR_NBR_OVERFLOW = 0x20
ERR_GRID_SYNC = 0x40
ERR_BUILD_INCOMPLETE = 0x80
ERR_WORKSPACE = 0x100
ERR_MESH_NEAR = 0x200
ERR_CYL_OVERFLOW = 0x400

PRECISION_FP64 = 0
PRECISION_MIXED = 1

FIELD = {"pos": (0, 4), "vel": (1, 3), "acc": (2, 3), "ep": (3, 4), "ev": (4, 4), "ea": (5, 4),
         "force": (6, 3), "moment": (7, 3), "mass": (8, 1), "inertia": (9, 3), "tmax": (10, 3),
         "rres": (11, 1)}
HISTORY_P2P = 0
HISTORY_PLANE = 1
HISTORY_MESH = 2
HISTORY_CYLINDER = 3
PART_HISTORY = 1
PART_POSITIONS = 2

SCENE_BOX_LATTICE = 1
SCENE_DAM_BREAK = 2
SCENE_WHEEL_BED = 3
SCENE_BATCH_MEMBER = 4
SCENE_PILE = 5


class XdemError(RuntimeError):
    """A C entry point returned a non-zero status."""

    def __init__(self, fn: str, rc: int):
        super().__init__(f"{fn} failed: {STATUS.get(rc, rc)} ({rc})")
        self.rc = rc


class double3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class uint3(C.Structure):
    _fields_ = [("x", C.c_uint), ("y", C.c_uint), ("z", C.c_uint)]


class DemParameters(C.Structure):
    """device_dem_parameters (decl:128-150)."""
    _fields_ = [("rollingCondition", C.c_ubyte), ("_pad0", C.c_ubyte * 3), ("contact_model", C.c_int),
                ("np", C.c_uint), ("nmp", C.c_uint), ("nCluster", C.c_uint), ("nClusterObject", C.c_uint),
                ("nTsdaConnection", C.c_uint), ("nTsdaConnectionList", C.c_uint),
                ("nTsdaConnectionBodyData", C.c_uint), ("nsphere", C.c_uint), ("nplane", C.c_uint),
                ("ncylinder", C.c_uint), ("ncell", C.c_uint), ("grid_size", uint3), ("dt", C.c_double),
                ("half2dt", C.c_double), ("cell_size", C.c_double), ("cohesion", C.c_double),
                ("gravity", double3), ("world_origin", double3)]


class ContactProperty(C.Structure):
    """device_contact_property (decl:230-243)."""
    _fields_ = [(n, C.c_double) for n in ("Ei", "Ej", "pri", "prj", "Gi", "Gj", "rest", "fric", "s_fric",
                                          "rfric", "coh", "sratio", "amp")]


class PlaneInfo(C.Structure):
    """device_plane_info (decl:161-174)."""
    _fields_ = [("id", C.c_uint), ("_pad0", C.c_uint), ("l1", C.c_double), ("l2", C.c_double)] + \
               [(n, double3) for n in ("u1", "u2", "uw", "xw", "pa", "pb", "w2", "w3", "w4")]


class TriangleInfo(C.Structure):
    """device_triangle_info (decl:152-159)."""
    _fields_ = [("id", C.c_int), ("tid", C.c_uint), ("P", double3), ("Q", double3), ("R", double3)]


class CylinderInfo(C.Structure):
    """xdem_cylinder_info (device_cylinder_info, decl:220-228)."""
    _fields_ = [("id", C.c_uint), ("empty_part", C.c_uint), ("thickness", C.c_double), ("len_rr", double3),
                ("pbase", double3), ("ptop", double3)]


class CubeParticleData(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("dx", "dy", "dz", "lx", "ly", "lz", "minr", "maxr")]


class PlaneParticleData(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("dx", "dy", "dz", "drx", "dry", "drz", "lx", "ly", "lz", "minr", "maxr")]


class CircleParticleData(C.Structure):
    _fields_ = [(f, C.c_double) for f in ("diameter", "sx", "sy", "sz", "dx", "dy", "dz", "minr", "maxr")]


class Quirks(C.Structure):
    _fields_ = [("gpu_force_law", C.c_int), ("velocity_not_updated", C.c_int), ("stale_ft", C.c_int),
                ("flat_inertia", C.c_int)]


class Scene(C.Structure):
    _fields_ = [("np", C.c_uint), ("pos", C.POINTER(C.c_double)), ("vel", C.POINTER(C.c_double)),
                ("mass", C.POINTER(C.c_double)), ("inertia", C.POINTER(C.c_double)), ("nplane", C.c_uint),
                ("planes", C.POINTER(PlaneInfo)), ("plane_dbi", C.POINTER(C.c_double)),
                ("plane_cp", C.POINTER(ContactProperty)), ("cp_pp", ContactProperty),
                ("contact_model", C.c_int), ("dt", C.c_double), ("nsteps", C.c_uint),
                ("gravity", C.c_double * 3), ("world_origin", C.c_double * 3), ("cell_size", C.c_double),
                ("density", C.c_double), ("min_radius", C.c_double), ("max_radius", C.c_double)]


class Mesh(C.Structure):
    """xdem_mesh: a rigid triangle-mesh body (vertices in the body frame)."""
    _fields_ = [("ntriangle", C.c_uint), ("vertices", C.POINTER(C.c_double)), ("dbi", C.c_double * 15),
                ("cp", ContactProperty)]


class HhtWheelConfig(C.Structure):
    """xdem_hht_wheel_config (include/xdem_mbd.h)."""
    _fields_ = [("mass", C.c_double), ("inertia", C.c_double * 3), ("dbi", C.c_double * 15),
                ("gravity", C.c_double * 3), ("load", C.c_double * 3), ("omega", C.c_double), ("dt", C.c_double),
                ("alpha", C.c_double), ("tolerance", C.c_double), ("max_iterations", C.c_int)]


class CtxConfig(C.Structure):
    _fields_ = [("precision", C.c_int), ("p2p_capacity", C.c_uint), ("quirks", Quirks),
                ("block_size", C.c_uint), ("grid_dim_max", C.c_uint), ("fuse_velocity", C.c_int),
                ("device", C.c_int), ("buckets_per_sphere", C.c_uint), ("skin", C.c_double),
                ("step_graph", C.c_int), ("rebuild_wait_limit", C.c_uint), ("full_blocks", C.c_uint),
                ("merge_pairs_below", C.c_uint), ("keep_order", C.c_int), ("skin_cap", C.c_double)]


class CtxInfo(C.Structure):
    _fields_ = [("nscene", C.c_uint), ("np", C.c_uint), ("nsorted", C.c_uint), ("nplane_total", C.c_uint),
                ("nbucket", C.c_uint), ("key_bits", C.c_uint), ("p2p_capacity", C.c_uint),
                ("device_bytes", C.c_ulonglong), ("steps_done", C.c_ulonglong),
                ("list_builds", C.c_ulonglong), ("list_spheres", C.c_uint), ("contact_spheres", C.c_uint),
                ("mesh_spheres", C.c_uint), ("ntriangle", C.c_uint)]


class CtxSceneInfo(C.Structure):
    _fields_ = [("np", C.c_uint), ("nplane", C.c_uint), ("p2p_capacity", C.c_uint), ("has_mesh", C.c_uint),
                ("first_particle", C.c_uint), ("ncylinder", C.c_uint)]


class AutotuneResult(C.Structure):
    _fields_ = [("full_blocks", C.c_uint), ("buckets_per_sphere", C.c_uint), ("ms_per_step", C.c_double),
                ("from_cache", C.c_int), ("ntrial", C.c_uint), ("trial_full_blocks", C.c_uint * 16),
                ("trial_buckets_per_sphere", C.c_uint * 16), ("trial_ms_per_step", C.c_double * 16),
                ("key", C.c_char * 256), ("skin_cap", C.c_double), ("trial_skin_cap", C.c_double * 16)]


assert C.sizeof(DemParameters) == 144
assert C.sizeof(ContactProperty) == 104
assert C.sizeof(PlaneInfo) == 240
assert C.sizeof(TriangleInfo) == 80
assert C.sizeof(CylinderInfo) == 88

P = C.c_void_p
U = C.c_uint
D = C.c_double
I = C.c_int

# name -> argument types (all return int)
LAUNCHER_SIGNATURES = {
    "xdem_setDEMSymbolicParameter": [C.POINTER(DemParameters)],
    "xdem_set_quirks": [C.POINTER(Quirks)],
    "xdem_set_stream": [P],
    "xdem_set_body_force_buffers": [P] * 6,
    "xdem_get_body_force_buffers": [C.POINTER(P)] * 6,
    "xdem_get_error_word": [C.POINTER(U), I],
    "xdem_workspace_size": [U, C.POINTER(C.c_ulonglong)],
    "xdem_reserve_workspace": [U],
    "xdem_set_workspace": [P, C.c_ulonglong, U],
    "xdem_vv_update_position": [P] * 6 + [U],
    "xdem_vv_update_velocity": [P] * 9 + [U],
    "xdem_cu_calculateHashAndIndex": [P, P, P, U, U],
    "xdem_cu_calculateHashAndIndexForPolygonSphere": [P, P, U, U, P],
    "xdem_cu_reorderDataAndFindCellStart": [P] * 5 + [U, U],
    "xdem_cu_reorderDataAndFindCellStart_noalloc": [P] * 5 + [U, U],
    "xdem_cu_calculate_p2p": [P] * 16 + [U],
    "xdem_cu_plane_contact_force": [P] * 15 + [U],
    "xdem_cu_particle_polygonObject_collision": [P] * 19 + [U, U],
    "xdem_cu_cylinder_contact_force": [P] * 15 + [U],
    "xdem_cu_decide_rolling_friction_moment": [P] * 6 + [U],
    "xdem_reduction": [P, U, C.POINTER(D)],
    "xdem_cu_update_meshObjectData": [P] * 5 + [U],
}

CTX_SIGNATURES = {
    "xdem_ctx_create": [C.POINTER(CtxConfig), C.POINTER(P)],
    "xdem_ctx_destroy": [P],
    "xdem_ctx_add_scene": [P, C.POINTER(Scene), C.POINTER(U)],
    "xdem_ctx_commit": [P],
    "xdem_ctx_step": [P, U],
    "xdem_ctx_collide": [P],
    "xdem_ctx_finish": [P],
    "xdem_ctx_synchronize": [P],
    "xdem_ctx_read": [P, I, U, P],
    "xdem_ctx_write": [P, I, U, P],
    "xdem_ctx_device_ptr": [P, I, C.POINTER(P)],
    "xdem_ctx_sphere_order": [P, U, C.POINTER(C.c_uint)],
    "xdem_ctx_reorder": [P],
    "xdem_ctx_read_history": [P, I, U, P, P, P],
    "xdem_ctx_write_history": [P, I, U, P, P, P],
    "xdem_ctx_scene_info_get": [P, U, C.POINTER(CtxSceneInfo)],
    "xdem_ctx_save_part": [P, U, C.c_char_p, D],
    "xdem_ctx_load_part": [P, U, C.c_char_p, I, C.POINTER(D)],
    "xdem_autotune": [C.POINTER(Scene), U, C.POINTER(CtxConfig), U, U, C.c_char_p, C.POINTER(AutotuneResult)],
    "xdem_ctx_error_word": [P, C.POINTER(U), I],
    "xdem_ctx_body_wrench": [P, U, P],
    "xdem_ctx_set_body_pose": [P, U, U, P],
    "xdem_ctx_add_mesh": [P, U, C.c_void_p, C.POINTER(U)],
    "xdem_ctx_set_mesh_pose": [P, U, P],
    "xdem_ctx_mesh_wrench": [P, U, P],
    "xdem_ctx_add_cylinder": [P, U, C.POINTER(CylinderInfo), P, C.POINTER(ContactProperty), C.POINTER(U)],
    "xdem_ctx_set_cylinder_pose": [P, U, P],
    "xdem_ctx_cylinder_wrench": [P, U, P],
    "xdem_ctx_info_get": [P, C.POINTER(CtxInfo)],
    "xdem_ctx_contact_count": [P, C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong)],
    "xdem_ctx_stream": [P, C.POINTER(P)],
    "xdem_ctx_graph_capture": [P, U],
    "xdem_ctx_graph_launch": [P, U],
    "xdem_ctx_set_timing": [P, I],
    "xdem_ctx_timing": [P, C.POINTER(D)],
}

# include/xdem_mbd.h entries that live in libxdem.so (they drive a context)
COSIM_SIGNATURES = {
    "xdem_cosim_wheel_steps": [P, P, U, U, P],
}

class CalibConfig(C.Structure):
    _fields_ = [("capacity", C.c_uint), ("checkpoint_every", C.c_uint), ("max_steps", C.c_uint), ("device", C.c_int),
                ("quirks", Quirks)]


# include/xdem_calib.h (libxdem.so)
CALIB_SIGNATURES = {
    "xdem_calib_create": [C.POINTER(Scene), C.POINTER(CalibConfig), P],
    "xdem_calib_destroy": [P],
    "xdem_calib_forward": [P, P, U, P],
    "xdem_calib_forward_from": [P, P, U, U, P],
    "xdem_calib_state": [P, P, P, P, P],
    "xdem_calib_backward": [P, P, P],
    "xdem_calib_error_word": [P, C.POINTER(U), I],
    "xdem_calib_pair_eval": [C.POINTER(ContactProperty), I, C.POINTER(Quirks), D, D, D, D, D, P, P, P],
}

HOST_SIGNATURES = {
    "xdem_scene_build": ([I, C.c_ulonglong, U, U, C.POINTER(Scene)], C.c_int),
    "xdem_scene_free": ([C.POINTER(Scene)], None),
    "xdem_make_plane": ([P, P, P, P, U, D, C.POINTER(PlaneInfo), P], C.c_int),
    "xdem_mass_inertia": ([P, U, D, P, P], None),
    "xdem_critical_time_step": ([D, D, D, D], C.c_double),
    "xdem_contact_property_make": ([D] * 9 + [C.POINTER(ContactProperty)], None),
    "xdem_fit_sphere_to_triangle": ([P, P, P, D, P], None),
    "xdem_rng_next": ([C.POINTER(C.c_ulonglong)], C.c_ulonglong),
    "xdem_rng_uniform": ([C.POINTER(C.c_ulonglong)], C.c_double),
    "xdem_wheel_mesh": ([D, D, U, U, U, D, D, U, P, C.POINTER(U)], C.c_int),
    "xdem_crt_rand": ([C.POINTER(U)], U),
    "xdem_crt_frand": ([C.POINTER(U)], D),
    "xdem_create_cube_particles": ([C.POINTER(CubeParticleData), C.POINTER(U), P, C.POINTER(U)], C.c_int),
    "xdem_create_plane_particles": ([C.POINTER(PlaneParticleData), C.POINTER(U), I, P, C.POINTER(U)], C.c_int),
    "xdem_create_circle_particles": ([C.POINTER(CircleParticleData), C.POINTER(U), P, C.POINTER(U)], C.c_int),
    "xdem_hht_wheel_create": ([C.POINTER(HhtWheelConfig), C.POINTER(P)], C.c_int),
    "xdem_hht_wheel_destroy": ([P], C.c_int),
    "xdem_hht_wheel_step": ([P, P, P, C.POINTER(I)], C.c_int),
    "xdem_hht_wheel_state": ([P, P, P, P, C.POINTER(D)], C.c_int),
}

_lib = None
_host = None


def _bind(lib, table, restype=C.c_int):
    for name, argtypes in table.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype


def host() -> C.CDLL:
    """The host-only helper library (scene builders, wall set-up); no GPU needed."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise ImportError(f"{HOST_LIB_PATH} is missing: run __graft_entry__.build() first")
        h = C.CDLL(HOST_LIB_PATH)
        for name, (argtypes, restype) in HOST_SIGNATURES.items():
            fn = getattr(h, name)
            fn.argtypes = argtypes
            fn.restype = restype
        _host = h
    return _host


def lib() -> C.CDLL:
    """The gfx950 library. Raises when it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = C.CDLL(LIB_PATH)
        _bind(L, LAUNCHER_SIGNATURES)
        _bind(L, CTX_SIGNATURES)
        _bind(L, COSIM_SIGNATURES)
        _bind(L, CALIB_SIGNATURES)
        for name, (argtypes, restype) in HOST_SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = restype
        _lib = L
    return _lib


def check(rc: int, fn: str) -> None:
    if rc != XDEM_OK:
        raise XdemError(fn, rc)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
