"""The C-ABI libraries load and export every entry point the public headers declare
(no device work: these run without a GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT
import xdynamics_dll_amd as xd

HEADERS = ["xdem_capi.h", "xdem_host.h", "xdem_mbd.h", "xdem_calib.h"]


def declared_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|double|unsigned long long)\s+\**(xdem_\w+)\s*\(", text, flags=re.M)))


def test_headers_declare_entry_points():
    names = declared_functions("xdem_capi.h")
    for expected in ("xdem_cu_calculate_p2p", "xdem_cu_plane_contact_force", "xdem_ctx_step", "xdem_reduction",
                     "xdem_setDEMSymbolicParameter", "xdem_cu_reorderDataAndFindCellStart"):
        assert expected in names


@pytest.mark.parametrize("header", HEADERS)
def test_gpu_library_exports_every_declared_symbol(header):
    lib = ctypes.CDLL(xd._lib.LIB_PATH)
    missing = [n for n in declared_functions(header) if not hasattr(lib, n)]
    assert not missing, missing


def test_host_library_exports_host_header():
    lib = ctypes.CDLL(xd._lib.HOST_LIB_PATH)
    names = declared_functions("xdem_host.h") + [n for n in declared_functions("xdem_mbd.h")
                                                 if n not in xd._lib.COSIM_SIGNATURES]
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    declared = set(declared_functions("xdem_capi.h"))
    bound = set(xd._lib.LAUNCHER_SIGNATURES) | set(xd._lib.CTX_SIGNATURES)
    assert declared == bound, (declared ^ bound)
    mbd = set(declared_functions("xdem_mbd.h"))
    mbd_bound = {n for n in xd._lib.HOST_SIGNATURES if n.startswith("xdem_hht")} | set(xd._lib.COSIM_SIGNATURES)
    assert mbd == mbd_bound, (mbd ^ mbd_bound)
    calib = set(declared_functions("xdem_calib.h"))
    assert calib == set(xd._lib.CALIB_SIGNATURES), (calib ^ set(xd._lib.CALIB_SIGNATURES))


def test_struct_layouts_match_reference_sizes():
    # device_dem_parameters, device_contact_property, device_plane_info, device_triangle_info
    assert ctypes.sizeof(xd._lib.DemParameters) == 144
    assert ctypes.sizeof(xd._lib.ContactProperty) == 104
    assert ctypes.sizeof(xd._lib.PlaneInfo) == 240
    assert ctypes.sizeof(xd._lib.TriangleInfo) == 80
    assert ctypes.sizeof(xd._lib.CylinderInfo) == 88       # device_cylinder_info
    assert xd._lib.CylinderInfo.len_rr.offset == 16 and xd._lib.CylinderInfo.ptop.offset == 64
    assert xd._lib.DemParameters.dt.offset == 64
    assert xd._lib.DemParameters.grid_size.offset == 52
    assert xd._lib.PlaneInfo.l1.offset == 8
    assert xd._lib.PlaneInfo.w4.offset == 216


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(xd._lib, "_lib", None)
    monkeypatch.setattr(xd._lib, "LIB_PATH", "/nonexistent/libxdem.so")
    with pytest.raises(ImportError):
        xd._lib.lib()


SHIM_CALLER = r"""
#include "xParallelDEM_decl.hpp"
#include <cstdio>
// Takes the address of every reference-named launcher, as a caller built against
// xParallelDEM_decl.cuh would reference them; nothing is called, so no GPU is needed.
int main()
{
    void* fns[] = {(void*)&setDEMSymbolicParameter, (void*)&vv_update_position, (void*)&vv_update_velocity,
                   (void*)&cu_calculateHashAndIndex, (void*)&cu_calculateHashAndIndexForPolygonSphere,
                   (void*)&cu_reorderDataAndFindCellStart, (void*)&cu_calculate_p2p,
                   (void*)&cu_plane_contact_force, (void*)&cu_particle_polygonObject_collision,
                   (void*)&cu_decide_rolling_friction_moment, (void*)&reduction, (void*)&cu_update_meshObjectData};
    static_assert(sizeof(device_dem_parameters) == 144, "device_dem_parameters");
    static_assert(sizeof(device_contact_property) == 104, "device_contact_property");
    static_assert(sizeof(device_plane_info) == 240, "device_plane_info");
    static_assert(sizeof(device_triangle_info) == 80, "device_triangle_info");
    std::printf("%d\n", (int)(sizeof(fns) / sizeof(fns[0])));
    return xdem_shim_status();
}
"""


def test_reference_named_shim_compiles_and_links(tmp_path):
    """include/xParallelDEM_decl.hpp: the reference's launcher names, forwarded to the C ABI."""
    import shutil
    import subprocess

    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "caller.cpp"
    src.write_text(SHIM_CALLER)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(xd._lib.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-O1", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe),
           f"-L{libdir}", "-lxdem", f"-Wl,-rpath,{libdir}", "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


CONFIG_LAYOUT = r"""
#include "xdem_capi.h"
#include <stddef.h>
#include <stdio.h>
int main(void)
{
    printf("%d %d %d\n", (int)sizeof(xdem_ctx_config), (int)offsetof(xdem_ctx_config, merge_pairs_below),
           (int)offsetof(xdem_ctx_config, keep_order));
    return 0;
}
"""


def test_context_config_layout_matches_binding(tmp_path):
    """xdem_ctx_config as C lays it out and as the ctypes binding does (size and the offsets of
    its last fields)."""
    import shutil
    import subprocess

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    src = tmp_path / "layout.c"
    src.write_text(CONFIG_LAYOUT)
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c11", f"-I{os.path.join(ROOT, 'include')}", str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    size, off_merge, off_keep = map(int, subprocess.run([str(exe)], capture_output=True, text=True).stdout.split())
    assert size == ctypes.sizeof(xd._lib.CtxConfig)
    assert off_merge == xd._lib.CtxConfig.merge_pairs_below.offset
    assert off_keep == xd._lib.CtxConfig.keep_order.offset
