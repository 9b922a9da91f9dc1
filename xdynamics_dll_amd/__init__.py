"""xdynamics_dll_amd -- the DEM time step of xdynamics (miniwooki/xdynamics_dll), MI355X-native.

Native code lives in xdynamics_dll_amd/_build (libxdem.so: gfx950 kernels + C ABI;
libxdem_host.so: host set-up helpers), built by __graft_entry__.build() or
`make -C xdynamics_dll_amd/csrc`. There is no CPU fallback: the GPU entry points raise when
the library is missing.
"""
from . import _lib
from ._lib import XdemError
from .scene import (PARITY_QUIRKS, SceneData, box_walls, contact_property, generate_circle, generate_cube,
                    generate_plane, make_cylinder, make_plane, wheel_mesh)
from .simulation import DemContext, autotune, xIntegratorVV

__all__ = ["_lib", "XdemError", "PARITY_QUIRKS", "SceneData", "box_walls", "contact_property", "generate_circle",
           "generate_cube", "generate_plane", "make_cylinder", "make_plane", "wheel_mesh", "DemContext", "autotune",
           "xIntegratorVV"]
