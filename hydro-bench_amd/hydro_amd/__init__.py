"""hydro_amd -- MI355X-native hydro-cal shallow-water timestep.

The kernels live in libhydro_amd.so (HIP for gfx950 behind the C ABI of
include/hydro.h); this package is the host side: ctypes binding, the seeded
synthetic meshes, and the reference's Python ``run()``/``step_fn`` contract
over PyTorch-ROCm device memory.
"""
from . import _lib
from ._lib import HydroError, default_tune, load
from .mesh import CONFIGS, HostMesh, make_channel_mesh, make_config, make_floodplain_mesh

__all__ = ["_lib", "HydroError", "default_tune", "load", "CONFIGS", "HostMesh", "make_channel_mesh",
           "make_config", "make_floodplain_mesh", "DeviceMesh", "run"]


def __getattr__(name):
    if name in ("DeviceMesh", "run"):
        from . import device

        return getattr(device, name)
    raise AttributeError(name)
