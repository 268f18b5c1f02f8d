"""marx_amd -- MI355X-native MARX photon ray-trace.

The compute path is libmarx_gpu.so (host C + HIP kernels for gfx950) behind
the reference's libmarx entry points; this package is its Python binding.
"""
from . import _lib
from ._lib import ATTR_DTYPE, DATA_DIR, LIB_PATH, PAR_FILE, load
from .sim import ParamFile, Photons, Tracer, offaxis_directions, on_axis_photons

__all__ = ["ATTR_DTYPE", "DATA_DIR", "LIB_PATH", "PAR_FILE", "load", "ParamFile",
           "Photons", "Tracer", "offaxis_directions", "on_axis_photons"]
