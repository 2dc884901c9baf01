"""MI355X (gfx950) re-implementation of the viss_optimize_ws visibility /
reconstruction hot path behind a C-ABI (include/vissws.h).

Python is a thin ctypes layer over libvissws.so; every stage runs in the
hand-written HIP kernels of csrc/.
"""
from . import _abi as abi  # noqa: F401
from .pipeline import Pipeline, Stages, VissError, default_config, lib_path, load_library, make_sky  # noqa: F401

__all__ = ["Pipeline", "Stages", "VissError", "default_config", "lib_path", "load_library", "make_sky", "abi"]
