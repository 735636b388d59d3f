"""Python host bindings of libelpaso_amd (MI355X FE assembly + harmonic sweep).

The compute path is the HIP library ``lib/libelpaso_amd.so`` reached through
its C-ABI (``include/elpaso_amd.h``); this package only marshals arrays.  It
raises if the library is missing: there is no CPU fallback.
"""
from .model import (BRICK8, BRICK20, FLUID8, FLUID20, ElementGroup, Material, Model)  # noqa: F401
from ._lib import (ElpError, Context, GpuModel, Server, SolveOpts, library_path, load_library,  # noqa: F401
                   EXPORTED_SYMBOLS)
