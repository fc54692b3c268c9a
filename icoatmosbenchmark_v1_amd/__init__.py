"""MI355X-native NICAM tracer-transport and vertical-implicit kernels.

Drop-in for the hot path of IcoAtmosBenchmark v1 (kernels/NICAM in the
reference tree): horizontal_flux, horizontal_limiter_thuburn,
vertical_limiter_thuburn and vi_rhow_solver(_putcoef), implemented as
hand-written HIP kernels for gfx950 behind the C ABI in include/nicam_tt.h.

* :mod:`.runtime`  -- ctypes binding and the :class:`Context` mirror of the
  reference interface
* :mod:`.drivers`  -- set-up and one-call drivers (the GPU counterparts of the
  reference's snapshot-replay programs)
* :mod:`.synth`    -- seeded synthetic inputs
"""
from .runtime import (  # noqa: F401
    LIB_PATH,
    PTR_DEVICE,
    PTR_HOST,
    Context,
    LibraryNotBuilt,
    NicamError,
    grd_input_vgrid,
    load_library,
)

__version__ = "0.1.0"
