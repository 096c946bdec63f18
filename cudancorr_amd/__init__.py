"""MI355X-native three-point triangle counting (threeD and rppi binning,
DDD/DDR/DRR/RRR).

The compute path is libncorr_hip.so (hand-written HIP kernels for gfx950,
C-ABI declared in include/ncorr_hip.h); this package only binds it with
ctypes (``_lib``) and provides seeded synthetic catalogues (``gen``).  There is
no CPU or PyTorch fallback: a missing library raises.
"""
from ._lib import (NC_ACCUM_EXACT, NC_ACCUM_FP32, NC_ACCUM_FP64, NC_BIN_RPPI, NC_BIN_THREED, NC_PATH_MIXED,  # noqa: F401
                   NC_PATH_PLAIN, Catalog, Context, NcError, lib)

__all__ = ["Context", "Catalog", "NcError", "NC_PATH_PLAIN", "NC_PATH_MIXED", "NC_BIN_THREED", "NC_BIN_RPPI",
           "NC_ACCUM_FP64", "NC_ACCUM_EXACT", "NC_ACCUM_FP32", "lib"]
