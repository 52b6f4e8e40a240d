"""MI355X (gfx950) Z2-gauge DQMC engine.

The compute path is libz2dqmc.so (hand-written HIP kernels behind the C ABI in
include/z2dqmc.h); this package is the Python host mirror of the reference's
sweep driver API.
"""
from ._lib import (Z2_BETA_PER_CHAIN, Z2_PREC_FP64, Z2_PREC_MIXED, Z2_QR_BLOCKED, Z2_QR_PIVOTED,  # noqa: F401
                   Z2_STAB_DSPLIT, Z2_STAB_REFERENCE, Z2Error, lib)
from ._lib import Z2_EXEC_EAGER, Z2_EXEC_GRAPH, Z2_FLIP_DELAYED, Z2_FLIP_REGISTER  # noqa: F401
from .engine import Batch, BatchChain, ChainSpec, Engine, OBS_NAMES  # noqa: F401
