"""MI355X CCZ4 right-hand-side and SSP-RK3 stage engine.

The compute path is the HIP library libgranite_ccz4.so (gfx950 kernels behind the C ABI
of include/granite_ccz4.h). This package only binds it and mirrors the host interface of
GRANITE's CCZ4Evolution / GridBlock (include/granite/spacetime/ccz4.hpp:49-111,
include/granite/core/grid.hpp:44-162). There is no CPU fallback: every call that needs
the engine raises if the library or a gfx950 device is unavailable.
"""
from .ccz4 import (  # noqa: F401
    NUM_SPACETIME_VARS,
    BC_OUTFLOW,
    BC_PERIODIC,
    BC_SOMMERFELD,
    BC_NONE,
    PREC_FP64,
    PREC_MIXED,
    TUNE_GHOST,
    TUNE_STAGE,
    TUNE_TRANSFER,
    CCZ4Evolution,
    CCZ4Params,
    Checkpoint,
    Engine,
    EngineError,
    GridBlock,
    Hierarchy,
    Level,
    SpacetimeVar,
    library,
    library_path,
    levels_rk3_step,
    tune_clear,
    tune_load,
    tune_save,
)
