"""MI355X (gfx950) Barnes-Hut connectivity update for RELeARN.

The compute path is librelearn_bh.so (hand-written HIP kernels behind the C ABI of
include/relearn_bh.h); this package holds its ctypes binding, the build recipe and the synthetic
workload generator.
"""
from ._lib import (RL_FLAG_COUNT_WALK, RL_FLAG_EXACT_ORDER, RL_FLAG_WALK_ALWAYS, RL_FLAG_WAVE_PER_SOURCE, RL_PREC_FP32_MIXED, RL_PREC_FP64, BarnesHutHIP, BHError, StepParams,  # noqa: F401
                   load)
