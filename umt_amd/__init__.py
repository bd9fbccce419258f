"""MI355X-native UCB Sn sweeps (Teton / LLNL UMT hot path): xyz and RZ.

The compute path is libumt_amd.so (hand-written HIP kernels for gfx950 behind
the C-ABI of include/umt_sweep.h and include/umt_rz.h); this package is the
Python host side.
"""
from ._lib import LIB_PATH, UmtError, load  # noqa: F401
from .sweep import (ZONE_LAYOUT, SweepContext, TiledProblem, draw_cross_sections,  # noqa: F401
                    solve_mixed, sweep_ucbxyz)
from .learn import Learner, train_on_iterates  # noqa: F401
from .batch import SweepBatch  # noqa: F401
from .rz import RzContext, RzProblem  # noqa: F401

__all__ = ["ZONE_LAYOUT", "SweepContext", "TiledProblem", "draw_cross_sections", "sweep_ucbxyz", "load",
           "UmtError", "LIB_PATH", "Learner", "train_on_iterates", "solve_mixed", "SweepBatch",
           "RzProblem", "RzContext"]
