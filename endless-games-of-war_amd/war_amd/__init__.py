"""MI355X (gfx950) engine for the War enumeration hot path.

Counts, over ranges of deals or canonical positions of single-suit War, the games that never end,
bit-exactly as the reference's gpu/war_gpu.py does, with hand-written HIP kernels behind a C ABI
(libwar_engine.so, include/war_engine.h).
"""
from ._lib import (KERNEL_PERSISTENT, KERNEL_PLAIN, MODE_CANON, MODE_FULL, MODE_HALF, WarError,  # noqa: F401
                   header_functions, load)
from .engine import MODES, Engine, make_request, total_positions  # noqa: F401

__all__ = ["Engine", "WarError", "MODES", "make_request", "total_positions", "load", "header_functions"]
