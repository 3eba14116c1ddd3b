"""Continuum Liquid FLIP/PIC time step on AMD Instinct MI355X (gfx950).

The package mirrors the reference add-on's solver interface:
`continuum_liquid_amd.Solver.General.main` keeps `main(config)` and the
message-callback channel, and `continuum_liquid_amd.Solver.Kernel` holds the
host step loop and stage wrappers.  All compute runs in the hand-written HIP
kernels of `csrc/`, loaded through the C-ABI by `runtime.load_library()`.
"""

from .runtime import ClqError, Context, load_library  # noqa: F401

__all__ = ["ClqError", "Context", "load_library"]
