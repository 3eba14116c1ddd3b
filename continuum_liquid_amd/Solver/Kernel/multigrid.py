"""Multigrid hierarchy and cycle (reference Solver/Kernel/multigrid.py).

The hierarchy is allocated by the library when the context is created
(`Context.level_shapes`); the W-cycle recursion runs natively in the library
so a cycle costs one call from Python.
"""

from __future__ import annotations


def coarse_grid_size(fine_size: int) -> int:
    """Cells of the next coarser level, keeping one ghost layer per side
    (reference multigrid.py:12-15)."""
    interior = max(0, int(fine_size) - 2)
    return (interior + 1) // 2 + 2


def level_shapes(shape, min_size: int = 8) -> list[tuple[int, int, int]]:
    """Shapes of the coarse levels below `shape` (reference multigrid.py:18-50)."""
    dims = [coarse_grid_size(n) for n in shape]
    shapes = []
    while all(n >= min_size for n in dims):
        shapes.append(tuple(dims))
        dims = [coarse_grid_size(n) for n in dims]
    return shapes


def w_cycle(ctx, pre_smooth: int = 2, post_smooth: int = 4, coarse_smooth_iterations: int = 64) -> None:
    """One W-cycle on the level-0 system (reference multigrid.py:53-187)."""
    ctx.w_cycle(pre_smooth, post_smooth, coarse_smooth_iterations)


def smooth(ctx, level: int, iterations: int) -> None:
    """Red-black Gauss-Seidel sweeps on one level (reference multigrid.py:190-220)."""
    ctx.smooth(level, iterations)


def coarse_smooth(ctx, iterations: int) -> None:
    """Sweeps on the coarsest level in one work-group (reference multigrid.py:223-249)."""
    ctx.coarse_smooth(iterations)
