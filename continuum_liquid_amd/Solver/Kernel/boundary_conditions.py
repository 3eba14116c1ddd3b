"""Boundary-condition stages (reference Solver/Kernel/boundary_conditions.py)."""

from __future__ import annotations


def neumann_domain_bc(ctx) -> None:
    """Ghost pressure <- nearest interior pressure (reference :188-199)."""
    ctx.neumann_bc()


def grid_domain_bc(ctx) -> None:
    """Zero wall-normal faces, copy tangential ghosts (reference :202-216)."""
    ctx.grid_domain_bc()


def grid_obstacle_bc(ctx) -> None:
    """Solid-cell fluid fraction 0, solid faces 0, then tangential
    extrapolation into solid cells (reference :219-245)."""
    ctx.grid_obstacle_bc()


def particle_domain_bc(ctx) -> None:
    """Reflect particles at [1.001 dx, (n - 1.001) dx] (reference :248-269)."""
    ctx.particle_domain_bc()


def particle_obstacle_bc(ctx) -> None:
    """Push particles out of solid voxels (reference :272-298)."""
    ctx.particle_obstacle_bc()
