"""Body forces and explicit viscosity (reference Solver/Kernel/velocity_update.py)."""

from __future__ import annotations


def update_velocity(ctx, gravity, force, rho: float, nu: float, dt: float) -> None:
    """u_work = u + dt (g + f / rho) on fluid cells + nu dt lap(u) on interior
    cells, then swap so that u is the updated field and u_old the input
    (reference velocity_update.py:8-52 and the swap of solver.py:374-376)."""
    ctx.update_velocity(gravity, force, rho, nu, dt)
