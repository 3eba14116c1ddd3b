"""Pressure projection stages (reference Solver/Kernel/pressure_solve.py)."""

from __future__ import annotations

from ... import runtime


def pressure_solve(ctx, dt: float, rho: float, solver_settings: dict, observer=None) -> dict:
    """RHS, finest operator, the solve and the Neumann ghost update.

    Reference pressure_solve.py:10-85.  `solver_settings` selects the solve:
    mode "fixed" runs `max_cycles` W-cycles (the reference: `iterations`
    cycles of W(2, 4) with 64 extra coarsest sweeps); "fp32" iterates W-cycles
    to `tolerance` relative residual.  `coarse_weight` scales every
    coarse-grid correction of the cycle (1: the reference's plain sum).
    Returns the solve statistics.
    """
    ctx.pressure_rhs(rho, dt)
    ctx.build_operator()
    mode = runtime.SOLVE_MODES[solver_settings.get("mode", "fixed")]
    if observer is not None:
        observer(ctx, None)
    stats = ctx.pressure_solve(
        mode,
        pre_smooth=int(solver_settings.get("pre_smooth", 2)),
        post_smooth=int(solver_settings.get("post_smooth", 4)),
        coarse_extra=int(solver_settings.get("coarse_extra", 64)),
        max_cycles=int(solver_settings.get("max_cycles", 1)),
        tolerance=float(solver_settings.get("tolerance", 1e-6)),
        coarse_dtype=int(solver_settings.get("coarse_dtype", 0)),
        vcycle_from=int(solver_settings.get("vcycle_from", 0)),
        coarse_weight=float(solver_settings.get("coarse_weight", 1.0)),
    )
    if observer is not None:
        observer(ctx, stats)
    return stats


def pressure_correction(ctx, dt: float, rho: float) -> None:
    """u -= dt / rho * grad p on interior faces (reference pressure_solve.py:88-114)."""
    ctx.pressure_correction(rho, dt)
