"""Particle stages of the step (reference Solver/Kernel/particles.py).

The reference wrappers take (kernels, queue, cl.Buffer..., scalars) and return
the swapped buffers; here the device buffers live in the `Context` and each
wrapper enqueues the matching C-ABI call on its stream.
"""

from __future__ import annotations


def particle_to_grid(ctx, particles_per_cell: float = 8.0) -> None:
    """Bin, stably sort and gather particles onto the MAC grid.

    Reference particles.py:73-171: zero counts, count_particles_per_cell,
    exclusive_prefix_sum, sort_particles, buffer swap and
    particle_to_grid_gather, which also writes the fluid fraction.
    """
    ctx.bin_particles()
    ctx.particle_to_grid(particles_per_cell)


def grid_to_particle(ctx, flip_ratio: float = 1.0) -> None:
    """FLIP update v += u(x) - u_old(x), blended with PIC when flip_ratio < 1.

    Reference particles.py:174-206 (pure FLIP, i.e. flip_ratio 1).
    """
    ctx.grid_to_particle(flip_ratio)


def advect_particles(ctx, dt: float) -> None:
    """Forward-Euler advection x += dt v (reference particles.py:209-226)."""
    ctx.advect_particles(dt)
