"""Host step loop of the FLIP/PIC solver (reference Solver/Kernel/solver.py).

`solver(config)` keeps the reference's configuration schema
(Internal/Core/export/export_config.py:121-173), time constants
(solver.py:96-106), seeding, output cadence, cancellation and messages
(solver.py:22-531).  The device work of one step is `step(ctx, params,
timings)`, which issues the stages in the reference order of
solver.py:308-510 against a `runtime.Context`:

    particle_to_grid -> grid_obstacle_bc -> grid_domain_bc -> update_velocity
    -> updated_grid_domain_bc (+ obstacle) -> pressure_solve
    -> pressure_correction -> projected_grid_domain_bc (+ obstacle)
    -> grid_to_particle -> previous positions -> advect_particles
    -> particle_domain_bc -> particle_obstacle_bc

With timing off, the last five run as one fused pass over the particles
(same kernel code, same results).

Settings added to the reference schema (all optional; defaults reproduce the
reference): `settings.gravity` (default (0, 0, -9.81)), `settings.flip_ratio`
(default 1.0, pure FLIP), and `settings.pressure_solver` with `mode`
("fixed" | "fp32" | "mixed" | "fcg" | "learned"), `tolerance`, `max_cycles`,
`pre_smooth`, `post_smooth`, `coarse_extra`, `coarse_dtype`, `vcycle_from` and
`coarse_weight` (defaults: fixed, `iterations` W(2, 4) cycles with 64 extra
coarsest sweeps, the W-cycle at every level and unweighted coarse-grid
corrections, reference pressure_solve.py:61-83).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from ... import runtime
from ..General.main import emit_message
from . import output, pressure_solve, source_particles, timing, voxelise_mesh

FIELD_DTYPE = np.float32
PARTICLES_PER_CELL = 8


@dataclass
class StepParams:
    dt: float
    rho: float = 1000.0
    nu: float = 1.0e-6
    gravity: tuple = (0.0, 0.0, -9.81)
    force: tuple = (0.0, 0.0, 0.0)
    flip_ratio: float = 1.0
    particles_per_cell: float = PARTICLES_PER_CELL
    solver: dict = field(default_factory=lambda: {"mode": "fixed", "max_cycles": 1})
    # called as observer(ctx, None) before and observer(ctx, stats) after the
    # pressure solve (the learned preconditioner's training data is dumped here)
    observer: object = None


def step(ctx, params: StepParams, timings=None) -> dict:
    """One time step on the device, in the reference order.  Returns the
    pressure-solve statistics.  Each stage is one C-ABI call enqueued on the
    context's stream (the reference's stage wrappers are cited per call)."""
    timings = timings or timing.DisabledTimings()
    with timings.section("solver", "particle_to_grid", ctx):
        ctx.bin_particles()  # particles.py:73-140: count, scan, sort, swap
        ctx.particle_to_grid(params.particles_per_cell)  # particles.py:142-171
    with timings.section("solver", "grid_obstacle_bc", ctx):
        ctx.grid_obstacle_bc()  # boundary_conditions.py:53-79
    with timings.section("solver", "grid_domain_bc", ctx):
        ctx.grid_domain_bc()  # boundary_conditions.py:36-50
    with timings.section("solver", "update_velocity", ctx):
        # velocity_update.py:8-52 and the buffer swap of solver.py:374-376
        ctx.update_velocity(params.gravity, params.force, params.rho, params.nu, params.dt)
    with timings.section("solver", "updated_grid_domain_bc", ctx):
        ctx.grid_domain_bc()
        ctx.grid_obstacle_bc()
    with timings.section("solver", "pressure_solve", ctx):
        stats = pressure_solve.pressure_solve(ctx, params.dt, params.rho, params.solver, params.observer)
    with timings.section("solver", "pressure_correction", ctx):
        pressure_solve.pressure_correction(ctx, params.dt, params.rho)
    with timings.section("solver", "projected_grid_domain_bc", ctx):
        ctx.grid_domain_bc()
        ctx.grid_obstacle_bc()
    if isinstance(timings, timing.DisabledTimings):
        ctx.grid_to_particle_advect(params.flip_ratio, params.dt)
        return stats
    with timings.section("solver", "grid_to_particle", ctx):
        ctx.grid_to_particle(params.flip_ratio)  # particles.py:174-206
    ctx.save_previous_positions()  # solver.py:469-474
    with timings.section("solver", "advect_particles", ctx):
        ctx.advect_particles(params.dt)  # particles.py:209-226
    with timings.section("solver", "particle_domain_bc", ctx):
        ctx.particle_domain_bc()  # boundary_conditions.py:82-103
    with timings.section("solver", "particle_obstacle_bc", ctx):
        ctx.particle_obstacle_bc()  # boundary_conditions.py:106-132
    return stats


def _solver_settings(settings: dict) -> dict:
    cycles = max(1, int(settings.get("iterations", 1)))  # reference solver.py:125
    chosen = {"mode": "fixed", "max_cycles": cycles, "pre_smooth": 2, "post_smooth": 4, "coarse_extra": 64,
              "tolerance": 1e-6}
    chosen.update(settings.get("pressure_solver") or {})
    return chosen


def _static_masks(ctx, simulation: dict, grid_shape, delta):
    """Source and obstacle masks (reference voxelise_mesh.py:138-163): one
    mask per source, one obstacle mask for every obstacle entry together;
    meshes are voxelised on the device."""
    origin = voxelise_mesh.grid_origin(simulation)

    def voxelise(meshes):
        mask = np.zeros(grid_shape, dtype=np.uint8)
        for triangles in meshes:
            if len(triangles):
                ctx.voxelise_mesh(triangles, origin, mask)
        return mask.astype(bool)

    bake_path = simulation.get("output_path")
    sources = []
    for entry in simulation.get("sources") or []:
        resolved = dict(entry)
        resolved["mask"] = source_particles.resolve_mask(entry, grid_shape, delta, voxelise, bake_path)
        sources.append(resolved)
    obstacle = np.zeros(grid_shape, bool)
    for entry in simulation.get("obstacles") or []:
        obstacle |= source_particles.resolve_mask(entry, grid_shape, delta, voxelise, bake_path)
    return sources, obstacle


@timing.profiled_run
def solver(config: dict, timings=None):
    simulation = config.get("simulation") or {}
    settings = simulation.get("settings") or {}
    domain = simulation.get("domain") or {}
    domain_size = domain.get("size") or {}
    physics = simulation.get("physics") or {}
    water = physics.get("water") or {}
    output_dir = Path(simulation["output_path"]).expanduser().resolve()
    cancel_event = (config.get("meta") or {}).get("cancel_event")

    lib = runtime.load_library()
    if runtime.device_count(lib) < 1:
        raise RuntimeError("No HIP device found: the Continuum Liquid step needs an AMD GPU")
    print("################################################################")

    # time (reference solver.py:96-106)
    start_frame = int(settings.get("start_frame", 1))
    end_frame = int(settings.get("end_frame", 250))
    fps = int(settings.get("fps", 24))
    steps_per_frame = max(0, int(settings.get("substeps", 1))) + 1
    total_frames = max(0, end_frame - start_frame)
    total_solver_steps = total_frames * steps_per_frame
    dt = 1.0 / (fps * steps_per_frame)

    # grid (reference solver.py:108-117)
    delta = float(domain["resolution"])
    grid_shape = tuple(int(float(domain_size[a]) / delta) for a in ("x", "y", "z"))

    params = StepParams(
        dt=dt,
        rho=float(water["density"]),
        nu=float(water["viscosity"]),
        gravity=tuple(float(g) for g in settings.get("gravity", (0.0, 0.0, -9.81))),
        flip_ratio=float(settings.get("flip_ratio", 1.0)),
        solver=_solver_settings(settings),
    )

    ctx = runtime.Context(lib, grid_shape, delta, capacity=1)
    try:
        sources, obstacle_mask = _static_masks(ctx, simulation, grid_shape, delta)
        rng = np.random.default_rng(0)
        positions = np.empty((0, 3), dtype=FIELD_DTYPE)
        velocities = np.empty((0, 3), dtype=FIELD_DTYPE)
        positions, velocities = source_particles.apply_sources(
            positions, velocities, sources, obstacle_mask, delta, rng, initial=True
        )
        count = len(positions)
        ctx.set_field("obstacle", obstacle_mask.astype(np.uint8))
        ctx.set_particles(positions, velocities)
        dynamic_sources = any(s.get("type") in {"INFLOW", "OUTFLOW"} for s in sources)
        if dynamic_sources:
            ctx.set_sources(sources)

        output_dir.mkdir(parents=True, exist_ok=True)
        cache = output.create_cache(output_dir, fps, start_frame, end_frame)
        writer = output.ParticleOutput(cache, ctx)
        try:
            writer.submit(positions, start_frame)
            emit_message({"type": "stats", "frame": 0, "total_frames": total_frames, "particle_count": count})
            for solver_step in range(total_solver_steps):
                if cancel_event is not None and cancel_event.is_set():
                    print("Bake cancellation requested. Stopping the simulation cleanly...")
                    break
                output_frame = (solver_step + 1) // steps_per_frame
                last_substep = solver_step % steps_per_frame == steps_per_frame - 1
                if dynamic_sources:
                    # the sources of solver.py:253-306 on the device: the same
                    # particles as the host seeding with the same generator,
                    # without moving every particle to the host and back
                    count = ctx.apply_sources(rng)
                if count == 0:
                    if last_substep:
                        writer.submit(np.empty((0, 3), FIELD_DTYPE), start_frame + output_frame)
                        emit_message({"type": "stats", "frame": output_frame, "total_frames": total_frames,
                                      "particle_count": 0})
                    continue
                step(ctx, params, timings)
                if last_substep:
                    with timings.section("transfer", "particles_to_host", ctx):
                        # pinned slot, copy in flight while the next steps run
                        writer.submit_device(start_frame + output_frame)
                    emit_message({"type": "stats", "frame": output_frame, "total_frames": total_frames,
                                  "particle_count": count})
        finally:
            writer.close()
    finally:
        ctx.close()
