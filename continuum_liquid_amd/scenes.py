"""Seeded synthetic scenes for the benchmark and the parity tests.

A scene is a grid (nx, ny, nz) of spacing dx, with one ghost layer per face,
and particles seeded in its interior cells: 8 per cell, one per 2x2x2
sub-cell, each uniformly jittered inside its sub-cell, FP32 positions.  A cell
is fluid when its centre lies inside a region.  Regions are given in
normalised coordinates (fractions of the domain extent along each axis).

The reference seeds uniformly in the whole cell with `default_rng(0)`
(Solver/Kernel/source_particles.py:42-75); the benchmark configurations ask
for sub-cell stratified jitter, which is what `stratified_fill` does.  The up
axis of the benchmark scenes is y (gravity -9.81 on y); the reference's own
configs put gravity on z (Solver/Kernel/solver.py:126).
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Sphere:
    centre: tuple[float, float, float]
    radius: float
    velocity: tuple[float, float, float] = (0.0, 0.0, 0.0)


@dataclass
class Scene:
    name: str
    shape: tuple[int, int, int]
    dx: float
    positions: np.ndarray
    velocities: np.ndarray
    gravity: tuple[float, float, float] = (0.0, -9.81, 0.0)
    fluid_cells: int = 0
    notes: dict = field(default_factory=dict)

    @property
    def particle_count(self) -> int:
        return len(self.positions)

    def digest(self) -> str:
        """SHA-256 of positions and velocities (NumPy does not promise the same
        Generator stream across versions, so inputs are pinned by hash)."""
        h = hashlib.sha256()
        h.update(np.ascontiguousarray(self.positions, np.float32).tobytes())
        h.update(np.ascontiguousarray(self.velocities, np.float32).tobytes())
        return h.hexdigest()


def cell_centres(shape, dx):
    """Normalised centre coordinates of every cell, as three broadcastable axes."""
    axes = []
    for a, n in enumerate(shape):
        extent = n * dx
        c = (np.arange(n, dtype=np.float64) + 0.5) * dx / extent
        view = [1, 1, 1]
        view[a] = n
        axes.append(c.reshape(view))
    return axes


def interior_mask(shape) -> np.ndarray:
    m = np.zeros(shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


def stratified_fill(mask: np.ndarray, dx: float, rng: np.random.Generator) -> np.ndarray:
    """8 FP32 particles per masked cell, one per 2x2x2 sub-cell, cell order."""
    cells = np.argwhere(mask).astype(np.float64)
    sub = np.array([[a, b, c] for a in (0, 1) for b in (0, 1) for c in (0, 1)], np.float64)
    jitter = rng.random((len(cells), 8, 3))
    pos = (cells[:, None, :] + (sub[None, :, :] + jitter) * 0.5) * dx
    return np.ascontiguousarray(pos.reshape(-1, 3), dtype=np.float32)


def dam_break(
    name: str,
    shape,
    dx: float,
    dam: tuple[float, float],
    spheres=(),
    seed: int = 0,
    up_axis: int = 1,
) -> Scene:
    """Dam (x < dam[0], up < dam[1], full depth) plus spheres with velocities."""
    shape = tuple(int(n) for n in shape)
    X = cell_centres(shape, dx)
    inside = interior_mask(shape)
    dam_mask = inside & (X[0] < dam[0]) & (X[up_axis] < dam[1])
    sphere_masks = []
    for s in spheres:
        extent = np.array(shape, np.float64) * dx
        d2 = sum(((X[a] * extent[a]) - s.centre[a]) ** 2 for a in range(3))
        sphere_masks.append(inside & (d2 < s.radius**2) & ~dam_mask)
    fluid = dam_mask.copy()
    for m in sphere_masks:
        fluid |= m
    rng = np.random.default_rng(seed)
    positions = stratified_fill(fluid, dx, rng)
    velocities = np.zeros_like(positions)
    # particles are generated in cell order; map each cell to its sphere velocity
    cell_vel = np.zeros(shape + (3,), np.float32)
    for s, m in zip(spheres, sphere_masks):
        cell_vel[m] = np.asarray(s.velocity, np.float32)
    velocities[:] = np.repeat(cell_vel[fluid], 8, axis=0)
    gravity = [0.0, 0.0, 0.0]
    gravity[up_axis] = -9.81
    return Scene(
        name=name,
        shape=shape,
        dx=dx,
        positions=positions,
        velocities=velocities,
        gravity=tuple(gravity),
        fluid_cells=int(fluid.sum()),
        notes={"seed": seed, "dam": dam, "spheres": [s.__dict__ for s in spheres]},
    )


def bench_scene(config: int, seed: int = 0) -> Scene:
    """The benchmark configurations 1, 2, 3 and 5 (4 is the learned solver,
    evaluated on the scene of configuration 2)."""
    if config == 1:
        return dam_break("dam-break 64^3", (64, 64, 64), 1.0 / 64, (0.4, 0.5), seed=seed)
    if config in (2, 4):
        sphere = Sphere((0.7, 0.8, 0.5), 0.12, (0.0, -2.0, 0.0))
        return dam_break("dam-break + sphere 128^3", (128, 128, 128), 1.0 / 128, (0.35, 0.6), [sphere], seed)
    if config == 3:
        sphere = Sphere((0.7, 0.8, 0.5), 0.12, (0.0, -2.0, 0.0))
        return dam_break("dam-break + sphere 256^3", (256, 256, 256), 1.0 / 256, (0.35, 0.6), [sphere], seed)
    if config == 5:
        rng = np.random.default_rng(seed + 1)
        spheres = []
        for _ in range(3):
            r = float(rng.uniform(0.08, 0.15))
            centre = (float(rng.uniform(0.8, 1.8)), float(rng.uniform(0.75, 0.85)), float(rng.uniform(0.25, 0.75)))
            spheres.append(Sphere(centre, r, (0.0, -1.0, 0.0)))
        return dam_break("wide dam-break + 3 spheres 512x256x256", (512, 256, 256), 1.0 / 256, (0.3, 0.7), spheres,
                         seed)
    raise ValueError(f"unknown configuration {config}")


def small_scene(n: int = 24, seed: int = 0, spheres: bool = True) -> Scene:
    """A small dam break (+ falling sphere) for parity tests."""
    sph = [Sphere((0.65, 0.7, 0.5), 0.18, (0.3, -2.0, 0.1))] if spheres else []
    return dam_break(f"dam-break {n}^3", (n, n, n), 1.0 / n, (0.4, 0.5), sph, seed)


def random_scene(shape, seed: int, dx: float | None = None) -> Scene:
    """Training scene of the learned preconditioner: 1-3 fluid boxes and 0-2
    spheres of random sizes and positions, each body moving with its own
    velocity drawn uniformly from [-3, 3] m/s per axis; gravity -9.81 on y."""
    shape = tuple(int(n) for n in shape)
    dx = dx if dx is not None else 1.0 / max(shape)
    rng = np.random.default_rng(seed)
    X = cell_centres(shape, dx)
    inside = interior_mask(shape)
    fluid = np.zeros(shape, bool)
    cell_vel = np.zeros(shape + (3,), np.float32)
    bodies = []
    for _ in range(int(rng.integers(1, 4))):
        lo = rng.uniform(0.0, 0.7, 3)
        hi = np.minimum(lo + rng.uniform(0.15, 0.6, 3), 1.0)
        lo[1] = 0.0 if rng.random() < 0.5 else lo[1]  # half of the boxes rest on the floor
        m = inside & ~fluid
        for a in range(3):
            m &= (X[a] >= lo[a]) & (X[a] < hi[a])
        vel = rng.uniform(-3.0, 3.0, 3).astype(np.float32)
        fluid |= m
        cell_vel[m] = vel
        bodies.append({"box": (lo.tolist(), hi.tolist()), "velocity": vel.tolist()})
    extent = np.array(shape, np.float64) * dx
    for _ in range(int(rng.integers(0, 3))):
        r = float(rng.uniform(0.08, 0.2)) * float(extent.min())
        c = rng.uniform(0.2, 0.8, 3) * extent
        d2 = sum((X[a] * extent[a] - c[a]) ** 2 for a in range(3))
        m = inside & ~fluid & (d2 < r * r)
        vel = rng.uniform(-3.0, 3.0, 3).astype(np.float32)
        fluid |= m
        cell_vel[m] = vel
        bodies.append({"sphere": (c.tolist(), r), "velocity": vel.tolist()})
    positions = stratified_fill(fluid, dx, rng)
    velocities = np.ascontiguousarray(np.repeat(cell_vel[fluid], 8, axis=0))
    return Scene(name=f"random scene {seed}", shape=shape, dx=dx, positions=positions, velocities=velocities,
                 gravity=(0.0, -9.81, 0.0), fluid_cells=int(fluid.sum()), notes={"seed": seed, "bodies": bodies})

