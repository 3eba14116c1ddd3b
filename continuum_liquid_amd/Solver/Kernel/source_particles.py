"""Host seeding of particles from source masks (reference
Solver/Kernel/source_particles.py).

Semantics kept: OUTFLOW removes particles whose cell lies in its mask, INFLOW
(every step) and GEOMETRY (first step only) top every masked, non-solid cell
up to `particle_density` particles placed uniformly at random in the cell,
with the source velocity (reference source_particles.py:42-99).

The reference builds masks by voxelising USD meshes (voxelise_mesh.py), which
needs `pxr`; here a source carries its mask directly: `mask` (bool array of
the grid shape) or `shapes`, a list of {"box": [[x0, y0, z0], [x1, y1, z1]]}
or {"sphere": [[cx, cy, cz], r]} in domain coordinates, marking cells whose
centre is inside.
"""

from __future__ import annotations

import numpy as np

FIELD_DTYPE = np.float32


def shape_mask(shapes, grid_shape, delta) -> np.ndarray:
    centres = [(np.arange(n) + 0.5) * delta for n in grid_shape]
    X, Y, Z = np.meshgrid(*centres, indexing="ij")
    mask = np.zeros(grid_shape, bool)
    for shape in shapes or ():
        if "box" in shape:
            lo, hi = (np.asarray(v, float) for v in shape["box"])
            mask |= (X >= lo[0]) & (X < hi[0]) & (Y >= lo[1]) & (Y < hi[1]) & (Z >= lo[2]) & (Z < hi[2])
        elif "sphere" in shape:
            c, r = shape["sphere"]
            mask |= (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 < float(r) ** 2
        else:
            raise ValueError(f"unknown source shape {shape}")
    return mask


def resolve_mask(entry: dict, grid_shape, delta) -> np.ndarray:
    if entry.get("mask") is not None:
        mask = np.asarray(entry["mask"], dtype=bool)
        if mask.shape != tuple(grid_shape):
            raise ValueError(f"source mask shape {mask.shape} != grid {tuple(grid_shape)}")
        return mask
    if entry.get("shapes"):
        return shape_mask(entry["shapes"], grid_shape, delta)
    if entry.get("geometry_inputs"):
        raise NotImplementedError(
            "USD mesh voxelisation is outside the HIP step; give the source a 'mask' or 'shapes' entry"
        )
    return np.zeros(grid_shape, bool)


def particle_cells(positions, delta, grid_shape) -> np.ndarray:
    if len(positions) == 0:
        return np.empty((0, 3), dtype=np.int32)
    cells = np.floor(np.asarray(positions) / delta).astype(np.int32)
    return np.clip(cells, 0, np.asarray(grid_shape, dtype=np.int32) - 1)


def mask_membership(positions, mask, delta) -> np.ndarray:
    cells = particle_cells(positions, delta, mask.shape)
    if len(cells) == 0:
        return np.zeros(0, dtype=bool)
    return mask[cells[:, 0], cells[:, 1], cells[:, 2]]


def remove_masked(particles, velocities, mask, delta):
    keep = ~mask_membership(particles, mask, delta)
    return particles[keep], velocities[keep]


def fill_source(particles, velocities, source, obstacle_mask, delta, rng):
    target = max(1, int(source.get("particle_density", 8)))
    region = np.asarray(source["mask"], dtype=bool) & ~obstacle_mask
    cells = np.argwhere(region)
    if len(cells) == 0:
        return particles, velocities
    counts = np.zeros(region.size, dtype=np.int64)
    if len(particles):
        flat = np.ravel_multi_index(particle_cells(particles, delta, region.shape).T, region.shape)
        counts = np.bincount(flat, minlength=region.size)
    missing = np.maximum(0, target - counts.reshape(region.shape)[tuple(cells.T)])
    spawn = np.repeat(cells, missing, axis=0)
    if len(spawn) == 0:
        return particles, velocities
    new_pos = (spawn + rng.random((len(spawn), 3))) * delta
    new_vel = np.broadcast_to(np.asarray(source.get("velocity", (0.0, 0.0, 0.0)), FIELD_DTYPE), new_pos.shape)
    return (
        np.ascontiguousarray(np.concatenate((particles, new_pos)), dtype=FIELD_DTYPE),
        np.ascontiguousarray(np.concatenate((velocities, new_vel)), dtype=FIELD_DTYPE),
    )


def apply_sources(particles, velocities, sources, obstacle_mask, delta, rng, initial=False):
    for source in sources:
        if source.get("type") == "OUTFLOW":
            particles, velocities = remove_masked(particles, velocities, source["mask"], delta)
    for source in sources:
        kind = source.get("type")
        if kind == "INFLOW" or (initial and kind == "GEOMETRY"):
            particles, velocities = fill_source(particles, velocities, source, obstacle_mask, delta, rng)
    return particles, velocities
