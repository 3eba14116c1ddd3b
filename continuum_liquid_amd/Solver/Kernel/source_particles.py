"""Particle sources and sinks applied on the host before a step.

Semantics (reference Solver/Kernel/source_particles.py:21-92, applied at
solver.py:253-306 and once at set-up):

* A particle at x belongs to the cell ``floor(x / delta)`` per axis, evaluated
  in the dtype of the position array and clamped into the grid.
* Every OUTFLOW source first deletes the particles whose cell it marks, in the
  order the sources are listed; surviving particles keep their order.
* Then every INFLOW source (on every call) and every GEOMETRY source (on the
  set-up call only), again in list order, tops up each cell it marks that is
  not solid to ``particle_density`` particles (at least 1, default 8),
  counting the particles present at that moment.  The new particles are
  appended in C order of their cells, a cell missing m particles
  contributing m consecutive ones.  Their positions are
  ``(cell + u) * delta`` with u from a single ``rng.random((count, 3))`` draw
  per source (float64, rounded to float32 once), and their velocity is the
  source's ``velocity``.  A source that adds nothing draws nothing, so the
  generator's stream, and with it every later position, matches the
  reference for the same seed.

A source carries its cells as ``mask`` (bool array of the grid shape),
``shapes`` (``{"box": [[x0, y0, z0], [x1, y1, z1]]}`` or ``{"sphere": [[cx, cy,
cz], r]}`` in domain coordinates, marking cells whose centre lies inside) or
``mesh`` (``{"vertices": (V, 3), "triangles": (T, 3)}`` in world coordinates,
optionally with a ``matrix_world``) or the reference's ``geometry_inputs``
(USD files with a ``matrix_world``, reference voxelise_mesh.py:111-163;
reading them needs the ``pxr`` package); meshes become masks by ray parity
on the GPU.
"""

from __future__ import annotations

import numpy as np

FIELD_DTYPE = np.float32
DEFAULT_DENSITY = 8


def shape_mask(shapes, grid_shape, delta) -> np.ndarray:
    """Cells whose centre lies inside any of the boxes / spheres."""
    axes = [(np.arange(n) + 0.5) * delta for n in grid_shape]
    cx, cy, cz = np.meshgrid(*axes, indexing="ij")
    inside = np.zeros(grid_shape, dtype=bool)
    for shape in shapes or ():
        if "box" in shape:
            lo, hi = (np.asarray(corner, dtype=float) for corner in shape["box"])
            inside |= (lo[0] <= cx) & (cx < hi[0]) & (lo[1] <= cy) & (cy < hi[1]) & (lo[2] <= cz) & (cz < hi[2])
        elif "sphere" in shape:
            centre, radius = shape["sphere"]
            inside |= (cx - centre[0]) ** 2 + (cy - centre[1]) ** 2 + (cz - centre[2]) ** 2 < float(radius) ** 2
        else:
            raise ValueError(f"unknown source shape {shape}")
    return inside


def resolve_mask(entry: dict, grid_shape, delta, voxelise=None, bake_path=None) -> np.ndarray:
    """The bool cell mask a source or obstacle entry describes.  `voxelise`
    (a list of (T, 3, 3) world-triangle arrays -> bool mask, the union of
    their ray-parity masks) serves mesh entries; `bake_path` is the folder
    the USD files of `geometry_inputs` are relative to."""
    from . import voxelise_mesh as vm

    if entry.get("mask") is not None:
        mask = np.asarray(entry["mask"], dtype=bool)
        if mask.shape != tuple(grid_shape):
            raise ValueError(f"source mask shape {mask.shape} != grid {tuple(grid_shape)}")
        return mask
    if entry.get("shapes"):
        return shape_mask(entry["shapes"], grid_shape, delta)
    meshes = []
    if entry.get("mesh") is not None:
        mesh = entry["mesh"]
        triangles = np.asarray(mesh["vertices"], dtype=np.float32)[np.asarray(mesh["triangles"], dtype=np.int64)]
        if mesh.get("matrix_world") is not None:
            triangles = vm.world_triangles(triangles, mesh["matrix_world"])
        meshes.append(triangles)
    for item in entry.get("geometry_inputs") or ():
        meshes.append(vm.entry_triangles(item, bake_path))
    if meshes:
        if voxelise is None:
            raise ValueError("a mesh entry needs the GPU voxeliser")
        return voxelise(meshes)
    return np.zeros(grid_shape, dtype=bool)


def cell_of(positions, delta, grid_shape) -> np.ndarray:
    """Integer cell (n, 3) of each particle, clamped into the grid."""
    if len(positions) == 0:
        return np.empty((0, 3), dtype=np.int32)
    upper = np.asarray(grid_shape, dtype=np.int32) - 1
    return np.clip(np.floor(np.asarray(positions) / delta).astype(np.int32), 0, upper)


def _flat_cells(positions, delta, grid_shape) -> np.ndarray:
    return np.ravel_multi_index(cell_of(positions, delta, grid_shape).T, grid_shape)


def drop_in(positions, velocities, mask, delta):
    """Particles (and their velocities) outside the cells `mask` marks."""
    if len(positions) == 0:
        return positions, velocities
    keep = ~mask.reshape(-1)[_flat_cells(positions, delta, mask.shape)]
    return positions[keep], velocities[keep]


def top_up(positions, velocities, source, solid, delta, rng):
    """Append the particles that bring every open source cell to the density."""
    density = max(1, int(source.get("particle_density", DEFAULT_DENSITY)))
    open_cells = np.asarray(source["mask"], dtype=bool) & ~solid
    flat_open = np.flatnonzero(open_cells)  # C order
    if flat_open.size == 0:
        return positions, velocities
    present = np.zeros(open_cells.size, dtype=np.int64)
    if len(positions):
        present = np.bincount(_flat_cells(positions, delta, open_cells.shape), minlength=open_cells.size)
    need = np.maximum(0, density - present[flat_open])
    total = int(need.sum())
    if total == 0:
        return positions, velocities
    corner = np.column_stack(np.unravel_index(np.repeat(flat_open, need), open_cells.shape))
    placed = ((corner + rng.random((total, 3))) * delta).astype(FIELD_DTYPE)
    speed = np.broadcast_to(np.asarray(source.get("velocity", (0.0, 0.0, 0.0)), dtype=FIELD_DTYPE), placed.shape)
    return (
        np.ascontiguousarray(np.concatenate((np.asarray(positions, FIELD_DTYPE), placed))),
        np.ascontiguousarray(np.concatenate((np.asarray(velocities, FIELD_DTYPE), speed))),
    )


def apply_sources(positions, velocities, sources, solid, delta, rng, initial=False):
    """Sinks first, then the top-ups, each in list order."""
    for source in sources:
        if source.get("type") == "OUTFLOW":
            positions, velocities = drop_in(positions, velocities, source["mask"], delta)
    fill_kinds = ("INFLOW", "GEOMETRY") if initial else ("INFLOW",)
    for source in sources:
        if source.get("type") in fill_kinds:
            positions, velocities = top_up(positions, velocities, source, solid, delta, rng)
    return positions, velocities
