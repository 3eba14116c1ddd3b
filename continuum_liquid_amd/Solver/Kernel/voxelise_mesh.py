"""Mesh geometry to cell masks (reference Solver/Kernel/voxelise_mesh.py).

Sources and obstacles given as meshes become cell masks by ray parity on the
device (runtime.Context.voxelise_mesh, the reference kernel
voxelise_mesh.cl:24-48).  The host side reads the triangles:

* `triangles_from_faces` walks a polygon mesh's faces (face-vertex counts and
  indices, hole faces skipped, faces with fewer than three vertices dropped)
  and splits polygons into triangles (reference voxelise_mesh.py:85-108);
* `triangulate_polygon` is the ear clipping of reference
  voxelise_mesh.py:17-71, kept step for step so that a mesh produces the
  same triangles, in the same order, as in the reference: the polygon is
  projected onto the plane of its two minor Newell-normal axes, and the first
  convex vertex (cross product above a tolerance) whose triangle holds no
  other remaining vertex is cut off, repeatedly, until three vertices remain;
* `load_usdc_triangles` reads the first UsdGeom.Mesh of a USD file; it needs
  the `pxr` package;
* `world_triangles` applies the entry's 4 x 4 `matrix_world`.

The grid's world origin is (-size_x / 2, -size_y / 2, 0), as in the reference
(voxelise_mesh.py:152).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np


def _turn(a, b, c) -> float:
    """z of (b - a) x (c - b): positive for a left turn a -> b -> c."""
    u, w = b - a, c - b
    return u[0] * w[1] - u[1] * w[0]


def _holds(point, a, b, c, orientation, tolerance) -> bool:
    """Whether `point` lies in triangle abc (edges included, within the
    tolerance) of the given orientation."""
    for start, end in ((a, b), (b, c), (c, a)):
        side = end - start
        offset = point - start
        if orientation * (side[0] * offset[1] - side[1] * offset[0]) < -tolerance:
            return False
    return True


def triangulate_polygon(points, face) -> list[list[int]]:
    """Triangles (as vertex-index triples) of one planar polygon face."""
    corners = np.asarray(points[face], dtype=np.float64)
    nxt = np.roll(corners, -1, axis=0)
    # Newell normal of the polygon
    normal = np.array([
        np.sum((corners[:, 1] - nxt[:, 1]) * (corners[:, 2] + nxt[:, 2])),
        np.sum((corners[:, 2] - nxt[:, 2]) * (corners[:, 0] + nxt[:, 0])),
        np.sum((corners[:, 0] - nxt[:, 0]) * (corners[:, 1] + nxt[:, 1])),
    ])
    flat = np.delete(corners, int(np.argmax(np.abs(normal))), axis=1)
    area2 = np.sum(flat[:, 0] * np.roll(flat[:, 1], -1) - np.roll(flat[:, 0], -1) * flat[:, 1])
    orientation = 1.0 if area2 >= 0.0 else -1.0
    span = max(float(np.ptp(flat, axis=0).max()), 1.0)
    tolerance = np.finfo(np.float64).eps * span * span * 32.0

    left = list(range(len(face)))
    out: list[list[int]] = []
    while len(left) > 3:
        for at in range(len(left)):
            before, here, after = left[at - 1], left[at], left[(at + 1) % len(left)]
            a, b, c = flat[before], flat[here], flat[after]
            if orientation * _turn(a, b, c) <= tolerance:
                continue  # reflex or flat corner
            if any(_holds(flat[o], a, b, c, orientation, tolerance)
                   for o in left if o not in (before, here, after)):
                continue  # another vertex inside the would-be ear
            out.append([face[before], face[here], face[after]])
            left.pop(at)
            break
        else:
            raise ValueError("Cannot triangulate a degenerate or self-intersecting USD face")
    out.append([face[o] for o in left])
    return out


def triangles_from_faces(points, indices, counts, holes=()) -> np.ndarray:
    """(T, 3, 3) float32 triangles of a polygon mesh in face-vertex form."""
    points = np.asarray(points, dtype=np.float32)
    indices = np.asarray(indices, dtype=np.int64)
    holes = {int(h) for h in holes}
    triples: list[list[int]] = []
    start = 0
    for number, size in enumerate(np.asarray(counts, dtype=np.int64)):
        face = [int(v) for v in indices[start:start + int(size)]]
        start += int(size)
        if number in holes or len(face) < 3:
            continue
        triples.extend([face] if len(face) == 3 else triangulate_polygon(points, face))
    if not triples:
        return np.empty((0, 3, 3), dtype=np.float32)
    return np.ascontiguousarray(points[np.asarray(triples, dtype=np.int64)])


def load_usdc_triangles(path) -> np.ndarray:
    """Triangles of the first mesh prim of a USD file (reference
    voxelise_mesh.py:74-108)."""
    try:
        from pxr import Usd, UsdGeom
    except ImportError as exc:
        raise NotImplementedError(
            f"reading {path} needs the pxr (USD) package; give the entry a 'mesh' "
            "(vertices, triangles), 'mask' or 'shapes' instead") from exc
    stage = Usd.Stage.Open(str(path))
    mesh = next((UsdGeom.Mesh(prim) for prim in stage.Traverse() if prim.IsA(UsdGeom.Mesh)), None)
    if mesh is None:
        raise ValueError(f"USD file contains no mesh: {path}")
    return triangles_from_faces(
        np.asarray(mesh.GetPointsAttr().Get(), dtype=np.float32),
        np.asarray(mesh.GetFaceVertexIndicesAttr().Get(), dtype=np.int32),
        np.asarray(mesh.GetFaceVertexCountsAttr().Get(), dtype=np.int32),
        mesh.GetHoleIndicesAttr().Get() or (),
    )


def world_triangles(triangles, matrix_world) -> np.ndarray:
    """Triangles moved by a 4 x 4 object-to-world matrix (column vectors)."""
    matrix = np.asarray(matrix_world, dtype=np.float32)
    flat = np.asarray(triangles, dtype=np.float32).reshape(-1, 3)
    moved = flat @ matrix[:3, :3].T + matrix[:3, 3]
    return np.ascontiguousarray(moved.reshape(-1, 3, 3), dtype=np.float32)


def grid_origin(simulation: dict) -> tuple[float, float, float]:
    size = simulation["domain"]["size"]
    return (-0.5 * float(size["x"]), -0.5 * float(size["y"]), 0.0)


def entry_triangles(entry: dict, bake_path) -> np.ndarray:
    """World triangles of one geometry input (reference voxelise_mesh.py:111-118)."""
    return world_triangles(load_usdc_triangles(Path(bake_path) / entry["mesh_file"]), entry["matrix_world"])


def voxelise_entries(ctx, entries, bake_path, origin) -> np.ndarray:
    """Union over `entries` of each entry's ray-parity mask (reference
    voxelise_mesh.py:121-135), as a bool grid array."""
    mask = np.zeros(ctx.shape, dtype=np.uint8)
    for entry in entries or ():
        triangles = entry_triangles(entry, bake_path)
        if triangles.size:
            ctx.voxelise_mesh(triangles, origin, mask)
    return mask.astype(bool)
