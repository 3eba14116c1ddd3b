"""Mesh voxelisation: the oracle's clq_voxelise_mesh (reference
Solver/Kernel/OpenCL/voxelise_mesh.cl:5-48) pinned against closed forms, and
the host-side mesh reading (reference Solver/Kernel/voxelise_mesh.py:17-118).

The ray-parity mask of a closed mesh is its inside: a box's cells are those
whose centre lies inside the box; a tessellated sphere's lie between the
inscribed and the circumscribed radius; a hollow shell's inside is empty; an
open triangle marks the cells whose ray (direction normalize(1, 0.3713907,
0.2178311)) crosses it.  Cells whose centre lies within a small margin of a
surface (where float rounding may decide) are excluded from each comparison.
"""

from __future__ import annotations

import numpy as np
import pytest

from continuum_liquid_amd import runtime
from continuum_liquid_amd.Solver.Kernel import source_particles as sp
from continuum_liquid_amd.Solver.Kernel import voxelise_mesh as vm

D = np.array([1.0, 0.3713907, 0.2178311])
D /= np.linalg.norm(D)


def box_mesh(lo, hi) -> np.ndarray:
    """12 outward-facing triangles of an axis-aligned box."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    quads = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (2, 3, 7, 6), (1, 2, 6, 5), (0, 4, 7, 3)]
    tris = [[q[0], q[1], q[2]] for q in quads] + [[q[0], q[2], q[3]] for q in quads]
    return v[np.array(tris)]


def sphere_mesh(centre, radius, rings=16, segments=32) -> np.ndarray:
    """UV sphere with vertices on the sphere (the polyhedron is inscribed)."""
    theta = np.linspace(0, np.pi, rings + 1)
    phi = np.linspace(0, 2 * np.pi, segments + 1)[:-1]
    t, p = np.meshgrid(theta, phi, indexing="ij")
    pts = np.stack([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)], -1) * radius + np.asarray(centre)
    tris = []
    for r in range(rings):
        for s in range(segments):
            a, b = pts[r, s], pts[r, (s + 1) % segments]
            c, d = pts[r + 1, s], pts[r + 1, (s + 1) % segments]
            if r > 0:
                tris.append([a, c, b])
            if r < rings - 1:
                tris.append([b, c, d])
    return np.asarray(tris, np.float32)


def centres(shape, delta, origin):
    axes = [origin[a] + (np.arange(shape[a]) + 0.5) * delta for a in range(3)]
    return np.stack(np.meshgrid(*axes, indexing="ij"), -1)


def interior(shape):
    m = np.zeros(shape, bool)
    m[1:-1, 1:-1, 1:-1] = True
    return m


def voxelise(lib, shape, delta, tris, origin, mask=None):
    with runtime.Context(lib, shape, delta) as ctx:
        return ctx.voxelise_mesh(tris, origin, mask).astype(bool)


SHAPE, DELTA, ORIGIN = (26, 22, 30), 0.05, (-0.65, -0.55, 0.0)


def test_box_mesh_marks_the_cells_inside(oracle_lib):
    lo, hi = np.array([-0.37, -0.26, 0.33]), np.array([0.29, 0.41, 1.12])
    got = voxelise(oracle_lib, SHAPE, DELTA, box_mesh(lo, hi), ORIGIN)
    c = centres(SHAPE, DELTA, ORIGIN)
    inside = np.all((c > lo) & (c < hi), axis=-1) & interior(SHAPE)
    assert inside.sum() > 1000
    assert np.array_equal(got, inside)


def test_boundary_cells_are_never_marked(oracle_lib):
    """A box enclosing the whole grid marks exactly the interior cells
    (voxelise_mesh.cl:36-37 skips index 0 and n - 1 on every axis)."""
    got = voxelise(oracle_lib, SHAPE, DELTA, box_mesh(np.array([-2.0, -2.0, -1.0]), np.array([2.0, 2.0, 3.0])),
                   ORIGIN)
    assert np.array_equal(got, interior(SHAPE))


def test_sphere_mesh_between_inscribed_and_circumscribed(oracle_lib):
    centre, radius, rings, segments = np.array([0.02, 0.03, 0.74]), 0.47, 16, 32
    got = voxelise(oracle_lib, SHAPE, DELTA, sphere_mesh(centre, radius, rings, segments), ORIGIN)
    dist = np.linalg.norm(centres(SHAPE, DELTA, ORIGIN) - centre, axis=-1)
    # every point of the inscribed polyhedron is within r, and every face is
    # at least r cos(pi/rings) cos(pi/segments) from the centre
    inner = radius * np.cos(np.pi / rings) * np.cos(np.pi / segments)
    sure_in = (dist < inner - 1e-4) & interior(SHAPE)
    sure_out = dist > radius + 1e-4
    assert sure_in.sum() > 2000
    assert got[sure_in].all()
    assert not got[sure_out].any()
    assert not got[~interior(SHAPE)].any()


def test_hollow_shell_and_union_of_calls(oracle_lib):
    """Two nested boxes in one entry: parity leaves the inner one empty.  Two
    calls on one mask give the union, and cells already set stay set."""
    outer = box_mesh(np.array([-0.52, -0.44, 0.12]), np.array([0.48, 0.43, 1.31]))
    inner_lo, inner_hi = np.array([-0.21, -0.17, 0.46]), np.array([0.18, 0.22, 0.93])
    shell = np.concatenate([outer, box_mesh(inner_lo, inner_hi)])
    got = voxelise(oracle_lib, SHAPE, DELTA, shell, ORIGIN)
    c = centres(SHAPE, DELTA, ORIGIN)
    in_inner = np.all((c > inner_lo) & (c < inner_hi), axis=-1)
    assert in_inner.sum() > 100 and not got[in_inner].any()
    # the union: the inner box in a second call fills the hole back in
    mask = np.zeros(SHAPE, np.uint8)
    mask[0, 0, 0] = 1  # a boundary cell set by the caller stays set
    with runtime.Context(oracle_lib, SHAPE, DELTA) as ctx:
        ctx.voxelise_mesh(shell, ORIGIN, mask)
        ctx.voxelise_mesh(box_mesh(inner_lo, inner_hi), ORIGIN, mask)
    solid = voxelise(oracle_lib, SHAPE, DELTA, outer, ORIGIN)
    expect = solid.copy()
    expect[0, 0, 0] = True
    assert np.array_equal(mask.astype(bool), expect)


def test_open_triangle_marks_the_cells_behind_it(oracle_lib):
    """A single triangle in the plane x = x0: a cell with centre left of the
    plane is marked iff its ray meets the plane inside the triangle."""
    x0 = 0.123
    a, b, c = np.array([x0, -0.4, 0.2]), np.array([x0, 0.45, 0.35]), np.array([x0, 0.05, 1.3])
    got = voxelise(oracle_lib, SHAPE, DELTA, np.array([[a, b, c]], np.float32), ORIGIN)
    p = centres(SHAPE, DELTA, ORIGIN)
    t = (x0 - p[..., 0]) / D[0]
    hit = p + t[..., None] * D
    # barycentric coordinates of the hit point in the (y, z) plane
    m = np.array([[b[1] - a[1], c[1] - a[1]], [b[2] - a[2], c[2] - a[2]]])
    uv = np.einsum("ij,...j->...i", np.linalg.inv(m), np.stack([hit[..., 1] - a[1], hit[..., 2] - a[2]], -1))
    w = 1 - uv[..., 0] - uv[..., 1]
    margin = np.minimum(np.minimum(uv[..., 0], uv[..., 1]), w)
    ahead = t > 1e-4
    sure_in = ahead & (margin > 1e-4) & interior(SHAPE)
    sure_out = (t < -1e-4) | (margin < -1e-4)
    assert sure_in.sum() > 100
    assert got[sure_in].all() and not got[sure_out].any()


def test_degenerate_triangles_change_nothing(oracle_lib):
    box = box_mesh(np.array([-0.37, -0.26, 0.33]), np.array([0.29, 0.41, 1.12]))
    line = np.array([[[-0.5, -0.5, 0.2], [0.0, 0.0, 0.7], [0.5, 0.5, 1.2]]], np.float32)  # collinear
    point = np.full((1, 3, 3), 0.1, np.float32)
    with_degenerate = voxelise(oracle_lib, SHAPE, DELTA, np.concatenate([box, line, point]), ORIGIN)
    assert np.array_equal(with_degenerate, voxelise(oracle_lib, SHAPE, DELTA, box, ORIGIN))


def test_mesh_entry_equals_shape_entry(oracle_lib):
    """A box given as a mesh (world coordinates, origin (-x/2, -y/2, 0)) and
    as a shape (domain coordinates) marks the same interior cells."""
    size = {"x": SHAPE[0] * DELTA, "y": SHAPE[1] * DELTA, "z": SHAPE[2] * DELTA}
    origin = vm.grid_origin({"domain": {"size": size}})
    lo, hi = np.array([0.31, 0.22, 0.43]), np.array([0.97, 0.83, 1.21])
    tris = box_mesh(lo + origin, hi + origin)
    vertices, index = np.unique(tris.reshape(-1, 3), axis=0, return_inverse=True)
    with runtime.Context(oracle_lib, SHAPE, DELTA) as ctx:
        def voxelise_fn(meshes):
            mask = np.zeros(SHAPE, np.uint8)
            for t in meshes:
                ctx.voxelise_mesh(t, origin, mask)
            return mask.astype(bool)

        from_mesh = sp.resolve_mask({"mesh": {"vertices": vertices, "triangles": index.reshape(-1, 3)}},
                                    SHAPE, DELTA, voxelise_fn)
    from_shape = sp.resolve_mask({"shapes": [{"box": [lo.tolist(), hi.tolist()]}]}, SHAPE, DELTA)
    assert np.array_equal(from_mesh, from_shape & interior(SHAPE))


def test_usd_entries_need_pxr():
    with pytest.raises(NotImplementedError, match="pxr"):
        sp.resolve_mask({"geometry_inputs": [{"mesh_file": "a.usdc", "matrix_world": np.eye(4).tolist()}]},
                        SHAPE, DELTA, lambda m: None, "/nonexistent")


# ---- host-side mesh reading --------------------------------------------------------------

def test_convex_polygon_is_cut_from_its_last_vertex():
    """The ear clipping takes the first convex corner with no vertex inside,
    scanning from index 0 with the previous vertex wrapping to the last: on a
    convex pentagon that is a fan around vertex 4 (worked by hand from the
    reference's steps, voxelise_mesh.py:50-71)."""
    angles = np.linspace(0, 2 * np.pi, 6)[:-1]
    pts = np.stack([np.cos(angles), np.sin(angles), np.zeros(5)], -1).astype(np.float32)
    assert vm.triangulate_polygon(pts, [0, 1, 2, 3, 4]) == [[4, 0, 1], [4, 1, 2], [2, 3, 4]]


@pytest.mark.parametrize("plane", ["xy", "yz", "xz"])
def test_concave_polygon_area_and_orientation(plane):
    """An L-shaped hexagon (one reflex corner) in each coordinate plane: the
    triangles cover its area exactly, with the polygon's orientation."""
    l2d = np.array([[0, 0], [2, 0], [2, 1], [1, 1], [1, 2], [0, 2]], float)
    pts = np.zeros((6, 3))
    axes = {"xy": (0, 1), "yz": (1, 2), "xz": (0, 2)}[plane]
    pts[:, axes[0]], pts[:, axes[1]] = l2d[:, 0], l2d[:, 1]
    face = [0, 1, 2, 3, 4, 5]
    tris = vm.triangulate_polygon(pts, face)
    assert len(tris) == 4
    area = 0.0
    for t in tris:
        a, b, c = l2d[t[0]], l2d[t[1]], l2d[t[2]]
        signed = 0.5 * ((b - a)[0] * (c - a)[1] - (b - a)[1] * (c - a)[0])
        assert signed > 0  # same orientation as the polygon (counter-clockwise)
        area += signed
    assert area == pytest.approx(3.0)


def test_self_intersecting_face_is_rejected():
    bow = np.array([[0, 0, 0], [1, 1, 0], [1, 0, 0], [0, 1, 0], [0.5, 2, 0]], np.float32)
    with pytest.raises(ValueError, match="Cannot triangulate"):
        vm.triangulate_polygon(bow, [0, 1, 2, 3, 0])


def test_faces_skip_holes_and_short_faces():
    pts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    counts = [3, 4, 2, 3]
    indices = [0, 1, 4, 0, 1, 2, 3, 1, 2, 2, 3, 4]
    tris = vm.triangles_from_faces(pts, indices, counts, holes=[3])
    # face 0 as is, face 1 (a quad) in two, face 2 (two vertices) and the hole dropped
    assert tris.shape == (3, 3, 3) and tris.dtype == np.float32
    assert np.array_equal(tris[0], pts[[0, 1, 4]])
    assert vm.triangles_from_faces(pts, [], [], ()).shape == (0, 3, 3)


def test_world_matrix_rotates_and_translates():
    tris = np.array([[[1, 0, 0], [0, 1, 0], [0, 0, 1]]], np.float32)
    m = np.eye(4)
    m[:3, :3] = [[0, -1, 0], [1, 0, 0], [0, 0, 1]]  # 90 degrees about z
    m[:3, 3] = [0.5, -0.25, 2.0]
    out = vm.world_triangles(tris, m)
    assert np.allclose(out[0, 0], [0.5, 0.75, 2.0]) and np.allclose(out[0, 1], [-0.5, -0.25, 2.0])
