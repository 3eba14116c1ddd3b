"""Mesh voxelisation on the device (reference
Solver/Kernel/OpenCL/voxelise_mesh.cl:24-48): the HIP kernel, which culls
triangles per tile of cells before the exact ray test, against the oracle's
brute force over every triangle, bit for bit; at 256^3 against the sphere's
closed form."""

from __future__ import annotations

import time

import numpy as np
import pytest

from continuum_liquid_amd import runtime

from .test_oracle_mesh import box_mesh, centres, interior, sphere_mesh

pytestmark = pytest.mark.gpu

D = np.array([1.0, 0.3713907, 0.2178311])
D /= np.linalg.norm(D)


def both(oracle_lib, hip_lib, shape, delta, tris, origin):
    out = []
    for lib in (oracle_lib, hip_lib):
        with runtime.Context(lib, shape, delta) as ctx:
            mask = np.zeros(shape, np.uint8)
            mask[1, 2, 3] = 1  # kept by both
            out.append(ctx.voxelise_mesh(tris, origin, mask))
    return out


def triangle_soup(rng, n, lo, hi):
    """Random triangles of every kind the cull has to get right: small and
    large, skinny, nearly parallel to the ray, degenerate, partly outside the
    grid, and a few with non-finite vertices."""
    lo, hi = np.asarray(lo), np.asarray(hi)
    centre = rng.uniform(lo - 0.2, hi + 0.2, (n, 1, 3))
    size = np.exp(rng.uniform(np.log(0.01), np.log(1.5), (n, 1, 1)))
    tris = centre + rng.normal(0, 1, (n, 3, 3)) * size
    k = n // 8
    # skinny: the third vertex near the first edge
    tris[:k, 2] = tris[:k, 0] + 0.5 * (tris[:k, 1] - tris[:k, 0]) + rng.normal(0, 1e-4, (k, 3))
    # nearly containing the ray direction: edges along D and a tiny offset
    tris[k:2 * k, 1] = tris[k:2 * k, 0] + D * rng.uniform(0.1, 1.0, (k, 1))
    tris[k:2 * k, 2] = tris[k:2 * k, 0] + D * rng.uniform(0.1, 1.0, (k, 1)) + rng.normal(0, 1e-3, (k, 3))
    # degenerate: a repeated vertex
    tris[2 * k:2 * k + 10, 2] = tris[2 * k:2 * k + 10, 1]
    tris[2 * k + 10, 0, 1] = np.nan
    tris[2 * k + 11, 2, 0] = np.inf
    return tris.astype(np.float32)


@pytest.mark.parametrize("case", ["box", "sphere", "shell", "soup"])
def test_mesh_mask_bit_exact(oracle_lib, hip_lib, case):
    shape, delta, origin = (48, 40, 56), 0.025, (-0.6, -0.5, 0.05)
    rng = np.random.default_rng(11)
    if case == "box":
        tris = box_mesh(np.array([-0.33, -0.27, 0.31]), np.array([0.41, 0.36, 1.02]))
    elif case == "sphere":
        tris = sphere_mesh(np.array([0.01, -0.02, 0.72]), 0.43, rings=48, segments=96)
    elif case == "shell":
        tris = np.concatenate([sphere_mesh(np.array([0.05, 0.0, 0.7]), 0.5, 24, 48),
                               box_mesh(np.array([-0.2, -0.15, 0.5]), np.array([0.25, 0.2, 0.9]))])
    else:
        hi = np.array(origin) + np.array(shape) * delta
        tris = triangle_soup(rng, 3000, origin, hi)
    print(f"mesh case {case}: {len(tris)} triangles on {shape}", flush=True)
    o, h = both(oracle_lib, hip_lib, shape, delta, tris, origin)
    assert np.array_equal(o, h)
    assert h[1, 2, 3] == 1
    if case != "soup":
        assert h[interior(shape)].sum() > 1000


def test_no_triangles_and_bad_shapes(hip_lib):
    with runtime.Context(hip_lib, (8, 8, 8), 0.1) as ctx:
        mask = ctx.voxelise_mesh(np.empty((0, 3, 3), np.float32), (0, 0, 0))
        assert not mask.any()
        with pytest.raises(ValueError):
            ctx.voxelise_mesh(np.zeros((4, 3), np.float32), (0, 0, 0))


def test_sphere_at_256_cubed(hip_lib):
    """A 20,000-triangle sphere on the 256^3 grid: the mask lies between the
    inscribed and circumscribed spheres (closed form) except where the
    reference's own ray test is blind (below)."""
    shape, delta = (256, 256, 256), 1.0 / 256
    origin = (-0.5, -0.5, 0.0)
    centre, radius, rings, segments = np.array([0.03, -0.04, 0.52]), 0.37, 100, 100
    tris = sphere_mesh(centre, radius, rings, segments)
    with runtime.Context(hip_lib, shape, delta) as ctx:
        ctx.voxelise_mesh(tris[:64], origin)  # first call: code objects loaded
        began = time.perf_counter()
        mask = ctx.voxelise_mesh(tris, origin).astype(bool)
        seconds = time.perf_counter() - began
    print(f"voxelise_mesh 256^3, {len(tris)} triangles: {seconds * 1e3:.1f} ms (with transfers)", flush=True)
    points = centres(shape, delta, origin)
    dist = np.linalg.norm(points - centre, axis=-1)
    inner = radius * np.cos(np.pi / rings) * np.cos(np.pi / segments)
    sure_in = (dist < inner - 1e-5) & interior(shape)
    wrong = np.argwhere((sure_in & ~mask) | ((dist > radius + 1e-5) & mask))
    print(f"cells against the closed form: {len(wrong)} of {int(sure_in.sum())} inside", flush=True)
    # The reference's ray test has two blind spots, both kept: it accepts a
    # hit on a triangle's edges (u, v >= 0, u + v <= 1 inclusive), so a ray
    # through an edge two triangles share counts twice; and it ignores every
    # triangle with |det| = |e1 . (D x e2)| < 1e-7 (an absolute threshold),
    # so a ray leaving through a small triangle seen edge-on is not counted.
    # Every disagreement must be one of these: its ray passes within 1e-5
    # (barycentric) of a triangle edge, or through a triangle with
    # |det| < 1e-7.
    assert len(wrong) <= 1e-5 * sure_in.sum()
    t64 = tris.astype(np.float64)
    e1, e2 = t64[:, 1] - t64[:, 0], t64[:, 2] - t64[:, 0]
    pvec = np.cross(D, e2)
    det = np.einsum("ij,ij->i", e1, pvec)
    for cell in wrong:
        tvec = points[tuple(cell)] - t64[:, 0]
        u = np.einsum("ij,ij->i", tvec, pvec) / det
        qvec = np.cross(tvec, e1)
        v = (qvec @ D) / det
        along = np.einsum("ij,ij->i", e2, qvec) / det
        bary = np.stack([u, v, 1 - u - v], -1)
        crossed = (along > 0) & (bary.min(-1) > -1e-5)
        grazed = crossed & ((np.abs(bary).min(-1) < 1e-5) | (np.abs(det) < 1.01e-7))
        assert grazed.any(), f"cell {tuple(cell)} is wrong with neither an edge nor a skipped triangle on its ray"
