"""Pins of the CPU oracle (oracle/cliquid_oracle.c) to mathematics it must obey.

The reference has no tests, fixtures or golden vectors, and its OpenCL path
cannot run here, so each oracle function is pinned by an independent check
chosen so that a plausible slip (a wrong face offset, a dropped neighbour bin,
a swapped sign, a transposed operand) fails it:

  binning (A1-A4)     closed-form keys in numpy, np.bincount, cumsum,
                      argsort(kind="stable"); adversarial boundary positions
  P2G + occupancy     FP64 brute force over every (particle, cell) pair, no
  (A5-A6)             neighbour-bin restriction; occupancy exactly count/8
  grid glue (A7, A8,  vectorised float32 closed forms (fused terms emulated
  A15)                in float64), exact or within 1 ulp
  RBGS (A9)           sequential red-black Gauss-Seidel in float64
  restriction (A10)   Galerkin identity A_c = P^T A P for piecewise-constant
                      P (scipy.sparse), coarse b = P^T r
  prolongation (A11)  closed form
  solve (A12-A14)     FP64 direct solve (scipy spsolve); residual norm in numpy
  G2P/advect (A16-17) trilinear reproduction of linear fields, FLIP/PIC limits,
                      reflection closed form
  voxelise            occupancy at scale 1 equals fluid_fraction > 0; density
                      mass balance
Ulp-level FP contraction inside the reference's CPU OpenCL runtime is not
pinned by any of these (no reference run exists); the HIP kernels are checked
against the oracle bit for bit in test_gpu_parity.py.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from continuum_liquid_amd import runtime, scenes

from .helpers import (F32, bin_keys, fma32, interior, make_context, operator_matrix, random_particles, run_to_pressure,
                      scene_from_arrays)


def random_scene(seed, shape=(10, 11, 9), n=4000, margin=0.0):
    rng = np.random.default_rng(seed)
    dx = 1.0 / shape[0]
    pos, vel = random_particles(rng, shape, dx, n, margin)
    return scene_from_arrays(f"random {shape}", shape, dx, pos, vel)


# ---- binning --------------------------------------------------------------------------------
@pytest.mark.parametrize("seed", [0, 1])
def test_binning_matches_closed_form(oracle_lib, seed):
    scene = random_scene(seed, n=20000)
    # add positions exactly on bin boundaries and outside the box
    dx = scene.dx
    edges = ((np.arange(0, 9, dtype=np.float64) + 0.5) * dx).astype(F32)
    extra = np.stack([edges, edges[::-1], edges], 1)
    outside = np.array([[-1, 0.5, 0.5], [2, 2, 2], [0.5, -0.01, 1.5]], F32)
    pos = np.concatenate([scene.positions, extra, outside])
    vel = np.random.default_rng(seed).normal(size=pos.shape).astype(F32)
    scene = scene_from_arrays("edges", scene.shape, dx, pos, vel)
    ctx = make_context(oracle_lib, scene)
    ctx.bin_particles()
    keys = bin_keys(pos, dx, scene.shape)
    counts = np.bincount(keys, minlength=int(np.prod(scene.shape)))
    assert np.array_equal(ctx.get_field("bin_counts").ravel(), counts)
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]])
    assert np.array_equal(ctx.get_field("bin_offsets").ravel(), offsets)
    order = np.argsort(keys, kind="stable")
    p, v = ctx.get_particles()
    assert np.array_equal(p, pos[order]) and np.array_equal(v, vel[order])


def test_bin_key_uses_fused_multiply_add(oracle_lib):
    """particles.cl:4 is contracted: the key is floor(fma(x, inv_dx, -0.5)).
    With dx = 1/100, x = 0.16499999f gives x*inv_dx = 16.5f after rounding
    but 16.4999... exactly, so the fused key is 15 and the unfused one 16.
    (Power-of-two spacings make the product exact and both keys agree.)"""
    n = 100
    dx = 1.0 / n
    x = F32(0.16499999)
    inv = F32(1.0 / dx)
    assert np.floor(F32(x * inv) - F32(0.5)) == 16 and np.floor(fma32(x, inv, F32(-0.5))) == 15
    pos = np.array([[x, 0.5, 0.5]], F32)
    ctx = make_context(oracle_lib, scene_from_arrays("fma key", (n, 8, 8), dx, pos, np.zeros_like(pos)))
    ctx.bin_particles()
    counts = ctx.get_field("bin_counts")
    assert counts[15].sum() == 1 and counts[16].sum() == 0


# ---- P2G -------------------------------------------------------------------------------------
def brute_force_p2g(pos, vel, shape, dx):
    """FP64 sums over all particles for every cell, reference weights."""
    inv = F32(1.0 / dx)
    p = (pos * inv).astype(np.float64)  # p = position * inv_dx in float32, as the kernel
    idx = np.indices(shape).reshape(3, -1).T.astype(np.float64)
    out = []
    for comp, off in enumerate(([1.0, 0.5, 0.5], [0.5, 1.0, 0.5], [0.5, 0.5, 1.0])):
        face = idx + np.asarray(off)
        w = np.prod(np.maximum(0.0, 1.0 - np.abs(p[None, :, :] - face[:, None, :])), axis=2)
        ws = w.sum(1)
        vs = (w * vel[None, :, comp].astype(np.float64)).sum(1)
        out.append(np.where(ws > 0, vs / np.where(ws > 0, ws, 1), 0.0).reshape(shape))
    cell = np.clip(np.trunc(pos * inv), 1, np.asarray(shape) - 2).astype(int)
    ff = np.zeros(shape)
    np.add.at(ff, tuple(cell.T), 1.0)
    return out, ff / 8.0


@pytest.mark.parametrize("seed", [0, 3])
def test_p2g_matches_brute_force(oracle_lib, seed):
    scene = random_scene(seed, shape=(9, 8, 10), n=2500)
    ctx = make_context(oracle_lib, scene)
    ctx.bin_particles()
    ctx.particle_to_grid(8.0)
    (u, v, w), ff = brute_force_p2g(scene.positions, scene.velocities, scene.shape, scene.dx)
    for name, ref in (("u", u), ("v", v), ("w", w)):
        got = ctx.get_field(name)
        scale = np.abs(ref).max()
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5 * scale, err_msg=name)
    assert np.array_equal(ctx.get_field("fluid_fraction"), ff.astype(F32))


# ---- grid glue -------------------------------------------------------------------------------
def _glue_context(oracle_lib, seed=5, shape=(9, 10, 8)):
    rng = np.random.default_rng(seed)
    scene = random_scene(seed, shape=shape, n=10)
    ctx = make_context(oracle_lib, scene)
    fields = {k: rng.normal(size=shape).astype(F32) for k in ("u", "v", "w")}
    ff = (rng.random(shape) < 0.6).astype(F32) * F32(0.5)
    for k, a in fields.items():
        ctx.set_field(k, a)
    ctx.set_field("fluid_fraction", ff)
    return ctx, fields, ff


def test_update_velocity_closed_form(oracle_lib):
    ctx, f, ff = _glue_context(oracle_lib)
    g, dt, nu, rho = (0.0, -9.81, 0.0), 1 / 240, 0.05, 1000.0
    ctx.update_velocity(g, (0.0, 2.0, 0.0), rho, nu, dt)
    inv2 = F32(1.0 / (ctx.dx * ctx.dx))
    acc = [F32(g[a]) + F32((0.0, 2.0, 0.0)[a]) / F32(rho) for a in range(3)]
    nudt = F32(nu) * F32(dt)
    inside = interior(ctx.shape)
    for a, name in enumerate("uvw"):
        x = f[name]
        r = np.where(ff > 0, fma32(F32(dt), acc[a], x), x).astype(F32)
        s = (np.roll(x, -1, 0) + np.roll(x, 1, 0)).astype(F32)
        for sh, ax in ((-1, 1), (1, 1), (-1, 2), (1, 2)):
            s = (s + np.roll(x, sh, ax)).astype(F32)
        lap = (fma32(F32(-6.0), x, s) * inv2).astype(F32)
        r = np.where(inside, fma32(nudt, lap, r), r).astype(F32)
        got = ctx.get_field(name)
        np.testing.assert_array_max_ulp(got, r, maxulp=1)
    # the inputs moved to the work buffers (swap of solver.py:374-376)
    assert np.array_equal(ctx.get_field("u_old"), f["u"])


def test_pressure_rhs_and_correction_closed_form(oracle_lib):
    ctx, f, ff = _glue_context(oracle_lib, seed=9)
    rho, dt = 1000.0, 1 / 240
    ctx.pressure_rhs(rho, dt)
    inv = F32(1.0 / ctx.dx)
    du = (f["u"] - np.roll(f["u"], 1, 0)).astype(F32)
    dv = (f["v"] - np.roll(f["v"], 1, 1)).astype(F32)
    dw = (f["w"] - np.roll(f["w"], 1, 2)).astype(F32)
    rhs = (F32(rho / dt) * (((du + dv).astype(F32) + dw).astype(F32) * inv)).astype(F32)
    rhs = np.where(interior(ctx.shape) & (ff > 0), rhs, 0).astype(F32)
    assert np.array_equal(ctx.get_field("rhs"), rhs)
    p = np.random.default_rng(2).normal(size=ctx.shape).astype(F32)
    ctx.set_field("pressure", p)
    ctx.pressure_correction(rho, dt)
    k = F32(dt / rho)
    for a, name in enumerate("uvw"):
        grad = ((np.roll(p, -1, a) - p).astype(F32) * inv).astype(F32)
        ref = np.where(interior(ctx.shape), fma32(-k, grad, f[name]), f[name]).astype(F32)
        np.testing.assert_array_max_ulp(ctx.get_field(name), ref, maxulp=1)


def finest_operator(mask):
    nx, ny, nz = mask.shape
    fluid = interior(mask.shape) & (mask > 0)
    i, j, k = np.indices(mask.shape)
    diag = 6.0 - (i == 1) - (i == nx - 2) - (j == 1) - (j == ny - 2) - (k == 1) - (k == nz - 2)
    diag = np.where(fluid, diag, 0).astype(F32)
    wx = np.where(fluid & (i < nx - 2) & (np.roll(mask, -1, 0) > 0), 1, 0).astype(F32)
    wy = np.where(fluid & (j < ny - 2) & (np.roll(mask, -1, 1) > 0), 1, 0).astype(F32)
    wz = np.where(fluid & (k < nz - 2) & (np.roll(mask, -1, 2) > 0), 1, 0).astype(F32)
    return diag, wx, wy, wz


def test_build_operator_closed_form(oracle_lib):
    ctx, _, ff = _glue_context(oracle_lib, seed=4)
    ctx.build_operator()
    for name, ref in zip(("diagonal", "weight_x", "weight_y", "weight_z"), finest_operator(ff)):
        assert np.array_equal(ctx.get_field(name), ref), name


def test_grid_domain_bc_and_neumann_closed_form(oracle_lib):
    ctx, f, _ = _glue_context(oracle_lib, seed=6)
    ctx.grid_domain_bc()
    nx, ny, nz = ctx.shape
    i, j, k = np.indices(ctx.shape)
    ci, cj, ck = np.clip(i, 1, nx - 2), np.clip(j, 1, ny - 2), np.clip(k, 1, nz - 2)
    ghost = ~interior(ctx.shape)
    walls = {"u": (i == 0) | (i >= nx - 2), "v": (j == 0) | (j >= ny - 2), "w": (k == 0) | (k >= nz - 2)}
    for name, wall in walls.items():
        x = f[name]
        ref = np.where(ghost, np.where(wall, 0, x[ci, cj, ck]), np.where(wall, 0, x)).astype(F32)
        assert np.array_equal(ctx.get_field(name), ref), name
    p = np.random.default_rng(1).normal(size=ctx.shape).astype(F32)
    ctx.set_field("pressure", p)
    ctx.neumann_bc()
    assert np.array_equal(ctx.get_field("pressure"), np.where(ghost, p[ci, cj, ck], p))


# ---- multigrid -------------------------------------------------------------------------------
def _pressure_context(oracle_lib, n=16, seed=0):
    scene = scenes.small_scene(n, seed=seed)
    ctx = make_context(oracle_lib, scene)
    run_to_pressure(ctx, scene)
    return ctx


def test_rbgs_sweep_matches_sequential_gauss_seidel(oracle_lib):
    ctx = _pressure_context(oracle_lib, 12)
    p0 = np.random.default_rng(0).normal(size=ctx.shape).astype(F32)
    ctx.set_field("pressure", p0)
    ctx.smooth(0, 1)
    mask = ctx.get_field("fluid_fraction")
    b = ctx.get_field("rhs").astype(np.float64)
    d, wx, wy, wz = (ctx.get_field(n).astype(np.float64) for n in ("diagonal", "weight_x", "weight_y", "weight_z"))
    p = np.where(mask > 0, p0, 0).astype(np.float64)
    h2 = np.float64(F32(ctx.dx)) ** 2
    for parity in (0, 1):
        for i, j, k in np.argwhere(interior(ctx.shape) & (mask > 0)):
            if (i + j + k) % 2 != parity or d[i, j, k] <= 0:
                continue
            s = (wx[i, j, k] * p[i + 1, j, k] + wx[i - 1, j, k] * p[i - 1, j, k] + wy[i, j, k] * p[i, j + 1, k]
                 + wy[i, j - 1, k] * p[i, j - 1, k] + wz[i, j, k] * p[i, j, k + 1] + wz[i, j, k - 1] * p[i, j, k - 1])
            p[i, j, k] = (s - h2 * b[i, j, k]) / d[i, j, k]
    got = ctx.get_field("pressure")
    np.testing.assert_allclose(got, p, rtol=1e-5, atol=1e-5 * np.abs(p).max())


def test_restriction_is_galerkin(oracle_lib):
    ctx = _pressure_context(oracle_lib, 16)
    p = np.random.default_rng(3).normal(size=ctx.shape).astype(F32)
    ctx.set_field("pressure", p)
    ctx.restrict(0)
    mask = ctx.get_field("fluid_fraction")
    fine = [ctx.get_field(n) for n in ("diagonal", "weight_x", "weight_y", "weight_z")]
    h2 = np.float64(F32(ctx.dx)) ** 2
    A, fluid, ids = operator_matrix(*fine, mask, h2)
    cshape = ctx.level_shapes[1]
    cmask = ctx.get_level_field(1, "mask")
    coarse = [ctx.get_level_field(1, n) for n in ("diagonal", "weight_x", "weight_y", "weight_z")]
    Ac, cfluid, cids = operator_matrix(*coarse, cmask, h2)
    # piecewise-constant prolongation P: fine fluid cell -> its parent
    fi = np.argwhere(fluid)
    parent = (fi - 1) // 2 + 1
    P = sp.csr_matrix((np.ones(len(fi)), (ids[tuple(fi.T)], cids[tuple(parent.T)])), shape=(fluid.sum(), cfluid.sum()))
    galerkin = (P.T @ A @ P).toarray()
    assert np.allclose(galerkin, Ac.toarray(), rtol=0, atol=1e-9 / h2)
    # coarse rhs = sum of child residuals, coarse p = 0
    pf = p.astype(np.float64)[fluid]
    r = ctx.get_field("rhs").astype(np.float64)[fluid] - A @ pf
    cb = ctx.get_level_field(1, "b").astype(np.float64)[cfluid]
    np.testing.assert_allclose(cb, P.T @ r, rtol=1e-5, atol=1e-5 * np.abs(cb).max())
    assert not ctx.get_level_field(1, "p").any()
    assert cmask.shape == cshape


def test_fully_fluid_block_coarse_stencil(oracle_lib):
    n = 14
    mask = np.zeros((n, n, n), F32)
    mask[1:-1, 1:-1, 1:-1] = 1
    scene = random_scene(0, shape=(n, n, n), n=1)
    ctx = make_context(oracle_lib, scene)
    ctx.set_field("fluid_fraction", mask)
    ctx.build_operator()
    ctx.restrict(0)
    d = ctx.get_level_field(1, "diagonal")
    wx = ctx.get_level_field(1, "weight_x")
    # an interior coarse cell away from walls: 8 children, diag 8*6 - 2*12 = 24, 4 couplings per face
    assert d[3, 3, 3] == 24 and wx[3, 3, 3] == 4
    assert ctx.get_level_field(1, "mask")[1:-1, 1:-1, 1:-1].all()


def test_prolongation_closed_form(oracle_lib):
    ctx = _pressure_context(oracle_lib, 16)
    ctx.restrict(0)
    pc = np.random.default_rng(8).normal(size=ctx.level_shapes[1]).astype(F32)
    ctx.set_level_field(1, "p", pc)
    pf = np.random.default_rng(9).normal(size=ctx.shape).astype(F32)
    ctx.set_field("pressure", pf)
    ctx.prolongate(0)
    mask = ctx.get_field("fluid_fraction")
    i, j, k = np.indices(ctx.shape)
    cn = np.asarray(ctx.level_shapes[1])
    a = np.minimum((i - 1) // 2 + 1, cn[0] - 2)
    b = np.minimum((j - 1) // 2 + 1, cn[1] - 2)
    c = np.minimum((k - 1) // 2 + 1, cn[2] - 2)
    ref = np.where(interior(ctx.shape) & (mask > 0), (pf + pc[a, b, c]).astype(F32), pf)
    assert np.array_equal(ctx.get_field("pressure"), ref)


def test_solve_matches_fp64_direct_solve(oracle_lib):
    ctx = _pressure_context(oracle_lib, 16, seed=2)
    stats = ctx.pressure_solve(runtime.SOLVE_FP32_TOLERANCE, post_smooth=2, max_cycles=60, tolerance=1e-7)
    mask = ctx.get_field("fluid_fraction")
    h2 = np.float64(F32(ctx.dx)) ** 2
    A, fluid, _ = operator_matrix(*[ctx.get_field(n) for n in ("diagonal", "weight_x", "weight_y", "weight_z")],
                                  mask, h2)
    b = ctx.get_field("rhs").astype(np.float64)[fluid]
    exact = spla.spsolve(A.tocsc(), b)
    p = ctx.get_field("pressure").astype(np.float64)[fluid]
    assert np.linalg.norm(p - exact) <= 1e-4 * np.linalg.norm(exact)
    # the FP64 residual norm of the library equals numpy's on the final iterate
    rr, bb = np.linalg.norm(b - A @ p), np.linalg.norm(b)
    assert stats["final_residual"] == pytest.approx(rr, rel=1e-9)
    assert stats["rhs_norm"] == pytest.approx(bb, rel=1e-12)
    assert stats["relative_residual"] < 1e-5


def test_w_cycle_contracts(oracle_lib):
    ctx = _pressure_context(oracle_lib, 24)
    r0, b = ctx.residual_norm()
    rates = []
    for _ in range(6):
        ctx.w_cycle(2, 2, 64)
        r1, _ = ctx.residual_norm()
        rates.append(r1 / r0)
        r0 = r1
    assert max(rates) < 0.6, rates


# ---- particles -------------------------------------------------------------------------------
def _linear_field(shape, coef, offset):
    idx = np.indices(shape).astype(np.float64)
    pos = [idx[a] + offset[a] for a in range(3)]  # face positions in grid units
    return (coef[0] + coef[1] * pos[0] + coef[2] * pos[1] + coef[3] * pos[2]).astype(F32)


@pytest.mark.parametrize("flip", [1.0, 0.0, 0.95])
def test_g2p_reproduces_linear_fields(oracle_lib, flip):
    shape = (10, 9, 11)
    scene = random_scene(2, shape=shape, n=3000, margin=1.6)
    ctx = make_context(oracle_lib, scene)
    coefs = {"u": (0.3, 0.1, -0.2, 0.05), "v": (-0.1, 0.02, 0.07, -0.3), "w": (0.2, -0.04, 0.0, 0.11)}
    offs = {"u": (1.0, 0.5, 0.5), "v": (0.5, 1.0, 0.5), "w": (0.5, 0.5, 1.0)}
    for name in "uvw":
        ctx.set_field(name, _linear_field(shape, coefs[name], offs[name]))
        ctx.set_field(name + "_old", np.zeros(shape, F32))
    ctx.grid_to_particle(flip)
    _, vel = ctx.get_particles()
    g = scene.positions.astype(np.float64) / scene.dx
    for a, name in enumerate("uvw"):
        c = coefs[name]
        exact = c[0] + c[1] * g[:, 0] + c[2] * g[:, 1] + c[3] * g[:, 2]
        v0 = scene.velocities[:, a].astype(np.float64)
        expect = flip * (v0 + exact) + (1 - flip) * exact
        np.testing.assert_allclose(vel[:, a], expect, rtol=0, atol=2e-5)


def test_advect_and_domain_bc_closed_form(oracle_lib):
    shape = (12, 12, 12)
    scene = random_scene(4, shape=shape, n=5000)
    ctx = make_context(oracle_lib, scene)
    dt = 0.05
    ctx.advect_particles(dt)
    ctx.particle_domain_bc()
    pos, vel = ctx.get_particles()
    x = fma32(F32(dt), scene.velocities, scene.positions)
    lo = F32(1.001 * scene.dx)
    hi = np.array([F32((n - 1.001) * scene.dx) for n in shape], F32)
    hit_lo, hit_hi = x < lo, x > hi
    y = np.where(hit_lo, (F32(2) * lo - x).astype(F32), x)
    y = np.where(hit_hi, (F32(2) * hi - y).astype(F32), y)
    y = np.minimum(np.maximum(y, lo), hi).astype(F32)
    assert np.array_equal(pos, y)
    assert np.array_equal(vel, np.where(hit_lo | hit_hi, 0, scene.velocities).astype(F32))


def test_obstacle_bc_moves_particles_out(oracle_lib):
    shape = (12, 12, 12)
    dx = 1 / 12
    obstacle = np.zeros(shape, np.uint8)
    obstacle[5:8, 5:8, 5:8] = 1
    prev = np.array([[4.5, 6.5, 6.5], [6.5, 6.5, 8.5]], F32) * F32(dx)  # outside, next to -x / +z faces
    now = np.array([[5.4, 6.5, 6.5], [6.5, 6.5, 7.6]], F32) * F32(dx)  # moved into the solid
    vel = np.array([[1.0, 0.3, 0.0], [0.2, 0.0, -1.0]], F32)
    scene = scene_from_arrays("obstacle", shape, dx, prev, vel)
    ctx = make_context(oracle_lib, scene)
    ctx.set_field("obstacle", obstacle)
    ctx.save_previous_positions()
    ctx.set_particles(now, vel)
    # keep the saved previous positions: set_particles does not touch them
    ctx.particle_obstacle_bc()
    pos, v = ctx.get_particles()
    cells = np.floor(pos / dx).astype(int)
    assert not obstacle[tuple(cells.T)].any()
    assert v[0, 0] == 0 and v[0, 1] == F32(0.3)  # inward x velocity removed, tangential kept
    assert v[1, 2] == 0 and v[1, 0] == F32(0.2)


def test_voxelise_occupancy_and_density(oracle_lib):
    scene = scenes.small_scene(16, seed=4)
    ctx = make_context(oracle_lib, scene)
    ctx.bin_particles()
    ctx.particle_to_grid(8.0)
    occ, dens = ctx.voxelise(1, 8.0, density=True)
    assert np.array_equal(occ.astype(bool), ctx.get_field("fluid_fraction") > 0)
    assert dens.sum() == pytest.approx(scene.particle_count / 8.0, rel=1e-5)
    occ2, dens2 = ctx.voxelise(2, 8.0, density=True)
    assert occ2.shape == tuple(2 * n for n in scene.shape)
    assert dens2.sum() == pytest.approx(scene.particle_count, rel=1e-5)
    # a refined cell is occupied only inside an occupied coarse cell
    coarse = occ2.reshape(16, 2, 16, 2, 16, 2).max(axis=(1, 3, 5))
    assert not (coarse & ~occ.astype(bool)).any()


@pytest.mark.parametrize("coarse_dtype", [0, 1], ids=["fp16", "bf16"])
def test_mixed_refinement_reaches_1e8_against_fp64_direct_solve(oracle_lib, coarse_dtype):
    """Mode 2 (FP64 refinement around the FP32/16-bit W-cycle) passes the FP32
    floor: 1e-8 relative residual, FP64-checked, and the iterate equals the
    scipy FP64 solution to that accuracy."""
    ctx = _pressure_context(oracle_lib, 16, seed=2)
    fp32 = _pressure_context(oracle_lib, 16, seed=2)
    s32 = fp32.pressure_solve(runtime.SOLVE_FP32_TOLERANCE, post_smooth=2, max_cycles=80, tolerance=1e-10)
    stats = ctx.pressure_solve(runtime.SOLVE_MIXED_REFINEMENT, post_smooth=2, max_cycles=80, tolerance=1e-8,
                               coarse_dtype=coarse_dtype)
    assert s32["relative_residual"] > 1e-7  # FP32 storage stalls above the target
    assert stats["relative_residual"] <= 1e-8
    mask = ctx.get_field("fluid_fraction")
    h2 = np.float64(F32(ctx.dx)) ** 2
    A, fluid, _ = operator_matrix(*[ctx.get_field(n) for n in ("diagonal", "weight_x", "weight_y", "weight_z")],
                                  mask, h2)
    b = ctx.get_field("rhs").astype(np.float64)[fluid]
    exact = spla.spsolve(A.tocsc(), b)
    p = ctx.get_field("pressure").astype(np.float64)[fluid]
    # the returned FP32 pressure is the FP64 iterate rounded once
    assert np.linalg.norm(p - exact) <= 1e-6 * np.linalg.norm(exact)


def test_coarse_weight_scales_every_coarse_correction(oracle_lib):
    """A fixed W-cycle with coarse_weight w on a three-level hierarchy equals
    the reference's stages with each coarse pressure multiplied by w just
    before its prolongation.  w = 0.5 makes that product exact, so the two
    must agree bit for bit; a weight applied to the fine pressure, to the
    sum, or to only one of the two prolongations would not."""
    pre, post, extra, w = 2, 3, 16, 0.5
    ref = _pressure_context(oracle_lib, 24, seed=4)
    last = len(ref.level_shapes) - 1
    assert last == 2

    def cycle(level):
        if level == last:
            ref.coarse_smooth(pre + extra)
            return
        ref.smooth(level, pre)
        ref.restrict(level)
        cycle(level + 1)
        cycle(level + 1)
        ref.set_level_field(level + 1, "p", ref.get_level_field(level + 1, "p") * F32(w))
        ref.prolongate(level)
        ref.smooth(level, post)

    cycle(0)
    ref.neumann_bc()
    ctx = _pressure_context(oracle_lib, 24, seed=4)
    ctx.pressure_solve(runtime.SOLVE_FIXED_CYCLES, pre_smooth=pre, post_smooth=post, coarse_extra=extra,
                       max_cycles=1, coarse_weight=w)
    assert np.array_equal(ctx.get_field("pressure"), ref.get_field("pressure"))
    # weight 0 selects 1, the plain sum of the reference
    plain = [_pressure_context(oracle_lib, 24, seed=4) for _ in range(2)]
    plain[0].pressure_solve(runtime.SOLVE_FIXED_CYCLES, max_cycles=1)
    plain[1].pressure_solve(runtime.SOLVE_FIXED_CYCLES, max_cycles=1, coarse_weight=0.0)
    assert np.array_equal(plain[0].get_field("pressure"), plain[1].get_field("pressure"))
    assert not np.array_equal(plain[0].get_field("pressure"), ctx.get_field("pressure"))
    # the stage entry points after a weighted solve keep the plain sum
    for c in (ctx, plain[0]):
        c.set_field("pressure", np.zeros(c.shape, F32))
        c.w_cycle(pre, post, extra)
    assert np.array_equal(ctx.get_field("pressure"), plain[0].get_field("pressure"))


@pytest.mark.parametrize("w", [2.0, -0.5, float("nan")])
def test_coarse_weight_outside_range_is_rejected(oracle_lib, w):
    ctx = _pressure_context(oracle_lib, 12)
    with pytest.raises(runtime.ClqError):
        ctx.pressure_solve(runtime.SOLVE_FCG, max_cycles=4, coarse_weight=w)


def test_coarse_weight_cuts_fcg_iterations(oracle_lib):
    """Piecewise-constant prolongation under-corrects smooth error; weighting
    the coarse corrections by 1.6 makes the FCG preconditioner stronger: the
    same tolerance in fewer iterations, to the scipy FP64 solution."""
    kw = dict(pre_smooth=1, post_smooth=1, coarse_extra=16, max_cycles=60, tolerance=1e-8, vcycle_from=2)
    runs = {}
    for w in (1.0, 1.6):
        ctx = _pressure_context(oracle_lib, 48, seed=1)
        runs[w] = (ctx, ctx.pressure_solve(runtime.SOLVE_FCG, coarse_weight=w, **kw))
    ctx, s = runs[1.6]
    assert s["relative_residual"] <= 1e-8
    assert s["cycles"] <= 0.8 * runs[1.0][1]["cycles"]
    mask = ctx.get_field("fluid_fraction")
    h2 = np.float64(F32(ctx.dx)) ** 2
    A, fluid, _ = operator_matrix(*[ctx.get_field(n) for n in ("diagonal", "weight_x", "weight_y", "weight_z")],
                                  mask, h2)
    b = ctx.get_field("rhs").astype(np.float64)[fluid]
    p64 = ctx.get_field("pressure_fp64")[fluid]
    assert np.linalg.norm(b - A @ p64) <= 1.0000001e-8 * np.linalg.norm(b)


def test_cycle_level_mask_matches_level_bound(oracle_lib):
    """vcycle_from = -m runs the coarse-grid correction twice below the levels
    whose bit is set in m: over three refinement cycles -1 and -3 give the
    cycles of vcycle_from 1 and 2 bit for bit, -2 (twice below level 1 only,
    once below level 0) a cycle of its own; FCG around that cycle, weighted by
    1.6, reaches the tolerance, to the scipy FP64 solution."""
    kw = dict(pre_smooth=1, post_smooth=1, coarse_extra=16)

    def refine(vcycle_from):
        ctx = _pressure_context(oracle_lib, 48, seed=1)
        s = ctx.pressure_solve(runtime.SOLVE_MIXED_REFINEMENT, vcycle_from=vcycle_from, max_cycles=3,
                               tolerance=1e-14, **kw)
        assert s["cycles"] == 3
        return ctx.get_field("pressure_fp64")

    p = {v: refine(v) for v in (1, -1, 2, -3, -2)}
    assert np.array_equal(p[-1], p[1]) and np.array_equal(p[-3], p[2])
    assert not np.array_equal(p[-2], p[2]) and not np.array_equal(p[-2], p[1])
    ctx = _pressure_context(oracle_lib, 48, seed=1)
    assert len(ctx.level_shapes) >= 4  # levels 0-3: the masks differ in what they run
    s = ctx.pressure_solve(runtime.SOLVE_FCG, vcycle_from=-2, max_cycles=60, tolerance=1e-8, coarse_weight=1.6, **kw)
    assert s["relative_residual"] <= 1e-8
    mask = ctx.get_field("fluid_fraction")
    h2 = np.float64(F32(ctx.dx)) ** 2
    A, fluid, _ = operator_matrix(*[ctx.get_field(n) for n in ("diagonal", "weight_x", "weight_y", "weight_z")],
                                  mask, h2)
    b = ctx.get_field("rhs").astype(np.float64)[fluid]
    p64 = ctx.get_field("pressure_fp64")[fluid]
    assert np.linalg.norm(b - A @ p64) <= 1.0000001e-8 * np.linalg.norm(b)


def test_weighted_cycle_breakdown_falls_back_to_weight_one(oracle_lib):
    """With the V-cycle below level 1 a weight of 1.9 over-corrects the
    inexact coarse solves until the preconditioner is indefinite: a step
    length alpha <= 0 appears within a few iterations.  The solve then goes on
    with weight 1 (restarting the directions) and still reaches the
    tolerance; without that switch this case stalls far above it."""
    kw = dict(pre_smooth=1, post_smooth=1, coarse_extra=16, max_cycles=60, tolerance=1e-8, vcycle_from=1)
    ctx = _pressure_context(oracle_lib, 48, seed=0)
    s = ctx.pressure_solve(runtime.SOLVE_FCG, coarse_weight=1.9, **kw)
    assert 0 < s["fallback_cycles"] < s["cycles"] < 30
    assert s["relative_residual"] <= 1e-8
    mask = ctx.get_field("fluid_fraction")
    h2 = np.float64(F32(ctx.dx)) ** 2
    A, fluid, _ = operator_matrix(*[ctx.get_field(n) for n in ("diagonal", "weight_x", "weight_y", "weight_z")],
                                  mask, h2)
    b = ctx.get_field("rhs").astype(np.float64)[fluid]
    p64 = ctx.get_field("pressure_fp64")[fluid]
    assert np.linalg.norm(b - A @ p64) <= 1.0000001e-8 * np.linalg.norm(b)


def test_fcg_reaches_tolerance_in_fewer_cycles(oracle_lib):
    """Mode 4 (FP64 flexible CG around the mixed-precision W-cycle) reaches the
    target with fewer W-cycles than refinement alone and matches the scipy
    FP64 solution."""
    kw = dict(pre_smooth=2, post_smooth=2, coarse_extra=16, max_cycles=80, tolerance=1e-9)
    ir = _pressure_context(oracle_lib, 16, seed=2)
    s_ir = ir.pressure_solve(runtime.SOLVE_MIXED_REFINEMENT, **kw)
    ctx = _pressure_context(oracle_lib, 16, seed=2)
    s = ctx.pressure_solve(runtime.SOLVE_FCG, **kw)
    assert s["relative_residual"] <= 1e-9
    assert s["cycles"] < s_ir["cycles"]
    mask = ctx.get_field("fluid_fraction")
    h2 = np.float64(F32(ctx.dx)) ** 2
    A, fluid, _ = operator_matrix(*[ctx.get_field(n) for n in ("diagonal", "weight_x", "weight_y", "weight_z")],
                                  mask, h2)
    b = ctx.get_field("rhs").astype(np.float64)[fluid]
    exact = spla.spsolve(A.tocsc(), b)
    p = ctx.get_field("pressure").astype(np.float64)[fluid]
    assert np.linalg.norm(p - exact) <= 1e-6 * np.linalg.norm(exact)
    p64 = ctx.get_field("pressure_fp64")[fluid]
    assert s["final_residual"] == pytest.approx(np.linalg.norm(b - A @ p64), rel=1e-6)


def test_fp64_check_solve_matches_direct_solve(oracle_lib):
    """The check solve (Jacobi-preconditioned CG in FP64) returns the scipy
    direct solution, and measures the error of an FCG solve against it: at a
    1e-6 FCG tolerance the error is small but not zero, at 1e-11 it vanishes."""
    ctx = _pressure_context(oracle_lib, 20, seed=3)
    h2 = np.float64(F32(ctx.dx)) ** 2
    A, fluid, _ = operator_matrix(*[ctx.get_field(n) for n in ("diagonal", "weight_x", "weight_y", "weight_z")],
                                  ctx.get_field("fluid_fraction"), h2)
    exact = spla.spsolve(A.tocsc(), ctx.get_field("rhs").astype(np.float64)[fluid])
    loose = ctx.pressure_solve(runtime.SOLVE_FCG, pre_smooth=1, post_smooth=1, coarse_extra=16, max_cycles=60,
                               tolerance=1e-6)
    check = ctx.fp64_check_solve(1e-12)
    assert check["converged"] and check["relative_residual"] <= 1e-12 and check["iterations"] > 10
    p = ctx.get_field("pressure_fp64")[fluid]
    want_l2 = np.linalg.norm(p - exact) / np.linalg.norm(exact)
    assert check["error_l2"] == pytest.approx(want_l2, rel=1e-3, abs=1e-12)
    assert check["error_max"] == pytest.approx(np.abs(p - exact).max() / np.abs(exact).max(), rel=1e-3, abs=1e-12)
    assert 1e-12 < check["error_l2"] < 1e-3 and loose["relative_residual"] <= 1e-6
    tight = _pressure_context(oracle_lib, 20, seed=3)
    tight.pressure_solve(runtime.SOLVE_FCG, pre_smooth=1, post_smooth=1, coarse_extra=16, max_cycles=100,
                         tolerance=1e-11)
    assert tight.fp64_check_solve(1e-12)["error_l2"] < 1e-8


def test_fcg_reports_the_recomputed_residual(oracle_lib):
    """Modes 2-4 report ||b - A x|| recomputed from the FP64 iterate (not the
    recurrence), checked here against numpy on the assembled operator."""
    ctx = _pressure_context(oracle_lib, 18, seed=6)
    h2 = np.float64(F32(ctx.dx)) ** 2
    A, fluid, _ = operator_matrix(*[ctx.get_field(n) for n in ("diagonal", "weight_x", "weight_y", "weight_z")],
                                  ctx.get_field("fluid_fraction"), h2)
    b = ctx.get_field("rhs").astype(np.float64)[fluid]
    stats = ctx.pressure_solve(runtime.SOLVE_FCG, pre_smooth=1, post_smooth=1, coarse_extra=16, max_cycles=60,
                               tolerance=1e-7, coarse_weight=1.6)
    x = ctx.get_field("pressure_fp64")[fluid]
    true = np.linalg.norm(b - A @ x)
    assert stats["final_residual"] == pytest.approx(true, rel=1e-6)
    assert stats["relative_residual"] <= 1e-7
    assert stats["recurrence_residual"] > 0 and stats["restarts"] >= 0


@pytest.mark.parametrize("mode", ["fixed", "fp32", "mixed", "fcg"])
def test_non_finite_solve_is_an_error(oracle_lib, mode):
    """A NaN in the right-hand side ends every solve mode with
    CLQ_ERR_NOT_CONVERGED instead of a silent NaN pressure."""
    ctx = _pressure_context(oracle_lib, 12, seed=1)
    # a NaN in an interior fluid cell (the system's rows)
    ff = ctx.get_field("fluid_fraction")
    cell = tuple(np.argwhere(ff[1:-1, 1:-1, 1:-1] > 0)[0] + 1)
    rhs = ctx.get_field("rhs")
    rhs[cell] = np.nan
    ctx.set_field("rhs", rhs)
    with pytest.raises(runtime.ClqError) as err:
        ctx.pressure_solve(runtime.SOLVE_MODES[mode], max_cycles=5, pre_smooth=1, post_smooth=1, coarse_extra=4)
    assert err.value.status == runtime.ERR_NOT_CONVERGED and "not finite" in str(err.value)


def test_binning_crowded_bin_matches_stable_sort(oracle_lib):
    """One bin holding 50,000 particles (plus NaN positions, which clamp to
    bin 0) is sorted stably like any other: the order is numpy's stable
    argsort of the closed-form keys, compared bit for bit."""
    rng = np.random.default_rng(12)
    shape, dx = (12, 10, 14), 1.0 / 12
    spread = (rng.random((3000, 3)) * np.array(shape) * dx).astype(F32)
    crowd = (np.full((50000, 3), 5.7, np.float32) * F32(dx) + rng.uniform(-0.1, 0.1, (50000, 3)).astype(F32) * F32(dx))
    nan = np.full((40, 3), np.nan, np.float32)
    pos = np.concatenate([spread[:1000], crowd, nan, spread[1000:]])
    order = rng.permutation(len(pos))
    pos = np.ascontiguousarray(pos[order])
    vel = rng.standard_normal(pos.shape).astype(F32)
    ctx = make_context(oracle_lib, scene_from_arrays("crowded", shape, dx, pos, vel))
    ctx.bin_particles()
    keys = bin_keys(np.nan_to_num(pos, nan=-1.0), dx, shape)
    want = np.argsort(keys, kind="stable")
    got_p, got_v = ctx.get_particles()
    assert np.array_equal(got_p.view(np.uint32), pos[want].view(np.uint32))
    assert np.array_equal(got_v, vel[want])
    counts = ctx.get_field("bin_counts").ravel()
    assert counts.max() >= 50000 and np.array_equal(counts, np.bincount(keys, minlength=counts.size))
