"""Host logic of the kept Python API, on the CPU.

`Solver.General.main` and `Solver.Kernel.solver` load the HIP library in
production; here `runtime.load_library` is replaced, inside each test only,
by the CPU oracle so that the whole host loop (config parsing, seeding,
output cadence, cache format, messages, cancellation, errors) runs without a
GPU.  Reference behaviour followed: Solver/General/main.py:13-52,
Solver/Kernel/solver.py:22-531, Solver/Kernel/output.py:19-53.
"""

from __future__ import annotations

import json
import threading

import numpy as np
import pytest

from continuum_liquid_amd import runtime
from continuum_liquid_amd.Solver.General import main as solver_main
from continuum_liquid_amd.Solver.Kernel import output, solver as host_solver


@pytest.fixture
def oracle_backend(monkeypatch, oracle_lib):
    monkeypatch.setattr(runtime, "load_library", lambda path=None: oracle_lib)
    return oracle_lib


def make_config(tmp_path, frames=3, substeps=1, cancel=None, **settings):
    n = 14
    cfg = {
        "simulation": {
            "output_path": str(tmp_path / "bake"),
            "settings": {"solver_backend": "GPU", "start_frame": 1, "end_frame": 1 + frames, "fps": 24,
                         "substeps": substeps, "iterations": 1, **settings},
            "domain": {"resolution": 1.0 / n, "size": {"x": 1.0, "y": 1.0, "z": 1.0}},
            "physics": {"water": {"density": 1000.0, "viscosity": 1e-6},
                        "air": {"density": 1.225, "viscosity": 1.5e-5}},
            "sources": [{"type": "GEOMETRY", "particle_density": 8, "velocity": [0.0, 0.0, 0.0],
                         "shapes": [{"box": [[0.08, 0.08, 0.08], [0.45, 0.9, 0.5]]}]}],
            "obstacles": [{"shapes": [{"box": [[0.6, 0.0, 0.0], [0.75, 1.0, 0.2]]}]}],
        },
        "meta": {"cancel_event": cancel},
    }
    return cfg


def test_main_runs_and_writes_the_viewer_cache(tmp_path, oracle_backend, capsys):
    messages = []
    solver_main.set_message_callback(messages.append)
    try:
        solver_main.main(make_config(tmp_path, frames=3))
    finally:
        solver_main.clear_message_callback()
    out = capsys.readouterr().out
    assert "Bake runtime:" in out
    meta = json.loads((tmp_path / "bake" / output.CACHE_FILENAME).read_text())
    assert meta == {"version": 2, "fps": 24, "start_frame": 1, "end_frame": 4, "variable_particle_count": True,
                    "dtype": "float32", "components": 3, "frame_directory": "particles",
                    "frame_pattern": "particles_{frame:06d}.bin"}
    frames = sorted((tmp_path / "bake" / "particles").glob("*.bin"))
    assert [f.name for f in frames] == [f"particles_{i:06d}.bin" for i in range(1, 5)]
    first = np.fromfile(frames[0], np.float32).reshape(-1, 3)
    last = np.fromfile(frames[-1], np.float32).reshape(-1, 3)
    assert len(first) == len(last) > 0
    assert last[:, 2].mean() < first[:, 2].mean()  # gravity on z pulls the block down
    assert not list((tmp_path / "bake" / "particles").glob("*.tmp"))
    assert [m["frame"] for m in messages] == [0, 1, 2, 3]
    assert all(m["type"] == "stats" and m["total_frames"] == 3 for m in messages)


def test_cancel_event_stops_cleanly(tmp_path, oracle_backend):
    cancel = threading.Event()
    seen = []

    def on_message(msg):
        seen.append(msg)
        if msg["frame"] == 1:
            cancel.set()

    solver_main.set_message_callback(on_message)
    try:
        solver_main.main(make_config(tmp_path, frames=5, cancel=cancel))
    finally:
        solver_main.clear_message_callback()
    assert [m["frame"] for m in seen] == [0, 1]
    assert len(list((tmp_path / "bake" / "particles").glob("*.bin"))) == 2


def test_debug_timing_reports_reference_sections(tmp_path, oracle_backend, capsys):
    solver_main.main(make_config(tmp_path, frames=1, debug_timing=True))
    out = capsys.readouterr().out
    for section in ("particle_to_grid", "pressure_solve", "grid_to_particle", "advect_particles",
                    "particle_domain_bc", "particle_obstacle_bc", "particles_to_host"):
        assert section in out


def test_cpu_backend_runs_the_same_step(tmp_path, oracle_backend):
    # the add-on's default backend string ("CPU", also assumed when the key is
    # missing) starts the same solver as "GPU" (reference main.py:41-46)
    frames = {}
    for backend in ("GPU", "cpu", None):
        cfg = make_config(tmp_path / str(backend), frames=2)
        if backend is None:
            del cfg["simulation"]["settings"]["solver_backend"]
        else:
            cfg["simulation"]["settings"]["solver_backend"] = backend
        solver_main.main(cfg)
        files = sorted((tmp_path / str(backend) / "bake" / "particles").glob("*.bin"))
        frames[backend] = [f.read_bytes() for f in files]
    assert len(frames["GPU"]) == 3
    assert frames["cpu"] == frames["GPU"] == frames[None]
    cfg = make_config(tmp_path / "other")
    cfg["simulation"]["settings"]["solver_backend"] = "VULKAN"
    with pytest.raises(ValueError, match="Unsupported solver backend"):
        solver_main.main(cfg)


def test_usd_geometry_needs_pxr(tmp_path, oracle_backend):
    cfg = make_config(tmp_path)
    cfg["simulation"]["sources"] = [{"type": "GEOMETRY", "geometry_inputs": [{"mesh_file": "geometry/a.usdc"}]}]
    with pytest.raises(NotImplementedError, match="pxr"):
        solver_main.main(cfg)


def _seed_by_loops(positions, velocities, sources, solid, delta, rng, initial):
    """The seeding rules restated with explicit loops over particles and cells
    (reference source_particles.py:21-92): sinks first, then each filling
    source tops its open cells up, cells in C order, one uniform draw per
    source."""
    shape = solid.shape
    pos = [tuple(p) for p in np.asarray(positions, np.float32)]
    vel = [tuple(v) for v in np.asarray(velocities, np.float32)]

    def cell(p):
        return tuple(min(max(int(np.floor(np.float32(c) / np.float32(delta))), 0), n - 1) for c, n in zip(p, shape))

    for src in sources:
        if src["type"] == "OUTFLOW":
            kept = [(p, v) for p, v in zip(pos, vel) if not src["mask"][cell(p)]]
            pos, vel = [p for p, _ in kept], [v for _, v in kept]
    for src in sources:
        if src["type"] == "INFLOW" or (initial and src["type"] == "GEOMETRY"):
            target = max(1, int(src.get("particle_density", 8)))
            counts = {}
            for p in pos:
                counts[cell(p)] = counts.get(cell(p), 0) + 1
            spawn = []
            for i in range(shape[0]):
                for j in range(shape[1]):
                    for k in range(shape[2]):
                        if src["mask"][i, j, k] and not solid[i, j, k]:
                            spawn += [(i, j, k)] * max(0, target - counts.get((i, j, k), 0))
            if not spawn:
                continue
            draw = rng.random((len(spawn), 3))
            for c, r in zip(spawn, draw):
                pos.append(tuple(np.float32((c[a] + r[a]) * delta) for a in range(3)))
                vel.append(tuple(np.float32(x) for x in src.get("velocity", (0.0, 0.0, 0.0))))
    return np.array(pos, np.float32).reshape(-1, 3), np.array(vel, np.float32).reshape(-1, 3)


def test_source_seeding_matches_the_rules():
    from continuum_liquid_amd.Solver.Kernel import source_particles as sp

    shape, delta = (6, 5, 7), 1.0 / 6
    r = np.random.default_rng(3)
    solid = r.random(shape) < 0.15
    box = np.zeros(shape, bool)
    box[1:4, 1:3, 2:6] = True
    inflow = np.zeros(shape, bool)
    inflow[4:6, 0:2, 0:3] = True
    sink = np.zeros(shape, bool)
    sink[0:3, 3:5, :] = True
    sources = [{"type": "GEOMETRY", "mask": box, "particle_density": 5, "velocity": (0.5, -1.0, 0.25)},
               {"type": "OUTFLOW", "mask": sink},
               {"type": "INFLOW", "mask": inflow, "particle_density": 3, "velocity": (0.0, 0.0, -2.0)},
               {"type": "INFLOW", "mask": np.zeros(shape, bool)}]
    # a few particles already inside the inflow, the sink and the box
    pos = (r.random((40, 3)) * np.array(shape) * delta).astype(np.float32)
    vel = r.standard_normal((40, 3)).astype(np.float32)
    for initial in (True, False):
        got = sp.apply_sources(pos, vel, sources, solid, delta, np.random.default_rng(7), initial=initial)
        want = _seed_by_loops(pos, vel, sources, solid, delta, np.random.default_rng(7), initial)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        assert got[0].dtype == np.float32 and got[1].dtype == np.float32
        assert len(got[0]) > len(pos) - 12
    # a second inflow call adds nothing and draws nothing: the generator stays put
    once = sp.apply_sources(pos, vel, sources, solid, delta, np.random.default_rng(7))
    rng = np.random.default_rng(11)
    again = sp.apply_sources(once[0], once[1], sources[2:], solid, delta, rng)
    assert np.array_equal(again[0], once[0])
    assert rng.random() == np.random.default_rng(11).random()


def test_stage_errors_propagate(tmp_path, oracle_backend):
    cfg = make_config(tmp_path, pressure_solver={"mode": "learned"})
    with pytest.raises(runtime.ClqError, match="solve mode"):
        solver_main.main(cfg)


def test_writer_errors_surface(tmp_path):
    writer = output.ParticleOutput(tmp_path / "missing" / "dir")
    writer.submit(np.zeros((4, 3), np.float32), 1)
    with pytest.raises(RuntimeError, match="writer failed"):
        writer.close()


def test_message_callback_is_thread_safe():
    got = []
    solver_main.set_message_callback(got.append)
    threads = [threading.Thread(target=solver_main.emit_message, args=({"i": i},)) for i in range(16)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    solver_main.clear_message_callback()
    solver_main.emit_message({"i": -1})
    assert sorted(m["i"] for m in got) == list(range(16))


def test_fused_and_sectioned_steps_agree(oracle_lib):
    from continuum_liquid_amd import scenes
    from continuum_liquid_amd.Solver.Kernel import timing

    scene = scenes.small_scene(16, seed=5)
    params = host_solver.StepParams(dt=1 / 240, gravity=scene.gravity, flip_ratio=0.95,
                                    solver={"mode": "fixed", "max_cycles": 1})
    out = []
    for timings in (None, timing.RunTimings()):
        with runtime.Context(oracle_lib, scene.shape, scene.dx, scene.particle_count) as ctx:
            ctx.set_particles(scene.positions, scene.velocities)
            host_solver.step(ctx, params, timings)
            out.append(ctx.get_particles())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def _source_scene(shape=(14, 12, 16), delta=1.0 / 14, seed=4):
    r = np.random.default_rng(seed)
    solid = np.zeros(shape, bool)
    solid[9:12, 0:4, 3:9] = True
    inflow = np.zeros(shape, bool)
    inflow[1:4, 8:11, 2:6] = True
    inflow[10, 1, 4] = True  # partly solid
    sink = np.zeros(shape, bool)
    sink[5:9, 0:3, :] = True
    sources = [{"type": "INFLOW", "mask": inflow, "particle_density": 6, "velocity": (0.5, -2.0, 0.25)},
               {"type": "OUTFLOW", "mask": sink},
               {"type": "INFLOW", "mask": sink[::-1].copy(), "particle_density": 3, "velocity": (0.0, 1.0, 0.0)},
               {"type": "GEOMETRY", "mask": np.roll(inflow, 5, axis=0), "particle_density": 8}]
    pos = (r.random((3000, 3)) * np.array(shape) * delta).astype(np.float32)
    vel = r.standard_normal((3000, 3)).astype(np.float32)
    return shape, delta, solid, sources, pos, vel


@pytest.mark.parametrize("initial", [False, True])
def test_device_sources_match_host_seeding(oracle_lib, initial):
    """clq_apply_sources (here the oracle's restatement) removes and emits the
    same particles, bit for bit, as source_particles.apply_sources with the
    same numpy generator, and leaves the generator in the same state."""
    from continuum_liquid_amd.Solver.Kernel import source_particles as sp

    shape, delta, solid, sources, pos, vel = _source_scene()
    host_rng, dev_rng = np.random.default_rng(9), np.random.default_rng(9)
    hp, hv = pos, vel
    with runtime.Context(oracle_lib, shape, delta, capacity=100) as ctx:  # small: grows while emitting
        ctx.set_field("obstacle", solid.astype(np.uint8))
        ctx.set_sources(sources)
        ctx.set_particles(pos, vel)
        for step in range(3):
            hp, hv = sp.apply_sources(hp, hv, sources, solid, delta, host_rng, initial=initial and step == 0)
            n = ctx.apply_sources(dev_rng, initial=initial and step == 0)
            dp, dv = ctx.get_particles()
            assert n == len(hp)
            assert np.array_equal(dp, hp) and np.array_equal(dv, hv)
            assert host_rng.random() == dev_rng.random()
            hp = hp + np.float32(0.3 * delta)  # move the particles between calls
            ctx.set_particles(hp, hv)


def test_bake_with_inflow_and_outflow_equals_host_seeding(tmp_path, oracle_backend, monkeypatch):
    """A bake with dynamic sources through the device entry points writes the
    same frames as the reference's host round trip (particles downloaded,
    seeded by source_particles.apply_sources, uploaded), reproduced here by
    swapping Context.apply_sources for it."""
    from continuum_liquid_amd.Solver.Kernel import source_particles as sp

    def bake(name):
        cfg = make_config(tmp_path / name, frames=3)
        cfg["simulation"]["sources"] += [
            {"type": "INFLOW", "particle_density": 4, "velocity": [1.0, 0.0, -0.5],
             "shapes": [{"box": [[0.1, 0.6, 0.55], [0.3, 0.8, 0.8]]}]},
            {"type": "OUTFLOW", "shapes": [{"box": [[0.0, 0.0, 0.0], [0.3, 0.3, 0.15]]}]}]
        solver_main.main(cfg)
        return [f.read_bytes() for f in sorted((tmp_path / name / "bake" / "particles").glob("*.bin"))]

    device = bake("device")
    state = {}

    def host_round_trip(ctx, rng, initial=False):
        pos, vel = ctx.get_particles() if ctx.particle_count else (np.empty((0, 3), np.float32),) * 2
        solid = ctx.get_field("obstacle").astype(bool)
        pos, vel = sp.apply_sources(pos, vel, state["sources"], solid, ctx.dx, rng, initial)
        ctx.set_particles(pos, vel)
        return len(pos)

    def keep_sources(ctx, sources):
        state["sources"] = sources

    monkeypatch.setattr(runtime.Context, "apply_sources", host_round_trip)
    monkeypatch.setattr(runtime.Context, "set_sources", keep_sources)
    host = bake("host")
    assert len(device) == 4 and device == host
    assert len(device[-1]) != len(device[0])


def test_device_frames_through_pinned_slots(tmp_path, oracle_lib):
    """submit_device writes the positions held when it was called, also when
    the particles move and grow before the writer runs; the download rejects
    a slot that is too small and an unknown ticket."""
    rng = np.random.default_rng(3)
    shape, delta = (10, 9, 8), 0.1
    pos = (rng.random((500, 3)) * np.array(shape) * delta).astype(np.float32)
    cache = output.create_cache(tmp_path, 24, 1, 4)
    with runtime.Context(oracle_lib, shape, delta, capacity=len(pos)) as ctx:
        ctx.set_particles(pos)
        writer = output.ParticleOutput(cache, ctx)
        expected = []
        for frame in range(1, 5):
            expected.append(ctx.get_positions())
            assert writer.submit_device(frame) == len(expected[-1])
            grown = np.concatenate([expected[-1] + np.float32(0.01), expected[-1][: 50 * frame]])
            ctx.set_particles(grown)
        writer.close()
        for frame, want in enumerate(expected, start=1):
            got = np.fromfile(output.frame_path(cache, frame), np.float32).reshape(-1, 3)
            assert np.array_equal(got, want)
        small = ctx.host_alloc(3)
        with pytest.raises(runtime.ClqError, match="host buffer holds"):
            ctx.positions_to_host_async(small)
        ctx.host_free(small)
        with pytest.raises(runtime.ClqError, match="unknown download ticket"):
            ctx.wait_host_copy(10**6)


def test_bake_with_a_mesh_obstacle_equals_the_same_box_as_a_shape(tmp_path, oracle_backend):
    """An obstacle given as a triangle mesh in world coordinates (grid origin
    (-x/2, -y/2, 0)) is voxelised by ray parity and bakes exactly like the
    same box given as a shape in domain coordinates."""
    from .test_oracle_mesh import box_mesh

    lo, hi = [0.55, 0.2, 0.1], [0.8, 0.7, 0.35]
    tris = box_mesh(np.array(lo) + [-0.5, -0.5, 0.0], np.array(hi) + [-0.5, -0.5, 0.0])
    vertices, index = np.unique(tris.reshape(-1, 3), axis=0, return_inverse=True)

    def bake(name, obstacle):
        cfg = make_config(tmp_path / name, frames=2)
        cfg["simulation"]["obstacles"] = [obstacle]
        solver_main.main(cfg)
        return [f.read_bytes() for f in sorted((tmp_path / name / "bake" / "particles").glob("*.bin"))]

    as_mesh = bake("mesh", {"mesh": {"vertices": vertices, "triangles": index.reshape(-1, 3)}})
    as_shape = bake("shape", {"shapes": [{"box": [lo, hi]}]})
    assert len(as_mesh) == 3 and as_mesh == as_shape
