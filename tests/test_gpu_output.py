"""Frame output through pinned slots on the device (reference
Solver/Kernel/output.py:56-75, 178-195): a frame's download is in flight
while the next steps run and move the particles, and the file still holds the
positions of the moment it was submitted."""

from __future__ import annotations

import numpy as np
import pytest

from continuum_liquid_amd import scenes
from continuum_liquid_amd.Solver.Kernel import output
from continuum_liquid_amd.Solver.Kernel import solver as host_solver

from .helpers import default_params, make_context

pytestmark = pytest.mark.gpu


def test_frames_in_flight_hold_their_step(hip_lib, tmp_path):
    scene = scenes.small_scene(32, seed=5)
    params = default_params(scene)
    steps = 6
    # the positions after each step, read synchronously
    with make_context(hip_lib, scene) as ctx:
        expected = []
        for _ in range(steps):
            host_solver.step(ctx, params)
            expected.append(ctx.get_positions())
    # the same steps (deterministic, bit for bit) with a frame submitted after
    # each step and not waited for; three slots, so up to three copies are in
    # flight while the next steps run
    cache = output.create_cache(tmp_path, 24, 1, steps)
    with make_context(hip_lib, scene) as ctx:
        writer = output.ParticleOutput(cache, ctx)
        try:
            for frame in range(1, steps + 1):
                host_solver.step(ctx, params)
                assert writer.submit_device(frame) == len(expected[frame - 1])
        finally:
            writer.close()
    for frame, want in enumerate(expected, start=1):
        got = np.fromfile(output.frame_path(cache, frame), np.float32).reshape(-1, 3)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_snapshot_survives_growth_and_reupload(hip_lib, tmp_path):
    """A download started before the particle arrays are reallocated (an
    upload larger than the capacity) still lands the old positions."""
    scene = scenes.small_scene(20, seed=6)
    cache = output.create_cache(tmp_path, 24, 1, 3)
    with make_context(hip_lib, scene) as ctx:
        writer = output.ParticleOutput(cache, ctx)
        want = []
        try:
            for frame in range(1, 4):
                want.append(ctx.get_positions())
                writer.submit_device(frame)
                grown = np.concatenate([want[-1], want[-1] * np.float32(0.5)])
                ctx.set_particles(grown)  # doubles the count: new device arrays
        finally:
            writer.close()
    for frame, w in enumerate(want, start=1):
        got = np.fromfile(output.frame_path(cache, frame), np.float32).reshape(-1, 3)
        assert np.array_equal(got, w)
