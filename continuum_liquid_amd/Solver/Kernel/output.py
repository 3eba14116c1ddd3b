"""Particle cache writer (format of the reference's Solver/Kernel/output.py).

On-disk contract kept byte-compatible with the reference viewer
(Internal/Core/load_result.py:76-112): `particle_cache.json` with version 2,
fps, start/end frame, `variable_particle_count: true`, dtype float32, 3
components, frame directory `particles` and pattern
`particles_{frame:06d}.bin`; each frame is raw float32 N x 3 in domain-local
coordinates, published by renaming a `.tmp` file (reference output.py:19-53).

Frames are written by two writer threads (reference output.py:99-116); the
step loop hands over a host copy and continues.  A writer exception is
re-raised on the next submit or at close (reference output.py:91-96).
"""

from __future__ import annotations

import json
import threading
from concurrent.futures import Future, ThreadPoolExecutor
from pathlib import Path

import numpy as np

CACHE_FILENAME = "particle_cache.json"
FRAME_DIRECTORY = "particles"
FRAME_PATTERN = "particles_{frame:06d}.bin"
WRITER_COUNT = 2


def create_cache(directory, fps, start_frame, end_frame) -> Path:
    directory = Path(directory)
    (directory / FRAME_DIRECTORY).mkdir(parents=True, exist_ok=True)
    metadata = {
        "version": 2,
        "fps": int(fps),
        "start_frame": int(start_frame),
        "end_frame": int(end_frame),
        "variable_particle_count": True,
        "dtype": "float32",
        "components": 3,
        "frame_directory": FRAME_DIRECTORY,
        "frame_pattern": FRAME_PATTERN,
    }
    (directory / CACHE_FILENAME).write_text(json.dumps(metadata, indent=2), encoding="utf-8")
    return directory


def frame_path(cache_directory, frame: int) -> Path:
    return Path(cache_directory) / FRAME_DIRECTORY / FRAME_PATTERN.format(frame=int(frame))


def write_particles(cache_directory, particles, frame) -> Path:
    """Write one frame atomically: the final name appears only when complete."""
    positions = np.ascontiguousarray(particles, dtype=np.float32)
    if positions.ndim != 2 or positions.shape[1] != 3:
        raise ValueError("Particle positions must have shape (N, 3)")
    final = frame_path(cache_directory, frame)
    partial = final.with_name(final.name + ".tmp")
    try:
        positions.tofile(partial)
        partial.replace(final)
    finally:
        partial.unlink(missing_ok=True)
    return final


class ParticleOutput:
    """Asynchronous frame writer with a bounded number of frames in flight."""

    def __init__(self, cache_directory, writers: int = WRITER_COUNT):
        self.directory = Path(cache_directory)
        self._pool = ThreadPoolExecutor(max_workers=writers, thread_name_prefix="ParticleWriter")
        self._slots = threading.BoundedSemaphore(writers + 1)
        self._pending: list[Future] = []
        self._closed = False

    def _raise_finished_errors(self) -> None:
        still = []
        for fut in self._pending:
            if fut.done():
                exc = fut.exception()
                if exc is not None:
                    raise RuntimeError(f"particle cache writer failed: {exc}") from exc
            else:
                still.append(fut)
        self._pending = still

    def submit(self, positions: np.ndarray, frame: int) -> None:
        if self._closed:
            raise RuntimeError("particle output is closed")
        self._raise_finished_errors()
        self._slots.acquire()
        data = np.array(positions, dtype=np.float32, copy=True)

        def job():
            try:
                write_particles(self.directory, data, frame)
            finally:
                self._slots.release()

        self._pending.append(self._pool.submit(job))

    def close(self) -> None:
        if self._closed:
            return
        self._closed = True
        self._pool.shutdown(wait=True)
        errors = [f.exception() for f in self._pending if f.exception() is not None]
        self._pending = []
        if errors:
            raise RuntimeError(f"particle cache writer failed: {errors[0]}") from errors[0]
