"""Particle cache writer (format of the reference's Solver/Kernel/output.py).

On-disk contract kept byte-compatible with the reference viewer
(Internal/Core/load_result.py:76-112): `particle_cache.json` with version 2,
fps, start/end frame, `variable_particle_count: true`, dtype float32, 3
components, frame directory `particles` and pattern
`particles_{frame:06d}.bin`; each frame is raw float32 N x 3 in domain-local
coordinates, published by renaming a `.tmp` file (reference output.py:19-53).

Frames are written by two writer threads (reference output.py:99-116) from
three frame slots (reference output.py:16, 124-135).  With a context the
slots are pinned host memory of its device and a frame is handed over as an
asynchronous device-to-host copy (reference output.py:178-195): the step loop
continues while the copy is in flight, and the writer thread waits for it
before writing.  A writer exception is re-raised on the next submit or at
close (reference output.py:91-96).
"""

from __future__ import annotations

import json
import queue
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


class _Slot:
    """One frame slot: a float32 (rows, 3) buffer, pinned when a context
    allocated it."""

    __slots__ = ("array", "pinned")

    def __init__(self):
        self.array = None
        self.pinned = False

    @property
    def rows(self) -> int:
        return 0 if self.array is None else self.array.shape[0]


class ParticleOutput:
    """Asynchronous frame writer with a bounded number of frames in flight.

    `ctx` (optional) is the runtime.Context whose positions submit_device
    downloads; its pinned host memory backs the slots."""

    def __init__(self, cache_directory, ctx=None, writers: int = WRITER_COUNT):
        self.directory = Path(cache_directory)
        self._ctx = ctx
        self._pool = ThreadPoolExecutor(max_workers=writers, thread_name_prefix="ParticleWriter")
        self._free: queue.Queue[_Slot] = queue.Queue()
        self._slots = [_Slot() for _ in range(writers + 1)]
        for slot in self._slots:
            self._free.put(slot)
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

    def _release(self, slot: _Slot) -> None:
        if slot.pinned and slot.array is not None:
            self._ctx.host_free(slot.array)
        slot.array, slot.pinned = None, False

    def _take_slot(self, rows: int) -> _Slot:
        if self._closed:
            raise RuntimeError("particle output is closed")
        self._raise_finished_errors()
        slot = self._free.get()  # blocks while every slot is in flight
        if slot.rows < rows:
            # counts grow with inflow: leave a quarter of headroom
            grown = max(rows, rows + rows // 4)
            try:
                self._release(slot)
                if self._ctx is not None:
                    slot.array, slot.pinned = self._ctx.host_alloc(grown), True
                else:
                    slot.array = np.empty((grown, 3), np.float32)
            except BaseException:
                self._free.put(slot)
                raise
        return slot

    def _write_later(self, slot: _Slot, count: int, frame: int, ticket: int) -> None:
        ctx = self._ctx

        def job():
            try:
                if ticket >= 0:
                    ctx.wait_host_copy(ticket)
                data = slot.array[:count] if count else np.empty((0, 3), np.float32)
                write_particles(self.directory, data, frame)
            finally:
                self._free.put(slot)

        self._pending.append(self._pool.submit(job))

    def submit(self, positions: np.ndarray, frame: int) -> None:
        """Queues a frame from host memory (reference output.py:164-173)."""
        positions = np.ascontiguousarray(positions, dtype=np.float32)
        if positions.ndim != 2 or positions.shape[1] != 3:
            raise ValueError("Particle positions must have shape (N, 3)")
        slot = self._take_slot(len(positions))
        if len(positions):
            slot.array[: len(positions)] = positions
        self._write_later(slot, len(positions), frame, -1)

    def submit_device(self, frame: int) -> int:
        """Queues a frame of the context's current positions, copied without
        blocking (reference output.py:176-195); returns the particle count."""
        if self._ctx is None:
            raise RuntimeError("submit_device needs the output to be created with a context")
        slot = self._take_slot(self._ctx.particle_count)
        try:
            count, ticket = self._ctx.positions_to_host_async(slot.array) if slot.rows else (0, -1)
        except BaseException:
            self._free.put(slot)
            raise
        self._write_later(slot, count, frame, ticket)
        return count

    def close(self) -> None:
        if self._closed:
            return
        self._closed = True
        self._pool.shutdown(wait=True)
        errors = [f.exception() for f in self._pending if f.exception() is not None]
        self._pending = []
        for slot in self._slots:
            self._release(slot)
        if errors:
            raise RuntimeError(f"particle cache writer failed: {errors[0]}") from errors[0]
