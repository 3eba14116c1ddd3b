"""Continuum Liquid solver entry point and runtime message channel.

Same public names and semantics as the reference's Solver/General/main.py:
`set_message_callback`, `clear_message_callback` and `emit_message` guard one
callback with a lock (reference main.py:13-32); `main(config)` dispatches on
`simulation.settings.solver_backend` and prints the bake runtime on exit
(reference main.py:35-52).  The "GPU" backend runs the gfx950 HIP step.  The
reference's "CPU" backend ran its OpenCL kernels on a host CPU device; this
build has no CPU compute path, so "CPU" raises instead of running elsewhere.
"""

from __future__ import annotations

import threading
from time import perf_counter
from typing import Any, Callable

MessageCallback = Callable[[dict[str, Any]], None]

_message_callback: MessageCallback | None = None
_message_callback_lock = threading.Lock()


def set_message_callback(callback: MessageCallback | None) -> None:
    """Set the callback used for solver runtime messages."""
    global _message_callback
    with _message_callback_lock:
        _message_callback = callback


def clear_message_callback() -> None:
    """Remove the current solver runtime callback."""
    global _message_callback
    with _message_callback_lock:
        _message_callback = None


def emit_message(message: dict[str, Any]) -> None:
    """Forward a runtime message to the active caller (outside the lock)."""
    with _message_callback_lock:
        callback = _message_callback
    if callback is not None:
        callback(message)


def main(config: dict[str, Any]) -> None:
    """Select the configured backend and execute the solver."""
    total_start_time = perf_counter()
    try:
        simulation_config = config.get("simulation") or {}
        settings = simulation_config.get("settings") or {}
        solver_backend = str(settings.get("solver_backend", "CPU")).strip().upper()

        if solver_backend == "GPU":
            from ..Kernel import solver as solver_kernel_module

            return solver_kernel_module.solver(config)
        if solver_backend == "CPU":
            raise RuntimeError(
                "The CPU backend is not part of the MI355X build: set "
                "simulation.settings.solver_backend to 'GPU' to run the HIP step"
            )
        raise ValueError(f"Unsupported solver backend: {solver_backend}")
    finally:
        total_runtime = perf_counter() - total_start_time
        print(f"Bake runtime: {total_runtime:.3f} s")
        print("################################################################")
