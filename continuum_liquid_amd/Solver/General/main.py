"""Bake entry point of the add-on's solver package and its progress channel.

Interface (reference Solver/General/main.py:13-52, used by the add-on's bake
operator on a worker thread):

* ``set_message_callback(cb)`` installs one callable that receives the solver's
  progress dictionaries, ``clear_message_callback()`` removes it and
  ``emit_message(msg)`` delivers a dictionary to whichever callable is
  installed at that moment.  Installing, removing and looking up the callable
  are serialised by one lock; the callable itself runs outside the lock, on
  the solver's thread, so it may install or remove callables itself.
* ``main(config)`` runs one bake described by the exported configuration and
  returns ``None``.  ``simulation.settings.solver_backend`` names the device
  class; the add-on offers "CPU" (its default, also assumed when the key is
  absent) and "GPU", and both names start the same solver, which here is the
  gfx950 HIP step.  Any other name is a ``ValueError``.  Whatever the outcome,
  the wall time of the call is printed as ``Bake runtime: <seconds> s``
  followed by a separator line.
"""

from __future__ import annotations

import threading
from time import perf_counter
from typing import Any, Callable, Optional

MessageCallback = Callable[[dict[str, Any]], None]

# The add-on writes these two names into the exported configuration.
BACKENDS = ("CPU", "GPU")
SEPARATOR = "#" * 64


class _ProgressChannel:
    """Holds the one installed progress callable behind a lock."""

    def __init__(self) -> None:
        self._guard = threading.Lock()
        self._target: Optional[MessageCallback] = None

    def install(self, target: Optional[MessageCallback]) -> None:
        with self._guard:
            self._target = target

    def current(self) -> Optional[MessageCallback]:
        with self._guard:
            return self._target


_channel = _ProgressChannel()


def set_message_callback(callback: MessageCallback | None) -> None:
    """Install `callback` as the receiver of progress messages (None removes it)."""
    _channel.install(callback)


def clear_message_callback() -> None:
    """Stop delivering progress messages."""
    _channel.install(None)


def emit_message(message: dict[str, Any]) -> None:
    """Hand `message` to the receiver installed now, if there is one."""
    receiver = _channel.current()
    if receiver is not None:
        receiver(message)


def _backend_of(config: dict[str, Any]) -> str:
    settings = (config.get("simulation") or {}).get("settings") or {}
    return str(settings.get("solver_backend", "CPU")).strip().upper()


def main(config: dict[str, Any]) -> None:
    """Run the bake described by `config` on the HIP step."""
    began = perf_counter()
    try:
        backend = _backend_of(config)
        if backend not in BACKENDS:
            raise ValueError(f"Unsupported solver backend: {backend}")
        # Both device classes of the add-on map onto the one HIP step.
        from ..Kernel.solver import solver as run_bake

        run_bake(config)
    finally:
        print(f"Bake runtime: {perf_counter() - began:.3f} s")
        print(SEPARATOR)
