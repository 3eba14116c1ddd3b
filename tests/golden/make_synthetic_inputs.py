"""Writes synthetic_inputs.json: SHA-256 digests of the benchmark scenes'
seeded particle arrays, and the first outputs of numpy's PCG64 generator
(the stream the particle sources reproduce on the device), so that a change
of numpy's Generator streams between versions is caught instead of silently
changing every input.

    python tests/golden/make_synthetic_inputs.py
"""

from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))  # the repository root

from continuum_liquid_amd import scenes  # noqa: E402


def scene_digest(config: int) -> dict:
    s = scenes.bench_scene(config)
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(s.positions, dtype=np.float32).tobytes())
    h.update(np.ascontiguousarray(s.velocities, dtype=np.float32).tobytes())
    return {"shape": list(s.shape), "particles": int(s.particle_count), "sha256": h.hexdigest()}


def generator_vectors() -> dict:
    rng = np.random.default_rng(0)
    state = rng.bit_generator.state["state"]
    first = np.random.default_rng(0).random(8).tolist()
    jumped = {}
    for k in (1, 3, 1000, 123456789):
        g = np.random.default_rng(0)
        g.bit_generator.advance(k)
        jumped[str(k)] = g.random(2).tolist()
    return {"seed0_state": str(state["state"]), "seed0_inc": str(state["inc"]), "seed0_random": first,
            "after_advance": jumped}


def main() -> None:
    out = {"numpy": np.__version__, "configs": {str(c): scene_digest(c) for c in (1, 2, 3, 5)},
           "pcg64": generator_vectors()}
    (HERE / "synthetic_inputs.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
