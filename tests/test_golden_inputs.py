"""Synthetic inputs and the random stream pinned by committed vectors
(tests/golden/synthetic_inputs.json, written by make_synthetic_inputs.py).

numpy does not promise that Generator streams stay the same across versions;
the benchmark scenes and the particle sources depend on PCG64 through
default_rng.  The digests catch a changed stream, and a big-integer
restatement of PCG64 (128-bit LCG with multiplier
0x2360ED051FC65DA44385DF649FCCF645, XSL-RR 128/64 output, (x >> 11) * 2^-53
doubles, jump-ahead by repeated squaring), the arithmetic the device
sources use, reproduces the recorded doubles from the recorded state."""

from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from continuum_liquid_amd import scenes

GOLDEN = json.loads((Path(__file__).parent / "golden" / "synthetic_inputs.json").read_text())
MULT = 0x2360ED051FC65DA44385DF649FCCF645
MASK = (1 << 128) - 1


def pcg_step(state: int, inc: int) -> int:
    return (state * MULT + inc) & MASK


def pcg_double(state: int) -> float:
    x = ((state >> 64) ^ state) & ((1 << 64) - 1)
    rot = state >> 122
    r = ((x >> rot) | (x << ((64 - rot) & 63))) & ((1 << 64) - 1)
    return (r >> 11) * 2.0 ** -53


def pcg_advance(state: int, inc: int, delta: int) -> int:
    acc_mult, acc_plus, cur_mult, cur_plus = 1, 0, MULT, inc
    while delta:
        if delta & 1:
            acc_mult = acc_mult * cur_mult & MASK
            acc_plus = (acc_plus * cur_mult + cur_plus) & MASK
        cur_plus = (cur_mult + 1) * cur_plus & MASK
        cur_mult = cur_mult * cur_mult & MASK
        delta >>= 1
    return (acc_mult * state + acc_plus) & MASK


@pytest.mark.parametrize("config", ["1", "2"])
def test_scene_inputs_match_their_digest(config):
    want = GOLDEN["configs"][config]
    s = scenes.bench_scene(int(config))
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(s.positions, dtype=np.float32).tobytes())
    h.update(np.ascontiguousarray(s.velocities, dtype=np.float32).tobytes())
    assert list(s.shape) == want["shape"] and s.particle_count == want["particles"]
    assert h.hexdigest() == want["sha256"]


def test_numpy_stream_matches_the_recorded_one():
    g = GOLDEN["pcg64"]
    rng = np.random.default_rng(0)
    st = rng.bit_generator.state["state"]
    assert str(st["state"]) == g["seed0_state"] and str(st["inc"]) == g["seed0_inc"]
    assert np.random.default_rng(0).random(8).tolist() == g["seed0_random"]
    for k, vals in g["after_advance"].items():
        r = np.random.default_rng(0)
        r.bit_generator.advance(int(k))
        assert r.random(2).tolist() == vals


def test_pcg64_restatement_reproduces_the_recorded_doubles():
    g = GOLDEN["pcg64"]
    state, inc = int(g["seed0_state"]), int(g["seed0_inc"])
    s = state
    for want in g["seed0_random"]:
        s = pcg_step(s, inc)  # a draw steps the state, then outputs it
        assert pcg_double(s) == want
    for k, vals in g["after_advance"].items():
        s = pcg_advance(state, inc, int(k))
        for want in vals:
            s = pcg_step(s, inc)
            assert pcg_double(s) == want
