"""The learned preconditioner: training data from the solver itself, training
on the GPU through the C-ABI (clq_net_train_step), and weight files.

A training pair is what one pressure solve of a running simulation sees and
returns: the right-hand side (scaled divergence), the fluid fraction (the
solve's mask), the pressure before the solve (the previous step's, the warm
start) and the FP64 solution of the solve rounded to FP32.  Pairs come from
`scenes.random_scene` simulations advanced with the FP64 FCG solve.

The network is fully convolutional: weights trained on one grid shape are
loaded into a `runtime.Network` of any other shape.
"""

from __future__ import annotations

import time
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import runtime, scenes
from .Solver.Kernel import solver as host_solver

WEIGHTS_DIR = Path(__file__).resolve().parent / "weights"
DEFAULT_WEIGHTS = WEIGHTS_DIR / "unet_pressure.npy"

# the solve that produces the targets
TARGET_SOLVER = {"mode": "fcg", "tolerance": 1e-8, "max_cycles": 60, "pre_smooth": 2, "post_smooth": 2,
                 "coarse_extra": 16}


@dataclass
class PairSet:
    rhs: np.ndarray     # [n, nx, ny, nz] float32
    mask: np.ndarray
    prev: np.ndarray
    target: np.ndarray

    def __len__(self) -> int:
        return len(self.rhs)

    def save(self, path) -> None:
        np.savez(path, rhs=self.rhs, mask=self.mask, prev=self.prev, target=self.target)

    @classmethod
    def load(cls, path) -> "PairSet":
        with np.load(path, allow_pickle=False) as f:
            return cls(f["rhs"], f["mask"], f["prev"], f["target"])

    def split(self, n_train: int) -> tuple["PairSet", "PairSet"]:
        a = PairSet(self.rhs[:n_train], self.mask[:n_train], self.prev[:n_train], self.target[:n_train])
        b = PairSet(self.rhs[n_train:], self.mask[n_train:], self.prev[n_train:], self.target[n_train:])
        return a, b


def collect_pairs(lib, shape, n_scenes: int, steps: int, seed: int = 0, dt: float = 1.0 / 240.0,
                  flip_ratio: float = 0.95, skip: int = 0, log=None) -> PairSet:
    """Simulate `n_scenes` random scenes for `skip + steps` steps each and keep
    one pair per solve after the first `skip` steps of each scene."""
    out = {k: [] for k in ("rhs", "mask", "prev", "target")}
    for s in range(n_scenes):
        scene = scenes.random_scene(shape, seed + s)
        if scene.fluid_cells == 0:
            continue
        ctx = runtime.Context(lib, scene.shape, scene.dx, capacity=scene.particle_count)
        for k, v in scene_pairs(ctx, scene, steps, skip, dt, flip_ratio).items():
            out[k].extend(v)
        ctx.close()
        if log:
            log(f"scene {s}: {scene.fluid_cells} fluid cells, {len(out['rhs'])} pairs so far")
    return PairSet(*(np.stack(out[k]) if out[k] else np.zeros((0,) + tuple(shape), np.float32)
                     for k in ("rhs", "mask", "prev", "target")))


def scene_pairs(ctx, scene, steps: int, skip: int, dt: float = 1.0 / 240.0, flip_ratio: float = 0.95) -> dict:
    """Run `skip + steps` steps of `scene` on the context (its particles
    replace the context's) and return the pairs of the last `steps` solves
    that converged, as lists of float32 arrays."""
    out = {k: [] for k in ("rhs", "mask", "prev", "target")}
    pending = {}
    step_index = [0]

    def observer(c, stats):
        if step_index[0] < skip:
            return
        if stats is None:
            pending["rhs"] = c.get_field("rhs")
            pending["mask"] = c.get_field("fluid_fraction")
            pending["prev"] = c.get_field("pressure")
        elif stats["relative_residual"] <= 1e-6:  # keep converged solves only
            for k in ("rhs", "mask", "prev"):
                out[k].append(pending[k])
            out["target"].append(c.get_field("pressure_fp64").astype(np.float32))

    ctx.set_particles(scene.positions, scene.velocities)
    params = host_solver.StepParams(dt=dt, gravity=scene.gravity, flip_ratio=flip_ratio,
                                    solver=dict(TARGET_SOLVER), observer=observer)
    for k in range(skip + steps):
        step_index[0] = k
        host_solver.step(ctx, params)
    return out


def train_streaming(net: "runtime.Network", lib, shape, total: int, window: int = 512, passes: int = 2,
                    batch: int = 4, lr: float = 1e-3, lr_end: float = 1e-5, steps: int = 36, skip: int = 4,
                    seed: int = 0, prev_dropout: float = 0.5, log=None, checkpoint=None) -> list[float]:
    """Training on `total` pairs generated on the fly, too many to hold at
    once: consecutive seeded random scenes (seeds seed, seed + 1, ...) fill a
    window of `window` pairs, the network takes `passes` shuffled passes of
    Adam steps over it (prev_dropout as in `train`), and the window is
    refilled.  The step size follows a cosine from `lr` to `lr_end` over the
    `total` pairs (at a constant 1e-3 the loss of 128^3 training rose again
    after about 10k pairs).  The next scene is built on a host thread while
    the device runs the current one.  `checkpoint(net, seen)`, when given, is
    called after every window.  Returns the mean loss of each window."""
    from concurrent.futures import ThreadPoolExecutor

    shape = tuple(int(n) for n in shape)
    keys = ("rhs", "mask", "prev", "target")
    buf = {k: np.empty((window,) + shape, np.float32) for k in keys}
    rng = np.random.default_rng(seed)
    losses: list[float] = []
    fill = seen = 0
    t0 = time.perf_counter()

    def train_window(n: int) -> None:
        total_loss, steps_done = 0.0, 0
        for _ in range(passes):
            order = rng.permutation(n)
            for i in range(0, n - batch + 1, batch):
                idx = np.sort(order[i:i + batch])
                prev = buf["prev"][idx].copy()
                prev[rng.random(len(idx)) < prev_dropout] = 0.0
                rate = lr_end + 0.5 * (lr - lr_end) * (1.0 + np.cos(np.pi * min(seen / total, 1.0)))
                total_loss += net.train_step(buf["rhs"][idx], buf["mask"][idx], prev, buf["target"][idx], rate)
                steps_done += 1
        losses.append(total_loss / max(steps_done, 1))
        if checkpoint is not None:
            checkpoint(net, seen)
        if log:
            log(f"{seen} pairs: window of {n}, loss {losses[-1]:.4e} ({time.perf_counter() - t0:.1f} s)")

    ctx = None
    scene_seed = seed
    try:
        with ThreadPoolExecutor(1) as pool:
            upcoming = pool.submit(scenes.random_scene, shape, scene_seed)
            while seen < total:
                scene = upcoming.result()
                scene_seed += 1
                upcoming = pool.submit(scenes.random_scene, shape, scene_seed)
                if scene.fluid_cells == 0:
                    continue
                if ctx is None:
                    ctx = runtime.Context(lib, scene.shape, scene.dx, capacity=scene.particle_count)
                pairs = scene_pairs(ctx, scene, steps, skip)
                for j in range(len(pairs["rhs"])):
                    for k in keys:
                        buf[k][fill] = pairs[k][j]
                    fill += 1
                    seen += 1
                    if fill == window or seen == total:
                        train_window(fill)
                        fill = 0
                    if seen == total:
                        break
    finally:
        if ctx is not None:
            ctx.close()
    return losses


def train(net: "runtime.Network", data: PairSet, epochs: int, batch: int = 4, lr: float = 1e-3, seed: int = 0,
          prev_dropout: float = 0.5, log=None) -> list[float]:
    """Adam on the mean squared error of the scaled pressure (clq_net_train_step).

    With probability `prev_dropout` a sample's previous-pressure channel is
    zeroed: inside FCG the network is applied to residuals with no previous
    solution, as a preconditioner, and to (rhs, previous pressure) as the
    initial guess; both uses are trained."""
    rng = np.random.default_rng(seed)
    losses = []
    n = len(data)
    t0 = time.perf_counter()
    for epoch in range(epochs):
        order = rng.permutation(n)
        epoch_loss, batches = 0.0, 0
        for i in range(0, n - batch + 1, batch):
            idx = np.sort(order[i:i + batch])
            prev = data.prev[idx].copy()
            prev[rng.random(len(idx)) < prev_dropout] = 0.0
            loss = net.train_step(data.rhs[idx], data.mask[idx], prev, data.target[idx], lr)
            epoch_loss += loss
            batches += 1
        losses.append(epoch_loss / max(batches, 1))
        if log:
            log(f"epoch {epoch}: loss {losses[-1]:.4e} ({time.perf_counter() - t0:.1f} s)")
    return losses


def relative_error(net: "runtime.Network", data: PairSet, with_prev: bool = True) -> float:
    """Mean over pairs of ||out - target|| / ||target|| on the interior fluid cells."""
    errs = []
    inner = scenes.interior_mask(data.rhs.shape[1:])
    for i in range(len(data)):
        out = net.forward(data.rhs[i], data.mask[i], data.prev[i] if with_prev else None)
        sel = inner & (data.mask[i] > 0)
        t = data.target[i][sel].astype(np.float64)
        errs.append(np.linalg.norm(out[sel] - t) / max(np.linalg.norm(t), 1e-30))
    return float(np.mean(errs))


def save_weights(net: "runtime.Network", path=DEFAULT_WEIGHTS) -> Path:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.save(path, net.get_params())
    return path


def load_weights(lib, shape, dx: float, path=DEFAULT_WEIGHTS) -> "runtime.Network":
    """A network for the (shape, dx) pressure grid with the parameters stored at `path`."""
    params = np.load(path, allow_pickle=False)
    net = runtime.Network(lib, shape, dx, seed=None)
    if params.size != net.param_count:
        raise ValueError(f"{path}: {params.size} parameters, expected {net.param_count}")
    net.set_params(params)
    return net


def main(argv=None) -> int:
    """Generate pairs on the GPU, train, and write the weights file."""
    import argparse

    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--shape", type=int, nargs=3, default=(64, 64, 64))
    ap.add_argument("--scenes", type=int, default=24)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--skip", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--holdout", type=int, default=2, help="scenes kept out of training for the report")
    ap.add_argument("--out", type=Path, default=DEFAULT_WEIGHTS)
    ap.add_argument("--pairs", type=Path, default=None, help="also save the pairs here (.npz)")
    ap.add_argument("--stream", type=int, default=0, metavar="N",
                    help="train on N pairs generated on the fly (window by window) instead of a stored set")
    ap.add_argument("--window", type=int, default=512, help="--stream: pairs held at once")
    ap.add_argument("--passes", type=int, default=2, help="--stream: passes over each window")
    args = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    lib = runtime.load_library()
    shape = tuple(args.shape)
    t0 = time.perf_counter()
    if args.stream:
        net = runtime.Network(lib, shape, 1.0 / max(shape), seed=args.seed)
        # held out: scenes of seeds the stream never reaches; the weights kept
        # are those of the window with the lowest held-out error (mean of the
        # errors with and without the previous pressure, the two uses)
        test_set = collect_pairs(lib, shape, args.holdout, 3, seed=10**6 + args.seed, skip=args.skip)

        def score(trained):
            return 0.5 * (relative_error(trained, test_set) + relative_error(trained, test_set, with_prev=False))

        before = relative_error(net, test_set)
        best = {"score": float("inf"), "seen": 0}

        def checkpoint(trained, seen):
            sc = score(trained)
            if sc < best["score"]:
                best.update(score=sc, seen=seen)
                save_weights(trained, args.out)
            log(f"  held-out score {sc:.3e} (best {best['score']:.3e} after {best['seen']} pairs)")

        train_streaming(net, lib, shape, args.stream, window=args.window, passes=args.passes, batch=args.batch,
                        lr=args.lr, steps=args.steps, skip=args.skip, seed=args.seed, log=log, checkpoint=checkpoint)
        kept = load_weights(lib, shape, 1.0 / max(shape), args.out)
        log(f"held-out relative error: untrained {before:.3e}, kept (after {best['seen']} pairs) "
            f"{relative_error(kept, test_set):.3e} (no previous pressure: "
            f"{relative_error(kept, test_set, with_prev=False):.3e}); {time.perf_counter() - t0:.1f} s")
        log(f"weights written to {args.out}")
        return 0
    data = collect_pairs(lib, shape, args.scenes, args.steps, seed=args.seed, skip=args.skip, log=log)
    log(f"{len(data)} pairs in {time.perf_counter() - t0:.1f} s")
    if args.pairs:
        data.save(args.pairs)
    per_scene = len(data) // max(args.scenes, 1)
    train_set, test_set = data.split(len(data) - args.holdout * per_scene)
    net = runtime.Network(lib, shape, 1.0 / max(shape), seed=args.seed)
    before = relative_error(net, test_set) if len(test_set) else float("nan")
    train(net, train_set, args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, log=log)
    after = relative_error(net, test_set) if len(test_set) else float("nan")
    after_np = relative_error(net, test_set, with_prev=False) if len(test_set) else float("nan")
    log(f"held-out relative error: untrained {before:.3e}, trained {after:.3e} (no previous pressure: {after_np:.3e})")
    path = save_weights(net, args.out)
    log(f"weights written to {path}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
