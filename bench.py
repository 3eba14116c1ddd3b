"""Benchmark of the FLIP/PIC time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 3] [--solver mixed]

A step is one full time step of the path on the resident particle set:
binning, P2G + occupancy, grid glue, the pressure solve to the relative
residual target (FP64-checked), pressure correction, and the fused G2P/FLIP
blend/advection/particle boundary pass.  The default workload is the 256^3
dam break + falling sphere (dx = 1/256, 8 stratified-jitter FP32 particles
per fluid cell, ~28.6M particles, dt = 1/240, FLIP 0.95, pressure solve by
FP64 flexible CG preconditioned with one mixed-precision (1, 1) cycle that
runs its coarse-grid correction twice below level 1 and once elsewhere, the
corrections weighted by 1.6).

Prints one JSON line: ms/step (`value`), particle-steps/s, pressure-solve ms,
a `roofline` object for the dominant kernel (algorithmic bytes per launch
divided by the launch duration measured with HIP events on the library's
stream during the timed steps) and a `cpu_baseline` object (the CPU oracle,
a C/OpenMP restatement of the reference's OpenCL kernels, timed on this
host's cores on a bounded sample of the same workload), plus the same oracle
at the reference's own solver settings per section.  After the timed steps
(outside the timed region) the last solve is checked against an independent
FP64 solve (`fp64_error_max`), and the run goes on to the configuration's
last step, timing the final 20 steps apart (`late_phase`).
"""

from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from continuum_liquid_amd import runtime, scenes  # noqa: E402
from continuum_liquid_amd.Solver.Kernel import solver as host_solver  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "ms/step @256^3 (~30M particles); pressure-solve ms to 1e-6 rel residual"


def algorithmic_bytes(n_particles: int, cells: int, swept_cells: float) -> dict:
    """Compulsory HBM bytes per launch unit in the reference FP32 layout (each
    array counted once): binning 60N + 16C, P2G + occupancy 24N + 24C, one
    red-black sweep of the finest level 28 bytes per swept cell (p, b, mask,
    diagonal and three weights; the sweep visits the rows holding an interior
    fluid cell), fused G2P + advection 48N + 24C."""
    N, C = n_particles, cells
    return {
        "binning": 60 * N + 16 * C,
        "p2g": 24 * N + 24 * C,
        "rbgs_fine_sweep": int(28 * swept_cells),
        "particle_update": 48 * N + 24 * C,
    }


def measured_traffic(config: int, solver: str, kernel: str, swept: float):
    """HBM bytes per launch of `kernel` measured with rocprofv3 (FETCH_SIZE
    scaled by the copy calibration, plus WRITE_SIZE) for this workload, read
    from the `bench-traffic:` lines of the committed profile summaries
    (profiles/*.md); None when no profile covers it.  The counters cannot be
    read from inside this process, so the figure is as current as the profile
    named in the returned source."""
    for path in sorted((ROOT / "profiles").glob("*.md")):
        for line in path.read_text().splitlines():
            if not line.startswith("bench-traffic:"):
                continue
            fields = dict(item.split("=", 1) for item in line.split()[1:])
            if fields.get("config") != str(config) or fields.get("solver") != solver or fields.get("kernel") != kernel:
                continue
            if "bytes_per_launch" in fields:
                return int(float(fields["bytes_per_launch"])), f"profiles/{path.name}"
            if "bytes_per_swept_cell" in fields:
                return int(float(fields["bytes_per_swept_cell"]) * swept), f"profiles/{path.name}"
    return None, None


# steps of each configuration's run (the timed steps cover the first ones;
# the late phase times the last ones)
CONFIG_STEPS = {1: 100, 2: 200, 3: 100, 4: 200, 5: 50}


def coarsest_cells(shape) -> int:
    """Cells of the coarsest multigrid level the library builds for `shape`
    (coarse size (n - 2 + 1) // 2 + 2 while every axis stays >= 8)."""
    dims = [int(n) for n in shape]
    while True:
        nxt = [(max(0, n - 2) + 1) // 2 + 2 for n in dims]
        if min(nxt) < 8:
            return dims[0] * dims[1] * dims[2]
        dims = nxt


def krylov_coarse_sweeps(shape) -> int:
    """Extra coarsest-level sweeps of the FCG preconditioner's cycle: 8 when
    the coarsest level has at most 1000 cells (the two last levels then run in
    one resident work-group), else 16.  At 256^3 (coarsest 10^3) 8 sweeps give
    the same 9.0 iterations per solve as 16 and 7.19 against 7.31 ms per step
    (three alternating runs each); at 512x256x256 (coarsest 18x10x10) 8 took
    9.85 iterations against 9.25."""
    return 8 if coarsest_cells(shape) <= 1000 else 16


def krylov_cycle_shape(shape) -> int:
    """vcycle_from of the FCG preconditioner's cycle: the coarse-grid
    correction twice below level 1 only and once below level 0 (-2) when the
    last two levels run in the resident work-group (coarsest level of at most
    1000 cells), else twice below levels 0 and 1 (2).  The first visits the
    levels of 66^3 and below half as often per cycle: at 256^3 12.2 FCG
    iterations instead of 9.0 but 6.71 against 7.00 ms per step (128^3 2.38
    against 2.55, 64^3 1.19 against 1.47; no breakdown over the 100 steps of
    each workload); at 512x256x256 (coarsest 18x10x10) it lost, 12.99 against
    12.82 ms."""
    return -2 if coarsest_cells(shape) <= 1000 else 2


def solver_settings(name: str, tolerance: float) -> dict:
    """W(2, 2) cycles for the plain FP32 and mixed-refinement solves.  The
    plain FP32 baseline keeps the reference's 64 extra coarsest sweeps and the
    W-cycle at every level; the refined solves use 16 extra sweeps, which
    converge in the same number of cycles (measured on the 64^3 dam break: 38
    cycles to 1e-8 with 64 or 16), and the W-cycle on the two finest levels
    with the V-cycle below.  The FCG solves precondition with W(1, 1) cycles:
    at 256^3 they take 16.8 iterations per step against 14.2 with W(2, 2), and
    15.5 against 17.2 ms/step (sweep over smoothing 0-3, V/W level and
    coarsest sweeps on MI355X; the level shape is then chosen by
    krylov_cycle_shape).  Their cycles weigh every coarse-grid
    correction by 1.6 (over-correction of the piecewise-constant
    prolongation): 9.0 FCG iterations per 256^3 step against 17.2 unweighted,
    no breakdown over 20-100 steps of the 64^3, 128^3, 256^3 and 512x256x256
    workloads; 1.7 is slower over 30 steps at 256^3, 1.8 breaks down in about
    half of the solves, and with the V-cycle from level 1 even 1.6 does."""
    base = {"pre_smooth": 2, "post_smooth": 2, "coarse_extra": 64, "tolerance": tolerance}
    if name == "fp32":
        return {**base, "mode": "fp32", "max_cycles": 60}
    if name == "mixed":
        return {**base, "mode": "mixed", "max_cycles": 60, "coarse_dtype": 0, "coarse_extra": 16, "vcycle_from": 2}
    if name == "fcg":
        return {**base, "mode": "fcg", "max_cycles": 60, "coarse_dtype": 0, "coarse_extra": 16, "vcycle_from": 2,
                "pre_smooth": 1, "post_smooth": 1, "coarse_weight": 1.6}
    if name == "learned":
        return {**base, "mode": "learned", "max_cycles": 60, "coarse_dtype": 0, "coarse_extra": 16, "vcycle_from": 2,
                "pre_smooth": 1, "post_smooth": 1, "coarse_weight": 1.6}
    if name == "fixed":
        return {**base, "mode": "fixed", "max_cycles": 1, "post_smooth": 4}
    raise SystemExit(f"unknown solver {name}")


def cpu_baseline(scene, params, steps: int) -> dict:
    """Time the CPU oracle on the same workload: one untimed step, then `steps`."""
    lib = runtime.bind_library(ctypes.CDLL(str(ROOT / "oracle" / "libcliquid_oracle.so")))
    cores = len(os.sched_getaffinity(0))
    with runtime.Context(lib, scene.shape, scene.dx, scene.particle_count) as ctx:
        ctx.set_particles(scene.positions, scene.velocities)
        host_solver.step(ctx, params)
        t0 = time.perf_counter()
        cycles = []
        for _ in range(steps):
            cycles.append(host_solver.step(ctx, params)["cycles"])
        ms = (time.perf_counter() - t0) * 1e3 / steps
    return {
        "value": round(ms, 3),
        "unit": "ms/step",
        "cores": cores,
        "kind": "port",
        "sample": f"{steps} step(s) of the same {scene.name} workload after 1 untimed step, "
                  f"W-cycles per step {cycles}; OpenMP over {cores} threads on {platform.processor() or 'host'}",
    }


def cpu_reference_settings(scene, warmup: int = 3) -> dict:
    """The oracle at the reference's own solver settings (one W(2, 4) cycle
    with 64 extra coarsest sweeps, pure FLIP), per section with the
    reference's section names (Solver/Kernel/timing.py), after `warmup`
    untimed steps; one timed step."""
    from continuum_liquid_amd.Solver.Kernel import timing

    lib = runtime.bind_library(ctypes.CDLL(str(ROOT / "oracle" / "libcliquid_oracle.so")))
    params = host_solver.StepParams(dt=1.0 / 240.0, gravity=scene.gravity, flip_ratio=1.0,
                                    solver=solver_settings("fixed", 1e-6))
    with runtime.Context(lib, scene.shape, scene.dx, scene.particle_count) as ctx:
        ctx.set_particles(scene.positions, scene.velocities)
        for _ in range(warmup):
            host_solver.step(ctx, params)
        timings = timing.RunTimings()
        t0 = time.perf_counter()
        host_solver.step(ctx, params, timings)
        ms = (time.perf_counter() - t0) * 1e3
    sections = {name: round(seconds * 1e3, 3) for (_, name), (_, seconds) in timings.entries.items()}
    return {"value": round(ms, 3), "unit": "ms/step", "cores": len(os.sched_getaffinity(0)), "kind": "port",
            "sample": f"1 step after {warmup} untimed steps, reference settings: one W(2,4) cycle with 64 extra "
                      "coarsest sweeps, pure FLIP; sections timed with device synchronisation",
            "sections_ms": sections}


def calibrate_copy(nbytes: int) -> int:
    """Copy nbytes between two device buffers three times (torch's copy
    kernel: nbytes read and nbytes written per launch)."""
    import torch

    n = nbytes // 4
    src = torch.arange(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": n * 4, "launches": 3}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=3, choices=(1, 2, 3, 4, 5),
                    help="1: 64^3, 2: 128^3, 3: 256^3 (default), 4: 128^3 with the learned solve, 5: 512x256x256")
    ap.add_argument("--solver", default="fcg", choices=("fp32", "mixed", "fcg", "learned", "fixed"))
    ap.add_argument("--weights", type=Path, default=None,
                    help="learned solver: parameter file (default: the package's trained weights)")
    ap.add_argument("--stall-ratio", type=float, default=0.5,
                    help="learned solver: residual ratio above which the W-cycle takes over")
    ap.add_argument("--tolerance", type=float, default=1e-6)
    ap.add_argument("--coarse-extra", type=int, default=None, help="extra sweeps on the coarsest level")
    ap.add_argument("--smooth", type=int, nargs=2, default=None, metavar=("PRE", "POST"),
                    help="smoothing sweeps before and after the coarse-grid correction")
    ap.add_argument("--vcycle-from", type=int, default=None,
                    help="refined solves: W-cycle above this level, V-cycle from it down (0: W everywhere)")
    ap.add_argument("--coarse-weight", type=float, default=None,
                    help="weight of the cycle's coarse-grid corrections (1: the reference's plain sum)")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 check solves after the timed steps")
    ap.add_argument("--no-late", dest="late", action="store_false",
                    help="skip running on to the configuration's last step (late-phase timing)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE to bytes) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy)

    if args.config == 4 and args.solver == "fcg":
        args.solver = "learned"
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): "
              "the step has no multi-GPU path in this build", file=sys.stderr)
        return 2

    lib = runtime.load_library()
    if runtime.device_count(lib) < 1:
        print("no HIP device visible", file=sys.stderr)
        return 1

    scene = scenes.bench_scene(args.config)
    settings = solver_settings(args.solver, args.tolerance)
    if settings["mode"] in ("fcg", "learned"):
        settings["coarse_extra"] = krylov_coarse_sweeps(scene.shape)
        settings["vcycle_from"] = krylov_cycle_shape(scene.shape)
    if args.vcycle_from is not None:
        settings["vcycle_from"] = args.vcycle_from
    if args.coarse_extra is not None:
        settings["coarse_extra"] = args.coarse_extra
    if args.smooth is not None:
        settings["pre_smooth"], settings["post_smooth"] = args.smooth
    if args.coarse_weight is not None:
        settings["coarse_weight"] = args.coarse_weight
    params = host_solver.StepParams(dt=1.0 / 240.0, gravity=scene.gravity, flip_ratio=0.95, solver=settings)
    print(f"# {scene.name}: {scene.particle_count} particles, grid {scene.shape}, solver {args.solver}",
          file=sys.stderr, flush=True)

    with runtime.Context(lib, scene.shape, scene.dx, scene.particle_count) as ctx:
        ctx.set_particles(scene.positions, scene.velocities)
        if args.solver == "learned":
            from continuum_liquid_amd import learned

            ctx.attach_net(learned.load_weights(lib, scene.shape, scene.dx, args.weights or learned.DEFAULT_WEIGHTS),
                           args.stall_ratio)
        fallback, swept = [], []
        for _ in range(args.warmup):
            host_solver.step(ctx, params)
        ctx.sync()
        ctx.kernel_timing(True)
        solve_ms, relres, cycles, recres, restarts = [], [], [], [], []
        t0 = time.perf_counter()
        for i in range(args.steps):
            stats = host_solver.step(ctx, params)
            solve_ms.append(stats["solve_ms"])
            relres.append(stats["relative_residual"])
            recres.append(stats["recurrence_residual"] / stats["rhs_norm"] if stats["rhs_norm"] > 0 else 0.0)
            restarts.append(stats["restarts"])
            cycles.append(stats["cycles"])
            fallback.append(stats["fallback_cycles"])
            swept.append(stats["active_cells"])
        ctx.sync()
        elapsed = time.perf_counter() - t0
        ktimes = ctx.kernel_times()
        ctx.kernel_timing(False)
        # untimed from here on: the FP64 check of the last timed solve, then
        # the rest of the configuration's steps with the last 20 timed apart
        checks = [] if args.no_check else [ctx.fp64_check_solve(1e-12)]
        late = None
        total = CONFIG_STEPS[args.config]
        done = args.warmup + args.steps
        if args.late and total > done:
            tail = min(20, total - done)
            for _ in range(total - done - tail):
                host_solver.step(ctx, params)
            ctx.sync()
            ctx.kernel_timing(True)
            late_solve, late_res = [], []
            t1 = time.perf_counter()
            for _ in range(tail):
                st = host_solver.step(ctx, params)
                late_solve.append(st["solve_ms"])
                late_res.append(st["relative_residual"])
            ctx.sync()
            late_elapsed = time.perf_counter() - t1
            late_k = ctx.kernel_times()
            ctx.kernel_timing(False)
            late = {"steps": [total - tail + 1, total], "ms_per_step": round(late_elapsed * 1e3 / tail, 3),
                    "pressure_solve_ms": round(float(np.mean(late_solve)), 3),
                    "relative_residual_max": float(np.max(late_res)),
                    "kernel_avg_ms": {k: round(t / u, 5) for k, (t, u) in late_k.items() if u}}
            relres.extend(late_res)
            if not args.no_check:
                checks.append(ctx.fp64_check_solve(1e-12))

    ms_per_step = elapsed * 1e3 / args.steps
    n, cells = scene.particle_count, int(np.prod(scene.shape))
    nbytes = algorithmic_bytes(n, cells, float(np.mean(swept)))
    kernels = {}
    for name, (total_ms, units) in ktimes.items():
        if units == 0:
            continue
        avg_ms = total_ms / units
        gbs = nbytes[name] / (avg_ms * 1e-3) / 1e9
        kernels[name] = {
            "avg_ms": round(avg_ms, 5),
            "launch_units": units,
            "share_of_step": round(total_ms / (elapsed * 1e3), 4),
            "algorithmic_bytes": nbytes[name],
            "achieved_GBs": round(gbs, 1),
            "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4),
        }
    dominant = max(kernels, key=lambda k: ktimes[k][0])
    d = kernels[dominant]
    traffic, traffic_source = measured_traffic(args.config, args.solver, dominant, float(np.mean(swept)))

    result = {
        "metric": METRIC,
        "value": round(ms_per_step, 3),
        "unit": "ms/step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.solver in ("fp32", "fixed") else "fp32+fp64",
        "data": "synthetic",
        "config": {
            "workload": f"{scene.name}, dx=1/{round(1 / scene.dx)}, {n} FP32 particles (8 per fluid cell, "
                        f"stratified jitter, seed 0), dt=1/240, FLIP 0.95, solver {args.solver} with ({settings['pre_smooth']},{settings['post_smooth']}) cycles, vcycle_from {settings.get('vcycle_from', 0)}, to "
                        f"{args.tolerance:g} relative residual",
            "grid": list(scene.shape),
            "particles": n,
            "solver": args.solver,
            "vcycle_from": settings.get("vcycle_from", 0),
            "smoothing": [settings["pre_smooth"], settings["post_smooth"]],
            "coarse_extra": settings["coarse_extra"],
            "coarse_weight": settings.get("coarse_weight", 1.0),
        },
        "particle_steps_per_s": round(n * args.steps / elapsed, 1),
        "pressure_solve_ms": round(float(np.mean(solve_ms)), 3),
        "w_cycles_per_step": round(float(np.mean(cycles)), 2),
        # the solver's reported residual is recomputed as ||b - A x|| from its
        # FP64 iterate; the recurrence value is the iteration's own estimate
        "true_relative_residual_max": float(np.max(relres)),
        **({"recurrence_relative_residual_max": float(np.max(recres)), "restarts": int(np.sum(restarts))}
           if args.solver in ("fcg", "learned") else {}),
        **({"fp64_error_max": float(max(c["error_max"] for c in checks)),
            "fp64_error_l2_max": float(max(c["error_l2"] for c in checks)),
            "fp64_check": [{k: (round(v, 3) if k == "solve_ms" else v) for k, v in c.items()} for c in checks]}
           if checks else {}),
        **({"late_phase": late} if late else {}),
        "swept_cells_per_sweep": int(np.mean(swept)),
        **({"w_cycle_fallback_per_step": round(float(np.mean(fallback)), 2)} if args.solver == "learned" else {}),
        # FCG iterations run with weight 1 after the weighted cycle broke down
        **({"unweighted_fallback_per_step": round(float(np.mean(fallback)), 2),
            "solves_with_fallback": int(np.count_nonzero(fallback))} if args.solver == "fcg" else {}),
        "roofline": {
            "kernel": dominant,
            "bound": "hbm",
            "achieved": d["achieved_GBs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": d["frac_hbm_peak"],
            "traffic": traffic,
            "traffic_source": traffic_source,
        },
        "kernels": kernels,
    }
    if not args.no_cpu:
        print("# timing the CPU oracle on the same workload", file=sys.stderr, flush=True)
        cpu_params = params
        if args.solver == "learned":  # the oracle has no network: it runs the FCG solve of the same workload
            cpu_params = host_solver.StepParams(dt=params.dt, gravity=params.gravity, flip_ratio=params.flip_ratio,
                                                solver={**solver_settings("fcg", args.tolerance),
                                                        "coarse_extra": settings["coarse_extra"]})
        result["cpu_baseline"] = cpu_baseline(scene, cpu_params, args.cpu_steps)
        result["cpu_baseline_reference_settings"] = cpu_reference_settings(scene)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
