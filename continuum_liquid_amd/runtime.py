"""ctypes binding of the C-ABI in include/cliquid.h.

`load_library()` loads the gfx950 HIP library built in-tree
(continuum_liquid_amd/libcliquid_hip.so) and fails loudly when it is missing
or when no GPU is visible: there is no CPU fallback.  `bind_library()` sets the
argument types on any CDLL that exports the ABI; `Context` owns one
simulation grid on the device and raises `ClqError` for every non-zero status.

The stage methods map one-to-one onto the reference's stage wrappers
(Solver/Kernel/particles.py, multigrid.py, pressure_solve.py,
velocity_update.py, boundary_conditions.py); see the C header for citations.
"""

from __future__ import annotations

import ctypes
import threading
from ctypes import POINTER, byref, c_char_p, c_double, c_int32, c_int64, c_uint8, c_void_p
from pathlib import Path

import numpy as np

PACKAGE_DIR = Path(__file__).resolve().parent
HIP_LIBRARY = PACKAGE_DIR / "libcliquid_hip.so"
ABI_VERSION = 4

# enum clq_status
OK, ERR_INVALID, ERR_DEVICE, ERR_NOMEM, ERR_UNSUPPORTED, ERR_NOT_CONVERGED = range(6)

# enum clq_field
FIELDS = {
    "u": 0, "v": 1, "w": 2, "u_old": 3, "v_old": 4, "w_old": 5, "fluid_fraction": 6, "pressure": 7, "rhs": 8,
    "diagonal": 9, "weight_x": 10, "weight_y": 11, "weight_z": 12, "bin_counts": 13, "bin_offsets": 14,
    "obstacle": 15, "pressure_fp64": 16,
}
FIELD_DTYPES = {"bin_counts": np.int32, "bin_offsets": np.int32, "obstacle": np.uint8, "pressure_fp64": np.float64}
LEVEL_FIELDS = {"p": 0, "b": 1, "mask": 2, "diagonal": 3, "weight_x": 4, "weight_y": 5, "weight_z": 6}

# enum clq_solve_mode
SOLVE_FIXED_CYCLES, SOLVE_FP32_TOLERANCE, SOLVE_MIXED_REFINEMENT, SOLVE_LEARNED_FCG, SOLVE_FCG = range(5)
SOLVE_MODES = {
    "fixed": SOLVE_FIXED_CYCLES,
    "fp32": SOLVE_FP32_TOLERANCE,
    "mixed": SOLVE_MIXED_REFINEMENT,
    "learned": SOLVE_LEARNED_FCG,
    "fcg": SOLVE_FCG,
}
STAGES = ("bin", "p2g", "pressure_solve", "g2p_advect")
# enum clq_source_type
SOURCE_TYPES = {"GEOMETRY": 0, "INFLOW": 1, "OUTFLOW": 2}
KERNEL_CLASSES = ("binning", "p2g", "rbgs_fine_sweep", "particle_update")


class ClqError(RuntimeError):
    """A non-zero status from the C-ABI, with clq_last_error()'s message."""

    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where} failed with status {status}: {message}")
        self.status = status


class SolveParams(ctypes.Structure):
    _fields_ = [
        ("mode", c_int32),
        ("pre_smooth", c_int32),
        ("post_smooth", c_int32),
        ("coarse_extra", c_int32),
        ("max_cycles", c_int32),
        ("coarse_dtype", c_int32),
        ("vcycle_from", c_int32),
        ("coarse_weight", ctypes.c_float),
        ("tolerance", c_double),
    ]


class SolveStats(ctypes.Structure):
    _fields_ = [
        ("cycles", c_int32),
        ("fallback_cycles", c_int32),
        ("rhs_norm", c_double),
        ("initial_residual", c_double),
        ("final_residual", c_double),
        ("relative_residual", c_double),
        ("solve_ms", c_double),
        ("active_cells", c_int64),
        ("recurrence_residual", c_double),
        ("restarts", c_int32),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class CheckStats(ctypes.Structure):
    _fields_ = [
        ("iterations", c_int32),
        ("converged", c_int32),
        ("relative_residual", c_double),
        ("error_l2", c_double),
        ("error_max", c_double),
        ("solve_ms", c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = c_void_p
_D3 = POINTER(c_double)
_SIGNATURES = {
    "clq_abi_version": ([], ctypes.c_int),
    "clq_device_count": ([POINTER(c_int32)], ctypes.c_int),
    "clq_create": ([c_int32, c_int32, c_int32, c_double, c_int64, POINTER(_P)], ctypes.c_int),
    "clq_destroy": ([_P], ctypes.c_int),
    "clq_sync": ([_P], ctypes.c_int),
    "clq_set_particles": ([_P, _P, _P, c_int64], ctypes.c_int),
    "clq_get_particles": ([_P, _P, _P, c_int64], ctypes.c_int),
    "clq_get_previous_positions": ([_P, _P, c_int64], ctypes.c_int),
    "clq_get_particle_count": ([_P, POINTER(c_int64)], ctypes.c_int),
    "clq_voxelise_mesh": ([_P, _P, c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double, _P], ctypes.c_int),
    "clq_host_alloc": ([_P, c_int64, POINTER(_P)], ctypes.c_int),
    "clq_host_free": ([_P, _P], ctypes.c_int),
    "clq_positions_to_host_async": ([_P, _P, c_int64, POINTER(c_int64), POINTER(c_int64)], ctypes.c_int),
    "clq_host_copy_wait": ([_P, c_int64], ctypes.c_int),
    "clq_set_field": ([_P, c_int32, _P, c_int64], ctypes.c_int),
    "clq_get_field": ([_P, c_int32, _P, c_int64], ctypes.c_int),
    "clq_level_count": ([_P, POINTER(c_int32)], ctypes.c_int),
    "clq_level_shape": ([_P, c_int32, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)], ctypes.c_int),
    "clq_set_level_field": ([_P, c_int32, c_int32, _P, c_int64], ctypes.c_int),
    "clq_get_level_field": ([_P, c_int32, c_int32, _P, c_int64], ctypes.c_int),
    "clq_bin_particles": ([_P], ctypes.c_int),
    "clq_particle_to_grid": ([_P, c_double], ctypes.c_int),
    "clq_grid_obstacle_bc": ([_P], ctypes.c_int),
    "clq_grid_domain_bc": ([_P], ctypes.c_int),
    "clq_update_velocity": ([_P, _D3, _D3, c_double, c_double, c_double], ctypes.c_int),
    "clq_pressure_rhs": ([_P, c_double, c_double], ctypes.c_int),
    "clq_build_operator": ([_P], ctypes.c_int),
    "clq_pressure_solve": ([_P, POINTER(SolveParams), POINTER(SolveStats)], ctypes.c_int),
    "clq_fp64_check_solve": ([_P, c_double, c_int32, POINTER(CheckStats)], ctypes.c_int),
    "clq_w_cycle": ([_P, c_int32, c_int32, c_int32], ctypes.c_int),
    "clq_smooth": ([_P, c_int32, c_int32], ctypes.c_int),
    "clq_restrict": ([_P, c_int32], ctypes.c_int),
    "clq_prolongate": ([_P, c_int32], ctypes.c_int),
    "clq_coarse_smooth": ([_P, c_int32], ctypes.c_int),
    "clq_neumann_bc": ([_P], ctypes.c_int),
    "clq_residual_norm": ([_P, POINTER(c_double), POINTER(c_double)], ctypes.c_int),
    "clq_pressure_correction": ([_P, c_double, c_double], ctypes.c_int),
    "clq_grid_to_particle": ([_P, c_double], ctypes.c_int),
    "clq_save_previous_positions": ([_P], ctypes.c_int),
    "clq_advect_particles": ([_P, c_double], ctypes.c_int),
    "clq_particle_domain_bc": ([_P], ctypes.c_int),
    "clq_particle_obstacle_bc": ([_P], ctypes.c_int),
    "clq_grid_to_particle_advect": ([_P, c_double, c_double], ctypes.c_int),
    "clq_voxelise": ([_P, c_int32, c_double, _P, _P], ctypes.c_int),
    "clq_set_sources": ([_P, c_int32, _P, _P, _P, _P], ctypes.c_int),
    "clq_apply_sources": ([_P, c_int32, _P, POINTER(c_int64)], ctypes.c_int),
    "clq_stage_times": ([_P, POINTER(c_double), c_int32], ctypes.c_int),
    "clq_kernel_timing": ([_P, c_int32], ctypes.c_int),
    "clq_kernel_times": ([_P, POINTER(c_double), POINTER(c_int64), c_int32], ctypes.c_int),
}
# The learned preconditioner's entry points (GPU library only: the CPU
# oracle restates the FLIP step, the network's CPU check is a torch model).
_NET_SIGNATURES = {
    "clq_net_param_count": ([], c_int64),
    "clq_net_create": ([c_int32, c_int32, c_int32, c_double, POINTER(_P)], ctypes.c_int),
    "clq_net_destroy": ([_P], ctypes.c_int),
    "clq_net_init": ([_P, ctypes.c_uint64], ctypes.c_int),
    "clq_net_set_params": ([_P, _P, c_int64], ctypes.c_int),
    "clq_net_get_params": ([_P, _P, c_int64], ctypes.c_int),
    "clq_net_forward": ([_P, _P, _P, _P, _P], ctypes.c_int),
    "clq_net_backward": ([_P, _P, _P, c_int64], ctypes.c_int),
    "clq_net_train_step": ([_P, c_int32, _P, _P, _P, _P, c_double, POINTER(c_double)], ctypes.c_int),
    "clq_attach_net": ([_P, _P, c_double], ctypes.c_int),
}
# Entry points on device pointers (GPU library only).
_DEVICE_SIGNATURES = {
    "clq_residual_norm_noalloc": ([_P, _P, _P, _P, _P, _P, _P, c_int32, c_int32, c_int32, c_double, _P, _P],
                                  ctypes.c_int),
}
EXPORTED_SYMBOLS = ("clq_last_error", *_SIGNATURES, *_NET_SIGNATURES, *_DEVICE_SIGNATURES)
NET_SYMBOLS = tuple(_NET_SIGNATURES)
GPU_ONLY_SYMBOLS = (*_NET_SIGNATURES, *_DEVICE_SIGNATURES)


def bind_library(lib: ctypes.CDLL) -> ctypes.CDLL:
    """Set restype/argtypes for every ABI entry point of `lib`."""
    lib.clq_last_error.restype = c_char_p
    lib.clq_last_error.argtypes = []
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    for name, (args, res) in {**_NET_SIGNATURES, **_DEVICE_SIGNATURES}.items():
        if hasattr(lib, name):
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
    version = lib.clq_abi_version()
    if version != ABI_VERSION:
        raise RuntimeError(f"library ABI version {version} != expected {ABI_VERSION}")
    return lib


_lock = threading.Lock()
_hip_library: ctypes.CDLL | None = None


def load_library(path: str | Path | None = None) -> ctypes.CDLL:
    """Load the HIP library (once per process).  Raises when it is absent."""
    global _hip_library
    with _lock:
        if _hip_library is not None and path is None:
            return _hip_library
        lib_path = Path(path) if path is not None else HIP_LIBRARY
        if not lib_path.exists():
            raise RuntimeError(
                f"{lib_path} is missing: build it with `python -m continuum_liquid_amd.build` "
                "(the Continuum Liquid HIP step has no CPU fallback)"
            )
        lib = bind_library(ctypes.CDLL(str(lib_path)))
        if path is None:
            _hip_library = lib
        return lib


def _ptr(array: np.ndarray) -> c_void_p:
    return c_void_p(array.ctypes.data)


class Context:
    """One simulation grid (and its particles) resident on the device."""

    def __init__(self, lib: ctypes.CDLL, shape, dx: float, capacity: int = 1):
        self.lib = lib
        self.shape = tuple(int(n) for n in shape)
        self.dx = float(dx)
        handle = c_void_p()
        self._check(lib.clq_create(*self.shape, self.dx, int(capacity), byref(handle)), "clq_create")
        self.handle = handle
        levels = c_int32()
        self._check(lib.clq_level_count(self.handle, byref(levels)), "clq_level_count")
        self.level_shapes = []
        for level in range(levels.value):
            a, b, d = c_int32(), c_int32(), c_int32()
            self._check(lib.clq_level_shape(self.handle, level, byref(a), byref(b), byref(d)), "clq_level_shape")
            self.level_shapes.append((a.value, b.value, d.value))

    # -- plumbing ---------------------------------------------------------------------------------
    def _check(self, status: int, where: str) -> None:
        if status != OK:
            message = self.lib.clq_last_error()
            raise ClqError(status, where, message.decode() if message else "")

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.clq_destroy(self.handle)
            self.handle = None
        self._net = None

    def __enter__(self) -> "Context":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    @property
    def cells(self) -> int:
        return self.shape[0] * self.shape[1] * self.shape[2]

    def sync(self) -> None:
        self._check(self.lib.clq_sync(self.handle), "clq_sync")

    # -- data movement ----------------------------------------------------------------------------
    def set_particles(self, positions: np.ndarray, velocities: np.ndarray | None = None) -> None:
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        vel = None if velocities is None else np.ascontiguousarray(velocities, dtype=np.float32).reshape(-1, 3)
        if vel is not None and vel.shape != pos.shape:
            raise ValueError("positions and velocities must have the same shape")
        self._check(
            self.lib.clq_set_particles(self.handle, _ptr(pos), None if vel is None else _ptr(vel), len(pos)),
            "clq_set_particles",
        )

    @property
    def particle_count(self) -> int:
        n = c_int64()
        self._check(self.lib.clq_get_particle_count(self.handle, byref(n)), "clq_get_particle_count")
        return n.value

    def get_particles(self, out_positions: np.ndarray | None = None, out_velocities: np.ndarray | None = None):
        n = self.particle_count
        pos = out_positions if out_positions is not None else np.empty((n, 3), np.float32)
        vel = out_velocities if out_velocities is not None else np.empty((n, 3), np.float32)
        self._check(self.lib.clq_get_particles(self.handle, _ptr(pos), _ptr(vel), n), "clq_get_particles")
        return pos, vel

    def get_positions(self, out: np.ndarray | None = None) -> np.ndarray:
        n = self.particle_count
        pos = out if out is not None else np.empty((n, 3), np.float32)
        self._check(self.lib.clq_get_particles(self.handle, _ptr(pos), None, n), "clq_get_particles")
        return pos

    def voxelise_mesh(self, triangles: np.ndarray, origin, mask: np.ndarray | None = None) -> np.ndarray:
        """Ray-parity voxelisation of one geometry entry (voxelise_mesh.cl:24-48):
        `triangles` (T, 3, 3) float32 world coordinates, `origin` the world
        position of the grid's corner.  The cells an odd number of triangles
        surround are set in `mask` (a bool or uint8 grid array, updated in
        place and returned; a new zero mask when None)."""
        tris = np.ascontiguousarray(triangles, dtype=np.float32)
        if tris.size and (tris.ndim != 3 or tris.shape[1:] != (3, 3)):
            raise ValueError(f"triangles must have shape (T, 3, 3), got {tris.shape}")
        if mask is None:
            mask = np.zeros(self.shape, np.uint8)
        if mask.shape != self.shape or mask.dtype not in (np.uint8, np.bool_) or not mask.flags.c_contiguous:
            raise ValueError("mask must be a C-contiguous uint8 or bool array of the grid's shape")
        ox, oy, oz = (float(v) for v in origin)
        self._check(self.lib.clq_voxelise_mesh(self.handle, _ptr(tris), len(tris), ox, oy, oz, mask.ctypes.data),
                    "clq_voxelise_mesh")
        return mask

    def host_alloc(self, rows: int) -> np.ndarray:
        """A float32 (rows, 3) array in pinned host memory of the context's
        device (output.py:56-75); release it with host_free."""
        ptr = ctypes.c_void_p()
        rows = max(1, int(rows))
        self._check(self.lib.clq_host_alloc(self.handle, rows * 12, byref(ptr)), "clq_host_alloc")
        flat = np.ctypeslib.as_array((ctypes.c_float * (rows * 3)).from_address(ptr.value))
        return flat.reshape(rows, 3)

    def host_free(self, array: np.ndarray) -> None:
        self._check(self.lib.clq_host_free(self.handle, array.ctypes.data), "clq_host_free")

    def positions_to_host_async(self, out: np.ndarray) -> tuple[int, int]:
        """Starts the copy of the current positions into the pinned array
        `out` (output.py:178-195) and returns (count, ticket) at once; the
        rows [0, count) are valid after wait_host_copy(ticket)."""
        if out.dtype != np.float32 or out.ndim != 2 or out.shape[1] != 3 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float32 (rows, 3) array")
        n, ticket = c_int64(), c_int64()
        self._check(self.lib.clq_positions_to_host_async(self.handle, out.ctypes.data, out.shape[0], byref(n),
                                                         byref(ticket)), "clq_positions_to_host_async")
        return n.value, ticket.value

    def wait_host_copy(self, ticket: int) -> None:
        self._check(self.lib.clq_host_copy_wait(self.handle, int(ticket)), "clq_host_copy_wait")

    def get_previous_positions(self) -> np.ndarray:
        n = self.particle_count
        pos = np.empty((n, 3), np.float32)
        self._check(self.lib.clq_get_previous_positions(self.handle, _ptr(pos), n), "clq_get_previous_positions")
        return pos

    def set_field(self, name: str, values: np.ndarray) -> None:
        arr = np.ascontiguousarray(values, dtype=FIELD_DTYPES.get(name, np.float32))
        if arr.size != self.cells:
            raise ValueError(f"field {name} needs {self.cells} values, got {arr.size}")
        self._check(self.lib.clq_set_field(self.handle, FIELDS[name], _ptr(arr), arr.nbytes), "clq_set_field")

    def get_field(self, name: str) -> np.ndarray:
        arr = np.empty(self.shape, FIELD_DTYPES.get(name, np.float32))
        self._check(self.lib.clq_get_field(self.handle, FIELDS[name], _ptr(arr), arr.nbytes), "clq_get_field")
        return arr

    def set_level_field(self, level: int, name: str, values: np.ndarray) -> None:
        arr = np.ascontiguousarray(values, dtype=np.float32)
        self._check(
            self.lib.clq_set_level_field(self.handle, level, LEVEL_FIELDS[name], _ptr(arr), arr.nbytes),
            "clq_set_level_field",
        )

    def get_level_field(self, level: int, name: str) -> np.ndarray:
        arr = np.empty(self.level_shapes[level], np.float32)
        self._check(
            self.lib.clq_get_level_field(self.handle, level, LEVEL_FIELDS[name], _ptr(arr), arr.nbytes),
            "clq_get_level_field",
        )
        return arr

    # -- stages -------------------------------------------------------------------------------------
    def bin_particles(self) -> None:
        self._check(self.lib.clq_bin_particles(self.handle), "clq_bin_particles")

    def particle_to_grid(self, particles_per_cell: float = 8.0) -> None:
        self._check(self.lib.clq_particle_to_grid(self.handle, float(particles_per_cell)), "clq_particle_to_grid")

    def grid_obstacle_bc(self) -> None:
        self._check(self.lib.clq_grid_obstacle_bc(self.handle), "clq_grid_obstacle_bc")

    def grid_domain_bc(self) -> None:
        self._check(self.lib.clq_grid_domain_bc(self.handle), "clq_grid_domain_bc")

    def update_velocity(self, gravity, force, rho: float, nu: float, dt: float) -> None:
        g = (c_double * 3)(*[float(x) for x in gravity])
        f = (c_double * 3)(*[float(x) for x in force])
        self._check(
            self.lib.clq_update_velocity(self.handle, g, f, float(rho), float(nu), float(dt)), "clq_update_velocity"
        )

    def pressure_rhs(self, rho: float, dt: float) -> None:
        self._check(self.lib.clq_pressure_rhs(self.handle, float(rho), float(dt)), "clq_pressure_rhs")

    def build_operator(self) -> None:
        self._check(self.lib.clq_build_operator(self.handle), "clq_build_operator")

    def pressure_solve(
        self,
        mode: int = SOLVE_FIXED_CYCLES,
        *,
        pre_smooth: int = 2,
        post_smooth: int = 4,
        coarse_extra: int = 64,
        max_cycles: int = 1,
        tolerance: float = 1e-6,
        coarse_dtype: int = 0,
        vcycle_from: int = 0,
        coarse_weight: float = 1.0,
    ) -> dict:
        params = SolveParams(mode=int(mode), pre_smooth=pre_smooth, post_smooth=post_smooth,
                             coarse_extra=coarse_extra, max_cycles=max_cycles, coarse_dtype=coarse_dtype,
                             vcycle_from=vcycle_from, coarse_weight=coarse_weight, tolerance=tolerance)
        stats = SolveStats()
        self._check(self.lib.clq_pressure_solve(self.handle, byref(params), byref(stats)), "clq_pressure_solve")
        return stats.as_dict()

    def fp64_check_solve(self, tolerance: float = 1e-12, max_iterations: int = 20000) -> dict:
        """Independent FP64 solve of the current level-0 system and the error of
        the last pressure solve against it (clq_fp64_check_solve)."""
        stats = CheckStats()
        self._check(self.lib.clq_fp64_check_solve(self.handle, tolerance, max_iterations, byref(stats)),
                    "clq_fp64_check_solve")
        return stats.as_dict()

    def attach_net(self, net: "Network | None", stall_ratio: float = 0.5) -> None:
        """Use `net` in SOLVE_LEARNED_FCG (None detaches).  The context keeps a
        reference so the network outlives its use here."""
        if not hasattr(self.lib, "clq_attach_net"):
            raise ClqError(ERR_UNSUPPORTED, "clq_attach_net", "this library has no learned preconditioner")
        self._check(self.lib.clq_attach_net(self.handle, net.handle if net is not None else None, stall_ratio),
                    "clq_attach_net")
        self._net = net

    def w_cycle(self, pre_smooth: int = 2, post_smooth: int = 4, coarse_extra: int = 64) -> None:
        self._check(self.lib.clq_w_cycle(self.handle, pre_smooth, post_smooth, coarse_extra), "clq_w_cycle")

    def smooth(self, level: int, iterations: int) -> None:
        self._check(self.lib.clq_smooth(self.handle, level, iterations), "clq_smooth")

    def restrict(self, level: int) -> None:
        self._check(self.lib.clq_restrict(self.handle, level), "clq_restrict")

    def prolongate(self, level: int) -> None:
        self._check(self.lib.clq_prolongate(self.handle, level), "clq_prolongate")

    def coarse_smooth(self, iterations: int) -> None:
        self._check(self.lib.clq_coarse_smooth(self.handle, iterations), "clq_coarse_smooth")

    def neumann_bc(self) -> None:
        self._check(self.lib.clq_neumann_bc(self.handle), "clq_neumann_bc")

    def residual_norm(self) -> tuple[float, float]:
        r, b = c_double(), c_double()
        self._check(self.lib.clq_residual_norm(self.handle, byref(r), byref(b)), "clq_residual_norm")
        return r.value, b.value

    def pressure_correction(self, rho: float, dt: float) -> None:
        self._check(self.lib.clq_pressure_correction(self.handle, float(rho), float(dt)), "clq_pressure_correction")

    def grid_to_particle(self, flip_ratio: float = 1.0) -> None:
        self._check(self.lib.clq_grid_to_particle(self.handle, float(flip_ratio)), "clq_grid_to_particle")

    def save_previous_positions(self) -> None:
        self._check(self.lib.clq_save_previous_positions(self.handle), "clq_save_previous_positions")

    def advect_particles(self, dt: float) -> None:
        self._check(self.lib.clq_advect_particles(self.handle, float(dt)), "clq_advect_particles")

    def particle_domain_bc(self) -> None:
        self._check(self.lib.clq_particle_domain_bc(self.handle), "clq_particle_domain_bc")

    def particle_obstacle_bc(self) -> None:
        self._check(self.lib.clq_particle_obstacle_bc(self.handle), "clq_particle_obstacle_bc")

    def grid_to_particle_advect(self, flip_ratio: float, dt: float) -> None:
        self._check(
            self.lib.clq_grid_to_particle_advect(self.handle, float(flip_ratio), float(dt)),
            "clq_grid_to_particle_advect",
        )

    def voxelise(self, scale: int = 1, particles_per_cell: float = 8.0, density: bool = False):
        dims = tuple(scale * n for n in self.shape)
        occ = np.empty(dims, np.uint8)
        dens = np.empty(dims, np.float32) if density else None
        self._check(
            self.lib.clq_voxelise(self.handle, int(scale), float(particles_per_cell), _ptr(occ),
                                  None if dens is None else _ptr(dens)),
            "clq_voxelise",
        )
        return occ, dens

    def set_sources(self, sources) -> None:
        """Device copies of the sources' cells (clq_set_sources): dicts with
        `type` (GEOMETRY, INFLOW, OUTFLOW), a bool `mask` of the grid shape,
        `particle_density` and `velocity`.  Set the obstacle field first."""
        kinds = [s for s in sources if s.get("type") in SOURCE_TYPES]
        types = np.array([SOURCE_TYPES[s["type"]] for s in kinds], np.int32)
        masks = np.zeros((max(1, len(kinds)),) + self.shape, np.uint8)
        for q, s in enumerate(kinds):
            masks[q] = np.asarray(s["mask"], bool)
        dens = np.array([int(s.get("particle_density", 8)) for s in kinds], np.int32)
        vel = np.array([s.get("velocity", (0.0, 0.0, 0.0)) for s in kinds], np.float32).reshape(-1, 3)
        self._check(self.lib.clq_set_sources(self.handle, len(kinds), _ptr(types), _ptr(masks), _ptr(dens),
                                             _ptr(np.ascontiguousarray(vel))), "clq_set_sources")

    def apply_sources(self, rng: np.random.Generator, initial: bool = False) -> int:
        """The sources' removal and top-up on the device (clq_apply_sources),
        drawing from `rng` (a PCG64 Generator), which is advanced by the draws
        used: the same particles, bit for bit, as source_particles.apply_sources
        with the same generator.  Returns the particle count."""
        state = rng.bit_generator.state
        if state.get("bit_generator") != "PCG64" or state.get("has_uint32"):
            raise ValueError("apply_sources needs a PCG64 generator without a buffered 32-bit draw")
        s, inc = int(state["state"]["state"]), int(state["state"]["inc"])
        m64 = (1 << 64) - 1
        words = np.array([s >> 64, s & m64, inc >> 64, inc & m64], np.uint64)
        draws = c_int64()
        self._check(self.lib.clq_apply_sources(self.handle, int(bool(initial)), _ptr(words), byref(draws)),
                    "clq_apply_sources")
        if draws.value:
            rng.bit_generator.advance(draws.value)
        return self.particle_count

    def stage_times(self) -> dict:
        ms = (c_double * len(STAGES))()
        n = self.lib.clq_stage_times(self.handle, ms, len(STAGES))
        if n < 0:
            self._check(n, "clq_stage_times")
        return {STAGES[i]: ms[i] for i in range(n)}


    def kernel_timing(self, enable: bool) -> None:
        self._check(self.lib.clq_kernel_timing(self.handle, int(bool(enable))), "clq_kernel_timing")

    def kernel_times(self) -> dict:
        """{class: (total_ms, launch_units)} accumulated since kernel_timing(True)."""
        k = len(KERNEL_CLASSES)
        ms = (c_double * k)()
        units = (c_int64 * k)()
        n = self.lib.clq_kernel_times(self.handle, ms, units, k)
        if n < 0:
            self._check(n, "clq_kernel_times")
        return {KERNEL_CLASSES[i]: (ms[i], units[i]) for i in range(n)}


def device_count(lib: ctypes.CDLL) -> int:
    count = c_int32()
    status = lib.clq_device_count(byref(count))
    return count.value if status == OK else 0


class Network:
    """The learned preconditioner's U-Net for one pressure grid (shape and cell
    size dx), on the device (include/cliquid.h, clq_net_*).  Arrays are float32
    grid-shaped."""

    def __init__(self, lib: ctypes.CDLL, shape, dx: float, seed: int | None = 0):
        if not hasattr(lib, "clq_net_create"):
            raise ClqError(ERR_UNSUPPORTED, "clq_net_create", "this library has no learned preconditioner")
        self.lib = lib
        self.shape = tuple(int(n) for n in shape)
        self.dx = float(dx)
        handle = c_void_p()
        self._check(lib.clq_net_create(*self.shape, self.dx, byref(handle)), "clq_net_create")
        self.handle = handle
        self.param_count = int(lib.clq_net_param_count())
        if seed is not None:
            self._check(lib.clq_net_init(self.handle, int(seed)), "clq_net_init")

    _check = Context._check

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            self.lib.clq_net_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def _grid(self, a, name: str) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.float32)
        if a.shape != self.shape:
            raise ValueError(f"{name} has shape {a.shape}, expected {self.shape}")
        return a

    def get_params(self) -> np.ndarray:
        out = np.empty(self.param_count, np.float32)
        self._check(self.lib.clq_net_get_params(self.handle, _ptr(out), out.size), "clq_net_get_params")
        return out

    def set_params(self, params) -> None:
        params = np.ascontiguousarray(params, dtype=np.float32)
        self._check(self.lib.clq_net_set_params(self.handle, _ptr(params), params.size), "clq_net_set_params")

    def forward(self, rhs, mask, prev=None) -> np.ndarray:
        rhs, mask = self._grid(rhs, "rhs"), self._grid(mask, "mask")
        prev = self._grid(prev, "prev") if prev is not None else None
        out = np.empty(self.shape, np.float32)
        self._check(self.lib.clq_net_forward(self.handle, _ptr(rhs), _ptr(mask),
                                             _ptr(prev) if prev is not None else None, _ptr(out)), "clq_net_forward")
        return out

    def backward(self, grad_out) -> np.ndarray:
        grad_out = self._grid(grad_out, "grad_out")
        grads = np.empty(self.param_count, np.float32)
        self._check(self.lib.clq_net_backward(self.handle, _ptr(grad_out), _ptr(grads), grads.size),
                    "clq_net_backward")
        return grads

    def train_step(self, rhs, mask, prev, target, lr: float = 1e-3) -> float:
        """One Adam step on a batch ([batch, *shape] arrays; prev may be None)."""
        def batch(a, name):
            a = np.ascontiguousarray(a, dtype=np.float32)
            if a.ndim == 3:
                a = a[None]
            if a.shape[1:] != self.shape:
                raise ValueError(f"{name} has shape {a.shape}, expected (batch, *{self.shape})")
            return a
        rhs, mask, target = batch(rhs, "rhs"), batch(mask, "mask"), batch(target, "target")
        prev = batch(prev, "prev") if prev is not None else None
        n = rhs.shape[0]
        if mask.shape[0] != n or target.shape[0] != n or (prev is not None and prev.shape[0] != n):
            raise ValueError("batch sizes differ")
        loss = c_double()
        self._check(self.lib.clq_net_train_step(self.handle, n, _ptr(rhs), _ptr(mask),
                                                _ptr(prev) if prev is not None else None, _ptr(target), lr,
                                                byref(loss)), "clq_net_train_step")
        return loss.value

