"""clq_residual_norm_noalloc, the entry point that allocates nothing, against
clq_residual_norm on the same level-0 system: same kernels, same block
count, so the two norms agree bit for bit.  Two calls in a row, the second on
a smaller grid (the shared device scratch keeps the first call's partials)."""

from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from continuum_liquid_amd import runtime, scenes

from .helpers import make_context, run_to_pressure

pytestmark = pytest.mark.gpu

ARRAYS = ("pressure", "rhs", "fluid_fraction", "diagonal", "weight_x", "weight_y", "weight_z")


class DeviceMemory:
    """Device buffers for the test, through the HIP runtime the library is
    linked against (the test allocates; the entry point under test does not)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so", mode=os.RTLD_GLOBAL | getattr(os, "RTLD_NOLOAD", 0))
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.hip.hipDeviceSynchronize.argtypes = []
        self.owned = []

    def alloc(self, nbytes: int) -> ctypes.c_void_p:
        ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(ptr), nbytes) == 0
        self.owned.append(ptr)
        return ptr

    def upload(self, array: np.ndarray) -> ctypes.c_void_p:
        a = np.ascontiguousarray(array)
        ptr = self.alloc(a.nbytes)
        assert self.hip.hipMemcpy(ptr, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return ptr

    def download(self, ptr, out: np.ndarray) -> np.ndarray:
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return out

    def close(self):
        for ptr in self.owned:
            self.hip.hipFree(ptr)
        self.owned = []


@pytest.fixture
def mem(hip_lib):
    m = DeviceMemory()
    yield m
    m.close()


def test_residual_norm_noalloc_matches_context_entry(hip_lib, mem):
    for n, seed in ((40, 3), (24, 5)):
        scene = scenes.small_scene(n, seed=seed)
        ctx = make_context(hip_lib, scene)
        run_to_pressure(ctx, scene)
        ctx.pressure_solve(runtime.SOLVE_FIXED_CYCLES, max_cycles=1, post_smooth=2)
        want = ctx.residual_norm()
        dev = [mem.upload(ctx.get_field(name)) for name in ARRAYS]
        out = mem.upload(np.full(2, -1.0))
        rc = hip_lib.clq_residual_norm_noalloc(*dev, *scene.shape, scene.dx, out, None)
        assert rc == runtime.OK, hip_lib.clq_last_error()
        got = mem.download(out, np.empty(2))
        print(f"{scene.shape}: noalloc {got}, context {want}", flush=True)
        assert got[0] == want[0] and got[1] == want[1] and got[0] > 0
        ctx.close()


def test_residual_norm_noalloc_rejects_bad_sizes(hip_lib, mem):
    ptr = mem.alloc(27 * 4)
    out = mem.alloc(16)
    rc = hip_lib.clq_residual_norm_noalloc(ptr, ptr, ptr, ptr, ptr, ptr, ptr, 3, 2, 3, 0.1, out, None)
    assert rc == runtime.ERR_INVALID and b"too small" in hip_lib.clq_last_error()
    rc = hip_lib.clq_residual_norm_noalloc(ptr, ptr, ptr, ptr, ptr, ptr, ptr, 2048, 2048, 1024, 0.1, out, None)
    assert rc == runtime.ERR_INVALID and b"maximum" in hip_lib.clq_last_error()
