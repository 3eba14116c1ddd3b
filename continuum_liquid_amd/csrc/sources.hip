// Particle sources and sinks on the device: the per-step work of the
// reference's host round trip (Solver/Kernel/solver.py:253-306 with
// source_particles.py:42-92) without moving the particles to the host.
//
//   outflow: particles whose cell floor(x / delta) (float32 division, clamped
//            into the grid) an OUTFLOW source marks are removed; the
//            survivors keep their order (flags, exclusive scan, scatter).
//   inflow:  each open cell (source mask and not solid) is topped up to the
//            source's density, counting the particles present; the new
//            particles are appended in C order of their cells, a cell
//            missing m particles taking m consecutive slots, at
//            float32(((double)cell + u) * delta) per axis with the source's
//            velocity.
//
// The uniforms u are the host's own numpy generator stream: PCG64
// (128-bit LCG, multiplier 0x2360ED051FC65DA44385DF649FCCF645, output
// XSL-RR 128/64; a draw steps the state, then outputs) turned into doubles
// as (x >> 11) * 2^-53, numpy's Generator.random().  New particle j of a
// source takes draws base + 3 j .. base + 3 j + 2, which is the order of
// rng.random((count, 3)); every thread jumps the generator ahead to its first
// draw in O(log) 128-bit steps.  The host passes the generator's state in and
// advances its own generator by the draws consumed, so device and host
// seeding produce the same bits.
#include "common.hpp"

namespace clq {

typedef unsigned __int128 u128;

constexpr u128 kPcgMult = ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull;

// the state after `delta` steps of s -> s * kPcgMult + inc (mod 2^128)
__device__ __forceinline__ u128 pcg_advance(u128 state, u128 inc, uint64_t delta) {
    u128 acc_mult = 1, acc_plus = 0, cur_mult = kPcgMult, cur_plus = inc;
    while (delta) {
        if (delta & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1;
    }
    return acc_mult * state + acc_plus;
}

__device__ __forceinline__ double pcg_uniform(u128 state) {
    const uint64_t x = (uint64_t)(state >> 64) ^ (uint64_t)state;
    const unsigned rot = (unsigned)(state >> 122);
    const uint64_t r = (x >> rot) | (x << ((64u - rot) & 63u));
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

// the particle's cell, floor(x / delta) in float32 clamped into the grid
__device__ __forceinline__ int64_t source_cell(const float *x, float delta, int nx, int ny, int nz) {
    const int i = floor_clamp(x[0] / delta, 0, nx - 1);
    const int j = floor_clamp(x[1] / delta, 0, ny - 1);
    const int k = floor_clamp(x[2] / delta, 0, nz - 1);
    return gidx(i, j, k, ny, nz);
}

__global__ __launch_bounds__(256) void outflow_flags_kernel(const float *__restrict__ pos, int n,
                                                            const uint8_t *__restrict__ sink, float delta, int nx,
                                                            int ny, int nz, int32_t *__restrict__ keep) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    keep[i] = sink[source_cell(pos + 3LL * i, delta, nx, ny, nz)] ? 0 : 1;
}

__global__ __launch_bounds__(256) void compact_kernel(const float *__restrict__ pos, const float *__restrict__ vel,
                                                      const int32_t *__restrict__ keep,
                                                      const int32_t *__restrict__ slot, int n,
                                                      float *__restrict__ pos_out, float *__restrict__ vel_out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n || !keep[i]) return;
    const int64_t s = slot[i];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        pos_out[3 * s + c] = pos[3LL * i + c];
        vel_out[3 * s + c] = vel[3LL * i + c];
    }
}

// present[c] += 1 for the particles whose cell is open for this source
__global__ __launch_bounds__(256) void count_open_kernel(const float *__restrict__ pos, int n,
                                                         const uint8_t *__restrict__ open, float delta, int nx,
                                                         int ny, int nz, int32_t *__restrict__ present) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t c = source_cell(pos + 3LL * i, delta, nx, ny, nz);
    if (open[c]) atomicAdd(present + c, 1);
}

// missing[c] = max(0, density - present[c]) on the open cells, 0 elsewhere
// (in place)
__global__ __launch_bounds__(256) void missing_kernel(int32_t *__restrict__ present,
                                                      const uint8_t *__restrict__ open, int density, int64_t C) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= C) return;
    present[c] = open[c] ? max(0, density - present[c]) : 0;
}

// one thread per cell: its missing particles, slots n0 + first[c] + t
__global__ __launch_bounds__(256) void emit_kernel(const int32_t *__restrict__ missing,
                                                   const int32_t *__restrict__ first, int64_t C, int ny, int nz,
                                                   double delta, uint64_t s_hi, uint64_t s_lo, uint64_t i_hi,
                                                   uint64_t i_lo, int64_t n0, float vx, float vy, float vz,
                                                   float *__restrict__ pos, float *__restrict__ vel) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= C) return;
    const int m = missing[c];
    if (m == 0) return;
    int i, j, k;
    coords(c, ny, nz, i, j, k);
    const double corner[3] = {(double)i, (double)j, (double)k};
    const u128 inc = ((u128)i_hi << 64) | i_lo;
    // the state just before this cell's first draw
    u128 st = pcg_advance(((u128)s_hi << 64) | s_lo, inc, 3ull * (uint64_t)first[c]);
    for (int t = 0; t < m; ++t) {
        const int64_t slot = n0 + first[c] + t;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            st = st * kPcgMult + inc;  // a draw steps the state, then outputs it
            pos[3 * slot + a] = (float)((corner[a] + pcg_uniform(st)) * delta);
        }
        vel[3 * slot] = vx;
        vel[3 * slot + 1] = vy;
        vel[3 * slot + 2] = vz;
    }
}

void launch_outflow_flags(const float *pos, int n, const uint8_t *sink, float delta, int nx, int ny, int nz,
                          int32_t *keep, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(outflow_flags_kernel, dim3((n + 255) / 256), dim3(256), 0, s, pos, n, sink, delta, nx, ny, nz,
                       keep);
}

void launch_compact(const float *pos, const float *vel, const int32_t *keep, const int32_t *slot, int n,
                    float *pos_out, float *vel_out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(compact_kernel, dim3((n + 255) / 256), dim3(256), 0, s, pos, vel, keep, slot, n, pos_out,
                       vel_out);
}

void launch_count_open(const float *pos, int n, const uint8_t *open, float delta, int nx, int ny, int nz,
                       int32_t *present, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(count_open_kernel, dim3((n + 255) / 256), dim3(256), 0, s, pos, n, open, delta, nx, ny, nz,
                       present);
}

void launch_missing(int32_t *present, const uint8_t *open, int density, int64_t C, hipStream_t s) {
    hipLaunchKernelGGL(missing_kernel, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, s, present, open, density, C);
}

void launch_emit(const int32_t *missing, const int32_t *first, int64_t C, int ny, int nz, double delta,
                 const uint64_t *rng, uint64_t draws_before, int64_t n0, const float *velocity, float *pos, float *vel,
                 hipStream_t s) {
    // the generator as it stands after the draws of earlier sources: the
    // host's state advanced on the host (u128 arithmetic), passed by value
    const u128 inc = ((u128)rng[2] << 64) | rng[3];
    u128 st = ((u128)rng[0] << 64) | rng[1];
    u128 acc_mult = 1, acc_plus = 0, cur_mult = kPcgMult, cur_plus = inc;
    for (uint64_t d = draws_before; d; d >>= 1) {
        if (d & 1) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1) * cur_plus;
        cur_mult *= cur_mult;
    }
    st = acc_mult * st + acc_plus;
    hipLaunchKernelGGL(emit_kernel, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, s, missing, first, C, ny, nz,
                       delta, (uint64_t)(st >> 64), (uint64_t)st, rng[2], rng[3], n0, velocity[0], velocity[1],
                       velocity[2], pos, vel);
}

}  // namespace clq
