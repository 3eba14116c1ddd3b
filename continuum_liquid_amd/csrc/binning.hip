// Particle binning: histogram, exclusive scan and a stable counting sort.
//
// Replaces count_particles_per_cell, scan_blocks_exclusive,
// add_scanned_block_offsets and sort_particles (reference
// Solver/Kernel/OpenCL/particles.cl:44-133, driven by
// Solver/Kernel/particles.py:39-140).
//
// Pipeline (all integer work, HBM-bound):
//   1. bin_count: one lane per particle computes its bin key and takes a
//      provisional slot with one atomicAdd per run of equal keys in the
//      wavefront (particles arrive nearly sorted from the previous step, so a
//      64-lane wave usually touches ~8 bins).
//   2. exclusive scan of the counts: tile reduce, scan of tile sums, tile scan.
//   3. bin_permute: perm[offset[key] + provisional] = particle index.
//   4. bin_gather: every output slot ranks its particle among the members of
//      its bin by original index (stable order), then moves position and
//      velocity.  The reference's atomic_inc scatter leaves the order within a
//      bin to chance; this one is the stable order, deterministic run to run.
#include "common.hpp"

namespace clq {

// interpolation_base (particles.cl:3-6): floor(fma(x, inv_dx, -0.5)) clamped
// to [0, n-2]; the reference expression is contracted (llvm.fmuladd).
__device__ __forceinline__ int bin_key(const float *p, float inv_dx, int nx, int ny, int nz) {
    const int bx = floor_clamp(fmaf(p[0], inv_dx, -0.5f), 0, nx - 2);
    const int by = floor_clamp(fmaf(p[1], inv_dx, -0.5f), 0, ny - 2);
    const int bz = floor_clamp(fmaf(p[2], inv_dx, -0.5f), 0, nz - 2);
    return (int)gidx(bx, by, bz, ny, nz);
}

// Runs of equal keys within the wave (lanes in order): whether this lane
// heads a run, the run's head lane and the lane after its end.
struct WaveRun {
    bool head;
    int head_lane, end_lane;
};
__device__ __forceinline__ WaveRun wave_run(int k, bool valid, int lane) {
    const int prev = __shfl_up(k, 1);
    const bool head = valid && (lane == 0 || prev != k);
    const unsigned long long heads = __ballot(head);
    const unsigned long long live = __ballot(valid);
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes <= lane (lane 63: all ones)
    const unsigned long long below = heads & upto;
    const unsigned long long above = heads & ~upto;
    WaveRun r;
    r.head = head;
    r.head_lane = below ? 63 - __clzll((long long)below) : lane;
    r.end_lane = above ? __ffsll((long long)above) - 1 : 64 - __clzll((long long)live);
    return r;
}

// One lane per particle: its bin key and a provisional slot in its bin (one
// atomic per run of equal keys in the wave, handed out in lane order).
__global__ __launch_bounds__(256) void bin_count_kernel(const float *__restrict__ pos, int n, float inv_dx, int nx,
                                                        int ny, int nz, int32_t *__restrict__ key,
                                                        int32_t *__restrict__ rank, int32_t *__restrict__ counts) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < n;
    int k = -1;
    if (valid) k = bin_key(pos + 3LL * i, inv_dx, nx, ny, nz);
    const WaveRun r = wave_run(k, valid, lane);
    int base = 0;
    if (r.head) base = atomicAdd(counts + k, r.end_lane - lane);
    base = __shfl(base, r.head_lane);
    if (valid) {
        key[i] = k;
        rank[i] = base + (lane - r.head_lane);
    }
}

void launch_bin_count(const float *pos, int n, float inv_dx, int nx, int ny, int nz, int32_t *key, int32_t *rank,
                      int32_t *counts, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(bin_count_kernel, dim3((n + 255) / 256), dim3(256), 0, s, pos, n, inv_dx, nx, ny, nz, key,
                       rank, counts);
}

// ---- exclusive scan ----------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 8;
constexpr int kScanTile = kScanThreads * kScanPerThread;  // 2048 elements

int64_t scan_scratch_elems(int64_t count) { return (count + kScanTile - 1) / kScanTile + 1; }

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ __forceinline__ int block_exclusive_scan(int v, int &total) {
    __shared__ int wave_sums[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nwaves = (blockDim.x + 63) >> 6;
    int incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    if (lane == 63) wave_sums[wave] = incl;
    __syncthreads();
    int wave_off = 0, tot = 0;
    for (int q = 0; q < nwaves; ++q) {
        const int ws = wave_sums[q];
        if (q < wave) wave_off += ws;
        tot += ws;
    }
    __syncthreads();
    total = tot;
    return wave_off + incl - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_reduce(const int32_t *__restrict__ in, int64_t count,
                                                                 int32_t *__restrict__ tile_sums) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
    int s = 0;
    if (base + kScanPerThread <= count) {
        const int4 a = *reinterpret_cast<const int4 *>(in + base);
        const int4 b = *reinterpret_cast<const int4 *>(in + base + 4);
        s = a.x + a.y + a.z + a.w + b.x + b.y + b.z + b.w;
    } else {
        for (int q = 0; q < kScanPerThread; ++q)
            if (base + q < count) s += in[base + q];
    }
    int total;
    block_exclusive_scan(s, total);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

// Single work-group exclusive scan of the tile sums, in chunks of 1024.
__global__ __launch_bounds__(1024) void scan_tile_sums(int32_t *__restrict__ tile_sums, int64_t tiles) {
    int carry = 0;
    for (int64_t start = 0; start < tiles; start += 1024) {
        const int64_t i = start + threadIdx.x;
        const int v = i < tiles ? tile_sums[i] : 0;
        int total;
        const int ex = block_exclusive_scan(v, total);
        if (i < tiles) tile_sums[i] = carry + ex;
        carry += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_apply(const int32_t *__restrict__ in, int64_t count,
                                                                const int32_t *__restrict__ tile_offsets,
                                                                int32_t *__restrict__ out) {
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanPerThread;
    int v[kScanPerThread];
    const bool full = base + kScanPerThread <= count;
    if (full) {
        const int4 a = *reinterpret_cast<const int4 *>(in + base);
        const int4 b = *reinterpret_cast<const int4 *>(in + base + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
        for (int q = 0; q < kScanPerThread; ++q) v[q] = base + q < count ? in[base + q] : 0;
    }
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanPerThread; ++q) s += v[q];
    int total;
    int run = block_exclusive_scan(s, total) + tile_offsets[blockIdx.x];
    int o[kScanPerThread];
#pragma unroll
    for (int q = 0; q < kScanPerThread; ++q) {
        o[q] = run;
        run += v[q];
    }
    if (full) {
        *reinterpret_cast<int4 *>(out + base) = make_int4(o[0], o[1], o[2], o[3]);
        *reinterpret_cast<int4 *>(out + base + 4) = make_int4(o[4], o[5], o[6], o[7]);
    } else {
        for (int q = 0; q < kScanPerThread; ++q)
            if (base + q < count) out[base + q] = o[q];
    }
}

// `in` and `out` must be 16-byte aligned (hipMalloc allocations are).
void launch_exclusive_scan(const int32_t *in, int32_t *out, int64_t count, int32_t *scratch, hipStream_t s) {
    if (count <= 0) return;
    const int64_t tiles = (count + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(scan_tile_reduce, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, in, count, scratch);
    hipLaunchKernelGGL(scan_tile_sums, dim3(1), dim3(1024), 0, s, scratch, tiles);
    hipLaunchKernelGGL(scan_tile_apply, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, in, count, scratch, out);
}

// ---- permutation and stable gather ---------------------------------------------
__global__ __launch_bounds__(256) void bin_permute_kernel(const int32_t *__restrict__ key,
                                                          const int32_t *__restrict__ rank,
                                                          const int32_t *__restrict__ offsets, int n,
                                                          int32_t *__restrict__ perm) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    perm[offsets[key[i]] + rank[i]] = i;
}

void launch_bin_permute(const int32_t *key, const int32_t *rank, const int32_t *offsets, int n, int32_t *perm,
                        hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(bin_permute_kernel, dim3((n + 255) / 256), dim3(256), 0, s, key, rank, offsets, n, perm);
}

__global__ __launch_bounds__(256) void bin_gather_kernel(const int32_t *__restrict__ perm,
                                                         const int32_t *__restrict__ key,
                                                         const int32_t *__restrict__ offsets,
                                                         const int32_t *__restrict__ counts,
                                                         const float *__restrict__ pos, const float *__restrict__ vel,
                                                         int n, float *__restrict__ pos_out,
                                                         float *__restrict__ vel_out) {
    const int slot = blockIdx.x * 256 + threadIdx.x;
    if (slot >= n) return;
    const int i = perm[slot];
    // the particle's data is loaded first: it depends on i alone
    const int64_t src = 3LL * i;
    const float p0 = pos[src], p1 = pos[src + 1], p2 = pos[src + 2];
    const float v0 = vel[src], v1 = vel[src + 1], v2 = vel[src + 2];
    const int b = key[i];
    const int start = offsets[b], cnt = counts[b];
    // rank among the bin's members by original index, eight members' loads
    // in flight at a time (the members are slots start .. start + cnt - 1;
    // cnt >= 1, and indices past the last member read the last one, whose
    // comparison is dropped)
    int r = 0;
    for (int t0 = 0; t0 < cnt; t0 += 8) {
        int m[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) m[u] = perm[start + min(t0 + u, cnt - 1)];
#pragma unroll
        for (int u = 0; u < 8; ++u) r += (t0 + u < cnt) & (m[u] < i);
    }
    const int64_t dst = 3LL * (start + r);
    pos_out[dst] = p0; pos_out[dst + 1] = p1; pos_out[dst + 2] = p2;
    vel_out[dst] = v0; vel_out[dst + 1] = v1; vel_out[dst + 2] = v2;
}

void launch_bin_gather(const int32_t *perm, const int32_t *key, const int32_t *offsets, const int32_t *counts,
                       const float *pos, const float *vel, int n, float *pos_out, float *vel_out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(bin_gather_kernel, dim3((n + 255) / 256), dim3(256), 0, s, perm, key, offsets, counts, pos,
                       vel, n, pos_out, vel_out);
}

}  // namespace clq
