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
//      64-lane wave usually touches ~8 bins); a wave handles four groups of
//      64 particles so that their atomics are in flight together.
//   2. exclusive scan of the counts: tile reduce, scan of tile sums, tile scan.
//   3. bin_permute: perm[offset[key] + provisional] = particle index (and
//      the big-bin counter of step 4 set to 0).
//   4. bin_order: a block takes 256 consecutive bins, stages their slice of
//      perm in LDS with a slot-to-bin table, ranks each member among its
//      bin's members by original index (stable order) and moves its
//      position and velocity to the ranked slot.  Bins of more than kBigBin members are listed instead and
//      sorted by a block-level merge sort (bin_big_sort, O(c log^2 c) for c
//      members, so no bin size makes the ranking quadratic), then moved
//      (bin_big_move).
//   The reference's atomic_inc scatter leaves the order within a bin to
//   chance; this one is the stable order, deterministic run to run.
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
// atomic per run of equal keys in each group of 64 particles, handed out in
// lane order).  Wave w of the grid handles particles 256 w .. 256 w + 255 as
// four groups.  The heads of all runs of the four groups are listed in LDS
// (at most 256, typically ~32) and their atomics issued together, one lane
// per head, so that the wave waits for one round trip instead of four.
constexpr int kCountGroups = 4;
__global__ __launch_bounds__(256) void bin_count_kernel(const float *__restrict__ pos, int n, float inv_dx, int nx,
                                                        int ny, int nz, int32_t *__restrict__ key,
                                                        int32_t *__restrict__ rank, int32_t *__restrict__ counts) {
    __shared__ int head_key[4][64 * kCountGroups], head_len[4][64 * kCountGroups], head_base[4][64 * kCountGroups];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t first = ((int64_t)blockIdx.x * 4 + wave) * (64 * kCountGroups) + lane;
    // every load issued first (lanes past the end read the last particle)
    float x[kCountGroups][3];
#pragma unroll
    for (int g = 0; g < kCountGroups; ++g) {
        const int64_t i = min(first + 64 * g, (int64_t)n - 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) x[g][c] = pos[3 * i + c];
    }
    int k[kCountGroups], slot[kCountGroups];
    WaveRun r[kCountGroups];
    int heads = 0;  // runs listed so far (wave-uniform)
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int g = 0; g < kCountGroups; ++g) {
        const bool valid = first + 64 * g < n;
        k[g] = valid ? bin_key(x[g], inv_dx, nx, ny, nz) : -1;
        r[g] = wave_run(k[g], valid, lane);
        const unsigned long long hm = __ballot(r[g].head);
        slot[g] = heads + __popcll(hm & below);  // this lane's list entry when it heads a run
        if (r[g].head) {
            head_key[wave][slot[g]] = k[g];
            head_len[wave][slot[g]] = r[g].end_lane - lane;
        }
        heads += __popcll(hm);
    }
    __builtin_amdgcn_wave_barrier();
    for (int j = lane; j < heads; j += 64)
        head_base[wave][j] = atomicAdd(counts + head_key[wave][j], head_len[wave][j]);
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int g = 0; g < kCountGroups; ++g) {
        const int64_t i = first + 64 * g;
        // the list entry of this lane's run head, from the head lane
        const int hs = __shfl(slot[g], r[g].head_lane);
        if (i < n) {
            key[i] = k[g];
            rank[i] = head_base[wave][hs] + (lane - r[g].head_lane);
        }
    }
}

void launch_bin_count(const float *pos, int n, float inv_dx, int nx, int ny, int nz, int32_t *key, int32_t *rank,
                      int32_t *counts, hipStream_t s) {
    if (n <= 0) return;
    constexpr int per_block = 256 * kCountGroups;
    hipLaunchKernelGGL(bin_count_kernel, dim3((n + per_block - 1) / per_block), dim3(256), 0, s, pos, n, inv_dx, nx,
                       ny, nz, key, rank, counts);
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

// ---- permutation and stable order ---------------------------------------------
constexpr int kPermPer = 4;
__global__ __launch_bounds__(256) void bin_permute_kernel(const int32_t *__restrict__ key,
                                                          const int32_t *__restrict__ rank,
                                                          const int32_t *__restrict__ offsets, int n,
                                                          int32_t *__restrict__ perm, int32_t *__restrict__ big_count) {
    const int base = blockIdx.x * (256 * kPermPer) + threadIdx.x;
    if (base == 0) *big_count = 0;  // read by bin_order, the next launch
    int k[kPermPer], r[kPermPer], o[kPermPer];
#pragma unroll
    for (int q = 0; q < kPermPer; ++q) {
        const int i = base + 256 * q;
        k[q] = i < n ? key[i] : 0;
        r[q] = i < n ? rank[i] : 0;
    }
#pragma unroll
    for (int q = 0; q < kPermPer; ++q) o[q] = offsets[k[q]];
#pragma unroll
    for (int q = 0; q < kPermPer; ++q) {
        const int i = base + 256 * q;
        if (i < n) perm[o[q] + r[q]] = i;
    }
}

void launch_bin_permute(const int32_t *key, const int32_t *rank, const int32_t *offsets, int n, int32_t *perm,
                        int32_t *big_count, hipStream_t s) {
    if (n <= 0) return;
    constexpr int per_block = 256 * kPermPer;
    hipLaunchKernelGGL(bin_permute_kernel, dim3((n + per_block - 1) / per_block), dim3(256), 0, s, key, rank, offsets,
                       n, perm, big_count);
}

constexpr int kOrderBins = 256;   // bins per block
constexpr int kOrderCap = 4096;   // perm entries staged in LDS
constexpr int kBigBin = 64;       // larger bins take the merge-sort path

int big_bin_capacity(int64_t n) { return (int)(n / (kBigBin + 1) + 1); }

// Block b: bins [256 b, 256 b + 256) and their slots [S0, S1).  Slot s holds
// provisionally particle i = perm[s]; its stable slot is start + (number of
// its bin's members with a smaller index), counted over the staged slice
// (read from perm when the slice exceeds kOrderCap).  A staged slot finds
// its bin in bin_s, which every bin's thread fills for its own slots (one
// LDS read instead of a binary search of eight dependent reads).  Bins of more than
// kBigBin members are appended to big_list and skipped.
__global__ __launch_bounds__(256) void bin_order_kernel(const int32_t *__restrict__ perm,
                                                        const int32_t *__restrict__ offsets,
                                                        const int32_t *__restrict__ counts, int64_t C, int n,
                                                        const float *__restrict__ pos, const float *__restrict__ vel,
                                                        float *__restrict__ pos_out, float *__restrict__ vel_out,
                                                        int32_t *__restrict__ big_list,
                                                        int32_t *__restrict__ big_count, int big_cap) {
    __shared__ int off_s[kOrderBins], cnt_s[kOrderBins];
    __shared__ int p_s[kOrderCap];
    __shared__ uint8_t bin_s[kOrderCap];  // block-local bin of each staged slot
    const int t = threadIdx.x;
    const int64_t b0 = (int64_t)blockIdx.x * kOrderBins;
    for (int q = t; q < kOrderBins; q += 256) {
        const int64_t b = b0 + q;
        const int c = b < C ? counts[b] : 0;
        off_s[q] = b < C ? offsets[b] : n;
        cnt_s[q] = c;
        if (c > kBigBin) {
            const int at = atomicAdd(big_count, 1);
            if (at < big_cap) big_list[at] = (int)b;
        }
    }
    __syncthreads();
    const int S0 = off_s[0], S1 = off_s[kOrderBins - 1] + cnt_s[kOrderBins - 1];
    const int span = S1 - S0;
    if (span == 0) return;  // block-uniform
    const bool staged = span <= kOrderCap;  // block-uniform
    if (staged) {
        for (int q = t; q < span; q += 256) p_s[q] = perm[S0 + q];
        for (int q = t; q < kOrderBins; q += 256)
            for (int m = off_s[q] - S0, e = m + cnt_s[q]; m < e; ++m) bin_s[m] = (uint8_t)q;
    }
    __syncthreads();
    // kMoves slots per thread at a time: ranks first, then every load, then
    // every store
    constexpr int kMoves = 4;
    for (int s0 = S0 + t; s0 < S1; s0 += 256 * kMoves) {
        int src[kMoves], dst[kMoves];
#pragma unroll
        for (int q = 0; q < kMoves; ++q) {
            const int s = s0 + 256 * q;
            src[q] = -1;
            dst[q] = 0;
            if (s >= S1) continue;
            // the bin holding slot s: the last j with off_s[j] <= s
            int lo = 0;
            if (staged) {
                lo = bin_s[s - S0];
            } else {
                int hi = kOrderBins;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (off_s[mid] <= s) lo = mid;
                    else hi = mid;
                }
            }
            const int start = off_s[lo], cnt = cnt_s[lo];
            if (cnt > kBigBin) continue;
            const int i = staged ? p_s[s - S0] : perm[s];
            int r = 0;
            if (staged) {
                const int *ms = p_s + (start - S0);
#pragma unroll 8
                for (int m = 0; m < cnt; ++m) r += ms[m] < i;
            } else {
                for (int m = 0; m < cnt; ++m) r += perm[start + m] < i;
            }
            src[q] = i;
            dst[q] = start + r;
        }
        float pv[kMoves][6];
#pragma unroll
        for (int q = 0; q < kMoves; ++q) {
            const int64_t i = src[q] < 0 ? 0 : src[q];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                pv[q][c] = pos[3 * i + c];
                pv[q][3 + c] = vel[3 * i + c];
            }
        }
#pragma unroll
        for (int q = 0; q < kMoves; ++q) {
            if (src[q] < 0) continue;
            const int64_t f = dst[q];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                pos_out[3 * f + c] = pv[q][c];
                vel_out[3 * f + c] = pv[q][3 + c];
            }
        }
    }
}

__device__ __forceinline__ void move_particle(const float *__restrict__ pos, const float *__restrict__ vel,
                                              float *__restrict__ pos_out, float *__restrict__ vel_out, int64_t i,
                                              int64_t f) {
    const float p0 = pos[3 * i], p1 = pos[3 * i + 1], p2 = pos[3 * i + 2];
    const float v0 = vel[3 * i], v1 = vel[3 * i + 1], v2 = vel[3 * i + 2];
    pos_out[3 * f] = p0; pos_out[3 * f + 1] = p1; pos_out[3 * f + 2] = p2;
    vel_out[3 * f] = v0; vel_out[3 * f + 1] = v1; vel_out[3 * f + 2] = v2;
}

// Number of the len sorted values a[0..len) below x.
__device__ __forceinline__ int count_below(const int32_t *__restrict__ a, int len, int x) {
    int lo = 0, hi = len;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Each listed bin's slice of perm sorted by value (particle index) with a
// bottom-up merge sort in global memory by one block, runs merged by
// ranking every element in the other run (values are distinct), scratch the
// same slots of `scratch`.
__global__ __launch_bounds__(256) void bin_big_sort_kernel(const int32_t *__restrict__ big_list,
                                                           const int32_t *__restrict__ big_count, int big_cap,
                                                           const int32_t *__restrict__ offsets,
                                                           const int32_t *__restrict__ counts, int32_t *perm,
                                                           int32_t *scratch) {
    const int nb = min(*big_count, big_cap);
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int bin = big_list[b], start = offsets[bin], cnt = counts[bin];
        int32_t *src = perm + start, *dst = scratch + start;
        for (int64_t w = 1; w < cnt; w *= 2) {
            for (int idx = threadIdx.x; idx < cnt; idx += 256) {
                const int64_t a = idx - idx % (2 * w), mid = min(a + w, (int64_t)cnt), end = min(a + 2 * w, (int64_t)cnt);
                const int x = src[idx];
                const int64_t to = idx < mid ? (idx - a) + count_below(src + mid, (int)(end - mid), x)
                                             : (idx - mid) + count_below(src + a, (int)(mid - a), x);
                dst[a + to] = x;
            }
            __syncthreads();
            int32_t *tmp = src;
            src = dst;
            dst = tmp;
        }
        if (src != perm + start)
            for (int idx = threadIdx.x; idx < cnt; idx += 256) perm[start + idx] = src[idx];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void bin_big_move_kernel(const int32_t *__restrict__ big_list,
                                                           const int32_t *__restrict__ big_count, int big_cap,
                                                           const int32_t *__restrict__ offsets,
                                                           const int32_t *__restrict__ counts,
                                                           const int32_t *__restrict__ perm,
                                                           const float *__restrict__ pos,
                                                           const float *__restrict__ vel, float *__restrict__ pos_out,
                                                           float *__restrict__ vel_out) {
    const int nb = min(*big_count, big_cap);
    for (int b = blockIdx.x; b < nb; b += gridDim.x) {
        const int bin = big_list[b], start = offsets[bin], cnt = counts[bin];
        for (int f = start + threadIdx.x; f < start + cnt; f += 256) move_particle(pos, vel, pos_out, vel_out, perm[f], f);
    }
}

void launch_bin_order(const int32_t *perm, const int32_t *offsets, const int32_t *counts, int64_t C, int n,
                      const float *pos, const float *vel, float *pos_out, float *vel_out, int32_t *big_list,
                      int32_t *big_count, int big_cap, int32_t *scratch, hipStream_t s) {
    if (n <= 0) return;
    const int64_t blocks = (C + kOrderBins - 1) / kOrderBins;
    hipLaunchKernelGGL(bin_order_kernel, dim3((unsigned)blocks), dim3(256), 0, s, perm, offsets, counts, C, n, pos, vel,
                       pos_out, vel_out, big_list, big_count, big_cap);
    hipLaunchKernelGGL(bin_big_sort_kernel, dim3(256), dim3(256), 0, s, big_list, big_count, big_cap, offsets, counts,
                       const_cast<int32_t *>(perm), scratch);
    hipLaunchKernelGGL(bin_big_move_kernel, dim3(256), dim3(256), 0, s, big_list, big_count, big_cap, offsets, counts,
                       perm, pos, vel, pos_out, vel_out);
}

}  // namespace clq
