// Learned preconditioner: a 3-level 3D U-Net on the pressure grid, trained on
// the GPU by the kernels below.  No reference counterpart (the reference's
// solve is the W-cycle alone, Solver/Kernel/pressure_solve.py:61-83).
//
// Architecture (channels; every conv is 3x3x3, zero padding, bias, ReLU):
//   input 3 (rhs / s, fluid indicator, previous pressure / k), zero-padded to 16
//   L0: c1 16->16, c2 16->16 = s0        pool 2x2x2 (average)
//   L1: c3 16->32, c4 32->32 = s1        pool
//   L2: c5 32->64, c6 64->64             nearest upsample, concat [up, s1] -> 96
//   L1: c7 96->32, c8 32->32             nearest upsample, concat [up, s0] -> 48
//   L0: c9 48->16, c10 16->16            head: 1x1x1 16->1, linear, FP32
// Scales: s = rms of rhs over the interior fluid cells, k = s (N h)^2 / 1024
// with N the largest grid dimension (for A = L / h^2, p = A^-1 b has the
// magnitude s h^2 on local modes and up to s (N h)^2 on the domain-wide ones;
// k is the 32^3 grid-unit scale carried to every resolution).  Inputs rhs / s
// and prev / k, output p = prev + k * head (prev = 0 when absent): the head
// predicts the change from the previous pressure, O(1) at every resolution,
// and the map is invariant to the scale of the right-hand side.
// The network grid is the pressure grid rounded up to multiples of 4 in each
// dimension (zero inputs in the padding).  Activations are BF16 with channels
// fastest ([voxel][channel]); gradients are FP32.  Convolutions are implicit
// GEMMs on the matrix cores (v_mfma_f32_16x16x32_bf16 over pairs of taps, or
// v_mfma_f32_16x16x16_bf16 over single taps; FP32 accumulation): a wave owns
// 64 voxels x C_out, the K loop runs over (tap, 16 input channels).
//   forward   Y[v][co] = relu(b[co] + sum_{tap,ci} X[v + d(tap)][ci] W[tap][ci][co])
//   data grad G[u][ci] = sum_{tap,co} Z[u - d(tap)][co] W[tap][ci][co],
//             Z = dL/dY masked by Y > 0, rounded to BF16 on load
//   weight grad dW[tap][ci][co] = sum_v X[v + d(tap)][ci] Z[v][co], one FP32
//             partial per block of voxels, summed over blocks in a fixed order
// Parameters (FP32 master copy, Adam state FP32), in this order: for each conv
// layer W[tap][ci][co] (tap = 9 (dx+1) + 3 (dy+1) + (dz+1)) then bias[co];
// then the head's 16 weights and its bias.  BF16 copies in the forward layout
// [tap][co][ci] and the data-gradient layout [tap][ci][co] are refreshed after
// every change of the master copy.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "common.hpp"

namespace clq {
namespace {

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct Dims {
    int X, Y, Z;
    int64_t V;
};

inline Dims make_dims(int x, int y, int z) { return Dims{x, y, z, (int64_t)x * y * z}; }

__device__ __forceinline__ void vox(const Dims &d, int64_t v, int &x, int &y, int &z) {
    const uint32_t q = (uint32_t)v, yz = (uint32_t)d.Y * (uint32_t)d.Z;
    const uint32_t a = q / yz, r = q - a * yz, b = r / (uint32_t)d.Z;
    x = (int)a;
    y = (int)b;
    z = (int)(r - b * (uint32_t)d.Z);
}

// ---- LDS-tiled convolution ------------------------------------------------------------------
// A block owns a 4 x 4 x 16 brick of output voxels: wave w the x-slice x0 + w,
// its M-tile mt the row y0 + mt, lane m of a tile the voxel z0 + m.  For each
// chunk of 16 input channels the block stages the 6 x 6 x 18 input brick with
// its halo in LDS once (zeros outside the grid; for the data gradient the ReLU
// mask and the BF16 rounding of dL/dY are applied while staging), and the 27
// taps read their A fragments from LDS: 32 contiguous bytes per voxel, so a
// wave's fragment read is one conflict-free 512-byte run.  B fragments (the
// weights of one tap and 16 output channels) come from global memory and
// serve all four M-tiles.  Two taps share one v_mfma_f32_16x16x32_bf16 (K =
// 16 channels of each), which halves the fragment reads and weight loads per
// MAC; the 16-channel convolution over the upsampled concatenation keeps one
// tap per v_mfma_f32_16x16x16_bf16, which measured faster for it (1.21
// against 1.40 ms at 256^3; every other shape 7-42 % faster with pairs).
constexpr int kBX = 4, kBY = 4, kBZ = 16;
constexpr int kHX = kBX + 2, kHY = kBY + 2, kHZ = kBZ + 2;
constexpr int kHalo = kHX * kHY * kHZ;
constexpr int kStageIters = (kHalo * 2 + 255) / 256;  // staging items per thread

// CA > 0 (forward only): the input is the concatenation [up(coarse), in] with
// CA channels from the coarse tensor (dims dc, read at the parent voxel) and
// CI - CA from `in`, built while staging instead of materialised.
template <int CI, int CO, int MODE, int CA = 0>
__global__ __launch_bounds__(256) void conv3_lds_kernel(const void *__restrict__ in, const uint16_t *__restrict__ ymask,
                                                        const uint16_t *__restrict__ wt,
                                                        const float *__restrict__ bias, void *__restrict__ out,
                                                        Dims d, const uint16_t *__restrict__ coarse = nullptr,
                                                        Dims dc = Dims{0, 0, 0, 0}) {
    constexpr int NT = CO / 16;
    __shared__ __align__(16) uint16_t tile[kHalo * 16];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15, kg = lane >> 4;
    const int z0 = blockIdx.x * kBZ, y0 = blockIdx.y * kBY, x0 = blockIdx.z * kBX;
    float4_t acc[kBY][NT];
#pragma unroll
    for (int mt = 0; mt < kBY; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[mt][t] = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
    // the bias of this lane's output channels, loaded once
    float bco[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) bco[t] = MODE == 0 ? bias[t * 16 + m] : 0.0f;
    // the 16 input channels c0.. of halo voxel hv, half `half` (8 channels)
    auto stage_value = [&](int c0, int hv, int half) -> uint4 {
        const int hz = hv % kHZ, hxy = hv / kHZ, hy = hxy % kHY, hx = hxy / kHY;
        const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
        uint4 val = make_uint4(0u, 0u, 0u, 0u);
        if (gx >= 0 && gx < d.X && gy >= 0 && gy < d.Y && gz >= 0 && gz < d.Z) {
            const int64_t vox_index = ((int64_t)gx * d.Y + gy) * d.Z + gz;
            const int64_t off = vox_index * CI + c0 + half * 8;
            if (CA > 0 && c0 < CA) {
                const int64_t pv = ((int64_t)(gx >> 1) * dc.Y + (gy >> 1)) * dc.Z + (gz >> 1);
                val = *reinterpret_cast<const uint4 *>(coarse + pv * CA + c0 + half * 8);
            } else if (CA > 0) {
                val = *reinterpret_cast<const uint4 *>(static_cast<const uint16_t *>(in) + vox_index * (CI - CA) +
                                                       (c0 - CA) + half * 8);
            } else if (MODE == 0) {
                val = *reinterpret_cast<const uint4 *>(static_cast<const uint16_t *>(in) + off);
            } else {
                const float4 g0 = *reinterpret_cast<const float4 *>(static_cast<const float *>(in) + off);
                const float4 g1 = *reinterpret_cast<const float4 *>(static_cast<const float *>(in) + off + 4);
                const uint4 y = *reinterpret_cast<const uint4 *>(ymask + off);
                auto z = [](uint32_t ypair, float ga, float gb) {
                    const uint32_t lo = bf2f((uint16_t)(ypair & 0xffffu)) > 0.0f ? f2bf(ga) : 0u;
                    const uint32_t hi = bf2f((uint16_t)(ypair >> 16)) > 0.0f ? f2bf(gb) : 0u;
                    return lo | (hi << 16);
                };
                val = make_uint4(z(y.x, g0.x, g0.y), z(y.y, g0.z, g0.w), z(y.z, g1.x, g1.y), z(y.w, g1.z, g1.w));
            }
        }
        return val;
    };
    for (int c0 = 0; c0 < CI; c0 += 16) {
        __syncthreads();  // the previous chunk's fragments have been read
        // every item's loads first, then the LDS stores (one memory round
        // trip for the chunk's staging)
        uint4 staged[kStageIters];
#pragma unroll
        for (int it = 0; it < kStageIters; ++it) {
            const int item = threadIdx.x + it * 256;
            staged[it] = item < kHalo * 2 ? stage_value(c0, item >> 1, item & 1) : make_uint4(0u, 0u, 0u, 0u);
        }
#pragma unroll
        for (int it = 0; it < kStageIters; ++it) {
            const int item = threadIdx.x + it * 256;
            if (item < kHalo * 2) *reinterpret_cast<uint4 *>(tile + (item >> 1) * 16 + (item & 1) * 8) = staged[it];
        }
        __syncthreads();
        if constexpr (CO == 16 && CA > 0) {  // forward only: one tap per MFMA, K = 16 channels
            for (int tap = 0; tap < 27; ++tap) {
                const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
                short4_t b[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t)
                    b[t] = *reinterpret_cast<const short4_t *>(wt + ((int64_t)tap * CO + t * 16 + m) * CI + c0 +
                                                               kg * 4);
                const int base = ((wave + 1 + dx) * kHY + 1 + dy) * kHZ + m + 1 + dz;
#pragma unroll
                for (int mt = 0; mt < kBY; ++mt) {
                    const short4_t a = *reinterpret_cast<const short4_t *>(tile + (base + mt * kHZ) * 16 + kg * 4);
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[t], acc[mt][t], 0, 0, 0);
                }
            }
        } else {
            // two taps per MFMA: K = (tap 2p, channels 0-15; tap 2p + 1,
            // channels 0-15); lane group kg holds K 8 kg .. 8 kg + 7, i.e. tap
            // 2p + kg / 2, channels 8 (kg % 2) .. 8 (kg % 2) + 7; the 28th tap
            // is zero (its lanes read tap 26's fragment and drop it)
            const int cpart = (kg & 1) * 8;
            for (int tp = 0; tp < 14; ++tp) {
                const int tap_raw = 2 * tp + (kg >> 1);
                const bool live = tap_raw < 27;
                const int tap = live ? tap_raw : 26;
                int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
                if (MODE == 1) {
                    dx = -dx;
                    dy = -dy;
                    dz = -dz;
                }
                uint4 b[NT];
#pragma unroll
                for (int t = 0; t < NT; ++t) {
                    b[t] = *reinterpret_cast<const uint4 *>(wt + ((int64_t)tap * CO + t * 16 + m) * CI + c0 + cpart);
                    if (!live) b[t] = make_uint4(0u, 0u, 0u, 0u);
                }
                const int base = ((wave + 1 + dx) * kHY + 1 + dy) * kHZ + m + 1 + dz;
#pragma unroll
                for (int mt = 0; mt < kBY; ++mt) {
                    uint4 a = *reinterpret_cast<const uint4 *>(tile + (base + mt * kHZ) * 16 + cpart);
                    if (!live) a = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
                    for (int t = 0; t < NT; ++t)
                        acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                                             __builtin_bit_cast(bf16x8_t, b[t]),
                                                                             acc[mt][t], 0, 0, 0);
                }
            }
        }
    }
    // D[row = voxel z0 + kg*4 + j][col = channel t*16 + m]
    const int ox = x0 + wave;
    if (ox >= d.X) return;
#pragma unroll
    for (int mt = 0; mt < kBY; ++mt) {
        const int oy = y0 + mt;
        if (oy >= d.Y) continue;
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int oz = z0 + kg * 4 + j;
                if (oz >= d.Z) continue;
                const int64_t v = ((int64_t)ox * d.Y + oy) * d.Z + oz;
                const int co = t * 16 + m;
                if (MODE == 0)
                    static_cast<uint16_t *>(out)[v * CO + co] = f2bf(fmaxf(acc[mt][t][j] + bco[t], 0.0f));
                else
                    static_cast<float *>(out)[v * CO + co] = acc[mt][t][j];
            }
    }
}

// ---- weight gradient ------------------------------------------------------------------------
// dW[tap][ci][co] = sum_v X[v + d(tap)][ci] Z[v][co] for the 16 input channels
// c0 .. c0 + 15 of one launch: per tap the GEMM D (ci x co) += A (ci x 32
// voxels) B (32 voxels x co) on v_mfma_f32_16x16x32_bf16, K = the 32 voxels
// of two z-rows of a 4 x 4 x 16 brick.  A persistent block walks the bricks
// blockIdx.x, blockIdx.x + gridDim.x, ...; for each it stages the 6 x 6 x 18
// input brick with its halo in LDS as [ci][voxel] (zeros off the grid) and
// Z of the brick's 256 voxels as [co][voxel] (dL/dY masked by Y > 0, BF16),
// every staging load of an item batch issued before its LDS stores, and
// wave w accumulates the taps w, w + 4, ... for all CO output channels in
// registers over all the block's bricks.  A lane's eight voxels of a tap
// start anywhere in a channel's row (the tap shifts them by one voxel), so
// they are read as the three aligned 8-byte words that hold them and
// funnel-shifted into place: the unaligned 8-byte reads the compiler made of
// them before stalled the LDS (3x slower at 128^3).  The block writes its
// partial dW[tap][ci - c0][co] at the end; the partials are summed over the
// blocks in a fixed order (reduce_wgrad_kernel).
constexpr int kWX = 4, kWY = 4, kWZ = 16;        // brick of one step
constexpr int kWHX = kWX + 2, kWHY = kWY + 2, kWHZ = kWZ + 2;
constexpr int kWHalo = kWHX * kWHY * kWHZ;
constexpr int kWBlocks = 512;                  // persistent blocks (at most)
constexpr int kXRow = kWHalo + 4;               // LDS row of one input channel (8-byte multiple, 6-bank skew)
constexpr int kZRow = kWX * kWY * kWZ + 8;     // LDS row of one output channel (16-byte multiple, 4-bank skew)
constexpr int kTapsPerWave = 7;                // ceil(27 / 4)

inline int wgrad_bricks(const Dims &d) {
    return ((d.X + kWX - 1) / kWX) * ((d.Y + kWY - 1) / kWY) * ((d.Z + kWZ - 1) / kWZ);
}
inline int wgrad_blocks(const Dims &d) { return std::min(kWBlocks, wgrad_bricks(d)); }

template <int CI, int CO>
__global__ __launch_bounds__(256) void conv3_wgrad_kernel(const uint16_t *__restrict__ x, const float *__restrict__ gy,
                                                          const uint16_t *__restrict__ y, float *__restrict__ partials,
                                                          Dims d, int c0) {
    constexpr int TO = CO / 16;
    __shared__ __align__(16) uint16_t xs[16 * kXRow];
    __shared__ __align__(16) uint16_t zs[CO * kZRow];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m = lane & 15, kg = lane >> 4;
    const int nbx = (d.X + kWX - 1) / kWX, nby = (d.Y + kWY - 1) / kWY, nbz = (d.Z + kWZ - 1) / kWZ;
    const int nb = nbx * nby * nbz;
    float4_t acc[kTapsPerWave][TO];
#pragma unroll
    for (int tt = 0; tt < kTapsPerWave; ++tt)
#pragma unroll
        for (int t = 0; t < TO; ++t) acc[tt][t] = float4_t{0.0f, 0.0f, 0.0f, 0.0f};
    for (int brick = blockIdx.x; brick < nb; brick += gridDim.x) {
        const int bz = brick % nbz, bxy = brick / nbz, by = bxy % nby, bx = bxy / nby;
        const int x0 = bx * kWX, y0 = by * kWY, z0 = bz * kWZ;
        __syncthreads();  // the previous brick's fragments have been read
        // input brick with halo, 8 channels per item, transposed to [ci][voxel];
        // every item's load is issued before the first LDS store
        constexpr int kXIters = (kWHalo * 2 + 255) / 256;
        uint4 xv[kXIters];
#pragma unroll
        for (int it = 0; it < kXIters; ++it) {
            const int item = threadIdx.x + it * 256;
            xv[it] = make_uint4(0u, 0u, 0u, 0u);
            if (item < kWHalo * 2) {
                const int hv = item >> 1, half = item & 1;
                const int hz = hv % kWHZ, hxy = hv / kWHZ, hy = hxy % kWHY, hx = hxy / kWHY;
                const int gx = x0 - 1 + hx, gyy = y0 - 1 + hy, gz = z0 - 1 + hz;
                if (gx >= 0 && gx < d.X && gyy >= 0 && gyy < d.Y && gz >= 0 && gz < d.Z)
                    xv[it] = *reinterpret_cast<const uint4 *>(x + (((int64_t)gx * d.Y + gyy) * d.Z + gz) * CI + c0 +
                                                              half * 8);
            }
        }
        // Z of the brick's voxels, 8 channels per item, transposed to
        // [co][voxel]; loads issued in batches of up to four items
        constexpr int kZIters = CO / 8;  // kWX * kWY * kWZ * (CO / 8) items over 256 threads
        constexpr int kZBatch = kZIters < 4 ? kZIters : 4;
        auto z_item = [&](int it, float4 &g0, float4 &g1, uint4 &yy) {
            const int item = threadIdx.x + it * 256;
            const int bv = item / (CO / 8), cg = item - bv * (CO / 8);
            const int lz = bv % kWZ, lxy = bv / kWZ, ly = lxy % kWY, lx = lxy / kWY;
            const int gx = x0 + lx, gyy = y0 + ly, gz = z0 + lz;
            g0 = g1 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            yy = make_uint4(0u, 0u, 0u, 0u);  // Y = +0: masked out
            if (gx < d.X && gyy < d.Y && gz < d.Z) {
                const int64_t off = (((int64_t)gx * d.Y + gyy) * d.Z + gz) * CO + cg * 8;
                g0 = *reinterpret_cast<const float4 *>(gy + off);
                g1 = *reinterpret_cast<const float4 *>(gy + off + 4);
                yy = *reinterpret_cast<const uint4 *>(y + off);
            }
        };
#pragma unroll
        for (int it = 0; it < kXIters; ++it) {
            const int item = threadIdx.x + it * 256;
            if (item < kWHalo * 2) {
                const int hv = item >> 1, half = item & 1;
                const uint32_t w4[4] = {xv[it].x, xv[it].y, xv[it].z, xv[it].w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    xs[(half * 8 + 2 * q) * kXRow + hv] = (uint16_t)(w4[q] & 0xffffu);
                    xs[(half * 8 + 2 * q + 1) * kXRow + hv] = (uint16_t)(w4[q] >> 16);
                }
            }
        }
#pragma unroll
        for (int i0 = 0; i0 < kZIters; i0 += kZBatch) {
            float4 g0[kZBatch], g1[kZBatch];
            uint4 yy[kZBatch];
#pragma unroll
            for (int q = 0; q < kZBatch; ++q) z_item(i0 + q, g0[q], g1[q], yy[q]);
#pragma unroll
            for (int q = 0; q < kZBatch; ++q) {
                const int item = threadIdx.x + (i0 + q) * 256;
                const int bv = item / (CO / 8), cg = item - bv * (CO / 8);
                const float gv[8] = {g0[q].x, g0[q].y, g0[q].z, g0[q].w, g1[q].x, g1[q].y, g1[q].z, g1[q].w};
                const uint32_t y4[4] = {yy[q].x, yy[q].y, yy[q].z, yy[q].w};
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const uint16_t ye = (uint16_t)(e & 1 ? y4[e >> 1] >> 16 : y4[e >> 1] & 0xffffu);
                    zs[(cg * 8 + e) * kZRow + bv] = bf2f(ye) > 0.0f ? f2bf(gv[e]) : (uint16_t)0;
                }
            }
        }
        __syncthreads();
        // z-rows 2 p and 2 p + 1 of the brick (r = 4 lx + ly): K = their 32
        // voxels on v_mfma_f32_16x16x32_bf16; lane group kg holds K 8 kg ..
        // 8 kg + 7, i.e. row 2 p + kg / 2, voxels 8 (kg % 2) .. 8 (kg % 2) + 7
        for (int p = 0; p < kWX * kWY / 2; ++p) {
            const int rr = 2 * p + (kg >> 1);
            const int lx = rr / kWY, ly = rr - lx * kWY;
            uint4 b[TO];
#pragma unroll
            for (int t = 0; t < TO; ++t)
                b[t] = *reinterpret_cast<const uint4 *>(zs + (t * 16 + m) * kZRow + rr * kWZ + (kg & 1) * 8);
#pragma unroll
            for (int tt = 0; tt < kTapsPerWave; ++tt) {
                const int tap = wave + 4 * tt;
                if (tap < 27) {  // wave-uniform
                    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
                    const int hv = ((lx + 1 + dx) * kWHY + (ly + 1 + dy)) * kWHZ + 1 + dz + (kg & 1) * 8;
                    // the eight voxels hv .. hv + 7 of channel m from the three
                    // aligned 8-byte words that hold them (an unaligned LDS
                    // read stalls); k = hv % 4
                    const int o = m * kXRow + hv;
                    const uint2 *w2 = reinterpret_cast<const uint2 *>(xs + (o & ~3));
                    const uint2 q0 = w2[0], q1 = w2[1], q2 = w2[2];
                    const bool hi = (hv & 2) != 0, odd = (hv & 1) != 0;
                    const uint32_t e0 = hi ? q0.y : q0.x, e1 = hi ? q1.x : q0.y, e2 = hi ? q1.y : q1.x,
                                   e3 = hi ? q2.x : q1.y, e4 = hi ? q2.y : q2.x;
                    const uint32_t r0 = odd ? __builtin_amdgcn_alignbit(e1, e0, 16) : e0;
                    const uint32_t r1 = odd ? __builtin_amdgcn_alignbit(e2, e1, 16) : e1;
                    const uint32_t r2 = odd ? __builtin_amdgcn_alignbit(e3, e2, 16) : e2;
                    const uint32_t r3 = odd ? __builtin_amdgcn_alignbit(e4, e3, 16) : e3;
                    const bf16x8_t a = __builtin_bit_cast(bf16x8_t, make_uint4(r0, r1, r2, r3));
#pragma unroll
                    for (int t = 0; t < TO; ++t)
                        acc[tt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, __builtin_bit_cast(bf16x8_t, b[t]),
                                                                             acc[tt][t], 0, 0, 0);
                }
            }
        }
    }
    // D[row = ci][col = co]: the lane holds column m, rows kg*4 + j
    float *dst = partials + (int64_t)blockIdx.x * 27 * 16 * CO;
#pragma unroll
    for (int tt = 0; tt < kTapsPerWave; ++tt) {
        const int tap = wave + 4 * tt;
        if (tap >= 27) continue;
#pragma unroll
        for (int t = 0; t < TO; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) dst[((int64_t)tap * 16 + kg * 4 + j) * CO + t * 16 + m] = acc[tt][t][j];
    }
}

// dW[tap][c0 + ci][co] (+)= sum_b partials[b][tap][ci][co], blocks in order
__global__ void reduce_wgrad_kernel(const float *__restrict__ partials, int blocks, int CI, int CO, int c0,
                                    float *__restrict__ dw, int accumulate) {
    const int64_t n = 27LL * 16 * CO;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int b = 0; b < blocks; ++b) s += partials[(int64_t)b * n + i];
    const int tap = (int)(i / (16 * CO)), rem = (int)(i % (16 * CO)), ci = rem / CO, co = rem % CO;
    float &o = dw[((int64_t)tap * CI + c0 + ci) * CO + co];
    o = accumulate ? o + s : s;
}

// Sums over voxels of a [V][C] tensor in a fixed order: block b adds the
// elements of its range of whole voxels, thread t the channel t % C (C
// divides 256), and writes partials[b][c]; reduce_channel_kernel adds the
// blocks.
constexpr int kSumBlocks = 256;

// dbias[co] partials: sum_v Z[v][co], Z = dL/dY where Y > 0 (FP32)
__global__ __launch_bounds__(256) void bias_grad_partials_kernel(const float *__restrict__ gy,
                                                                 const uint16_t *__restrict__ y, int64_t V, int CO,
                                                                 float *__restrict__ partials) {
    __shared__ float sh[256];
    const int64_t per = (V + gridDim.x - 1) / gridDim.x;
    const int64_t e0 = (int64_t)blockIdx.x * per * CO, e1 = min((int64_t)(blockIdx.x + 1) * per, V) * CO;
    float s = 0.0f;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) s += bf2f(y[e]) > 0.0f ? gy[e] : 0.0f;
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off >= CO; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < CO) partials[(int64_t)blockIdx.x * CO + threadIdx.x] = sh[threadIdx.x];
}

// dst[c] (+)= sum_b partials[b][c] for c < C, blocks in order
__global__ void reduce_channel_kernel(const float *__restrict__ partials, int blocks, int C, float *__restrict__ dst,
                                      int accumulate) {
    const int c = threadIdx.x;
    if (c >= C) return;
    float s = 0.0f;
    for (int b = 0; b < blocks; ++b) s += partials[(int64_t)b * C + c];
    dst[c] = accumulate ? dst[c] + s : s;
}


__device__ __forceinline__ float block_sum(float s) {
    __shared__ float sh[256];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    return sh[0];
}

// ---- pooling, upsampling with concatenation ------------------------------------------------
// fine dims df, coarse dims dc = df / 2
__device__ __forceinline__ int64_t child(const Dims &df, int x, int y, int z, int q) {
    return ((int64_t)(2 * x + (q >> 2)) * df.Y + (2 * y + ((q >> 1) & 1))) * df.Z + 2 * z + (q & 1);
}
__device__ __forceinline__ int64_t parent(const Dims &df, const Dims &dc, int64_t v) {
    int x, y, z;
    vox(df, v, x, y, z);
    return ((int64_t)(x >> 1) * dc.Y + (y >> 1)) * dc.Z + (z >> 1);
}

__global__ void pool_fwd_kernel(const uint16_t *__restrict__ in, uint16_t *__restrict__ out, Dims df, Dims dc, int C) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= dc.V * C) return;
    const int64_t v = i / C;
    const int c = (int)(i - v * C);
    int x, y, z;
    vox(dc, v, x, y, z);
    float s = 0.0f;
    for (int q = 0; q < 8; ++q) s += bf2f(in[child(df, x, y, z, q) * C + c]);
    out[i] = f2bf(s * 0.125f);
}

// gin (+)= gout[parent] / 8
__global__ void pool_bwd_kernel(const float *__restrict__ gout, float *__restrict__ gin, Dims df, Dims dc, int C,
                                int accumulate) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= df.V * C) return;
    const int64_t v = i / C;
    const int c = (int)(i - v * C);
    const float g = gout[parent(df, dc, v) * C + c] * 0.125f;
    gin[i] = accumulate ? gin[i] + g : g;
}

// out[v] = [coarse[parent(v)][0:CA], skip[v][0:CB]]
__global__ void upcat_fwd_kernel(const uint16_t *__restrict__ coarse, const uint16_t *__restrict__ skip,
                                 uint16_t *__restrict__ out, Dims df, Dims dc, int CA, int CB) {
    const int C = CA + CB;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= df.V * C) return;
    const int64_t v = i / C;
    const int c = (int)(i - v * C);
    out[i] = c < CA ? coarse[parent(df, dc, v) * CA + c] : skip[v * CB + (c - CA)];
}

// coarse gradient = sum over the 8 children of gout[:, 0:CA]
__global__ void upcat_bwd_coarse_kernel(const float *__restrict__ gout, float *__restrict__ gcoarse, Dims df, Dims dc,
                                        int CA, int CB) {
    const int C = CA + CB;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= dc.V * CA) return;
    const int64_t v = i / CA;
    const int c = (int)(i - v * CA);
    int x, y, z;
    vox(dc, v, x, y, z);
    float s = 0.0f;
    for (int q = 0; q < 8; ++q) s += gout[child(df, x, y, z, q) * C + c];
    gcoarse[i] = s;
}

// skip gradient = gout[:, CA:]
__global__ void upcat_bwd_skip_kernel(const float *__restrict__ gout, float *__restrict__ gskip, int64_t V, int CA,
                                      int CB) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= V * CB) return;
    const int64_t v = i / CB;
    gskip[i] = gout[v * (CA + CB) + CA + (i - v * CB)];
}

// ---- input normalisation, packing, head, loss ----------------------------------------------
// network voxel v -> pressure-grid cell g; false in the padding
__device__ __forceinline__ bool grid_cell(const Dims &G, const Dims &d0, int64_t v, int64_t &g, bool &inner) {
    int x, y, z;
    vox(d0, v, x, y, z);
    if (x >= G.X || y >= G.Y || z >= G.Z) return false;
    g = ((int64_t)x * G.Y + y) * G.Z + z;
    inner = x > 0 && y > 0 && z > 0 && x < G.X - 1 && y < G.Y - 1 && z < G.Z - 1;
    return true;
}

constexpr int kNormBlocks = 1024;

// partial sums of rhs^2 and of the count over interior cells with mask > 0
__global__ __launch_bounds__(256) void norm_partials_kernel(const float *__restrict__ rhs, const float *__restrict__ mask,
                                                            Dims G, double *__restrict__ partials) {
    __shared__ double sh[2][256];
    double s = 0.0, n = 0.0;
    for (int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x; g < G.V; g += (int64_t)gridDim.x * 256) {
        int x, y, z;
        vox(G, g, x, y, z);
        if (x > 0 && y > 0 && z > 0 && x < G.X - 1 && y < G.Y - 1 && z < G.Z - 1 && mask[g] > 0.0f) {
            const double r = rhs[g];
            s += r * r;
            n += 1.0;
        }
    }
    sh[0][threadIdx.x] = s;
    sh[1][threadIdx.x] = n;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = sh[0][0];
        partials[kNormBlocks + blockIdx.x] = sh[1][0];
    }
}

// scal[0] = 1/s, scal[1] = k = s * len2, scal[2] = 1/count, scal[3] = 1/k
// (all 0 when there is no fluid cell or rhs vanishes there)
__global__ __launch_bounds__(256) void norm_final_kernel(const double *__restrict__ partials, int blocks, double len2,
                                                         float *__restrict__ scal) {
    __shared__ double sh[2][256];
    double s = 0.0, n = 0.0;
    for (int b = threadIdx.x; b < blocks; b += 256) {
        s += partials[b];
        n += partials[kNormBlocks + b];
    }
    sh[0][threadIdx.x] = s;
    sh[1][threadIdx.x] = n;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double cnt = sh[1][0];
        const double rms = cnt > 0.0 ? sqrt(sh[0][0] / cnt) : 0.0;
        scal[0] = rms > 0.0 ? (float)(1.0 / rms) : 0.0f;
        scal[1] = (float)(rms * len2);
        scal[2] = cnt > 0.0 ? (float)(1.0 / cnt) : 0.0f;
        scal[3] = rms > 0.0 ? (float)(1.0 / (rms * len2)) : 0.0f;
    }
}

// in16[v] = bf16 [rhs / s, mask > 0, prev / k, 0 ... 0]
__global__ void pack_input_kernel(const float *__restrict__ rhs, const float *__restrict__ mask,
                                  const float *__restrict__ prev, const float *__restrict__ scal, Dims G, Dims d0,
                                  uint16_t *__restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= d0.V) return;
    int64_t g = 0;
    bool inner = false;
    uint32_t w0 = 0, w1 = 0;
    if (grid_cell(G, d0, v, g, inner)) {
        const float is = scal[0];
        const uint16_t c0 = f2bf(rhs[g] * is), c1 = mask[g] > 0.0f ? (uint16_t)0x3f80 : (uint16_t)0;
        const uint16_t c2 = prev ? f2bf(prev[g] * scal[3]) : (uint16_t)0;
        w0 = (uint32_t)c0 | ((uint32_t)c1 << 16);
        w1 = c2;
    }
    uint4 *o = reinterpret_cast<uint4 *>(out + v * 16);
    o[0] = make_uint4(w0, w1, 0u, 0u);
    o[1] = make_uint4(0u, 0u, 0u, 0u);
}

// pred[v] = sum_c w[c] y[v][c] + b   (normalised prediction, FP32)
__global__ void head_fwd_kernel(const uint16_t *__restrict__ y, const float *__restrict__ w, int64_t V,
                                float *__restrict__ pred) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= V) return;
    float s = w[16];
    for (int c = 0; c < 16; ++c) s += w[c] * bf2f(y[v * 16 + c]);
    pred[v] = s;
}

// out[g] = prev[g] + k * pred[v(g)] on the pressure grid (prev may be NULL)
__global__ void unpack_output_kernel(const float *__restrict__ pred, const float *__restrict__ scal,
                                     const float *__restrict__ prev, Dims G, Dims d0, float *__restrict__ out) {
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (g >= G.V) return;
    int x, y, z;
    vox(G, g, x, y, z);
    const float change = scal[1] * pred[((int64_t)x * d0.Y + y) * d0.Z + z];
    out[g] = prev ? prev[g] + change : change;
}

// gpred[v] = k * grad_out[g(v)] (dL/d pred from dL/d out), zero in the padding
__global__ void pack_grad_kernel(const float *__restrict__ gout, const float *__restrict__ scal, Dims G, Dims d0,
                                 float *__restrict__ gpred) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= d0.V) return;
    int64_t g = 0;
    bool inner = false;
    gpred[v] = grid_cell(G, d0, v, g, inner) ? gout[g] * scal[1] : 0.0f;
}

// Mean squared error of the normalised prediction over the interior fluid
// cells: e = pred - (target - prev) / k, L = sum e^2 / count,
// gpred = 2 e / count / batch (prev may be NULL).
__global__ __launch_bounds__(256) void mse_kernel(const float *__restrict__ pred, const float *__restrict__ target,
                                                  const float *__restrict__ prev, const float *__restrict__ mask,
                                                  const float *__restrict__ scal, Dims G, Dims d0, float inv_batch,
                                                  float *__restrict__ gpred, float *__restrict__ partial) {
    float s = 0.0f;
    const float is = scal[3], inv_n = scal[2];
    for (int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x; v < d0.V; v += (int64_t)gridDim.x * 256) {
        int64_t g = 0;
        bool inner = false;
        float gv = 0.0f;
        if (grid_cell(G, d0, v, g, inner) && inner && mask[g] > 0.0f) {
            const float e = pred[v] - (prev ? target[g] - prev[g] : target[g]) * is;
            gv = 2.0f * e * inv_n * inv_batch;
            s += e * e * inv_n;
        }
        gpred[v] = gv;
    }
    s = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// gy[v][c] = g[v] * w[c]
__global__ void head_bwd_data_kernel(const float *__restrict__ g, const float *__restrict__ w, int64_t V,
                                     float *__restrict__ gy) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= V * 16) return;
    gy[i] = g[i / 16] * w[i % 16];
}

// head weight gradient partials: sum_v g[v] y[v][c] (c < 16) and sum_v g[v]
// (c = 16), summed like bias_grad_partials_kernel (thread t: channel t % 16)
__global__ __launch_bounds__(256) void head_bwd_weight_partials_kernel(const float *__restrict__ g,
                                                                       const uint16_t *__restrict__ y, int64_t V,
                                                                       float *__restrict__ partials) {
    __shared__ float sh[256], shb[256];
    const int64_t per = (V + gridDim.x - 1) / gridDim.x;
    const int64_t e0 = (int64_t)blockIdx.x * per * 16, e1 = min((int64_t)(blockIdx.x + 1) * per, V) * 16;
    float s = 0.0f, sb = 0.0f;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) {
        const float gv = g[e >> 4];
        s += gv * bf2f(y[e]);
        if ((threadIdx.x & 15) == 0) sb += gv;
    }
    sh[threadIdx.x] = s;
    shb[threadIdx.x] = sb;
    __syncthreads();
    for (int off = 128; off >= 16; off >>= 1) {
        if (threadIdx.x < off) {
            sh[threadIdx.x] += sh[threadIdx.x + off];
            shb[threadIdx.x] += shb[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x < 16) partials[(int64_t)blockIdx.x * 17 + threadIdx.x] = sh[threadIdx.x];
    if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * 17 + 16] = shb[0];
}

// Adam with bias correction, FP32 master weights.
__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m1,
                            float *__restrict__ m2, int64_t n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.0f - b1) * gi;
    const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// BF16 copies of one conv layer from the master W[tap][ci][co]:
// fwd[tap][co][ci] and dgrad[tap][ci][co].
__global__ void pack_weights_kernel(const float *__restrict__ w, int CI, int CO, uint16_t *__restrict__ fwd,
                                    uint16_t *__restrict__ bwd) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t n = (int64_t)27 * CI * CO;
    if (i >= n) return;
    const int tap = (int)(i / ((int64_t)CI * CO));
    const int rem = (int)(i % ((int64_t)CI * CO));
    const int ci = rem / CO, co = rem % CO;
    const uint16_t b = f2bf(w[i]);
    bwd[i] = b;
    fwd[((int64_t)tap * CO + co) * CI + ci] = b;
}

// ---- host side ------------------------------------------------------------------------------
constexpr int kLayers = 10;
constexpr int kLayerIO[kLayers][2] = {{16, 16}, {16, 16}, {16, 32}, {32, 32}, {32, 64},
                                      {64, 64}, {96, 32}, {32, 32}, {48, 16}, {16, 16}};
constexpr int kLayerLevel[kLayers] = {0, 0, 1, 1, 2, 2, 1, 1, 0, 0};
constexpr int kLossBlocks = 512;

int64_t param_count() {
    int64_t n = 0;
    for (const auto &io : kLayerIO) n += 27LL * io[0] * io[1] + io[1];
    return n + 17;
}

inline dim3 g1(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace
}  // namespace clq

using namespace clq;

struct clq_net {
    int device = 0;
    hipStream_t stream = nullptr;
    Dims G{}, d0{}, d1{}, d2{};
    double dx = 0.0;
    int64_t nparams = 0;
    int64_t woff[kLayers] = {}, boff[kLayers] = {}, head_off = 0;
    float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    uint16_t *wf[kLayers] = {}, *wb[kLayers] = {};
    uint16_t *in16 = nullptr, *a1 = nullptr, *s0 = nullptr, *p0 = nullptr, *a3 = nullptr, *s1 = nullptr, *p1 = nullptr,
             *a5 = nullptr, *a6 = nullptr, *u1 = nullptr, *a7 = nullptr, *a8 = nullptr, *u0 = nullptr, *a9 = nullptr,
             *a10 = nullptr;
    float *pred = nullptr, *scal = nullptr;
    double *norm_partials = nullptr;
    // training state, allocated on first use
    bool train_ready = false;
    float *gpred = nullptr, *gbuf[3] = {}, *gs0 = nullptr, *gs1 = nullptr, *partials = nullptr;
    float *loss_partial = nullptr, *host_loss = nullptr;
    float *stage[4] = {};  // device copies of host inputs (rhs, mask, prev, target or output)
    int64_t adam_steps = 0;
    bool forward_done = false;
    std::vector<void *> allocs;

    int alloc(void **p, size_t bytes) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, bytes ? bytes : 16);
        if (e != hipSuccess) return set_error(CLQ_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        allocs.push_back(q);
        *p = q;
        e = hipMemsetAsync(q, 0, bytes ? bytes : 16, stream);
        if (e != hipSuccess) return set_error(CLQ_ERR_DEVICE, "hipMemsetAsync: %s", hipGetErrorString(e));
        return CLQ_OK;
    }
    template <typename T>
    int alloc_n(T **p, int64_t count) {
        return alloc(reinterpret_cast<void **>(p), (size_t)count * sizeof(T));
    }
};

namespace clq {
namespace {

#define NET_TRY(expr)                                                                                          \
    do {                                                                                                       \
        hipError_t e_ = (expr);                                                                                \
        if (e_ != hipSuccess)                                                                                  \
            return set_error(CLQ_ERR_DEVICE, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
    } while (0)

int net_check(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(CLQ_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
    return CLQ_OK;
}

inline dim3 conv_grid(const Dims &d) {
    return dim3((unsigned)((d.Z + kBZ - 1) / kBZ), (unsigned)((d.Y + kBY - 1) / kBY), (unsigned)((d.X + kBX - 1) / kBX));
}
template <int CI, int CO>
void conv_fwd(const uint16_t *x, const uint16_t *wf, const float *b, uint16_t *y, Dims d, hipStream_t s) {
    hipLaunchKernelGGL((conv3_lds_kernel<CI, CO, 0>), conv_grid(d), dim3(256), 0, s, x, nullptr, wf, b, y, d, nullptr,
                       Dims{0, 0, 0, 0});
}
// input [up(coarse) (CA channels, dims dc), skip (CI - CA channels)]
template <int CI, int CO, int CA>
void conv_fwd_cat(const uint16_t *coarse, Dims dc, const uint16_t *skip, const uint16_t *wf, const float *b,
                  uint16_t *y, Dims d, hipStream_t s) {
    hipLaunchKernelGGL((conv3_lds_kernel<CI, CO, 0, CA>), conv_grid(d), dim3(256), 0, s, skip, nullptr, wf, b, y, d,
                       coarse, dc);
}
// data gradient of a CI -> CO layer: maps dL/dY (CO channels) to dL/dX (CI)
template <int CI, int CO>
void conv_dgrad(const float *gy, const uint16_t *y, const uint16_t *wb, float *gx, Dims d, hipStream_t s) {
    hipLaunchKernelGGL((conv3_lds_kernel<CO, CI, 1>), conv_grid(d), dim3(256), 0, s, gy, y, wb, nullptr, gx, d,
                       nullptr, Dims{0, 0, 0, 0});
}
template <int CI, int CO>
void conv_wgrad(const uint16_t *x, const float *gy, const uint16_t *y, float *partials, float *dw, float *db, Dims d,
                int accumulate, hipStream_t s) {
    const int blocks = wgrad_blocks(d);
    for (int c0 = 0; c0 < CI; c0 += 16) {
        hipLaunchKernelGGL((conv3_wgrad_kernel<CI, CO>), dim3(blocks), dim3(256), 0, s, x, gy, y, partials, d, c0);
        hipLaunchKernelGGL(reduce_wgrad_kernel, g1(27LL * 16 * CO), dim3(256), 0, s, partials, blocks, CI, CO, c0, dw,
                           accumulate);
    }
    hipLaunchKernelGGL(bias_grad_partials_kernel, dim3(kSumBlocks), dim3(256), 0, s, gy, y, d.V, CO, partials);
    hipLaunchKernelGGL(reduce_channel_kernel, dim3(1), dim3(64), 0, s, partials, kSumBlocks, CO, db, accumulate);
}

void normalise_and_pack(clq_net *n, const float *rhs, const float *mask, const float *prev, hipStream_t s) {
    hipLaunchKernelGGL(norm_partials_kernel, dim3(kNormBlocks), dim3(256), 0, s, rhs, mask, n->G, n->norm_partials);
    const double h = (double)(float)n->dx;
    const double extent = h * (double)std::max(n->G.X, std::max(n->G.Y, n->G.Z));
    hipLaunchKernelGGL(norm_final_kernel, dim3(1), dim3(256), 0, s, n->norm_partials, kNormBlocks,
                       extent * extent / 1024.0, n->scal);
    hipLaunchKernelGGL(pack_input_kernel, g1(n->d0.V), dim3(256), 0, s, rhs, mask, prev, n->scal, n->G, n->d0,
                       n->in16);
}

// Forward pass from n->in16 to the normalised prediction n->pred.
void forward(clq_net *n, hipStream_t s) {
    const Dims d0 = n->d0, d1 = n->d1, d2 = n->d2;
    const float *w = n->w;
    conv_fwd<16, 16>(n->in16, n->wf[0], w + n->boff[0], n->a1, d0, s);
    conv_fwd<16, 16>(n->a1, n->wf[1], w + n->boff[1], n->s0, d0, s);
    hipLaunchKernelGGL(pool_fwd_kernel, g1(d1.V * 16), dim3(256), 0, s, n->s0, n->p0, d0, d1, 16);
    conv_fwd<16, 32>(n->p0, n->wf[2], w + n->boff[2], n->a3, d1, s);
    conv_fwd<32, 32>(n->a3, n->wf[3], w + n->boff[3], n->s1, d1, s);
    hipLaunchKernelGGL(pool_fwd_kernel, g1(d2.V * 32), dim3(256), 0, s, n->s1, n->p1, d1, d2, 32);
    conv_fwd<32, 64>(n->p1, n->wf[4], w + n->boff[4], n->a5, d2, s);
    conv_fwd<64, 64>(n->a5, n->wf[5], w + n->boff[5], n->a6, d2, s);
    conv_fwd_cat<96, 32, 64>(n->a6, d2, n->s1, n->wf[6], w + n->boff[6], n->a7, d1, s);
    conv_fwd<32, 32>(n->a7, n->wf[7], w + n->boff[7], n->a8, d1, s);
    conv_fwd_cat<48, 16, 32>(n->a8, d1, n->s0, n->wf[8], w + n->boff[8], n->a9, d0, s);
    conv_fwd<16, 16>(n->a9, n->wf[9], w + n->boff[9], n->a10, d0, s);
    hipLaunchKernelGGL(head_fwd_kernel, g1(d0.V), dim3(256), 0, s, n->a10, w + n->head_off, d0.V, n->pred);
}

// Backward pass from n->gpred (dL/d pred); parameter gradients into n->g,
// added to what is there when `acc`.  Gradient buffers hold dL/dY of the
// ReLU outputs; the consuming kernels apply the ReLU mask.
void backward(clq_net *n, int acc, hipStream_t s) {
    const Dims d0 = n->d0, d1 = n->d1, d2 = n->d2;
    float *gw = n->g, *A = n->gbuf[0], *B = n->gbuf[1], *Cb = n->gbuf[2], *P = n->partials;
    // the weight gradients of c7 and c9 read their concatenated inputs
    hipLaunchKernelGGL(upcat_fwd_kernel, g1(d1.V * 96), dim3(256), 0, s, n->a6, n->s1, n->u1, d1, d2, 64, 32);
    hipLaunchKernelGGL(upcat_fwd_kernel, g1(d0.V * 48), dim3(256), 0, s, n->a8, n->s0, n->u0, d0, d1, 32, 16);
    hipLaunchKernelGGL(head_bwd_weight_partials_kernel, dim3(kSumBlocks), dim3(256), 0, s, n->gpred, n->a10, d0.V, P);
    hipLaunchKernelGGL(reduce_channel_kernel, dim3(1), dim3(64), 0, s, P, kSumBlocks, 17, gw + n->head_off, acc);
    hipLaunchKernelGGL(head_bwd_data_kernel, g1(d0.V * 16), dim3(256), 0, s, n->gpred, n->w + n->head_off, d0.V, A);
    // c10: a9 -> a10
    conv_wgrad<16, 16>(n->a9, A, n->a10, P, gw + n->woff[9], gw + n->boff[9], d0, acc, s);
    conv_dgrad<16, 16>(A, n->a10, n->wb[9], B, d0, s);
    // c9: u0 -> a9; dL/du0 splits into the skip (s0) and the upsampled a8
    conv_wgrad<48, 16>(n->u0, B, n->a9, P, gw + n->woff[8], gw + n->boff[8], d0, acc, s);
    conv_dgrad<48, 16>(B, n->a9, n->wb[8], A, d0, s);
    hipLaunchKernelGGL(upcat_bwd_skip_kernel, g1(d0.V * 16), dim3(256), 0, s, A, n->gs0, d0.V, 32, 16);
    hipLaunchKernelGGL(upcat_bwd_coarse_kernel, g1(d1.V * 32), dim3(256), 0, s, A, B, d0, d1, 32, 16);
    // c8: a7 -> a8
    conv_wgrad<32, 32>(n->a7, B, n->a8, P, gw + n->woff[7], gw + n->boff[7], d1, acc, s);
    conv_dgrad<32, 32>(B, n->a8, n->wb[7], Cb, d1, s);
    // c7: u1 -> a7; dL/du1 splits into the skip (s1) and the upsampled a6
    conv_wgrad<96, 32>(n->u1, Cb, n->a7, P, gw + n->woff[6], gw + n->boff[6], d1, acc, s);
    conv_dgrad<96, 32>(Cb, n->a7, n->wb[6], A, d1, s);
    hipLaunchKernelGGL(upcat_bwd_skip_kernel, g1(d1.V * 32), dim3(256), 0, s, A, n->gs1, d1.V, 64, 32);
    hipLaunchKernelGGL(upcat_bwd_coarse_kernel, g1(d2.V * 64), dim3(256), 0, s, A, B, d1, d2, 64, 32);
    // c6: a5 -> a6, c5: p1 -> a5
    conv_wgrad<64, 64>(n->a5, B, n->a6, P, gw + n->woff[5], gw + n->boff[5], d2, acc, s);
    conv_dgrad<64, 64>(B, n->a6, n->wb[5], Cb, d2, s);
    conv_wgrad<32, 64>(n->p1, Cb, n->a5, P, gw + n->woff[4], gw + n->boff[4], d2, acc, s);
    conv_dgrad<32, 64>(Cb, n->a5, n->wb[4], A, d2, s);
    hipLaunchKernelGGL(pool_bwd_kernel, g1(d1.V * 32), dim3(256), 0, s, A, n->gs1, d1, d2, 32, 1);
    // c4: a3 -> s1, c3: p0 -> a3
    conv_wgrad<32, 32>(n->a3, n->gs1, n->s1, P, gw + n->woff[3], gw + n->boff[3], d1, acc, s);
    conv_dgrad<32, 32>(n->gs1, n->s1, n->wb[3], B, d1, s);
    conv_wgrad<16, 32>(n->p0, B, n->a3, P, gw + n->woff[2], gw + n->boff[2], d1, acc, s);
    conv_dgrad<16, 32>(B, n->a3, n->wb[2], A, d1, s);
    hipLaunchKernelGGL(pool_bwd_kernel, g1(d0.V * 16), dim3(256), 0, s, A, n->gs0, d0, d1, 16, 1);
    // c2: a1 -> s0, c1: in16 -> a1
    conv_wgrad<16, 16>(n->a1, n->gs0, n->s0, P, gw + n->woff[1], gw + n->boff[1], d0, acc, s);
    conv_dgrad<16, 16>(n->gs0, n->s0, n->wb[1], B, d0, s);
    conv_wgrad<16, 16>(n->in16, B, n->a1, P, gw + n->woff[0], gw + n->boff[0], d0, acc, s);
}

void pack_weights(clq_net *n) {
    for (int i = 0; i < kLayers; ++i) {
        const int64_t cnt = 27LL * kLayerIO[i][0] * kLayerIO[i][1];
        hipLaunchKernelGGL(pack_weights_kernel, g1(cnt), dim3(256), 0, n->stream, n->w + n->woff[i], kLayerIO[i][0],
                           kLayerIO[i][1], n->wf[i], n->wb[i]);
    }
}

int ensure_training(clq_net *n) {
    if (n->train_ready) return CLQ_OK;
    const int64_t V0 = n->d0.V, V1 = n->d1.V, V2 = n->d2.V;
    const int64_t gmax = std::max(std::max(V0 * 48, V1 * 96), V2 * 64);
    const Dims *lvl[3] = {&n->d0, &n->d1, &n->d2};
    int64_t pmax = (int64_t)kSumBlocks * 64;  // bias and head partials
    for (int i = 0; i < kLayers; ++i)
        pmax = std::max(pmax, (int64_t)wgrad_blocks(*lvl[kLayerLevel[i]]) * 27 * 16 * kLayerIO[i][1]);
    if (int rc = n->alloc_n(&n->g, n->nparams)) return rc;
    if (int rc = n->alloc_n(&n->m1, n->nparams)) return rc;
    if (int rc = n->alloc_n(&n->m2, n->nparams)) return rc;
    if (int rc = n->alloc_n(&n->gpred, V0)) return rc;
    for (auto &b : n->gbuf)
        if (int rc = n->alloc_n(&b, gmax)) return rc;
    if (int rc = n->alloc_n(&n->gs0, V0 * 16)) return rc;
    if (int rc = n->alloc_n(&n->gs1, V1 * 32)) return rc;
    if (int rc = n->alloc_n(&n->u0, V0 * 48)) return rc;  // concatenated inputs of c9 and c7
    if (int rc = n->alloc_n(&n->u1, V1 * 96)) return rc;
    if (int rc = n->alloc_n(&n->partials, pmax)) return rc;
    if (int rc = n->alloc_n(&n->loss_partial, kLossBlocks)) return rc;
    for (auto &b : n->stage)
        if (int rc = n->alloc_n(&b, n->G.V)) return rc;
    NET_TRY(hipHostMalloc((void **)&n->host_loss, kLossBlocks * sizeof(float), hipHostMallocDefault));
    n->train_ready = true;
    return CLQ_OK;
}

// splitmix64 + Box-Muller: a fixed, platform-independent normal sequence
struct InitRng {
    uint64_t state;
    uint64_t next() {
        uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return z ^ (z >> 31);
    }
    double uniform() { return ((double)(next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

}  // namespace

int unet_check_grid(const clq_net *n, int nx, int ny, int nz, double dx) {
    if (!n) return set_error(CLQ_ERR_INVALID, "no network attached");
    if (n->G.X != nx || n->G.Y != ny || n->G.Z != nz)
        return set_error(CLQ_ERR_INVALID, "network grid %dx%dx%d does not match the context grid %dx%dx%d", n->G.X,
                         n->G.Y, n->G.Z, nx, ny, nz);
    if ((float)n->dx != (float)dx)
        return set_error(CLQ_ERR_INVALID, "network cell size %g does not match the context's %g", n->dx, dx);
    return CLQ_OK;
}

int unet_device(const clq_net *n) { return n->device; }

// out = network(rhs, mask, prev) on the pressure grid, enqueued on stream s;
// all pointers are device arrays of the pressure grid, prev may be NULL.
void unet_predict(clq_net *n, const float *rhs, const float *mask, const float *prev, float *out, hipStream_t s) {
    normalise_and_pack(n, rhs, mask, prev, s);
    forward(n, s);
    hipLaunchKernelGGL(unpack_output_kernel, g1(n->G.V), dim3(256), 0, s, n->pred, n->scal, prev, n->G, n->d0, out);
}

}  // namespace clq

// ---- C-ABI ----------------------------------------------------------------------------------
extern "C" {

int64_t clq_net_param_count(void) { return param_count(); }

int clq_net_destroy(clq_net *n) {
    if (!n) return CLQ_OK;
    (void)hipSetDevice(n->device);
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    for (void *p : n->allocs) (void)hipFree(p);
    if (n->host_loss) (void)hipHostFree(n->host_loss);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return CLQ_OK;
}

int clq_net_create(int32_t nx, int32_t ny, int32_t nz, double dx, clq_net **out) {
    if (!out) return set_error(CLQ_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!(dx > 0.0)) return set_error(CLQ_ERR_INVALID, "dx must be positive");
    if (nx < 4 || ny < 4 || nz < 4) return set_error(CLQ_ERR_INVALID, "grid must be at least 4 cells in each dimension");
    const int64_t vol = (int64_t)((nx + 3) / 4 * 4) * ((ny + 3) / 4 * 4) * ((nz + 3) / 4 * 4);
    if (vol >= ((int64_t)1 << 31) / 48) return set_error(CLQ_ERR_INVALID, "grid too large for the network");
    clq_net *n = new clq_net;
    (void)hipGetDevice(&n->device);
    if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
        delete n;
        return set_error(CLQ_ERR_DEVICE, "hipStreamCreate failed");
    }
    n->G = make_dims(nx, ny, nz);
    n->dx = dx;
    n->d0 = make_dims((nx + 3) / 4 * 4, (ny + 3) / 4 * 4, (nz + 3) / 4 * 4);
    n->d1 = make_dims(n->d0.X / 2, n->d0.Y / 2, n->d0.Z / 2);
    n->d2 = make_dims(n->d0.X / 4, n->d0.Y / 4, n->d0.Z / 4);
    int64_t off = 0;
    for (int i = 0; i < kLayers; ++i) {
        n->woff[i] = off;
        off += 27LL * kLayerIO[i][0] * kLayerIO[i][1];
        n->boff[i] = off;
        off += kLayerIO[i][1];
    }
    n->head_off = off;
    n->nparams = off + 17;
    const int64_t V0 = n->d0.V, V1 = n->d1.V, V2 = n->d2.V;
    int rc = n->alloc_n(&n->w, n->nparams);
    for (int i = 0; i < kLayers && !rc; ++i) {
        const int64_t cnt = 27LL * kLayerIO[i][0] * kLayerIO[i][1];
        rc = n->alloc_n(&n->wf[i], cnt);
        if (!rc) rc = n->alloc_n(&n->wb[i], cnt);
    }
    struct {
        uint16_t **p;
        int64_t count;
    } acts[] = {{&n->in16, V0 * 16}, {&n->a1, V0 * 16}, {&n->s0, V0 * 16}, {&n->p0, V1 * 16}, {&n->a3, V1 * 32},
                {&n->s1, V1 * 32},   {&n->p1, V2 * 32}, {&n->a5, V2 * 64}, {&n->a6, V2 * 64},
                {&n->a7, V1 * 32},   {&n->a8, V1 * 32}, {&n->a9, V0 * 16}, {&n->a10, V0 * 16}};
    for (auto &a : acts)
        if (!rc) rc = n->alloc_n(a.p, a.count);
    if (!rc) rc = n->alloc_n(&n->pred, V0);
    if (!rc) rc = n->alloc_n(&n->scal, 4);
    if (!rc) rc = n->alloc_n(&n->norm_partials, 2 * kNormBlocks);
    if (!rc && hipStreamSynchronize(n->stream) != hipSuccess) rc = set_error(CLQ_ERR_DEVICE, "clearing the buffers failed");
    if (rc) {
        clq_net_destroy(n);
        return rc;
    }
    *out = n;
    return CLQ_OK;
}

int clq_net_set_params(clq_net *n, const float *params, int64_t count) {
    if (!n || !params) return set_error(CLQ_ERR_INVALID, "NULL argument");
    if (count != n->nparams) return set_error(CLQ_ERR_INVALID, "expected %lld parameters", (long long)n->nparams);
    (void)hipSetDevice(n->device);
    NET_TRY(hipMemcpyAsync(n->w, params, (size_t)count * 4, hipMemcpyHostToDevice, n->stream));
    pack_weights(n);
    n->adam_steps = 0;
    if (n->train_ready) {
        NET_TRY(hipMemsetAsync(n->m1, 0, (size_t)count * 4, n->stream));
        NET_TRY(hipMemsetAsync(n->m2, 0, (size_t)count * 4, n->stream));
    }
    NET_TRY(hipStreamSynchronize(n->stream));
    return net_check("clq_net_set_params");
}

int clq_net_init(clq_net *n, uint64_t seed) {
    if (!n) return set_error(CLQ_ERR_INVALID, "net is NULL");
    std::vector<float> p((size_t)n->nparams, 0.0f);
    InitRng rng{seed};
    for (int i = 0; i < kLayers; ++i) {  // He initialisation for the ReLU layers, zero biases
        const double sd = sqrt(2.0 / (27.0 * kLayerIO[i][0]));
        for (int64_t q = 0; q < 27LL * kLayerIO[i][0] * kLayerIO[i][1]; ++q)
            p[(size_t)(n->woff[i] + q)] = (float)(sd * rng.normal());
    }
    for (int c = 0; c < 16; ++c) p[(size_t)(n->head_off + c)] = (float)(0.25 * rng.normal());
    return clq_net_set_params(n, p.data(), n->nparams);
}

int clq_net_get_params(clq_net *n, float *params, int64_t count) {
    if (!n || !params) return set_error(CLQ_ERR_INVALID, "NULL argument");
    if (count != n->nparams) return set_error(CLQ_ERR_INVALID, "expected %lld parameters", (long long)n->nparams);
    (void)hipSetDevice(n->device);
    NET_TRY(hipMemcpyAsync(params, n->w, (size_t)count * 4, hipMemcpyDeviceToHost, n->stream));
    NET_TRY(hipStreamSynchronize(n->stream));
    return CLQ_OK;
}

int clq_net_forward(clq_net *n, const float *rhs, const float *mask, const float *prev, float *out) {
    if (!n || !rhs || !mask || !out) return set_error(CLQ_ERR_INVALID, "NULL argument");
    (void)hipSetDevice(n->device);
    if (int rc = ensure_training(n)) return rc;
    const size_t bytes = (size_t)n->G.V * 4;
    NET_TRY(hipMemcpyAsync(n->stage[0], rhs, bytes, hipMemcpyHostToDevice, n->stream));
    NET_TRY(hipMemcpyAsync(n->stage[1], mask, bytes, hipMemcpyHostToDevice, n->stream));
    if (prev) NET_TRY(hipMemcpyAsync(n->stage[2], prev, bytes, hipMemcpyHostToDevice, n->stream));
    unet_predict(n, n->stage[0], n->stage[1], prev ? n->stage[2] : nullptr, n->stage[3], n->stream);
    NET_TRY(hipMemcpyAsync(out, n->stage[3], bytes, hipMemcpyDeviceToHost, n->stream));
    NET_TRY(hipStreamSynchronize(n->stream));
    n->forward_done = true;
    return net_check("clq_net_forward");
}

int clq_net_backward(clq_net *n, const float *grad_out, float *grads, int64_t count) {
    if (!n || !grad_out || !grads) return set_error(CLQ_ERR_INVALID, "NULL argument");
    if (count != n->nparams) return set_error(CLQ_ERR_INVALID, "expected %lld gradients", (long long)n->nparams);
    if (!n->forward_done) return set_error(CLQ_ERR_INVALID, "clq_net_backward needs a preceding clq_net_forward");
    (void)hipSetDevice(n->device);
    NET_TRY(hipMemcpyAsync(n->stage[3], grad_out, (size_t)n->G.V * 4, hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(pack_grad_kernel, g1(n->d0.V), dim3(256), 0, n->stream, n->stage[3], n->scal, n->G, n->d0,
                       n->gpred);
    backward(n, 0, n->stream);
    NET_TRY(hipMemcpyAsync(grads, n->g, (size_t)count * 4, hipMemcpyDeviceToHost, n->stream));
    NET_TRY(hipStreamSynchronize(n->stream));
    return net_check("clq_net_backward");
}

int clq_net_train_step(clq_net *n, int32_t batch, const float *rhs, const float *mask, const float *prev,
                       const float *target, double lr, double *loss) {
    if (!n || !rhs || !mask || !target) return set_error(CLQ_ERR_INVALID, "NULL argument");
    if (batch < 1) return set_error(CLQ_ERR_INVALID, "batch must be >= 1");
    if (!(lr >= 0.0)) return set_error(CLQ_ERR_INVALID, "lr must be >= 0");
    (void)hipSetDevice(n->device);
    if (int rc = ensure_training(n)) return rc;
    const int64_t cells = n->G.V;
    const size_t bytes = (size_t)cells * 4;
    double total = 0.0;
    for (int b = 0; b < batch; ++b) {
        NET_TRY(hipMemcpyAsync(n->stage[0], rhs + b * cells, bytes, hipMemcpyHostToDevice, n->stream));
        NET_TRY(hipMemcpyAsync(n->stage[1], mask + b * cells, bytes, hipMemcpyHostToDevice, n->stream));
        if (prev) NET_TRY(hipMemcpyAsync(n->stage[2], prev + b * cells, bytes, hipMemcpyHostToDevice, n->stream));
        NET_TRY(hipMemcpyAsync(n->stage[3], target + b * cells, bytes, hipMemcpyHostToDevice, n->stream));
        normalise_and_pack(n, n->stage[0], n->stage[1], prev ? n->stage[2] : nullptr, n->stream);
        forward(n, n->stream);
        hipLaunchKernelGGL(mse_kernel, dim3(kLossBlocks), dim3(256), 0, n->stream, n->pred, n->stage[3],
                           prev ? n->stage[2] : nullptr, n->stage[1], n->scal, n->G, n->d0, 1.0f / (float)batch,
                           n->gpred, n->loss_partial);
        backward(n, b > 0, n->stream);
        NET_TRY(hipMemcpyAsync(n->host_loss, n->loss_partial, kLossBlocks * 4, hipMemcpyDeviceToHost, n->stream));
        NET_TRY(hipStreamSynchronize(n->stream));
        double l = 0.0;
        for (int q = 0; q < kLossBlocks; ++q) l += n->host_loss[q];
        total += l;
    }
    n->adam_steps++;
    const float c1 = (float)(1.0 - pow(0.9, (double)n->adam_steps));
    const float c2 = (float)(1.0 - pow(0.999, (double)n->adam_steps));
    hipLaunchKernelGGL(adam_kernel, g1(n->nparams), dim3(256), 0, n->stream, n->w, n->g, n->m1, n->m2, n->nparams,
                       (float)lr, 0.9f, 0.999f, 1e-8f, c1, c2);
    pack_weights(n);
    NET_TRY(hipStreamSynchronize(n->stream));
    n->forward_done = false;
    if (loss) *loss = total / batch;
    return net_check("clq_net_train_step");
}

}  // extern "C"
