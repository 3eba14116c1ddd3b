// P2G gather onto the MAC grid plus occupancy voxelisation.
//
// Replaces particle_to_grid_gather (reference
// Solver/Kernel/OpenCL/particles.cl:135-227, launched at
// Solver/Kernel/particles.py:142-163 with 8 lanes per cell).
//
// No float atomics.  Each cell keeps eight partial sums per component, one
// per lane of the reference's 8-lane work-group (neighbour bin n of the 27
// goes to partial n % 8), sees each bin's particles in sorted order and
// combines the partials with the reference's tree (0+4, 2+6, 1+5, 3+7, then
// pairs), so the summation order is the reference's and the result is
// bit-identical to the CPU oracle.
//
// Bins whose tent weight is provably zero for a component are skipped for
// that component: with bin b = floor(fma(x, inv_dx, -0.5)) and p = x*inv_dx,
// p - b lies in [0.5, 1.5] (also under the clamps at the domain edges), so a
// u face (offset (1, .5, .5)) only sees bins with y and z offsets in {-1, 0},
// v only x, z offsets in {-1, 0}, w only x, y in {-1, 0}; the occupancy cell
// trunc(p) only sees offsets in {-1, 0}^3.  Skipped terms are exact zeros, so
// skipping does not change a bit.  Of the 27 bins, 8 are visited for all three
// components and 12 for one.
#include <type_traits>
#include <utility>

#include "common.hpp"

namespace clq {

// ---- row-scan gather -----------------------------------------------------------------
// A wave owns the cells (i, j, k0 .. k0 + 61) and lane l the bin k0 - 1 + l of
// each neighbour bin row (i + ox, j + oy, *).  Rows are visited in the
// reference's neighbour order (ox, then oy); in a row, lane l walks its bin's
// particles once and computes the tent weights of each particle for its own
// cell (oz = 0), the cell above (the bin is that cell's oz = -1 neighbour) and
// the cell below (oz = +1).  The neighbours' weights and velocities are handed
// over with wavefront shifts, and each cell adds them with the same fmaf into the
// partial sum of that neighbour index.  The three oz of a row feed three
// different partials, and every partial still receives its bins in increasing
// neighbour index and each bin's particles in sorted order, so the sums are
// bit-identical to the reference's per-cell, per-partial order.  Only the
// component/bin pairs that can carry weight are evaluated (the oz = +1 bin
// only for w).  Each row's positions are first staged in LDS (below), which
// takes the scattered per-lane reads off the vector memory path.
constexpr int kSeg = 62;

// Lane l receives lane l - 1's value (lane_up) or lane l + 1's (lane_down)
// through a DPP wavefront shift (wave_shr:1 / wave_shl:1): a VALU move
// instead of an LDS permute.  Lanes 0 and 63 receive 0 (bound_ctrl).
__device__ __forceinline__ int lane_up(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int lane_down(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ float lane_up(float x) { return __int_as_float(lane_up(__float_as_int(x))); }
__device__ __forceinline__ float lane_down(float x) { return __int_as_float(lane_down(__float_as_int(x))); }

// max(0, 1 - |p - c|), the reference's tent: 1 - |p - c| <= 1, so clamping
// it to [0, 1] gives the same bits, and the clamp is the output modifier of
// the subtraction (one instruction instead of a subtraction and a max)
__device__ __forceinline__ float tent(float p, float c) {
    return __builtin_amdgcn_fmed3f(1.0f - fabsf(p - c), 0.0f, 1.0f);
}

// a particle of the row scan: position in cells (x * inv_dx) and velocity
struct RowParticle {
    float x, y, z, u, v, w;
};

// The bins of one neighbour row are contiguous in the sorted particle
// arrays.  A wave stages the row's positions (in cells) in LDS with
// coalesced loads, one particle per lane, before its lanes walk their bins:
// the walk's position reads then come from LDS instead of 64 scattered
// global reads per step (velocities stay global: staging them too costs more
// occupancy than it saves).  Slot q + q / 8 spreads consecutive bins over
// the LDS banks.  A row segment of 64 bins holds 512 particles at 8 per bin
// and up to ~575 once the fluid compresses (256^3 dam break); kStageCap = 592
// keeps five 4-wave blocks per CU within 160 KiB of LDS.  Rows with more
// particles read global memory.
constexpr int kStageCap = 592;
constexpr int kStageSlots = kStageCap + kStageCap / 8 + 8;
__device__ __forceinline__ int stage_slot(int q) { return q + (q >> 3); }

// Branch-free row visit: a lane whose bin has no particle left (has false)
// carries zero weights, so every accumulation runs on every lane.  Adding a
// zero term changes no bit: the term is +-0 (finite velocities), and a sum
// that starts at +0 and only ever rounds to nearest is never -0, so
// s + (+-0) == s.
template <int OX, int OY>
__device__ __forceinline__ void rowscan_row(const float *__restrict__ pos, const float *__restrict__ vel, int start,
                                            int cnt, int lane, int fl, int ll, float (*stage)[kStageSlots], int ci,
                                            int cj, int ck, int nx, int ny, int nz, float inv_dx, float (&vs)[8][3],
                                            float (&ws)[8][3], int (&occ)[8]) {
    constexpr int base = (OX + 1) * 9 + (OY + 1) * 3;
    constexpr int L_UP = (base + 0) & 7;    // this lane's bin is oz = -1 for cell ck + 1
    constexpr int L_OWN = (base + 1) & 7;   // oz = 0
    constexpr int L_DOWN = (base + 2) & 7;  // oz = +1 for cell ck - 1
    constexpr bool U = OY <= 0, V = OX <= 0, W = OX <= 0 && OY <= 0, OCC = OX <= 0 && OY <= 0;
    if constexpr (!(U || V || W)) {
        return;
    } else {
        const float gx = (float)ci, gy = (float)cj;
        const float gz_own = (float)ck, gz_up = (float)(ck + 1), gz_dn = (float)(ck - 1);
        // the row's particles: lanes fl..ll hold its bins, in bin order
        const int r0 = __builtin_amdgcn_readlane(start, fl), r1 = __builtin_amdgcn_readlane(start + cnt, ll);
        const int n = r1 - r0;
        const bool staged = n <= kStageCap;  // uniform
        if (staged)
            for (int q = lane; q < n; q += 64) {
                const int64_t t = (int64_t)r0 + q;
                const float x = pos[3 * t], y = pos[3 * t + 1], z = pos[3 * t + 2];
                const int sl = stage_slot(q);
                stage[0][sl] = x * inv_dx;
                stage[1][sl] = y * inv_dx;
                stage[2][sl] = z * inv_dx;
            }
        auto visit = [&](const RowParticle &q, bool has) {
            const float p0 = q.x, p1 = q.y, p2 = q.z;
            // x and y factors are shared by the three cells of the column
            const float ay = tent(p1, gy + 0.5f);
            const float uxy = has ? tent(p0, gx + 1.0f) * ay : 0.0f;             // u face
            const float vxy = has ? tent(p0, gx + 0.5f) * tent(p1, gy + 1.0f) : 0.0f;  // v face
            const float wxy = has ? tent(p0, gx + 0.5f) * ay : 0.0f;             // w face
            int tocc = -1000;
            if constexpr (OCC) {
                const int tx = trunc_clamp(p0, 1, nx - 2);
                const int ty = trunc_clamp(p1, 1, ny - 2);
                const int tz = trunc_clamp(p2, 1, nz - 2);
                if (has && tx == ci && ty == cj) tocc = tz;
            }
            // own cell, oz = 0
            if constexpr (U) {
                const float wg = uxy * tent(p2, gz_own + 0.5f);
                vs[L_OWN][0] = fmaf(wg, q.u, vs[L_OWN][0]);
                ws[L_OWN][0] = ws[L_OWN][0] + wg;
            }
            if constexpr (V) {
                const float wg = vxy * tent(p2, gz_own + 0.5f);
                vs[L_OWN][1] = fmaf(wg, q.v, vs[L_OWN][1]);
                ws[L_OWN][1] = ws[L_OWN][1] + wg;
            }
            if constexpr (W) {
                const float wg = wxy * tent(p2, gz_own + 1.0f);
                vs[L_OWN][2] = fmaf(wg, q.w, vs[L_OWN][2]);
                ws[L_OWN][2] = ws[L_OWN][2] + wg;
            }
            if constexpr (OCC) occ[L_OWN] += (tocc == ck);
            // the cell above (ck + 1) takes this bin as its oz = -1 neighbour
            // from the lane below (lane 0's sums are dropped)
            if constexpr (U) {
                const float wg = lane_up(uxy * tent(p2, gz_up + 0.5f));
                vs[L_UP][0] = fmaf(wg, lane_up(q.u), vs[L_UP][0]);
                ws[L_UP][0] = ws[L_UP][0] + wg;
            }
            if constexpr (V) {
                const float wg = lane_up(vxy * tent(p2, gz_up + 0.5f));
                vs[L_UP][1] = fmaf(wg, lane_up(q.v), vs[L_UP][1]);
                ws[L_UP][1] = ws[L_UP][1] + wg;
            }
            if constexpr (W) {
                const float wg = lane_up(wxy * tent(p2, gz_up + 1.0f));
                vs[L_UP][2] = fmaf(wg, lane_up(q.w), vs[L_UP][2]);
                ws[L_UP][2] = ws[L_UP][2] + wg;
                // the cell below (ck - 1) takes it as its oz = +1 neighbour (w only)
                const float wd = lane_down(wxy * tent(p2, gz_dn + 1.0f));
                vs[L_DOWN][2] = fmaf(wd, lane_down(q.w), vs[L_DOWN][2]);
                ws[L_DOWN][2] = ws[L_DOWN][2] + wd;
            }
            if constexpr (OCC) occ[L_UP] += lane_up(tocc == ck + 1 ? 1 : 0);
        };
        if (staged) {
            // positions from LDS when they are used; the velocities, from
            // global memory, in two buffers in turn (the next one loads while
            // the current one is used, with no register copies between them)
            auto load = [&](int sidx, float (&vv)[3]) {
                if (sidx < cnt) {
                    const int64_t t = (int64_t)start + sidx;
                    vv[0] = vel[3 * t];
                    vv[1] = vel[3 * t + 1];
                    vv[2] = vel[3 * t + 2];
                }
            };
            auto step = [&](int sidx, const float (&vv)[3]) {
                const int sl = stage_slot(min(start - r0 + sidx, kStageCap - 1));
                const RowParticle q = {stage[0][sl], stage[1][sl], stage[2][sl], vv[0], vv[1], vv[2]};
                visit(q, sidx < cnt);
            };
            float va[3] = {0.0f, 0.0f, 0.0f}, vb[3] = {0.0f, 0.0f, 0.0f};
            load(0, va);
            for (int sidx = 0; __any(sidx < cnt); sidx += 2) {
                load(sidx + 1, vb);
                step(sidx, va);
                if (!__any(sidx + 1 < cnt)) break;
                load(sidx + 2, va);
                step(sidx + 1, vb);
            }
        } else {  // a crowded row (rare): global memory, one particle at a time
            RowParticle a = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
            for (int sidx = 0; __any(sidx < cnt); ++sidx) {
                if (sidx < cnt) {
                    const int64_t t = (int64_t)start + sidx;
                    a.x = pos[3 * t] * inv_dx;
                    a.y = pos[3 * t + 1] * inv_dx;
                    a.z = pos[3 * t + 2] * inv_dx;
                    a.u = vel[3 * t];
                    a.v = vel[3 * t + 1];
                    a.w = vel[3 * t + 2];
                }
                visit(a, sidx < cnt);
            }
        }
    }
}

__global__ __launch_bounds__(256) void p2g_rowscan_kernel(const float *__restrict__ pos, const float *__restrict__ vel,
                                                          const int32_t *__restrict__ offsets,
                                                          const int32_t *__restrict__ counts, float *__restrict__ u,
                                                          float *__restrict__ v, float *__restrict__ w,
                                                          float *__restrict__ ff, float inv_dx, float inv_ppc, int nx,
                                                          int ny, int nz) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    // Blocks keep the launch order (segment, row group, plane): the blocks
    // resident together on an XCD then cover a compact region, which measured
    // fewer fetched bytes than per-XCD plane ranges or plane-ordered columns.
    const int ci = blockIdx.z, cj = blockIdx.y * 4 + wave;
    if (cj >= ny) return;  // whole wave
    const int ck0 = (int)blockIdx.x * kSeg - 1, ck = ck0 + lane;  // this lane's bin; also its cell
    // lanes fl..ll hold the bins 0..nz-2 of the segment
    const int fl = ck0 < 0 ? -ck0 : 0, ll = min(63, nz - 2 - ck0);
    __shared__ float stage[4][3][kStageSlots];
    float vs[8][3], ws[8][3];
    int occ[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        occ[l] = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            vs[l][a] = 0.0f;
            ws[l][a] = 0.0f;
        }
    }
    // the bin range of each neighbour row is fetched one row ahead, so the
    // waves over empty rows do not wait for eight round trips in turn
    auto fetch = [&](auto ox_c, auto oy_c, int &start, int &cnt) {
        constexpr int OX = decltype(ox_c)::value, OY = decltype(oy_c)::value;
        const int bx = ci + OX, by = cj + OY;
        start = 0;
        cnt = 0;
        if (bx >= 0 && bx < nx - 1 && by >= 0 && by < ny - 1 && ck >= 0 && ck < nz - 1) {
            const int64_t bin = gidx(bx, by, ck, ny, nz);
            start = offsets[bin];
            cnt = counts[bin];
        }
    };
    auto scan = [&](auto ox_c, auto oy_c, int start, int cnt) {
        constexpr int OX = decltype(ox_c)::value, OY = decltype(oy_c)::value;
        rowscan_row<OX, OY>(pos, vel, start, cnt, lane, fl, ll, stage[wave], ci, cj, ck, nx, ny, nz, inv_dx, vs, ws,
                            occ);
    };
    using m1 = std::integral_constant<int, -1>;
    using z0 = std::integral_constant<int, 0>;
    using p1 = std::integral_constant<int, 1>;
    // rows in the reference's neighbour order; (+1, +1) carries no weight
    int s0, c0, s1, c1;
    fetch(m1{}, m1{}, s0, c0);
    fetch(m1{}, z0{}, s1, c1);
    scan(m1{}, m1{}, s0, c0);
    fetch(m1{}, p1{}, s0, c0);
    scan(m1{}, z0{}, s1, c1);
    fetch(z0{}, m1{}, s1, c1);
    scan(m1{}, p1{}, s0, c0);
    fetch(z0{}, z0{}, s0, c0);
    scan(z0{}, m1{}, s1, c1);
    fetch(z0{}, p1{}, s1, c1);
    scan(z0{}, z0{}, s0, c0);
    fetch(p1{}, m1{}, s0, c0);
    scan(z0{}, p1{}, s1, c1);
    fetch(p1{}, z0{}, s1, c1);
    scan(p1{}, m1{}, s0, c0);
    scan(p1{}, z0{}, s1, c1);
    if (lane == 0 || lane == 63 || ck >= nz) return;
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) {
#pragma unroll
        for (int l = 0; l < off; ++l) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                vs[l][a] = vs[l][a] + vs[l + off][a];
                ws[l][a] = ws[l][a] + ws[l + off][a];
            }
            occ[l] += occ[l + off];
        }
    }
    const int64_t index = gidx(ci, cj, ck, ny, nz);
    u[index] = ws[0][0] > 0.0f ? vs[0][0] / ws[0][0] : 0.0f;
    v[index] = ws[0][1] > 0.0f ? vs[0][1] / ws[0][1] : 0.0f;
    w[index] = ws[0][2] > 0.0f ? vs[0][2] / ws[0][2] : 0.0f;
    ff[index] = (float)occ[0] * inv_ppc;
}

void launch_p2g(const float *pos, const float *vel, const int32_t *offsets, const int32_t *counts, float *u, float *v,
                float *w, float *ff, float inv_dx, float inv_ppc, int nx, int ny, int nz, hipStream_t s) {
    const dim3 grid((unsigned)((nz + kSeg - 1) / kSeg), (unsigned)((ny + 3) / 4), (unsigned)nx);
    hipLaunchKernelGGL(p2g_rowscan_kernel, grid, dim3(256), 0, s, pos, vel, offsets, counts, u, v, w, ff, inv_dx,
                       inv_ppc, nx, ny, nz);
}

}  // namespace clq
