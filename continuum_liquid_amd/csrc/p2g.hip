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
// only for w).
constexpr int kSeg = 62;

// Lane l receives lane l - 1's value (lane_up) or lane l + 1's (lane_down)
// through a DPP wavefront shift (wave_shr:1 / wave_shl:1): a VALU move
// instead of an LDS permute.  Lanes 0 and 63 receive 0; the callers drop
// those values.
__device__ __forceinline__ int lane_up(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int lane_down(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ float lane_up(float x) { return __int_as_float(lane_up(__float_as_int(x))); }
__device__ __forceinline__ float lane_down(float x) { return __int_as_float(lane_down(__float_as_int(x))); }

template <int OX, int OY>
__device__ __forceinline__ void rowscan_row(const float *__restrict__ pos, const float *__restrict__ vel, int start,
                                            int cnt, int lane, int ci, int cj, int ck, int nx, int ny, int nz,
                                            float inv_dx, float (&vs)[8][3], float (&ws)[8][3], int (&occ)[8]) {
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
        // particle sidx + 1 is loaded while particle sidx is processed
        float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f, m0 = 0.0f, m1 = 0.0f, m2 = 0.0f;
        if (cnt > 0) {
            const int64_t t = start;
            n0 = pos[3 * t];
            n1 = pos[3 * t + 1];
            n2 = pos[3 * t + 2];
            m0 = vel[3 * t];
            m1 = vel[3 * t + 1];
            m2 = vel[3 * t + 2];
        }
        for (int sidx = 0; __any(sidx < cnt); ++sidx) {
            const bool has = sidx < cnt;
            const float p0 = n0 * inv_dx, p1 = n1 * inv_dx, p2 = n2 * inv_dx;
            const float v0 = m0, v1 = m1, v2 = m2;
            if (sidx + 1 < cnt) {
                const int64_t t = (int64_t)start + sidx + 1;
                n0 = pos[3 * t];
                n1 = pos[3 * t + 1];
                n2 = pos[3 * t + 2];
                m0 = vel[3 * t];
                m1 = vel[3 * t + 1];
                m2 = vel[3 * t + 2];
            }
            // x and y factors are shared by the three cells of the column
            const float ax = fmaxf(0.0f, 1.0f - fabsf(p0 - (gx + 1.0f)));  // u face x
            const float ay = fmaxf(0.0f, 1.0f - fabsf(p1 - (gy + 0.5f)));
            const float bx = fmaxf(0.0f, 1.0f - fabsf(p0 - (gx + 0.5f)));  // v, w faces x
            const float by = fmaxf(0.0f, 1.0f - fabsf(p1 - (gy + 1.0f)));  // v face y
            const float uxy = ax * ay, vxy = bx * by, wxy = bx * ay;
            int tocc = 0;
            if constexpr (OCC) {
                const int tx = trunc_clamp(p0, 1, nx - 2);
                const int ty = trunc_clamp(p1, 1, ny - 2);
                const int tz = trunc_clamp(p2, 1, nz - 2);
                tocc = (tx == ci && ty == cj) ? tz : -1000;
            }
            // own cell, oz = 0
            if (has) {
                if constexpr (U) {
                    const float wg = uxy * fmaxf(0.0f, 1.0f - fabsf(p2 - (gz_own + 0.5f)));
                    vs[L_OWN][0] = fmaf(wg, v0, vs[L_OWN][0]);
                    ws[L_OWN][0] = ws[L_OWN][0] + wg;
                }
                if constexpr (V) {
                    const float wg = vxy * fmaxf(0.0f, 1.0f - fabsf(p2 - (gz_own + 0.5f)));
                    vs[L_OWN][1] = fmaf(wg, v1, vs[L_OWN][1]);
                    ws[L_OWN][1] = ws[L_OWN][1] + wg;
                }
                if constexpr (W) {
                    const float wg = wxy * fmaxf(0.0f, 1.0f - fabsf(p2 - (gz_own + 1.0f)));
                    vs[L_OWN][2] = fmaf(wg, v2, vs[L_OWN][2]);
                    ws[L_OWN][2] = ws[L_OWN][2] + wg;
                }
                if constexpr (OCC) occ[L_OWN] += (tocc == ck);
            }
            // for the cell above (ck + 1): this bin is its oz = -1 neighbour
            float su = 0.0f, sv = 0.0f, sw = 0.0f;
            if constexpr (U) su = uxy * fmaxf(0.0f, 1.0f - fabsf(p2 - (gz_up + 0.5f)));
            if constexpr (V) sv = vxy * fmaxf(0.0f, 1.0f - fabsf(p2 - (gz_up + 0.5f)));
            if constexpr (W) sw = wxy * fmaxf(0.0f, 1.0f - fabsf(p2 - (gz_up + 1.0f)));
            const int up_flags = has ? (1 | ((OCC && tocc == ck + 1) ? 2 : 0)) : 0;
            // for the cell below (ck - 1): this bin is its oz = +1 neighbour (w only)
            float sd = 0.0f;
            if constexpr (W) sd = wxy * fmaxf(0.0f, 1.0f - fabsf(p2 - (gz_dn + 1.0f)));
            const int rf = lane_up(up_flags);
            const int rd = lane_down(has ? 1 : 0);
            if constexpr (U) {
                const float wg = lane_up(su), vv = lane_up(v0);
                if (lane > 0 && (rf & 1)) {
                    vs[L_UP][0] = fmaf(wg, vv, vs[L_UP][0]);
                    ws[L_UP][0] = ws[L_UP][0] + wg;
                }
            }
            if constexpr (V) {
                const float wg = lane_up(sv), vv = lane_up(v1);
                if (lane > 0 && (rf & 1)) {
                    vs[L_UP][1] = fmaf(wg, vv, vs[L_UP][1]);
                    ws[L_UP][1] = ws[L_UP][1] + wg;
                }
            }
            if constexpr (W) {
                const float wg = lane_up(sw), vv = lane_up(v2);
                if (lane > 0 && (rf & 1)) {
                    vs[L_UP][2] = fmaf(wg, vv, vs[L_UP][2]);
                    ws[L_UP][2] = ws[L_UP][2] + wg;
                }
                const float wd = lane_down(sd), vd = lane_down(v2);
                if (lane < 63 && rd) {
                    vs[L_DOWN][2] = fmaf(wd, vd, vs[L_DOWN][2]);
                    ws[L_DOWN][2] = ws[L_DOWN][2] + wd;
                }
            }
            if constexpr (OCC)
                if (lane > 0 && (rf & 2)) occ[L_UP] += 1;
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
    const int ck = blockIdx.x * kSeg - 1 + lane;  // this lane's bin; also its cell
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
    auto row = [&](auto ox_c, auto oy_c) {
        constexpr int OX = decltype(ox_c)::value, OY = decltype(oy_c)::value;
        const int bx = ci + OX, by = cj + OY;
        if (bx < 0 || bx >= nx - 1 || by < 0 || by >= ny - 1) return;  // uniform
        int start = 0, cnt = 0;
        if (ck >= 0 && ck < nz - 1) {
            const int64_t bin = gidx(bx, by, ck, ny, nz);
            start = offsets[bin];
            cnt = counts[bin];
        }
        rowscan_row<OX, OY>(pos, vel, start, cnt, lane, ci, cj, ck, nx, ny, nz, inv_dx, vs, ws, occ);
    };
    using m1 = std::integral_constant<int, -1>;
    using z0 = std::integral_constant<int, 0>;
    using p1 = std::integral_constant<int, 1>;
    row(m1{}, m1{}); row(m1{}, z0{}); row(m1{}, p1{});
    row(z0{}, m1{}); row(z0{}, z0{}); row(z0{}, p1{});
    row(p1{}, m1{}); row(p1{}, z0{}); row(p1{}, p1{});
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
