// P2G gather onto the MAC grid plus occupancy voxelisation.
//
// Replaces particle_to_grid_gather (reference
// Solver/Kernel/OpenCL/particles.cl:135-227, launched at
// Solver/Kernel/particles.py:142-163 with 8 lanes per cell).
//
// One lane owns one cell and gathers from the neighbour bins of the sorted
// particle array; no float atomics.  The lane keeps eight partial sums, one
// per lane of the reference's 8-lane work-group (neighbour bin n goes to
// partial n % 8), visits each bin's particles in sorted order and combines the
// partials with the reference's tree (0+4, 2+6, 1+5, 3+7, then pairs), so the
// summation order is the reference's and the result is bit-identical to the
// CPU oracle.
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

template <int NB>
__device__ __forceinline__ void gather_bin(const float *__restrict__ pos, const float *__restrict__ vel,
                                           const int32_t *__restrict__ offsets, const int32_t *__restrict__ counts,
                                           int ci, int cj, int ck, int nx, int ny, int nz, float inv_dx,
                                           const float *upos, const float *vpos, const float *wpos, float (&vs)[8][3],
                                           float (&ws)[8][3], int (&cnt)[8]) {
    constexpr int ox = NB / 9 - 1, oy = (NB / 3) % 3 - 1, oz = NB % 3 - 1;
    constexpr bool doU = oy <= 0 && oz <= 0;
    constexpr bool doV = ox <= 0 && oz <= 0;
    constexpr bool doW = ox <= 0 && oy <= 0;
    constexpr bool doOcc = ox <= 0 && oy <= 0 && oz <= 0;
    constexpr int lane = NB & 7;
    if constexpr (!(doU || doV || doW)) {
        return;
    } else {
        const int bx = ci + ox, by = cj + oy, bz = ck + oz;
        if (bx < 0 || bx >= nx - 1 || by < 0 || by >= ny - 1 || bz < 0 || bz >= nz - 1) return;
        const int64_t bin = gidx(bx, by, bz, ny, nz);
        const int begin = offsets[bin];
        const int end = begin + counts[bin];
        for (int pi = begin; pi < end; ++pi) {
            const float x0 = pos[3LL * pi], x1 = pos[3LL * pi + 1], x2 = pos[3LL * pi + 2];
            const float p0 = x0 * inv_dx, p1 = x1 * inv_dx, p2 = x2 * inv_dx;
            if constexpr (doOcc) {
                const int tx = trunc_clamp(p0, 1, nx - 2);
                const int ty = trunc_clamp(p1, 1, ny - 2);
                const int tz = trunc_clamp(p2, 1, nz - 2);
                cnt[lane] += (tx == ci && ty == cj && tz == ck);
            }
            if constexpr (doU) {
                const float a = fmaxf(0.0f, 1.0f - fabsf(p0 - upos[0]));
                const float b = fmaxf(0.0f, 1.0f - fabsf(p1 - upos[1]));
                const float c = fmaxf(0.0f, 1.0f - fabsf(p2 - upos[2]));
                const float wgt = a * b * c;
                vs[lane][0] = fmaf(wgt, vel[3LL * pi], vs[lane][0]);
                ws[lane][0] = ws[lane][0] + wgt;
            }
            if constexpr (doV) {
                const float a = fmaxf(0.0f, 1.0f - fabsf(p0 - vpos[0]));
                const float b = fmaxf(0.0f, 1.0f - fabsf(p1 - vpos[1]));
                const float c = fmaxf(0.0f, 1.0f - fabsf(p2 - vpos[2]));
                const float wgt = a * b * c;
                vs[lane][1] = fmaf(wgt, vel[3LL * pi + 1], vs[lane][1]);
                ws[lane][1] = ws[lane][1] + wgt;
            }
            if constexpr (doW) {
                const float a = fmaxf(0.0f, 1.0f - fabsf(p0 - wpos[0]));
                const float b = fmaxf(0.0f, 1.0f - fabsf(p1 - wpos[1]));
                const float c = fmaxf(0.0f, 1.0f - fabsf(p2 - wpos[2]));
                const float wgt = a * b * c;
                vs[lane][2] = fmaf(wgt, vel[3LL * pi + 2], vs[lane][2]);
                ws[lane][2] = ws[lane][2] + wgt;
            }
        }
    }
}

template <int... NBs>
__device__ __forceinline__ void gather_all(std::integer_sequence<int, NBs...>, const float *pos, const float *vel,
                                           const int32_t *offsets, const int32_t *counts, int ci, int cj, int ck,
                                           int nx, int ny, int nz, float inv_dx, const float *upos, const float *vpos,
                                           const float *wpos, float (&vs)[8][3], float (&ws)[8][3], int (&cnt)[8]) {
    // Lane l of the reference visits l, l+8, l+16, l+24; the partials are
    // independent, so visiting bins 0..26 in order adds every term to its
    // partial in the reference order.
    (gather_bin<NBs>(pos, vel, offsets, counts, ci, cj, ck, nx, ny, nz, inv_dx, upos, vpos, wpos, vs, ws, cnt), ...);
}

__global__ __launch_bounds__(256) void p2g_kernel(const float *__restrict__ pos, const float *__restrict__ vel,
                                                  const int32_t *__restrict__ offsets,
                                                  const int32_t *__restrict__ counts, float *__restrict__ u,
                                                  float *__restrict__ v, float *__restrict__ w,
                                                  float *__restrict__ ff, float inv_dx, float inv_ppc, int nx,
                                                  int ny, int nz) {
    const int64_t C = (int64_t)nx * ny * nz;
    const int64_t index = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (index >= C) return;
    int ci, cj, ck;
    coords(index, ny, nz, ci, cj, ck);
    float vs[8][3], ws[8][3];
    int cnt[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        cnt[l] = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            vs[l][a] = 0.0f;
            ws[l][a] = 0.0f;
        }
    }
    const float gx = (float)ci, gy = (float)cj, gz = (float)ck;
    const float upos[3] = {gx + 1.0f, gy + 0.5f, gz + 0.5f};
    const float vpos[3] = {gx + 0.5f, gy + 1.0f, gz + 0.5f};
    const float wpos[3] = {gx + 0.5f, gy + 0.5f, gz + 1.0f};
    gather_all(std::make_integer_sequence<int, 27>{}, pos, vel, offsets, counts, ci, cj, ck, nx, ny, nz, inv_dx,
               upos, vpos, wpos, vs, ws, cnt);
    // Tree of particles.cl:204-211: offsets 4, 2, 1.
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) {
#pragma unroll
        for (int l = 0; l < off; ++l) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                vs[l][a] = vs[l][a] + vs[l + off][a];
                ws[l][a] = ws[l][a] + ws[l + off][a];
            }
            cnt[l] += cnt[l + off];
        }
    }
    u[index] = ws[0][0] > 0.0f ? vs[0][0] / ws[0][0] : 0.0f;
    v[index] = ws[0][1] > 0.0f ? vs[0][1] / ws[0][1] : 0.0f;
    w[index] = ws[0][2] > 0.0f ? vs[0][2] / ws[0][2] : 0.0f;
    ff[index] = (float)cnt[0] * inv_ppc;
}

// ---- LDS-staged gather ----------------------------------------------------------------
// A work-group of 4 waves owns cells (i, j0..j0+3, k0..k0+63), one lane per
// cell.  The bins a cell gathers from lie in rows (i+ox, j+oy, k-1..k+1); for
// the whole work-group that is 3 x 6 bin rows, and each row's particles are
// one contiguous range of the sorted array (bins k0-1..k0+64).  Rows are
// streamed through LDS in chunks with coalesced loads, in the order (ox, row),
// so every wave meets its neighbour bins in increasing reference neighbour
// index and each partial sum sees its terms in the reference order: the
// result is bit-identical to the per-cell kernel above.
constexpr int kTileJ = 4, kTileK = 64, kChunk = 1024;

struct StagedParticle {
    float p0, p1, p2, v0, v1, v2;
};

template <int NB>
__device__ __forceinline__ void staged_bin(const StagedParticle *__restrict__ lds, int lo, int hi, int ci, int cj,
                                           int ck, int nx, int ny, int nz, const float *upos, const float *vpos,
                                           const float *wpos, float (&vs)[8][3], float (&ws)[8][3], int (&cnt)[8]) {
    constexpr int ox = NB / 9 - 1, oy = (NB / 3) % 3 - 1, oz = NB % 3 - 1;
    constexpr bool doU = oy <= 0 && oz <= 0;
    constexpr bool doV = ox <= 0 && oz <= 0;
    constexpr bool doW = ox <= 0 && oy <= 0;
    constexpr bool doOcc = ox <= 0 && oy <= 0 && oz <= 0;
    constexpr int lane = NB & 7;
    if constexpr (doU || doV || doW) {
        for (int t = lo; t < hi; ++t) {
            const StagedParticle q = lds[t];
            if constexpr (doOcc) {
                const int tx = trunc_clamp(q.p0, 1, nx - 2);
                const int ty = trunc_clamp(q.p1, 1, ny - 2);
                const int tz = trunc_clamp(q.p2, 1, nz - 2);
                cnt[lane] += (tx == ci && ty == cj && tz == ck);
            }
            if constexpr (doU) {
                const float wgt = fmaxf(0.0f, 1.0f - fabsf(q.p0 - upos[0])) * fmaxf(0.0f, 1.0f - fabsf(q.p1 - upos[1])) *
                                  fmaxf(0.0f, 1.0f - fabsf(q.p2 - upos[2]));
                vs[lane][0] = fmaf(wgt, q.v0, vs[lane][0]);
                ws[lane][0] = ws[lane][0] + wgt;
            }
            if constexpr (doV) {
                const float wgt = fmaxf(0.0f, 1.0f - fabsf(q.p0 - vpos[0])) * fmaxf(0.0f, 1.0f - fabsf(q.p1 - vpos[1])) *
                                  fmaxf(0.0f, 1.0f - fabsf(q.p2 - vpos[2]));
                vs[lane][1] = fmaf(wgt, q.v1, vs[lane][1]);
                ws[lane][1] = ws[lane][1] + wgt;
            }
            if constexpr (doW) {
                const float wgt = fmaxf(0.0f, 1.0f - fabsf(q.p0 - wpos[0])) * fmaxf(0.0f, 1.0f - fabsf(q.p1 - wpos[1])) *
                                  fmaxf(0.0f, 1.0f - fabsf(q.p2 - wpos[2]));
                vs[lane][2] = fmaf(wgt, q.v2, vs[lane][2]);
                ws[lane][2] = ws[lane][2] + wgt;
            }
        }
    }
}

template <int... NBs>
__device__ __forceinline__ void staged_dispatch(std::integer_sequence<int, NBs...>, int nb,
                                                const StagedParticle *lds, int lo, int hi, int ci, int cj, int ck,
                                                int nx, int ny, int nz, const float *upos, const float *vpos,
                                                const float *wpos, float (&vs)[8][3], float (&ws)[8][3],
                                                int (&cnt)[8]) {
    // nb is uniform across the wave: exactly one instantiation runs.
    ((nb == NBs ? staged_bin<NBs>(lds, lo, hi, ci, cj, ck, nx, ny, nz, upos, vpos, wpos, vs, ws, cnt) : void()), ...);
}

__global__ __launch_bounds__(256) void p2g_staged_kernel(const float *__restrict__ pos, const float *__restrict__ vel,
                                                         const int32_t *__restrict__ offsets,
                                                         const int32_t *__restrict__ counts, float *__restrict__ u,
                                                         float *__restrict__ v, float *__restrict__ w,
                                                         float *__restrict__ ff, float inv_dx, float inv_ppc, int nx,
                                                         int ny, int nz) {
    __shared__ StagedParticle lds[kChunk];
    const int wave = threadIdx.x >> 6, lane_id = threadIdx.x & 63;
    const int k0 = blockIdx.x * kTileK, j0 = blockIdx.y * kTileJ, ci = blockIdx.z;
    const int cj = j0 + wave, ck = k0 + lane_id;
    const bool valid = cj < ny && ck < nz;
    float vs[8][3], ws[8][3];
    int cnt[8];
#pragma unroll
    for (int l = 0; l < 8; ++l) {
        cnt[l] = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            vs[l][a] = 0.0f;
            ws[l][a] = 0.0f;
        }
    }
    const float gx = (float)ci, gy = (float)cj, gz = (float)ck;
    const float upos[3] = {gx + 1.0f, gy + 0.5f, gz + 0.5f};
    const float vpos[3] = {gx + 0.5f, gy + 1.0f, gz + 0.5f};
    const float wpos[3] = {gx + 0.5f, gy + 0.5f, gz + 1.0f};
    const int bz_lo = max(k0 - 1, 0), bz_hi = min(k0 + kTileK, nz - 2);  // bin range of the rows
    for (int ox = -1; ox <= 1; ++ox) {
        const int bx = ci + ox;
        if (bx < 0 || bx > nx - 2) continue;  // uniform
        for (int row = 0; row < kTileJ + 2; ++row) {
            const int by = j0 - 1 + row;
            if (by < 0 || by > ny - 2 || bz_lo > bz_hi) continue;  // uniform
            const int64_t first = gidx(bx, by, bz_lo, ny, nz), last = gidx(bx, by, bz_hi, ny, nz);
            const int s0 = offsets[first], s1 = offsets[last] + counts[last];
            const int oy = by - cj;  // this wave's y offset for the row
            const bool mine = valid && oy >= -1 && oy <= 1;
            for (int c0 = s0; c0 < s1; c0 += kChunk) {
                const int c1 = min(c0 + kChunk, s1);
                __syncthreads();
                for (int t = c0 + (int)threadIdx.x; t < c1; t += 256) {
                    StagedParticle q;
                    q.p0 = pos[3LL * t] * inv_dx;
                    q.p1 = pos[3LL * t + 1] * inv_dx;
                    q.p2 = pos[3LL * t + 2] * inv_dx;
                    q.v0 = vel[3LL * t];
                    q.v1 = vel[3LL * t + 1];
                    q.v2 = vel[3LL * t + 2];
                    lds[t - c0] = q;
                }
                __syncthreads();
                if (!mine) continue;
                for (int oz = -1; oz <= 1; ++oz) {
                    const int bz = ck + oz;
                    if (bz < 0 || bz > nz - 2) continue;
                    const int64_t bin = gidx(bx, by, bz, ny, nz);
                    const int b0 = offsets[bin], b1 = b0 + counts[bin];
                    const int lo = max(b0, c0) - c0, hi = min(b1, c1) - c0;
                    if (lo >= hi) continue;
                    const int nb = (ox + 1) * 9 + (oy + 1) * 3 + (oz + 1);
                    staged_dispatch(std::make_integer_sequence<int, 27>{}, nb, lds, lo, hi, ci, cj, ck, nx, ny, nz,
                                    upos, vpos, wpos, vs, ws, cnt);
                }
            }
        }
    }
    if (!valid) return;
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) {
#pragma unroll
        for (int l = 0; l < off; ++l) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                vs[l][a] = vs[l][a] + vs[l + off][a];
                ws[l][a] = ws[l][a] + ws[l + off][a];
            }
            cnt[l] += cnt[l + off];
        }
    }
    const int64_t index = gidx(ci, cj, ck, ny, nz);
    u[index] = ws[0][0] > 0.0f ? vs[0][0] / ws[0][0] : 0.0f;
    v[index] = ws[0][1] > 0.0f ? vs[0][1] / ws[0][1] : 0.0f;
    w[index] = ws[0][2] > 0.0f ? vs[0][2] / ws[0][2] : 0.0f;
    ff[index] = (float)cnt[0] * inv_ppc;
}

// ---- row-scan gather -----------------------------------------------------------------
// A wave owns the cells (i, j, k0 .. k0 + 61) and lane l the bin k0 - 1 + l of
// each neighbour bin row (i + ox, j + oy, *).  Rows are visited in the
// reference's neighbour order (ox, then oy); in a row, lane l walks its bin's
// particles once and computes the tent weights of each particle for its own
// cell (oz = 0), the cell above (the bin is that cell's oz = -1 neighbour) and
// the cell below (oz = +1).  The neighbours' weights and velocities are handed
// over with wave shuffles, and each cell adds them with the same fmaf into the
// partial sum of that neighbour index.  The three oz of a row feed three
// different partials, and every partial still receives its bins in increasing
// neighbour index and each bin's particles in sorted order, so the sums are
// bit-identical to the per-cell kernel.  Only the component/bin pairs that can
// carry weight are evaluated (the oz = +1 bin only for w).
constexpr int kSeg = 62;

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
        for (int sidx = 0; __any(sidx < cnt); ++sidx) {
            const bool has = sidx < cnt;
            float p0 = 0.0f, p1 = 0.0f, p2 = 0.0f, v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
            if (has) {
                const int64_t t = (int64_t)start + sidx;
                p0 = pos[3 * t] * inv_dx;
                p1 = pos[3 * t + 1] * inv_dx;
                p2 = pos[3 * t + 2] * inv_dx;
                v0 = vel[3 * t];
                v1 = vel[3 * t + 1];
                v2 = vel[3 * t + 2];
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
            const int rf = __shfl_up(up_flags, 1);
            const int rd = __shfl_down(has ? 1 : 0, 1);
            if constexpr (U) {
                const float wg = __shfl_up(su, 1), vv = __shfl_up(v0, 1);
                if (lane > 0 && (rf & 1)) {
                    vs[L_UP][0] = fmaf(wg, vv, vs[L_UP][0]);
                    ws[L_UP][0] = ws[L_UP][0] + wg;
                }
            }
            if constexpr (V) {
                const float wg = __shfl_up(sv, 1), vv = __shfl_up(v1, 1);
                if (lane > 0 && (rf & 1)) {
                    vs[L_UP][1] = fmaf(wg, vv, vs[L_UP][1]);
                    ws[L_UP][1] = ws[L_UP][1] + wg;
                }
            }
            if constexpr (W) {
                const float wg = __shfl_up(sw, 1), vv = __shfl_up(v2, 1);
                if (lane > 0 && (rf & 1)) {
                    vs[L_UP][2] = fmaf(wg, vv, vs[L_UP][2]);
                    ws[L_UP][2] = ws[L_UP][2] + wg;
                }
                const float wd = __shfl_down(sd, 1), vd = __shfl_down(v2, 1);
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
