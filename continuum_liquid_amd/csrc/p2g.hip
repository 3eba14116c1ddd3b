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
// v only x, z offsets in {-1, 0}, w only x, y in {-1, 0}.  Skipped terms are
// exact zeros, so skipping does not change a bit.  Of the 27 bins, 8 are
// visited for all three components and 12 for one.
#include <type_traits>
#include <utility>

#include "common.hpp"

namespace clq {

// ---- block-staged row gather ------------------------------------------------------------
// A block of 4 waves owns a 2 x 2 patch of cell rows, wave w the row
// (ci0 + w / 2, cj0 + w % 2, *), over one segment of kSeg cells in z; lane l
// of every wave owns the bin (and cell) k = ck0 - 1 + l.  The 16 bin rows
// (ci0 - 1 .. ci0 + 2, cj0 - 1 .. cj0 + 2) the patch reads are staged one
// after the other into LDS (positions in cells and velocities, SoA; the next
// item's loads are in flight while the waves walk the current one), in
// lexicographic (bi, bj) order, which for every cell row is the reference's
// neighbour order (ox, then oy).  A row with more particles than one buffer
// holds is staged in consecutive chunks; a lane walks its bin's part of each
// chunk, so every partial sum still receives its bin's particles in sorted
// order.  For each staged row, each wave whose cell row has it as a
// neighbour ((ox, oy) in -1..1, and not (+1, +1), which carries no weight)
// walks it: lane l takes the particles of its bin one per step, computes the
// tent weights for its own cell (oz = 0), the cell above (the bin is that
// cell's oz = -1 neighbour) and the cell below (oz = +1, w only), and adds
// each with the same fmaf into the partial sum of that neighbour index (n =
// 9(ox+1) + 3(oy+1) + oz+1, partial n % 8), as the reference's 8-lane
// work-group does.  The cells above' and below's running sums of the walk's
// partials are moved into the particles' lanes with DPP wavefront shifts
// when the walk starts and moved back when it ends (two shifts per sum and
// walk instead of two per term and step), so every sum receives its terms in
// the same order.  A lane past its part of the chunk reads a sentinel
// particle far outside the grid, whose weights are 0.
//
// The kernel is bound by the chain of staged items each block runs through
// (one memory round trip and one barrier per item), so nothing on that chain
// waits for LDS it need not: every wave lists the items in its own registers
// (row ranges in lanes 0..15, chunk counts summed with DPP, the current item
// found with a ballot and readlanes), every lane keeps the start and count
// of its own bin of the next item in registers, loaded with the item's
// particles, and a walk reads each step's particle while it computes the
// previous one.
constexpr int kSeg = 62;
constexpr int kThreads = 256;
constexpr int kCap = 640;                            // particles per staged chunk
constexpr int kSlots = kCap + kCap / 8 + 2;         // skewed slots + sentinel
constexpr int kSentinel = kSlots - 1;
constexpr int kRowsStaged = 16;                      // 4 x 4 bin rows around the patch
constexpr int kPerThread = (kCap + kThreads - 1) / kThreads;

// Lane l receives lane l - 1's value (lane_up) or lane l + 1's (lane_down)
// through a DPP wavefront shift (wave_shr:1 / wave_shl:1): a VALU move
// instead of an LDS permute.  Lanes 0 and 63 receive 0 (bound_ctrl).
__device__ __forceinline__ int lane_up(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int lane_down(int x) { return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, true); }
__device__ __forceinline__ float lane_up(float x) { return __int_as_float(lane_up(__float_as_int(x))); }
__device__ __forceinline__ float lane_down(float x) { return __int_as_float(lane_down(__float_as_int(x))); }

// max(0, 1 - |p - c|), the reference's tent: 1 - |p - c| <= 1, so clamping
// it to [0, 1] gives the same bits, and the clamp is the output modifier of
// the subtraction
__device__ __forceinline__ float tent(float p, float c) {
    return __builtin_amdgcn_fmed3f(1.0f - fabsf(p - c), 0.0f, 1.0f);
}

// Slot q + q / 8 spreads consecutive bins over the LDS banks.
__device__ __forceinline__ int stage_slot(int q) { return q + (q >> 3); }

struct StageBuf {
    float a[6][kSlots];  // x, y, z in cells; u, v, w
};

// 16-lane inclusive prefix sum (lanes 0..15) with DPP row shifts
__device__ __forceinline__ int row_scan16(int x) {
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);
    return x;
}

// One walk of a staged chunk by one wave: the cell row (ci, cj) reads bin row
// (ci + OX, cj + OY).  Adding a zero term changes no bit: the term is +-0
// (finite velocities), and a sum that starts at +0 and only ever rounds to
// nearest is never -0, so s + (+-0) == s.
struct WalkParticle {
    float p0, p1, p2, qu, qv, qw;
};

// the particle in LDS slot sl (only the velocity components the walk uses)
template <bool U, bool V, bool W>
__device__ __forceinline__ WalkParticle walk_load(const StageBuf &sb, int sl) {
    WalkParticle q;
    q.p0 = sb.a[0][sl];
    q.p1 = sb.a[1][sl];
    q.p2 = sb.a[2][sl];
    q.qu = U ? sb.a[3][sl] : 0.0f;
    q.qv = V ? sb.a[4][sl] : 0.0f;
    q.qw = W ? sb.a[5][sl] : 0.0f;
    return q;
}

template <int OX, int OY>
__device__ __forceinline__ void walk_step(const WalkParticle &q, float gx, float gy, float gz_own, float gz_up,
                                          float gz_dn, float (&vs)[8][3], float (&ws)[8][3]) {
    constexpr int base = (OX + 1) * 9 + (OY + 1) * 3;
    constexpr int L_UP = (base + 0) & 7;    // this lane's bin is oz = -1 for cell ck + 1
    constexpr int L_OWN = (base + 1) & 7;   // oz = 0
    constexpr int L_DOWN = (base + 2) & 7;  // oz = +1 for cell ck - 1
    constexpr bool U = OY <= 0, V = OX <= 0, W = OX <= 0 && OY <= 0;
    const float ay = tent(q.p1, gy + 0.5f);
    // vs/ws[L_UP] hold the cell above's sums and vs/ws[L_DOWN] the cell
    // below's during the walk (walk moves them in and out)
    if constexpr (U) {
        const float uxy = tent(q.p0, gx + 1.0f) * ay;
        const float wo = uxy * tent(q.p2, gz_own + 0.5f);
        vs[L_OWN][0] = fmaf(wo, q.qu, vs[L_OWN][0]);
        ws[L_OWN][0] = ws[L_OWN][0] + wo;
        const float wu = uxy * tent(q.p2, gz_up + 0.5f);
        vs[L_UP][0] = fmaf(wu, q.qu, vs[L_UP][0]);
        ws[L_UP][0] = ws[L_UP][0] + wu;
    }
    if constexpr (V) {
        const float vxy = tent(q.p0, gx + 0.5f) * tent(q.p1, gy + 1.0f);
        const float wo = vxy * tent(q.p2, gz_own + 0.5f);
        vs[L_OWN][1] = fmaf(wo, q.qv, vs[L_OWN][1]);
        ws[L_OWN][1] = ws[L_OWN][1] + wo;
        const float wu = vxy * tent(q.p2, gz_up + 0.5f);
        vs[L_UP][1] = fmaf(wu, q.qv, vs[L_UP][1]);
        ws[L_UP][1] = ws[L_UP][1] + wu;
    }
    if constexpr (W) {
        const float wxy = tent(q.p0, gx + 0.5f) * ay;
        const float wo = wxy * tent(q.p2, gz_own + 1.0f);
        vs[L_OWN][2] = fmaf(wo, q.qw, vs[L_OWN][2]);
        ws[L_OWN][2] = ws[L_OWN][2] + wo;
        const float wu = wxy * tent(q.p2, gz_up + 1.0f);
        vs[L_UP][2] = fmaf(wu, q.qw, vs[L_UP][2]);
        ws[L_UP][2] = ws[L_UP][2] + wu;
        // the cell below (ck - 1) takes it as its oz = +1 neighbour
        const float wd = wxy * tent(q.p2, gz_dn + 1.0f);
        vs[L_DOWN][2] = fmaf(wd, q.qw, vs[L_DOWN][2]);
        ws[L_DOWN][2] = ws[L_DOWN][2] + wd;
    }
}

// Moves the running sums of the cell above (L_UP) into the lane below and
// those of the cell below (L_DOWN) into the lane above (IN), or back (!IN),
// so that during a walk each lane adds its particles' terms for the cells
// above and below to registers of its own; a sum's terms still arrive in the
// same order.  Lanes 0 and 63 hold the segment's halo cells, whose sums are
// dropped, so the zeros DPP brings in at the wave's ends change no result.
template <int OX, int OY, bool IN>
__device__ __forceinline__ void move_neighbour_sums(float (&vs)[8][3], float (&ws)[8][3]) {
    constexpr int base = (OX + 1) * 9 + (OY + 1) * 3;
    constexpr int L_UP = (base + 0) & 7, L_DOWN = (base + 2) & 7;
    constexpr bool U = OY <= 0, V = OX <= 0, W = OX <= 0 && OY <= 0;
    auto up_sum = [](float x) { return IN ? lane_down(x) : lane_up(x); };
    auto down_sum = [](float x) { return IN ? lane_up(x) : lane_down(x); };
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        if ((c == 0 && U) || (c == 1 && V) || (c == 2 && W)) {
            vs[L_UP][c] = up_sum(vs[L_UP][c]);
            ws[L_UP][c] = up_sum(ws[L_UP][c]);
        }
    }
    if constexpr (W) {
        vs[L_DOWN][2] = down_sum(vs[L_DOWN][2]);
        ws[L_DOWN][2] = down_sum(ws[L_DOWN][2]);
    }
}

template <int OX, int OY>
__device__ __forceinline__ void walk(const StageBuf &sb, int lo, int hi, int b0, int bc, int ci, int cj, int ck,
                                     float (&vs)[8][3], float (&ws)[8][3]) {
    constexpr bool U = OY <= 0, V = OX <= 0, W = OX <= 0 && OY <= 0;
    const int b1 = b0 + bc;
    const int a = max(b0, lo), e = min(b1, hi);
    const int n = max(e - a, 0), first = a - lo;
    const float gx = (float)ci, gy = (float)cj;
    const float gz_own = (float)ck, gz_up = (float)(ck + 1), gz_dn = (float)(ck - 1);
    auto slot_of = [&](int s) { return s < n ? stage_slot(first + s) : kSentinel; };
    // two steps per trip, each step's particle read while the previous step
    // is computed
    WalkParticle A = walk_load<U, V, W>(sb, slot_of(0));
    move_neighbour_sums<OX, OY, true>(vs, ws);
    for (int s = 0; __any(s < n); s += 2) {
        const WalkParticle B = walk_load<U, V, W>(sb, slot_of(s + 1));
        __builtin_amdgcn_sched_barrier(0);  // keep the reads ahead of the step
        walk_step<OX, OY>(A, gx, gy, gz_own, gz_up, gz_dn, vs, ws);
        if (!__any(s + 1 < n)) break;
        A = walk_load<U, V, W>(sb, slot_of(s + 2));
        __builtin_amdgcn_sched_barrier(0);
        walk_step<OX, OY>(B, gx, gy, gz_own, gz_up, gz_dn, vs, ws);
    }
    move_neighbour_sums<OX, OY, false>(vs, ws);
}

__global__ __launch_bounds__(kThreads) void p2g_staged_kernel(const float *__restrict__ pos,
                                                             const float *__restrict__ vel,
                                                             const int32_t *__restrict__ offsets,
                                                             const int32_t *__restrict__ counts, float *__restrict__ u,
                                                             float *__restrict__ v, float *__restrict__ w,
                                                             float *__restrict__ ff, float inv_dx, float inv_ppc,
                                                             int nx, int ny, int nz) {
    __shared__ StageBuf buf[2];
    // occupancy counts of the patch's cells, [wave][lane]
    __shared__ int occ_s[4][64];
    const int tid = threadIdx.x, lane = tid & 63;
    occ_s[tid >> 6][lane] = 0;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform per wave
    const int ci0 = 2 * (int)blockIdx.z, cj0 = 2 * (int)blockIdx.y;
    const int ci = ci0 + (wave >> 1), cj = cj0 + (wave & 1);
    const int ck0 = (int)blockIdx.x * kSeg - 1, ck = ck0 + lane;  // this lane's bin; also its cell
    // lanes fl..ll hold the bins 0..nz-2 of the segment
    const int fl = ck0 < 0 ? -ck0 : 0, ll = min(63, nz - 2 - ck0);
    // Every wave lists the items itself, in registers: lane r < 16 holds the
    // particle range [rl, rh) of bin row r = 4 (bi - ci0 + 1) + (bj - cj0 + 1)
    // and pend, the number of kCap chunks of rows 0..r; item k is chunk
    // k - (pend[r] - chunks[r]) of the row r with pend[r - 1] <= k < pend[r]
    // (a ballot and readlanes, all scalar).
    int rl = 0, rh = 0;
    if (lane < kRowsStaged) {
        const int bi = ci0 - 1 + (lane >> 2), bj = cj0 - 1 + (lane & 3);
        if (bi >= 0 && bi < nx - 1 && bj >= 0 && bj < ny - 1 && fl <= ll) {
            const int64_t f = gidx(bi, bj, ck0 + fl, ny, nz), l = gidx(bi, bj, ck0 + ll, ny, nz);
            rl = offsets[f];
            rh = offsets[l] + counts[l];
        }
        // (ci0 + 2, cj0 + 2) is the (+1, +1) neighbour of one cell row only
        if (lane == kRowsStaged - 1) rh = rl;
    }
    const int chunks = rh > rl ? (rh - rl + kCap - 1) / kCap : 0;
    const int pend = row_scan16(chunks), pbeg = pend - chunks;
    const int items = __builtin_amdgcn_readlane(pend, kRowsStaged - 1);
    struct Item {
        int r, lo, hi;
    };
    auto item_at = [&](int k) {
        Item it;
        it.r = (int)__popcll(__ballot(lane < kRowsStaged && pend <= k));
        const int j = k - __builtin_amdgcn_readlane(pbeg, it.r);
        it.lo = __builtin_amdgcn_readlane(rl, it.r) + j * kCap;
        it.hi = min(it.lo + kCap, __builtin_amdgcn_readlane(rh, it.r));
        return it;
    };
    float vs[8][3], ws[8][3];
#pragma unroll
    for (int l = 0; l < 8; ++l)
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            vs[l][a] = 0.0f;
            ws[l][a] = 0.0f;
        }
    if (items == 0) {  // no particle within reach: every sum is +0 (block-uniform)
        if (ci >= nx || cj >= ny || lane == 0 || lane == 63 || ck >= nz) return;
        const int64_t index = gidx(ci, cj, ck, ny, nz);
        u[index] = 0.0f;
        v[index] = 0.0f;
        w[index] = 0.0f;
        ff[index] = 0.0f;
        return;
    }
    // staging registers: thread t holds the particles t, t + kThreads, ... of
    // an item; every lane the start and count of its bin in the item's row
    float rp[kPerThread][3], rv[kPerThread][3];
    int nb0 = 0, nbc = 0;
    auto load_item = [&](const Item &it) {
        const int n = it.hi - it.lo;
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const int t = tid + q * kThreads;
            const int64_t src = 3LL * (it.lo + min(t, n - 1));  // in bounds (n >= 1); unused when t >= n
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                rp[q][c] = pos[src + c];
                rv[q][c] = vel[src + c];
            }
        }
        const int bi = ci0 - 1 + (it.r >> 2), bj = cj0 - 1 + (it.r & 3);
        nb0 = 0;
        nbc = 0;
        if (ck >= 0 && ck < nz - 1) {
            const int64_t bin = gidx(bi, bj, ck, ny, nz);
            nb0 = offsets[bin];
            nbc = counts[bin];
        }
    };
    auto store_item = [&](const Item &it, StageBuf &sb) {
        const int n = it.hi - it.lo;
        // occupancy (particles.cl:186-188): the cell clamp(trunc(x inv_dx),
        // 1, n - 2) of a particle is its bin or the bin above on each axis
        // (clamps and NaN included), so every particle whose cell the patch
        // owns is staged by this block once, from one of the rows
        // (ci0 - 1 .. ci0 + 1, cj0 - 1 .. cj0 + 1); integer counts, exact
        const bool occ_row = (it.r >> 2) <= 2 && (it.r & 3) <= 2;  // block-uniform
#pragma unroll
        for (int q = 0; q < kPerThread; ++q) {
            const int t = tid + q * kThreads;
            if (t < n) {
                const int sl = stage_slot(t);
                float pc[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    pc[c] = rp[q][c] * inv_dx;
                    sb.a[c][sl] = pc[c];
                    sb.a[3 + c][sl] = rv[q][c];
                }
                if (occ_row) {
                    const int tx = trunc_clamp(pc[0], 1, nx - 2) - ci0;
                    const int ty = trunc_clamp(pc[1], 1, ny - 2) - cj0;
                    const int tz = trunc_clamp(pc[2], 1, nz - 2) - ck0;
                    if ((unsigned)tx < 2u && (unsigned)ty < 2u && tz >= 1 && tz <= 62)
                        atomicAdd(&occ_s[2 * tx + ty][tz], 1);
                }
            }
        }
        if (tid < 6) sb.a[tid][kSentinel] = tid < 3 ? -1.0e9f : 0.0f;
    };
    Item cur = item_at(0);
    load_item(cur);
    __syncthreads();  // occ_s cleared
    store_item(cur, buf[0]);
    int cb0 = nb0, cbc = nbc;
    __syncthreads();
    for (int k = 0; k < items; ++k) {
        const bool more = k + 1 < items;  // uniform
        Item nxt = cur;
        if (more) {
            nxt = item_at(k + 1);
            load_item(nxt);  // in flight during the walk
        }
        const StageBuf &sb = buf[k & 1];
        const int ox = ci0 - 1 + (cur.r >> 2) - ci, oy = cj0 - 1 + (cur.r & 3) - cj;
        if (ci < nx && cj < ny && ox >= -1 && ox <= 1 && oy >= -1 && oy <= 1) {
            const int lo = cur.lo, hi = cur.hi;
            // uniform per wave: the (ox, oy) of this walk
            switch ((ox + 1) * 3 + (oy + 1)) {
            case 0: walk<-1, -1>(sb, lo, hi, cb0, cbc, ci, cj, ck, vs, ws); break;
            case 1: walk<-1, 0>(sb, lo, hi, cb0, cbc, ci, cj, ck, vs, ws); break;
            case 2: walk<-1, 1>(sb, lo, hi, cb0, cbc, ci, cj, ck, vs, ws); break;
            case 3: walk<0, -1>(sb, lo, hi, cb0, cbc, ci, cj, ck, vs, ws); break;
            case 4: walk<0, 0>(sb, lo, hi, cb0, cbc, ci, cj, ck, vs, ws); break;
            case 5: walk<0, 1>(sb, lo, hi, cb0, cbc, ci, cj, ck, vs, ws); break;
            case 6: walk<1, -1>(sb, lo, hi, cb0, cbc, ci, cj, ck, vs, ws); break;
            case 7: walk<1, 0>(sb, lo, hi, cb0, cbc, ci, cj, ck, vs, ws); break;
            default: break;  // (+1, +1): no weight
            }
        }
        // the other buffer was last read by the walk of the previous item,
        // which every wave finished before the previous barrier
        if (more) {
            store_item(nxt, buf[(k + 1) & 1]);
            cb0 = nb0;
            cbc = nbc;
        }
        __syncthreads();
        cur = nxt;
    }
    if (ci >= nx || cj >= ny || lane == 0 || lane == 63 || ck >= nz) return;
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) {
#pragma unroll
        for (int l = 0; l < off; ++l) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                vs[l][a] = vs[l][a] + vs[l + off][a];
                ws[l][a] = ws[l][a] + ws[l + off][a];
            }
        }
    }
    const int64_t index = gidx(ci, cj, ck, ny, nz);
    u[index] = ws[0][0] > 0.0f ? vs[0][0] / ws[0][0] : 0.0f;
    v[index] = ws[0][1] > 0.0f ? vs[0][1] / ws[0][1] : 0.0f;
    w[index] = ws[0][2] > 0.0f ? vs[0][2] / ws[0][2] : 0.0f;
    ff[index] = (float)occ_s[wave][lane] * inv_ppc;  // particles.cl:226
}

void launch_p2g(const float *pos, const float *vel, const int32_t *offsets, const int32_t *counts, float *u, float *v,
                float *w, float *ff, float inv_dx, float inv_ppc, int nx, int ny, int nz, hipStream_t s) {
    const dim3 grid((unsigned)((nz + kSeg - 1) / kSeg), (unsigned)((ny + 1) / 2), (unsigned)((nx + 1) / 2));
    hipLaunchKernelGGL(p2g_staged_kernel, grid, dim3(kThreads), 0, s, pos, vel, offsets, counts, u, v, w, ff, inv_dx,
                       inv_ppc, nx, ny, nz);
}

}  // namespace clq
