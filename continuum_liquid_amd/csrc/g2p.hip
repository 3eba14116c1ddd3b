// G2P with FLIP/PIC blend, advection and particle boundary conditions in one
// pass over the particles (one lane per particle).
//
// Restates grid_to_particle (particles.cl:229-262), the previous-position copy
// of solver.py:469-474, advect_particles (particles.cl:264-276),
// particle_domain_bc (boundary_conditions.cl:69-97) and particle_obstacle_bc
// (boundary_conditions.cl:175-264).  The reference runs them as four kernels
// and a buffer copy; fused, position and velocity are read and written once.
// Every sub-step is selectable by a flag so each stage entry point of the
// C-ABI runs the same code.
//
// The 48 grid samples of a particle (eight corners of u, v, w and of their
// old values) come from one of two places:
//   * with the particles in bin order (the binning of this step is current),
//     a block takes the particles of one bin-row segment, (bi, bj, k0 ..
//     k0 + 61), and first stages the 3 x 3 cell rows (bi - 1 .. bi + 1,
//     bj - 1 .. bj + 1) x (k0 - 1 .. k0 + 62) of all six arrays in LDS.  A
//     particle's bin is floor(fma(x, 1/dx, -0.5)) per axis, clamped; its
//     stencil base for the offset-1/2 axes is that bin and for the offset-1
//     axis the bin or the bin below (fl(r - 1) <= fl(r - 0.5) and
//     fl(r - 1) >= fl(r - 0.5) - 1 for every float r, the clamps included),
//     so every corner lies in the staged rows.  The samples are then LDS
//     reads instead of 48 scattered loads through the texture units, which
//     bound the per-particle form;
//   * otherwise each lane loads its corners from global memory.
// Both forms perform the same arithmetic on the same values.
#include "common.hpp"

namespace clq {

// sample_field at the staggered base and fraction of particles.cl:12-42.
struct Stencil {
    int b0, b1, b2;
    float f0, f1, f2;
};

__device__ __forceinline__ Stencil stagger(const float *x, float inv_dx, float o0, float o1, float o2, int nx, int ny,
                                           int nz) {
    const float t0 = fmaf(x[0], inv_dx, -o0), t1 = fmaf(x[1], inv_dx, -o1), t2 = fmaf(x[2], inv_dx, -o2);
    const int b0 = floor_clamp(t0, 0, nx - 2), b1 = floor_clamp(t1, 0, ny - 2), b2 = floor_clamp(t2, 0, nz - 2);
    Stencil s;
    s.b0 = b0;
    s.b1 = b1;
    s.b2 = b2;
    s.f0 = fminf(fmaxf(t0 - (float)b0, 0.0f), 1.0f);
    s.f1 = fminf(fmaxf(t1 - (float)b1, 0.0f), 1.0f);
    s.f2 = fminf(fmaxf(t2 - (float)b2, 0.0f), 1.0f);
    return s;
}

// The eight corner values of field `a` (u, v, w, u_old, v_old, w_old) around
// a stencil, corner c = 4 ox + 2 oy + oz: from global memory ...
struct GlobalFetch {
    const float *f[6];
    int ny, nz;
    __device__ __forceinline__ void operator()(int a, const Stencil &s, float (&c)[8]) const {
        const int64_t sx = (int64_t)ny * nz, sy = nz, base = gidx(s.b0, s.b1, s.b2, ny, nz);
#pragma unroll
        for (int n = 0; n < 8; ++n) c[n] = f[a][base + (n >> 2) * sx + ((n >> 1) & 1) * sy + (n & 1)];
    }
};

// ... or from the staged rows of a bin-row segment (layout [a][3][3][kStageZ]
// over x = bi - 1 .., y = bj - 1 .., z = zlo ..).  The bin argument above puts
// every corner in them; the clamps only keep each LDS address in the array.
constexpr int kG2PSeg = 62;             // bins per block along z
constexpr int kStageZ = kG2PSeg + 2;    // staged cells along z
struct StagedFetch {
    const float *st;
    int bi, bj, zlo;
    __device__ __forceinline__ void operator()(int a, const Stencil &s, float (&c)[8]) const {
        const int rx = min(max(s.b0 - (bi - 1), 0), 1), ry = min(max(s.b1 - (bj - 1), 0), 1);
        const int rz = min(max(s.b2 - zlo, 0), kStageZ - 2);
        const float *q = st + ((a * 3 + rx) * 3 + ry) * kStageZ + rz;
#pragma unroll
        for (int n = 0; n < 8; ++n) c[n] = q[(n >> 2) * 3 * kStageZ + ((n >> 1) & 1) * kStageZ + (n & 1)];
    }
};

// the trilinear sum over the corners, in the reference's corner order
__device__ __forceinline__ float blend(const float (&c)[8], const Stencil &s) {
    float value = 0.0f;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
        const float wx = (n >> 2) ? s.f0 : 1.0f - s.f0;
        const float wy = ((n >> 1) & 1) ? s.f1 : 1.0f - s.f1;
        const float wz = (n & 1) ? s.f2 : 1.0f - s.f2;
        const float w = wx * wy * wz;
        value = fmaf(w, c[n], value);  // particles.cl:37 (fmuladd)
    }
    return value;
}

__device__ __forceinline__ void obstacle_bc(float *x, float *v, const float *pv, const uint8_t *__restrict__ ob,
                                            float inv_dx, int nx, int ny, int nz) {
    const int cx = trunc_clamp(x[0] * inv_dx, 0, nx - 1), cy = trunc_clamp(x[1] * inv_dx, 0, ny - 1),
              cz = trunc_clamp(x[2] * inv_dx, 0, nz - 1);
    if (!ob[gidx(cx, cy, cz, ny, nz)]) return;
    const int px = trunc_clamp(pv[0] * inv_dx, 0, nx - 1), py = trunc_clamp(pv[1] * inv_dx, 0, ny - 1),
              pz = trunc_clamp(pv[2] * inv_dx, 0, nz - 1);
    float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f;
    if ((px != cx || py != cy || pz != cz) && !ob[gidx(px, py, pz, ny, nz)]) {
        n0 = (float)(px - cx); n1 = (float)(py - cy); n2 = (float)(pz - cz);
    } else {
        if (cx > 0 && !ob[gidx(cx - 1, cy, cz, ny, nz)]) n0 -= 1.0f;
        if (cx + 1 < nx && !ob[gidx(cx + 1, cy, cz, ny, nz)]) n0 += 1.0f;
        if (cy > 0 && !ob[gidx(cx, cy - 1, cz, ny, nz)]) n1 -= 1.0f;
        if (cy + 1 < ny && !ob[gidx(cx, cy + 1, cz, ny, nz)]) n1 += 1.0f;
        if (cz > 0 && !ob[gidx(cx, cy, cz - 1, ny, nz)]) n2 -= 1.0f;
        if (cz + 1 < nz && !ob[gidx(cx, cy, cz + 1, ny, nz)]) n2 += 1.0f;
    }
    const float len = sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
    if (len > 0.0f) {
        n0 = n0 / len; n1 = n1 / len; n2 = n2 / len;
    } else {
        const float f0 = pv[0] - x[0], f1 = pv[1] - x[1], f2 = pv[2] - x[2];
        const float fl = sqrtf(f0 * f0 + f1 * f1 + f2 * f2);
        if (fl > 0.0f) { n0 = f0 / fl; n1 = f1 / fl; n2 = f2 / fl; }
        else { n0 = 0.0f; n1 = 0.0f; n2 = 1.0f; }
    }
    const float vn = v[0] * n0 + v[1] * n1 + v[2] * n2;
    if (vn < 0.0f) {  // boundary_conditions.cl:237 (fmuladd)
        v[0] = fmaf(-vn, n0, v[0]); v[1] = fmaf(-vn, n1, v[1]); v[2] = fmaf(-vn, n2, v[2]);
    }
    const float dx = 1.0f / inv_dx;
    const int cell[3] = {cx, cy, cz};
    const float nn[3] = {n0, n1, n2};
    float corr[3] = {x[0], x[1], x[2]};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (nn[a] < 0.0f) corr[a] = fmaf(2.0f, (float)cell[a] * dx, -x[a]);
        else if (nn[a] > 0.0f) corr[a] = fmaf(2.0f, (float)(cell[a] + 1) * dx, -x[a]);
    }
    const int qx = trunc_clamp(corr[0] * inv_dx, 0, nx - 1), qy = trunc_clamp(corr[1] * inv_dx, 0, ny - 1),
              qz = trunc_clamp(corr[2] * inv_dx, 0, nz - 1);
    if (ob[gidx(qx, qy, qz, ny, nz)]) { corr[0] = pv[0]; corr[1] = pv[1]; corr[2] = pv[2]; }
    x[0] = corr[0]; x[1] = corr[1]; x[2] = corr[2];
}

template <int FLAGS, typename Fetch>
__device__ __forceinline__ void update_particle(int64_t i, float *__restrict__ pos, float *__restrict__ vel,
                                                float *__restrict__ prev, const Fetch &fetch,
                                                const uint8_t *__restrict__ ob, int nx, int ny, int nz,
                                                const ParticleUpdate &pu) {
    float x[3] = {pos[3LL * i], pos[3LL * i + 1], pos[3LL * i + 2]};
    float vv[3] = {vel[3LL * i], vel[3LL * i + 1], vel[3LL * i + 2]};
    if constexpr (FLAGS & kG2P) {
        const Stencil su = stagger(x, pu.inv_dx, 1.0f, 0.5f, 0.5f, nx, ny, nz);
        const Stencil sv = stagger(x, pu.inv_dx, 0.5f, 1.0f, 0.5f, nx, ny, nz);
        const Stencil sw = stagger(x, pu.inv_dx, 0.5f, 0.5f, 1.0f, nx, ny, nz);
        // all 48 corner values are loaded before any is used, so they are in
        // flight together (sampled one field at a time they were waited for
        // in pairs)
        float cu[8], cv[8], cw[8], cuo[8], cvo[8], cwo[8];
        fetch(0, su, cu);
        fetch(1, sv, cv);
        fetch(2, sw, cw);
        fetch(3, su, cuo);
        fetch(4, sv, cvo);
        fetch(5, sw, cwo);
        const float nu = blend(cu, su), nv = blend(cv, sv), nw = blend(cw, sw);
        const float f0 = vv[0] + (nu - blend(cuo, su));
        const float f1 = vv[1] + (nv - blend(cvo, sv));
        const float f2 = vv[2] + (nw - blend(cwo, sw));
        if (pu.pure_flip) {
            vv[0] = f0; vv[1] = f1; vv[2] = f2;
        } else {
            vv[0] = fmaf(pu.a, f0, pu.b * nu);
            vv[1] = fmaf(pu.a, f1, pu.b * nv);
            vv[2] = fmaf(pu.a, f2, pu.b * nw);
        }
    }
    float pv[3] = {x[0], x[1], x[2]};
    if constexpr (FLAGS & kSavePrev) {
        prev[3LL * i] = x[0]; prev[3LL * i + 1] = x[1]; prev[3LL * i + 2] = x[2];
    } else if constexpr (FLAGS & kObstacleBC) {
        pv[0] = prev[3LL * i]; pv[1] = prev[3LL * i + 1]; pv[2] = prev[3LL * i + 2];
    }
    if constexpr (FLAGS & kAdvect) {  // particles.cl:275 (fmuladd)
        x[0] = fmaf(pu.dt, vv[0], x[0]); x[1] = fmaf(pu.dt, vv[1], x[1]); x[2] = fmaf(pu.dt, vv[2], x[2]);
    }
    if constexpr (FLAGS & kDomainBC) {
        bool hit[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const bool lo = x[a] < pu.lower, hi = x[a] > pu.upper[a];
            hit[a] = lo || hi;
            float y = x[a];
            if (lo) y = fmaf(2.0f, pu.lower, -y);
            if (hi) y = fmaf(2.0f, pu.upper[a], -y);
            x[a] = fminf(fmaxf(y, pu.lower), pu.upper[a]);
        }
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (hit[a]) vv[a] = 0.0f;
    }
    if constexpr (FLAGS & kObstacleBC) obstacle_bc(x, vv, pv, ob, pu.inv_dx, nx, ny, nz);
    if constexpr (FLAGS & (kAdvect | kDomainBC | kObstacleBC)) {
        pos[3LL * i] = x[0]; pos[3LL * i + 1] = x[1]; pos[3LL * i + 2] = x[2];
    }
    if constexpr (FLAGS & (kG2P | kDomainBC | kObstacleBC)) {
        vel[3LL * i] = vv[0]; vel[3LL * i + 1] = vv[1]; vel[3LL * i + 2] = vv[2];
    }
}

template <int FLAGS>
__global__ __launch_bounds__(256) void particle_update_kernel(float *__restrict__ pos, float *__restrict__ vel,
                                                              float *__restrict__ prev, GlobalFetch fetch,
                                                              const uint8_t *__restrict__ ob, int n, int nx, int ny,
                                                              int nz, ParticleUpdate pu) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    update_particle<FLAGS>(i, pos, vel, prev, fetch, ob, nx, ny, nz, pu);
}

// One block per bin-row segment (bi = blockIdx.z, bj = blockIdx.y, bins
// k0 = blockIdx.x * kG2PSeg ..): stage, then update its particles.
template <int FLAGS>
__global__ __launch_bounds__(256) void particle_update_staged_kernel(float *__restrict__ pos, float *__restrict__ vel,
                                                                     float *__restrict__ prev, GlobalFetch fetch,
                                                                     const int32_t *__restrict__ offsets,
                                                                     const int32_t *__restrict__ counts,
                                                                     const uint8_t *__restrict__ ob, int nx, int ny,
                                                                     int nz, ParticleUpdate pu) {
    __shared__ float st[6 * 9 * kStageZ];
    const int bi = blockIdx.z, bj = blockIdx.y, k0 = blockIdx.x * kG2PSeg;
    const int k1 = min(k0 + kG2PSeg, nz - 1);  // bins k0 .. k1 - 1
    const int64_t first = gidx(bi, bj, k0, ny, nz), last = gidx(bi, bj, k1 - 1, ny, nz);
    const int lo = offsets[first], hi = offsets[last] + counts[last];
    if (lo >= hi) return;  // no particle in the segment (block-uniform)
    const int zlo = k0 - 1;
    // Only the rows an array is sampled in are staged (the bin argument
    // above): u rows x = bi - 1 .. bi + 1, y = bj, bj + 1; v rows x = bi,
    // bi + 1, y = bj - 1 .. bj + 1; w rows x = bi, bi + 1, y = bj, bj + 1.
    // Every load is issued before the first store, so the block waits for
    // one memory round trip; the array loop is unrolled so each array's
    // pointer is a scalar.
    constexpr int kPer = (6 * kStageZ + 255) / 256;
    float val[6][kPer];
    int slot[6][kPer];
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        const float *__restrict__ f = fetch.f[a];
        const int c = a % 3, rows = c == 2 ? 4 : 6;
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int e = threadIdx.x + q * 256, ri = e / kStageZ, z = zlo + e % kStageZ;
            const int rx = c == 0 ? ri / 2 : 1 + (c == 1 ? ri / 3 : ri / 2);
            const int ry = c == 1 ? ri % 3 : 1 + ri % 2;
            const int x = bi - 1 + rx, y = bj - 1 + ry;
            // unconditional load of a clamped cell; the value of a cell
            // outside the grid (never sampled) is dropped
            const bool inside = e < rows * kStageZ && x >= 0 && x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
            const float loaded =
                f[gidx(min(max(x, 0), nx - 1), min(max(y, 0), ny - 1), min(max(z, 0), nz - 1), ny, nz)];
            val[a][q] = inside ? loaded : 0.0f;
            slot[a][q] = e < rows * kStageZ ? ((a * 3 + rx) * 3 + ry) * kStageZ + e % kStageZ : -1;
        }
    }
#pragma unroll
    for (int a = 0; a < 6; ++a)
#pragma unroll
        for (int q = 0; q < kPer; ++q)
            if (slot[a][q] >= 0) st[slot[a][q]] = val[a][q];
    __syncthreads();
    const StagedFetch sf{st, bi, bj, zlo};
    for (int i = lo + (int)threadIdx.x; i < hi; i += 256)
        update_particle<FLAGS>(i, pos, vel, prev, sf, ob, nx, ny, nz, pu);
}

template <int F>
static void launch_one(float *pos, float *vel, float *prev, const GlobalFetch &g, const int32_t *offsets,
                       const int32_t *counts, const uint8_t *ob, int n, int nx, int ny, int nz,
                       const ParticleUpdate &pu, hipStream_t s) {
    if ((F & kG2P) && offsets) {
        const dim3 grid((unsigned)((nz - 1 + kG2PSeg - 1) / kG2PSeg), (unsigned)(ny - 1), (unsigned)(nx - 1));
        hipLaunchKernelGGL(particle_update_staged_kernel<F>, grid, dim3(256), 0, s, pos, vel, prev, g, offsets, counts,
                           ob, nx, ny, nz, pu);
        return;
    }
    hipLaunchKernelGGL(particle_update_kernel<F>, dim3((n + 255) / 256), dim3(256), 0, s, pos, vel, prev, g, ob, n, nx,
                       ny, nz, pu);
}

void launch_particle_update(float *pos, float *vel, float *prev, const float *u, const float *v, const float *w,
                            const float *uo, const float *vo, const float *wo, const int32_t *offsets,
                            const int32_t *counts, const uint8_t *ob, int n, int nx, int ny, int nz,
                            const ParticleUpdate &pu, hipStream_t s) {
    if (n <= 0) return;
    const GlobalFetch g{{u, v, w, uo, vo, wo}, ny, nz};
#define CLQ_CASE(F)                                                                          \
    case F:                                                                                  \
        launch_one<F>(pos, vel, prev, g, offsets, counts, ob, n, nx, ny, nz, pu, s);         \
        return;
    switch (pu.flags) {
        CLQ_CASE(kG2P)
        CLQ_CASE(kSavePrev)
        CLQ_CASE(kAdvect)
        CLQ_CASE(kDomainBC)
        CLQ_CASE(kObstacleBC)
        CLQ_CASE(kG2P | kAdvect | kDomainBC)
        CLQ_CASE(kG2P | kSavePrev | kAdvect | kDomainBC | kObstacleBC)
    default:
        break;
    }
#undef CLQ_CASE
}

}  // namespace clq
