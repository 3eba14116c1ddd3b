// G2P with FLIP/PIC blend, advection and particle boundary conditions in one
// pass over the particles (one lane per particle).
//
// Restates grid_to_particle (particles.cl:229-262), the previous-position copy
// of solver.py:469-474, advect_particles (particles.cl:264-276),
// particle_domain_bc (boundary_conditions.cl:69-97) and particle_obstacle_bc
// (boundary_conditions.cl:175-264).  The reference runs them as four kernels
// and a buffer copy; fused, position and velocity are read and written once
// (48 B per particle) and the 48 grid samples hit L2 because the particles are
// sorted by bin.  Every sub-step is selectable by a flag so each stage entry
// point of the C-ABI runs the same code.
#include "common.hpp"

namespace clq {

// sample_field at the staggered base and fraction of particles.cl:12-42.
struct Stencil {
    int64_t base;
    float f0, f1, f2;
};

__device__ __forceinline__ Stencil stagger(const float *x, float inv_dx, float o0, float o1, float o2, int nx, int ny,
                                           int nz) {
    const float t0 = fmaf(x[0], inv_dx, -o0), t1 = fmaf(x[1], inv_dx, -o1), t2 = fmaf(x[2], inv_dx, -o2);
    const int b0 = floor_clamp(t0, 0, nx - 2), b1 = floor_clamp(t1, 0, ny - 2), b2 = floor_clamp(t2, 0, nz - 2);
    Stencil s;
    s.base = gidx(b0, b1, b2, ny, nz);
    s.f0 = fminf(fmaxf(t0 - (float)b0, 0.0f), 1.0f);
    s.f1 = fminf(fmaxf(t1 - (float)b1, 0.0f), 1.0f);
    s.f2 = fminf(fmaxf(t2 - (float)b2, 0.0f), 1.0f);
    return s;
}

// The eight corner values of a field around a stencil, corner c = 4 ox + 2 oy + oz.
__device__ __forceinline__ void corners(const float *__restrict__ f, const Stencil &s, int64_t sx, int64_t sy,
                                        float (&c)[8]) {
#pragma unroll
    for (int n = 0; n < 8; ++n) c[n] = f[s.base + (n >> 2) * sx + ((n >> 1) & 1) * sy + (n & 1)];
}

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

template <int FLAGS>
__global__ __launch_bounds__(256) void particle_update_kernel(float *__restrict__ pos, float *__restrict__ vel,
                                                              float *__restrict__ prev, const float *__restrict__ u,
                                                              const float *__restrict__ v, const float *__restrict__ w,
                                                              const float *__restrict__ uo,
                                                              const float *__restrict__ vo,
                                                              const float *__restrict__ wo,
                                                              const uint8_t *__restrict__ ob, int n, int nx, int ny,
                                                              int nz, ParticleUpdate pu) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float x[3] = {pos[3LL * i], pos[3LL * i + 1], pos[3LL * i + 2]};
    float vv[3] = {vel[3LL * i], vel[3LL * i + 1], vel[3LL * i + 2]};
    if constexpr (FLAGS & kG2P) {
        const int64_t sx = (int64_t)ny * nz, sy = nz;
        const Stencil su = stagger(x, pu.inv_dx, 1.0f, 0.5f, 0.5f, nx, ny, nz);
        const Stencil sv = stagger(x, pu.inv_dx, 0.5f, 1.0f, 0.5f, nx, ny, nz);
        const Stencil sw = stagger(x, pu.inv_dx, 0.5f, 0.5f, 1.0f, nx, ny, nz);
        // all 48 corner values are loaded before any is used, so they are in
        // flight together (sampled one field at a time they were waited for
        // in pairs)
        float cu[8], cv[8], cw[8], cuo[8], cvo[8], cwo[8];
        corners(u, su, sx, sy, cu);
        corners(v, sv, sx, sy, cv);
        corners(w, sw, sx, sy, cw);
        corners(uo, su, sx, sy, cuo);
        corners(vo, sv, sx, sy, cvo);
        corners(wo, sw, sx, sy, cwo);
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

template <int F>
static void launch_one(float *pos, float *vel, float *prev, const float *u, const float *v, const float *w,
                       const float *uo, const float *vo, const float *wo, const uint8_t *ob, int n, int nx, int ny,
                       int nz, const ParticleUpdate &pu, hipStream_t s) {
    hipLaunchKernelGGL(particle_update_kernel<F>, dim3((n + 255) / 256), dim3(256), 0, s, pos, vel, prev, u, v, w, uo,
                       vo, wo, ob, n, nx, ny, nz, pu);
}

void launch_particle_update(float *pos, float *vel, float *prev, const float *u, const float *v, const float *w,
                            const float *uo, const float *vo, const float *wo, const uint8_t *ob, int n, int nx,
                            int ny, int nz, const ParticleUpdate &pu, hipStream_t s) {
    if (n <= 0) return;
#define CLQ_CASE(F)                                                                  \
    case F:                                                                          \
        launch_one<F>(pos, vel, prev, u, v, w, uo, vo, wo, ob, n, nx, ny, nz, pu, s); \
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
