// Particle voxelisation at s x the grid resolution: occupancy and smoothed
// density.  No reference counterpart beyond the occupancy of
// particle_to_grid_gather (particles.cl:186-188, 226), which is the s = 1 case.
//
// Occupancy: one lane per particle stores 1 into its refined cell
// clamp(trunc(x * s * inv_dx), s, s*(n-1)-1); concurrent stores write the same
// value, so the result is deterministic.
//
// Density: a gather, no float atomics.  One lane per coarse cell owns its s^3
// refined cells, scans the particles of the 27 neighbour bins of the sorted
// array (every particle whose trilinear splat at refined resolution reaches
// the cell lies in them) and accumulates the refined-cell-centred trilinear
// weights in FP64, scaled by s^3 / particles_per_cell and rounded once.
#include "common.hpp"

namespace clq {

__global__ __launch_bounds__(256) void occupancy_kernel(const float *__restrict__ pos, int n, float sinv, int s,
                                                        int nx, int ny, int nz, uint8_t *__restrict__ occ) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int64_t sy = (int64_t)s * ny, sz = (int64_t)s * nz;
    const int a = trunc_clamp(pos[3LL * i] * sinv, s, s * (nx - 1) - 1);
    const int b = trunc_clamp(pos[3LL * i + 1] * sinv, s, s * (ny - 1) - 1);
    const int c = trunc_clamp(pos[3LL * i + 2] * sinv, s, s * (nz - 1) - 1);
    occ[((int64_t)a * sy + b) * sz + c] = 1;
}

constexpr int kMaxScale = 2;

__global__ __launch_bounds__(256) void density_kernel(const float *__restrict__ pos, const int32_t *__restrict__ offsets,
                                                      const int32_t *__restrict__ counts, float sinv, int s, int nx,
                                                      int ny, int nz, double scale, float *__restrict__ dens) {
    const int64_t cell = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (cell >= (int64_t)nx * ny * nz) return;
    int ci, cj, ck;
    coords(cell, ny, nz, ci, cj, ck);
    const int64_t dims[3] = {(int64_t)s * nx, (int64_t)s * ny, (int64_t)s * nz};
    double acc[kMaxScale * kMaxScale * kMaxScale];
    for (int q = 0; q < s * s * s; ++q) acc[q] = 0.0;
    const int base_cell[3] = {ci * s, cj * s, ck * s};
    for (int ox = -1; ox <= 1; ++ox)
        for (int oy = -1; oy <= 1; ++oy)
            for (int oz = -1; oz <= 1; ++oz) {
                const int bx = ci + ox, by = cj + oy, bz = ck + oz;
                if (bx < 0 || bx >= nx - 1 || by < 0 || by >= ny - 1 || bz < 0 || bz >= nz - 1) continue;
                const int64_t bin = gidx(bx, by, bz, ny, nz);
                const int begin = offsets[bin], end = begin + counts[bin];
                for (int pi = begin; pi < end; ++pi) {
                    int base[3];
                    double fr[3];
                    bool touches = true;
                    for (int a = 0; a < 3; ++a) {
                        const double t = (double)pos[3LL * pi + a] * (double)sinv - 0.5;
                        double fl = floor(t);
                        if (fl < 0) fl = 0;
                        if (fl > (double)(dims[a] - 2)) fl = (double)(dims[a] - 2);
                        base[a] = (int)fl;
                        const double f = t - fl;
                        fr[a] = f < 0 ? 0 : (f > 1 ? 1 : f);
                        // the particle reaches refined cells base, base+1 on this axis
                        if (base[a] + 1 < base_cell[a] || base[a] >= base_cell[a] + s) touches = false;
                    }
                    if (!touches) continue;
                    for (int q = 0; q < s * s * s; ++q) {
                        const int r[3] = {base_cell[0] + q / (s * s), base_cell[1] + (q / s) % s, base_cell[2] + q % s};
                        double w = 1.0;
                        for (int a = 0; a < 3; ++a) {
                            const int d = r[a] - base[a];
                            w *= d == 0 ? 1.0 - fr[a] : (d == 1 ? fr[a] : 0.0);
                        }
                        acc[q] += w;
                    }
                }
            }
    for (int q = 0; q < s * s * s; ++q) {
        const int64_t r0 = base_cell[0] + q / (s * s), r1 = base_cell[1] + (q / s) % s, r2 = base_cell[2] + q % s;
        dens[(r0 * dims[1] + r1) * dims[2] + r2] = (float)(acc[q] * scale);
    }
}

int voxelise_max_scale() { return kMaxScale; }

void launch_voxelise(const float *pos, int n, const int32_t *offsets, const int32_t *counts, int s, float sinv,
                     double scale, int nx, int ny, int nz, uint8_t *occ, float *dens, hipStream_t st) {
    if (occ && n > 0)
        hipLaunchKernelGGL(occupancy_kernel, dim3((n + 255) / 256), dim3(256), 0, st, pos, n, sinv, s, nx, ny, nz, occ);
    if (dens) {
        const int64_t C = (int64_t)nx * ny * nz;
        hipLaunchKernelGGL(density_kernel, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, st, pos, offsets, counts,
                           sinv, s, nx, ny, nz, scale, dens);
    }
}

}  // namespace clq
