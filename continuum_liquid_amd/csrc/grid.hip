// Grid glue kernels of the step: boundary conditions, forces and viscosity,
// pressure right-hand side, finest operator, pressure correction.  All are
// streaming per-cell kernels, HBM-bound, one lane per cell.
//
// Each restates one reference kernel (Solver/Kernel/OpenCL/...):
//   grid_obstacle_bc + extrapolate_obstacle_velocity  boundary_conditions.cl:99-173
//   grid_domain_bc                                    boundary_conditions.cl:33-67
//   update_velocity                                   velocity_update.cl:3-43
//   pressure_rhs                                      pressure_solve.cl:3-24
//   build_finest_operator                             multigrid.cl:63-88
//   neumann_domain_bc                                 boundary_conditions.cl:18-31
//   pressure_correction                               pressure_solve.cl:26-47
#include "common.hpp"

namespace clq {

static inline dim3 cells_grid(int64_t C) { return dim3((unsigned)((C + 255) / 256)); }

__global__ __launch_bounds__(256) void grid_obstacle_kernel(float *__restrict__ u, float *__restrict__ v,
                                                            float *__restrict__ w, float *__restrict__ ff,
                                                            const uint8_t *__restrict__ ob, int nx, int ny, int nz) {
    const int64_t index = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (index >= (int64_t)nx * ny * nz) return;
    int i, j, k;
    coords(index, ny, nz, i, j, k);
    const bool self = ob[index];
    if (self) ff[index] = 0.0f;
    const bool xs = self || (i + 1 < nx && ob[gidx(i + 1, j, k, ny, nz)]);
    const bool ys = self || (j + 1 < ny && ob[gidx(i, j + 1, k, ny, nz)]);
    const bool zs = self || (k + 1 < nz && ob[gidx(i, j, k + 1, ny, nz)]);
    if (xs) u[index] = 0.0f;
    if (ys) v[index] = 0.0f;
    if (zs) w[index] = 0.0f;
}

// Solid cells read only fluid neighbours, which this kernel never writes.
__global__ __launch_bounds__(256) void extrapolate_obstacle_kernel(float *__restrict__ u, float *__restrict__ v,
                                                                   float *__restrict__ w,
                                                                   const uint8_t *__restrict__ ob, int nx, int ny,
                                                                   int nz) {
    const int64_t index = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (index >= (int64_t)nx * ny * nz || !ob[index]) return;
    int i, j, k;
    coords(index, ny, nz, i, j, k);
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f;
    int nf = 0;
    bool ax = false, ay = false, az = false;
    auto acc = [&](int a, int b, int c, int axis) {
        const int64_t nb = gidx(a, b, c, ny, nz);
        if (!ob[nb]) {
            t0 = t0 + u[nb];
            t1 = t1 + v[nb];
            t2 = t2 + w[nb];
            nf += 1;
            if (axis == 0) ax = true;
            if (axis == 1) ay = true;
            if (axis == 2) az = true;
        }
    };
    if (i > 0) acc(i - 1, j, k, 0);
    if (i + 1 < nx) acc(i + 1, j, k, 0);
    if (j > 0) acc(i, j - 1, k, 1);
    if (j + 1 < ny) acc(i, j + 1, k, 1);
    if (k > 0) acc(i, j, k - 1, 2);
    if (k + 1 < nz) acc(i, j, k + 1, 2);
    if (nf == 0) return;
    const float fn = (float)nf;
    u[index] = ax ? 0.0f : t0 / fn;
    v[index] = ay ? 0.0f : t1 / fn;
    w[index] = az ? 0.0f : t2 / fn;
}

void launch_grid_obstacle_bc(float *u, float *v, float *w, float *ff, const uint8_t *ob, int nx, int ny, int nz,
                             hipStream_t s) {
    const int64_t C = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(grid_obstacle_kernel, cells_grid(C), dim3(256), 0, s, u, v, w, ff, ob, nx, ny, nz);
    hipLaunchKernelGGL(extrapolate_obstacle_kernel, cells_grid(C), dim3(256), 0, s, u, v, w, ob, nx, ny, nz);
}

// The cells within S - 1 of a face: i, j or k in {0, n-S+1, ..., n-1} (S = 3:
// {0, n-2, n-1}, the cells the domain condition touches; S = 2: {0, n-1}, the
// ghost layer).  Item t enumerates them without repeats: first the S x planes,
// then the S y planes of the remaining x range [1, n-S], then the S z planes
// of the remaining (x, y) range, the S cells of one (x, y) row in a row.
template <int S>
__device__ __forceinline__ bool face_cell(int64_t t, int nx, int ny, int nz, int &i, int &j, int &k) {
    const int mx = nx - S, my = ny - S;  // remaining ranges [1, n-S]
    const int64_t a = (int64_t)S * ny * nz, b = (int64_t)S * mx * nz, c = (int64_t)S * mx * my;
    auto side = [](int64_t s, int n) { return s == 0 ? 0 : (int)(n - S + s); };
    if (t < a) {
        const int64_t pl = t / ((int64_t)ny * nz), r = t - pl * ny * nz;
        i = side(pl, nx);
        j = (int)(r / nz);
        k = (int)(r - (int64_t)j * nz);
        return true;
    }
    t -= a;
    if (t < b) {
        const int64_t pl = t / ((int64_t)mx * nz), r = t - pl * mx * nz;
        j = side(pl, ny);
        i = 1 + (int)(r / nz);
        k = (int)(r - (int64_t)(i - 1) * nz);
        return true;
    }
    t -= b;
    if (t < c) {  // plane fastest: a row's S cells (and the cells they copy) share cache lines
        const int64_t r = t / S;
        k = side(t - r * S, nz);
        i = 1 + (int)(r / my);
        j = 1 + (int)(r - (int64_t)(i - 1) * my);
        return true;
    }
    return false;
}
template <int S>
static int64_t face_cells(int nx, int ny, int nz) {
    return (int64_t)S * ny * nz + (int64_t)S * (nx - S) * nz + (int64_t)S * (nx - S) * (ny - S);
}

// A ghost cell copies only components that are not walls at its clamped
// interior cell, and interior lanes write only wall components: race free.
// Launched over the cells that face_cell<3> enumerates (every other cell is
// neither a ghost nor a wall face).
__global__ __launch_bounds__(256) void grid_domain_kernel(float *__restrict__ u, float *__restrict__ v,
                                                          float *__restrict__ w, int nx, int ny, int nz) {
    int i, j, k;
    if (!face_cell<3>((int64_t)blockIdx.x * 256 + threadIdx.x, nx, ny, nz, i, j, k)) return;
    const int64_t index = gidx(i, j, k, ny, nz);
    const bool ghost = !interior(i, j, k, nx, ny, nz);
    const bool wu = i == 0 || i >= nx - 2;
    const bool wv = j == 0 || j >= ny - 2;
    const bool ww = k == 0 || k >= nz - 2;
    if (!ghost && !wu && !wv && !ww) return;
    if (ghost) {
        const int ii = min(max(i, 1), nx - 2), jj = min(max(j, 1), ny - 2), kk = min(max(k, 1), nz - 2);
        const int64_t in = gidx(ii, jj, kk, ny, nz);
        u[index] = wu ? 0.0f : u[in];
        v[index] = wv ? 0.0f : v[in];
        w[index] = ww ? 0.0f : w[in];
        return;
    }
    if (wu) u[index] = 0.0f;
    if (wv) v[index] = 0.0f;
    if (ww) w[index] = 0.0f;
}

void launch_grid_domain_bc(float *u, float *v, float *w, int nx, int ny, int nz, hipStream_t s) {
    hipLaunchKernelGGL(grid_domain_kernel, cells_grid(face_cells<3>(nx, ny, nz)), dim3(256), 0, s, u, v, w, nx, ny,
                       nz);
}

// laplacian (helper.cl:54-66): fma(-6, centre, neighbour_sum) * inv_dx2.
__device__ __forceinline__ float laplacian(const float *__restrict__ f, int64_t idx, int64_t sx, int64_t sy,
                                           float inv_dx2) {
    const float centre = f[idx];
    float s = f[idx + sx] + f[idx - sx];
    s = s + f[idx + sy];
    s = s + f[idx - sy];
    s = s + f[idx + 1];
    s = s + f[idx - 1];
    return fmaf(-6.0f, centre, s) * inv_dx2;
}

__global__ __launch_bounds__(256) void update_velocity_kernel(const float *__restrict__ u, const float *__restrict__ v,
                                                              const float *__restrict__ w, float *__restrict__ ru,
                                                              float *__restrict__ rv, float *__restrict__ rw,
                                                              const float *__restrict__ ff, float ax, float ay,
                                                              float az, float dt, float nudt, float inv_dx2, int nx,
                                                              int ny, int nz) {
    const int64_t index = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (index >= (int64_t)nx * ny * nz) return;
    int i, j, k;
    coords(index, ny, nz, i, j, k);
    float r0 = u[index], r1 = v[index], r2 = w[index];
    if (ff[index] > 0.0f) {  // velocity_update.cl:31 (fmuladd)
        r0 = fmaf(dt, ax, r0);
        r1 = fmaf(dt, ay, r1);
        r2 = fmaf(dt, az, r2);
    }
    if (interior(i, j, k, nx, ny, nz)) {  // velocity_update.cl:35-37 (fmuladd)
        const int64_t sx = (int64_t)ny * nz, sy = nz;
        r0 = fmaf(nudt, laplacian(u, index, sx, sy, inv_dx2), r0);
        r1 = fmaf(nudt, laplacian(v, index, sx, sy, inv_dx2), r1);
        r2 = fmaf(nudt, laplacian(w, index, sx, sy, inv_dx2), r2);
    }
    ru[index] = r0;
    rv[index] = r1;
    rw[index] = r2;
}

void launch_update_velocity(const float *u, const float *v, const float *w, float *ru, float *rv, float *rw,
                            const float *ff, float ax, float ay, float az, float dt, float nudt, float inv_dx2, int nx,
                            int ny, int nz, hipStream_t s) {
    const int64_t C = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(update_velocity_kernel, cells_grid(C), dim3(256), 0, s, u, v, w, ru, rv, rw, ff, ax, ay, az,
                       dt, nudt, inv_dx2, nx, ny, nz);
}

__global__ __launch_bounds__(256) void pressure_rhs_kernel(const float *__restrict__ u, const float *__restrict__ v,
                                                           const float *__restrict__ w, const float *__restrict__ ff,
                                                           float *__restrict__ rhs, float rho_over_dt, float inv_dx,
                                                           int nx, int ny, int nz) {
    const int64_t index = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (index >= (int64_t)nx * ny * nz) return;
    int i, j, k;
    coords(index, ny, nz, i, j, k);
    if (!interior(i, j, k, nx, ny, nz) || ff[index] <= 0.0f) {
        rhs[index] = 0.0f;
        return;
    }
    const int64_t sx = (int64_t)ny * nz, sy = nz;
    const float du = u[index] - u[index - sx];
    const float dv = v[index] - v[index - sy];
    const float dw = w[index] - w[index - 1];
    rhs[index] = rho_over_dt * ((du + dv + dw) * inv_dx);
}

void launch_pressure_rhs(const float *u, const float *v, const float *w, const float *ff, float *rhs,
                         float rho_over_dt, float inv_dx, int nx, int ny, int nz, hipStream_t s) {
    const int64_t C = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(pressure_rhs_kernel, cells_grid(C), dim3(256), 0, s, u, v, w, ff, rhs, rho_over_dt, inv_dx,
                       nx, ny, nz);
}

__global__ __launch_bounds__(256) void build_finest_operator_kernel(const float *__restrict__ mask,
                                                                    float *__restrict__ diag, float *__restrict__ wx,
                                                                    float *__restrict__ wy, float *__restrict__ wz,
                                                                    int nx, int ny, int nz) {
    const int64_t index = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (index >= (int64_t)nx * ny * nz) return;
    int i, j, k;
    coords(index, ny, nz, i, j, k);
    float d = 0.0f, a = 0.0f, b = 0.0f, c = 0.0f;
    if (interior(i, j, k, nx, ny, nz) && mask[index] > 0.0f) {
        d = 6.0f - (float)(i == 1) - (float)(i == nx - 2) - (float)(j == 1) - (float)(j == ny - 2) -
            (float)(k == 1) - (float)(k == nz - 2);
        if (i < nx - 2) a = mask[index + (int64_t)ny * nz] > 0.0f ? 1.0f : 0.0f;
        if (j < ny - 2) b = mask[index + nz] > 0.0f ? 1.0f : 0.0f;
        if (k < nz - 2) c = mask[index + 1] > 0.0f ? 1.0f : 0.0f;
    }
    diag[index] = d;
    wx[index] = a;
    wy[index] = b;
    wz[index] = c;
}

// build_finest_operator_kernel by rows (block = row (i, j), thread = k), also
// writing the packed byte of each cell (bits 0-2: wx, wy, wz != 0, bit 3:
// mask > 0, bits 4-7: the diagonal, which is an integer 0..6 here) and the
// row flag (1 when the row holds an interior cell with mask > 0).
__global__ __launch_bounds__(256) void build_finest_operator_rows_kernel(
    const float *__restrict__ mask, float *__restrict__ diag, float *__restrict__ wx, float *__restrict__ wy,
    float *__restrict__ wz, uint8_t *__restrict__ op, int32_t *__restrict__ flags, int nx, int ny, int nz) {
    const int64_t row = blockIdx.x;
    const int i = (int)(row / ny), j = (int)(row - (int64_t)i * ny);
    const bool inner_row = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
    const int64_t sx = inner_row && i < nx - 2 ? (int64_t)ny * nz : 0, sy = inner_row && j < ny - 2 ? nz : 0;
    bool any = false;
    for (int k = threadIdx.x; k < nz; k += 256) {
        const int64_t index = row * nz + k;
        const bool inner = inner_row && k >= 1 && k <= nz - 2;
        const float m = mask[index], mx = mask[index + sx], my = mask[index + sy];
        const float mz = mask[index + (inner && k < nz - 2 ? 1 : 0)];
        const bool fluid = inner && m > 0.0f;
        int d = 0, a = 0, b = 0, c = 0;
        if (fluid) {
            d = 6 - (i == 1) - (i == nx - 2) - (j == 1) - (j == ny - 2) - (k == 1) - (k == nz - 2);
            a = i < nx - 2 && mx > 0.0f;
            b = j < ny - 2 && my > 0.0f;
            c = k < nz - 2 && mz > 0.0f;
        }
        any |= fluid;
        diag[index] = (float)d;
        wx[index] = (float)a;
        wy[index] = (float)b;
        wz[index] = (float)c;
        op[index] = (uint8_t)(a | (b << 1) | (c << 2) | (m > 0.0f ? 8 : 0) | (d << 4));
    }
    const int active = __syncthreads_or(any);
    if (threadIdx.x == 0) flags[row] = active;
}

// build_finest_operator_rows_kernel with four consecutive cells per lane
// (nz % 4 == 0: 16-byte loads and stores, the four packed bytes as one word)
// and one wave per row (the row flag by a ballot).
__global__ __launch_bounds__(256) void build_finest_operator_rows4_kernel(
    const float *__restrict__ mask, float *__restrict__ diag, float *__restrict__ wx, float *__restrict__ wy,
    float *__restrict__ wz, uint8_t *__restrict__ op, int32_t *__restrict__ flags, int nx, int ny, int nz) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= (int64_t)nx * ny) return;  // wave-uniform
    const int i = (int)(row / ny), j = (int)(row - (int64_t)i * ny);
    const bool inner_row = i >= 1 && i <= nx - 2 && j >= 1 && j <= ny - 2;
    const int64_t sx = inner_row && i < nx - 2 ? (int64_t)ny * nz : 0, sy = inner_row && j < ny - 2 ? nz : 0;
    const int dij = 6 - (i == 1) - (i == nx - 2) - (j == 1) - (j == ny - 2);
    bool any = false;
    for (int k0 = 4 * lane; k0 < nz; k0 += 256) {
        const int64_t base = row * nz + k0;
        const float4 m4 = *(const float4 *)(mask + base), mx4 = *(const float4 *)(mask + base + sx);
        const float4 my4 = *(const float4 *)(mask + base + sy);
        const float mn = mask[row * nz + min(k0 + 4, nz - 1)];
        const float m[5] = {m4.x, m4.y, m4.z, m4.w, mn}, mx[4] = {mx4.x, mx4.y, mx4.z, mx4.w},
                    my[4] = {my4.x, my4.y, my4.z, my4.w};
        float dv[4], av[4], bv[4], cv[4];
        uint32_t word = 0;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = k0 + u;
            const bool fluid = inner_row && k >= 1 && k <= nz - 2 && m[u] > 0.0f;
            int d = 0, a = 0, b = 0, c = 0;
            if (fluid) {
                d = dij - (k == 1) - (k == nz - 2);
                a = i < nx - 2 && mx[u] > 0.0f;
                b = j < ny - 2 && my[u] > 0.0f;
                c = k < nz - 2 && m[u + 1] > 0.0f;
            }
            any |= fluid;
            dv[u] = (float)d;
            av[u] = (float)a;
            bv[u] = (float)b;
            cv[u] = (float)c;
            word |= (uint32_t)(a | (b << 1) | (c << 2) | (m[u] > 0.0f ? 8 : 0) | (d << 4)) << (8 * u);
        }
        *(float4 *)(diag + base) = make_float4(dv[0], dv[1], dv[2], dv[3]);
        *(float4 *)(wx + base) = make_float4(av[0], av[1], av[2], av[3]);
        *(float4 *)(wy + base) = make_float4(bv[0], bv[1], bv[2], bv[3]);
        *(float4 *)(wz + base) = make_float4(cv[0], cv[1], cv[2], cv[3]);
        *(uint32_t *)(op + base) = word;
    }
    const bool active = __ballot(any) != 0;
    if (lane == 0) flags[row] = active ? 1 : 0;
}

void launch_build_finest_operator(const float *mask, float *diag, float *wx, float *wy, float *wz, uint8_t *op,
                                  int32_t *flags, int nx, int ny, int nz, hipStream_t s) {
    const int64_t C = (int64_t)nx * ny * nz;
    if (op && flags && nz % 4 == 0)
        hipLaunchKernelGGL(build_finest_operator_rows4_kernel, dim3((unsigned)(((int64_t)nx * ny + 3) / 4)),
                           dim3(256), 0, s, mask, diag, wx, wy, wz, op, flags, nx, ny, nz);
    else if (op && flags)
        hipLaunchKernelGGL(build_finest_operator_rows_kernel, dim3((unsigned)((int64_t)nx * ny)), dim3(256), 0, s,
                           mask, diag, wx, wy, wz, op, flags, nx, ny, nz);
    else
        hipLaunchKernelGGL(build_finest_operator_kernel, cells_grid(C), dim3(256), 0, s, mask, diag, wx, wy, wz, nx,
                           ny, nz);
}

// launched over the ghost layer (face_cell<2>): interior cells are only read
__global__ __launch_bounds__(256) void neumann_kernel(float *__restrict__ p, int nx, int ny, int nz) {
    int i, j, k;
    if (!face_cell<2>((int64_t)blockIdx.x * 256 + threadIdx.x, nx, ny, nz, i, j, k)) return;
    const int ii = min(max(i, 1), nx - 2), jj = min(max(j, 1), ny - 2), kk = min(max(k, 1), nz - 2);
    p[gidx(i, j, k, ny, nz)] = p[gidx(ii, jj, kk, ny, nz)];
}

void launch_neumann_bc(float *p, int nx, int ny, int nz, hipStream_t s) {
    hipLaunchKernelGGL(neumann_kernel, cells_grid(face_cells<2>(nx, ny, nz)), dim3(256), 0, s, p, nx, ny, nz);
}

__global__ __launch_bounds__(256) void pressure_correction_kernel(float *__restrict__ u, float *__restrict__ v,
                                                                  float *__restrict__ w, const float *__restrict__ p,
                                                                  float dt_over_rho, float inv_dx, int nx, int ny,
                                                                  int nz) {
    const int64_t index = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (index >= (int64_t)nx * ny * nz) return;
    int i, j, k;
    coords(index, ny, nz, i, j, k);
    if (!interior(i, j, k, nx, ny, nz)) return;
    const float centre = p[index];
    const float gx = (p[index + (int64_t)ny * nz] - centre) * inv_dx;
    const float gy = (p[index + nz] - centre) * inv_dx;
    const float gz = (p[index + 1] - centre) * inv_dx;
    u[index] = fmaf(-dt_over_rho, gx, u[index]);  // pressure_solve.cl:44-46 (fmuladd)
    v[index] = fmaf(-dt_over_rho, gy, v[index]);
    w[index] = fmaf(-dt_over_rho, gz, w[index]);
}

void launch_pressure_correction(float *u, float *v, float *w, const float *p, float dt_over_rho, float inv_dx, int nx,
                                int ny, int nz, hipStream_t s) {
    const int64_t C = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(pressure_correction_kernel, cells_grid(C), dim3(256), 0, s, u, v, w, p, dt_over_rho, inv_dx,
                       nx, ny, nz);
}

}  // namespace clq

namespace clq {

// Pressure correction from the FP64 pressure of the refined solves: the
// +1 neighbour index is clamped to the interior (the Neumann ghost value),
// the gradient and update run in FP64, u is rounded once.
__global__ __launch_bounds__(256) void pressure_correction64_kernel(float *__restrict__ u, float *__restrict__ v,
                                                                    float *__restrict__ w, const double *__restrict__ p,
                                                                    double dt_over_rho, double inv_dx, int nx, int ny,
                                                                    int nz) {
    const int64_t index = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (index >= (int64_t)nx * ny * nz) return;
    int i, j, k;
    coords(index, ny, nz, i, j, k);
    if (!interior(i, j, k, nx, ny, nz)) return;
    const double centre = p[index];
    const double gx = (p[gidx(min(i + 1, nx - 2), j, k, ny, nz)] - centre) * inv_dx;
    const double gy = (p[gidx(i, min(j + 1, ny - 2), k, ny, nz)] - centre) * inv_dx;
    const double gz = (p[gidx(i, j, min(k + 1, nz - 2), ny, nz)] - centre) * inv_dx;
    u[index] = (float)((double)u[index] - dt_over_rho * gx);
    v[index] = (float)((double)v[index] - dt_over_rho * gy);
    w[index] = (float)((double)w[index] - dt_over_rho * gz);
}

// pressure_correction64_kernel by rows, given the level-0 row flags (rows
// holding an interior fluid cell) of a pressure that is 0 at every other
// cell: a row whose flag and whose +x and +y neighbour rows' flags (indices
// clamped as the gradient clamps them) are all 0 sees only zero gradients,
// u - dt/rho * 0 == u, so its block returns at once.  Same values as the
// kernel over all cells.
__global__ __launch_bounds__(256) void pressure_correction64_rows_kernel(float *__restrict__ u, float *__restrict__ v,
                                                                         float *__restrict__ w,
                                                                         const double *__restrict__ p,
                                                                         const int32_t *__restrict__ flags,
                                                                         double dt_over_rho, double inv_dx, int nx,
                                                                         int ny, int nz) {
    const int64_t row = blockIdx.x;
    const int i = (int)(row / ny), j = (int)(row - (int64_t)i * ny);
    if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2) return;  // no interior cell in the row
    if (!flags[row] && !flags[(int64_t)min(i + 1, nx - 2) * ny + j] && !flags[row - j + min(j + 1, ny - 2)]) return;
    for (int k = 1 + threadIdx.x; k <= nz - 2; k += 256) {
        const int64_t index = row * nz + k;
        const double centre = p[index];
        const double gx = (p[gidx(min(i + 1, nx - 2), j, k, ny, nz)] - centre) * inv_dx;
        const double gy = (p[gidx(i, min(j + 1, ny - 2), k, ny, nz)] - centre) * inv_dx;
        const double gz = (p[gidx(i, j, min(k + 1, nz - 2), ny, nz)] - centre) * inv_dx;
        u[index] = (float)((double)u[index] - dt_over_rho * gx);
        v[index] = (float)((double)v[index] - dt_over_rho * gy);
        w[index] = (float)((double)w[index] - dt_over_rho * gz);
    }
}

void launch_pressure_correction64(float *u, float *v, float *w, const double *p, const int32_t *flags,
                                  double dt_over_rho, double inv_dx, int nx, int ny, int nz, hipStream_t s) {
    if (flags) {
        hipLaunchKernelGGL(pressure_correction64_rows_kernel, dim3((unsigned)((int64_t)nx * ny)), dim3(256), 0, s, u,
                           v, w, p, flags, dt_over_rho, inv_dx, nx, ny, nz);
        return;
    }
    const int64_t C = (int64_t)nx * ny * nz;
    hipLaunchKernelGGL(pressure_correction64_kernel, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, s, u, v, w, p,
                       dt_over_rho, inv_dx, nx, ny, nz);
}

}  // namespace clq
