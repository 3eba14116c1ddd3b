// Flexible conjugate gradients in FP64 on the level-0 pressure system, with a
// preconditioner applied by the caller (the mixed-precision W-cycle, or the
// learned network followed by it).  No reference counterpart: the reference
// runs a fixed number of stationary W-cycles (pressure_solve.py:61-83).
//
// Vectors over all C cells, FP64: x (iterate), r (residual), d (direction),
// q = A d, w = A z.  The preconditioned residual z arrives as FP32 in the
// level-0 correction array e.  The matrix-vector product is taken of z, right
// after the preconditioner and before beta is known, so the direction update
// streams: d = z + beta d, q = A z + beta q (= A d).  Scalars live in device
// memory so that alpha and beta never round-trip through the host:
//   s[0] rz = r.z   s[1] d.q   s[2] alpha   s[3] z.q   s[4] z.r   s[5] beta
//   s[6] r.r        s[7] b.b
// beta = -alpha (z_{k+1} . q_k) / (z_k . r_k), the flexible (Polak-Ribiere)
// form written with q_k = -(r_{k+1} - r_k) / alpha, so r_k need not be kept.
// The initialisation writes all cells (0 outside the interior fluid cells);
// the per-iteration kernels visit the active rows only (rows holding an
// interior fluid cell, or all cells when no row list is given): outside them
// x, r, z, d and q are zero or never read.
#include "common.hpp"

namespace clq {

constexpr int kT = 256;
constexpr int kB = 2048;

__device__ __forceinline__ bool fluid_cell(const Level &L, int64_t q) {
    int i, j, k;
    coords(q, L.ny, L.nz, i, j, k);
    return interior(i, j, k, L.nx, L.ny, L.nz) && L.mask[q] > 0.0f;
}

// A v at an interior cell, FP64, left-to-right over the non-zero weights (the
// order of the refinement residual).  Every load is issued before the weight
// tests (all seven cells are in bounds at an interior cell).
template <typename F>
__device__ __forceinline__ double apply_a(const Level &L, int64_t q, F v, double inv_h2) {
    const int64_t sx = (int64_t)L.ny * L.nz, sy = L.nz;
    const float w0 = L.wx[q], w1 = L.wx[q - sx], w2 = L.wy[q], w3 = L.wy[q - sy], w4 = L.wz[q], w5 = L.wz[q - 1];
    const double v0 = v(q + sx), v1 = v(q - sx), v2 = v(q + sy), v3 = v(q - sy), v4 = v(q + 1), v5 = v(q - 1);
    const double vc = v(q), d = (double)L.diag[q];
    double s = 0.0;
    if (w0 != 0.0f) s = s + (double)w0 * v0;
    if (w1 != 0.0f) s = s + (double)w1 * v1;
    if (w2 != 0.0f) s = s + (double)w2 * v2;
    if (w3 != 0.0f) s = s + (double)w3 * v3;
    if (w4 != 0.0f) s = s + (double)w4 * v4;
    if (w5 != 0.0f) s = s + (double)w5 * v5;
    return (s - d * vc) * inv_h2;
}

// cell of work item t: row-major over the listed rows, or t itself
__device__ __forceinline__ int64_t cell_of(const int32_t *__restrict__ rows, int nz, int64_t t) {
    if (!rows) return t;
    const int64_t ar = t / nz;
    return (int64_t)rows[ar] * nz + (t - ar * nz);
}

// apply_a with the operator decoded from the packed level-0 bytes
template <typename F>
__device__ __forceinline__ double apply_a_packed(const Level &L, const uint8_t *__restrict__ op, int64_t q, F v,
                                                 double inv_h2) {
    const int64_t sx = (int64_t)L.ny * L.nz, sy = L.nz;
    const uint32_t o = op[q], ox = op[q - sx], oy = op[q - sy], oz = op[q - 1];
    const double v0 = v(q + sx), v1 = v(q - sx), v2 = v(q + sy), v3 = v(q - sy), v4 = v(q + 1), v5 = v(q - 1);
    const double vc = v(q), d = (double)(o >> 4);
    double s = 0.0;
    if (o & 1) s = s + v0;
    if (ox & 1) s = s + v1;
    if (o & 2) s = s + v2;
    if (oy & 2) s = s + v3;
    if (o & 4) s = s + v4;
    if (oz & 4) s = s + v5;
    return (s - d * vc) * inv_h2;
}

template <int N>
__device__ __forceinline__ void block_sums(double (&v)[N], double *__restrict__ partials) {
    __shared__ double sh[N][kT];
#pragma unroll
    for (int a = 0; a < N; ++a) sh[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
#pragma unroll
            for (int a = 0; a < N; ++a) sh[a][threadIdx.x] += sh[a][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0)
#pragma unroll
        for (int a = 0; a < N; ++a) partials[a * kB + blockIdx.x] = sh[a][0];
}

// x = p32 on fluid cells, r = b - A x, r32 = float(r); partial r.r, b.b.
__global__ __launch_bounds__(kT) void fcg_init_kernel(Level L, const float *__restrict__ p32,
                                                     const float *__restrict__ b, double inv_h2, double *__restrict__ x,
                                                     double *__restrict__ r, float *__restrict__ r32,
                                                     double *__restrict__ partials) {
    double acc[2] = {0.0, 0.0};
    for (int64_t q = (int64_t)blockIdx.x * kT + threadIdx.x; q < L.C; q += (int64_t)gridDim.x * kT) {
        double xv = 0.0, rv = 0.0;
        if (fluid_cell(L, q)) {
            xv = (double)p32[q];
            rv = (double)b[q] - apply_a(L, q, [&](int64_t t) { return (double)p32[t]; }, inv_h2);
        }
        x[q] = xv;
        r[q] = rv;
        r32[q] = (float)rv;
        acc[0] += rv * rv;
        acc[1] += (double)b[q] * (double)b[q];
    }
    block_sums<2>(acc, partials);
}

// d = z + beta d, q = w + beta q on the interior fluid cells (0 elsewhere);
// partial d.q.  first: beta = 0.
__global__ __launch_bounds__(kT) void fcg_direction_kernel(Level L, const uint8_t *__restrict__ op,
                                                          const float *__restrict__ z, double *__restrict__ d,
                                                          const double *__restrict__ w, double *__restrict__ qv,
                                                          const double *__restrict__ s, int first,
                                                          const int32_t *__restrict__ rows, int64_t work,
                                                          double *__restrict__ partials) {
    const double beta = first ? 0.0 : s[5];
    double acc[1] = {0.0};
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < work; t += (int64_t)gridDim.x * kT) {
        const int64_t q = cell_of(rows, L.nz, t);
        int i, j, k;
        coords(q, L.ny, L.nz, i, j, k);
        const bool fluid = interior(i, j, k, L.nx, L.ny, L.nz) && (op ? (op[q] & 8) != 0 : L.mask[q] > 0.0f);
        const double zv = (double)z[q], dold = d[q], wv = w[q], qold = qv[q];
        double dv = 0.0, qn = 0.0;
        if (fluid) {
            dv = first ? zv : zv + beta * dold;
            qn = first ? wv : wv + beta * qold;
        }
        d[q] = dv;
        qv[q] = qn;
        acc[0] += dv * qn;
    }
    block_sums<1>(acc, partials);
}

// x += alpha d, r -= alpha q, r32 = float(r); partial r.r.  With e (the
// packed level-0 operator op given), also the correction array of the next
// preconditioner cycle after its first red half-sweep from e = 0 on r32:
// e = fmaf(-dd, r32, +0) / diag on the red interior fluid cells with diag > 0
// (every neighbour term w * 0 is +0), 0 elsewhere, as zero_rows and
// rbgs_rows_packed_kernel would leave it.
__global__ __launch_bounds__(kT) void fcg_update_kernel(Level L, double *__restrict__ x, double *__restrict__ r,
                                                       float *__restrict__ r32, const double *__restrict__ d,
                                                       const double *__restrict__ qv, const double *__restrict__ s,
                                                       const int32_t *__restrict__ rows, int64_t work,
                                                       const uint8_t *__restrict__ op, float *__restrict__ e, float dd,
                                                       double *__restrict__ partials) {
    const double alpha = s[2];
    double acc[1] = {0.0};
    for (int64_t t = (int64_t)blockIdx.x * kT + threadIdx.x; t < work; t += (int64_t)gridDim.x * kT) {
        const int64_t q = cell_of(rows, L.nz, t);
        const double xv = x[q] + alpha * d[q];
        const double rv = r[q] - alpha * qv[q];
        const float rf = (float)rv;
        x[q] = xv;
        r[q] = rv;
        r32[q] = rf;
        acc[0] += rv * rv;
        if (e) {
            int i, j, k;
            coords(q, L.ny, L.nz, i, j, k);
            const uint32_t o = op[q];
            const float dg = (float)(o >> 4);
            float ev = 0.0f;
            if (((i + j + k) & 1) == 0 && (o & 8) && interior(i, j, k, L.nx, L.ny, L.nz) && dg > 0.0f)
                ev = fmaf(-dd, rf, 0.0f) / dg;  // multigrid.cl:227 on e = 0
            e[q] = ev;
        }
    }
    block_sums<1>(acc, partials);
}

// w = A z on the interior fluid cells (0 elsewhere); partial z.q, z.r.  One
// contiguous chunk of work items per block, chunks laid out per XCD (the
// stencil's x-neighbour rows stay in that XCD's L2).
__global__ __launch_bounds__(kT) void fcg_dots_kernel(Level L, const uint8_t *__restrict__ op,
                                                     const float *__restrict__ z, const double *__restrict__ qv,
                                                     const double *__restrict__ r, double *__restrict__ w,
                                                     double inv_h2, const int32_t *__restrict__ rows, int64_t work,
                                                     double *__restrict__ partials) {
    double acc[2] = {0.0, 0.0};
    const int64_t chunk = (work + (int64_t)gridDim.x * kT - 1) / ((int64_t)gridDim.x * kT) * kT;
    const int64_t t0 = (int64_t)xcd_item(blockIdx.x, gridDim.x) * chunk, t1 = min(t0 + chunk, work);
    for (int64_t t = t0 + threadIdx.x; t < t1; t += kT) {
        const int64_t q = cell_of(rows, L.nz, t);
        double wv = 0.0;
        int i, j, k;
        coords(q, L.ny, L.nz, i, j, k);
        if (interior(i, j, k, L.nx, L.ny, L.nz)) {
            // neighbours outside the fluid carry zero weight: their values are
            // loaded with the rest but never enter the sum
            auto zv = [&](int64_t u) { return (double)z[u]; };
            const bool fluid = op ? (op[q] & 8) != 0 : L.mask[q] > 0.0f;
            const double aq = op ? apply_a_packed(L, op, q, zv, inv_h2) : apply_a(L, q, zv, inv_h2);
            if (fluid) wv = aq;
        }
        w[q] = wv;
        const double zq = (double)z[q];
        acc[0] += zq * qv[q];
        acc[1] += zq * r[q];
    }
    block_sums<2>(acc, partials);
}

// Reduce `n` partial arrays of `blocks` entries and update the scalars.
// stage 0: init (s6 = r.r, s7 = b.b); 1: d.q -> alpha; 2: r.r;
// 3: z.q, z.r -> beta (unless first), rz.
__global__ __launch_bounds__(kT) void fcg_scalars_kernel(const double *__restrict__ partials, int blocks, int stage,
                                                        int first, double *__restrict__ s) {
    __shared__ double sh[2][kT];
    const int n = (stage == 0 || stage == 3) ? 2 : 1;
    for (int a = 0; a < n; ++a) {
        double v = 0.0;
        for (int q = threadIdx.x; q < blocks; q += kT) v += partials[a * kB + q];
        sh[a][threadIdx.x] = v;
    }
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int a = 0; a < n; ++a) sh[a][threadIdx.x] += sh[a][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (stage == 0) {
        s[6] = sh[0][0];
        s[7] = sh[1][0];
    } else if (stage == 1) {
        s[1] = sh[0][0];
        s[2] = s[1] != 0.0 ? s[0] / s[1] : 0.0;
    } else if (stage == 2) {
        s[6] = sh[0][0];
    } else {
        s[3] = sh[0][0];
        s[4] = sh[1][0];
        s[5] = (first || s[0] == 0.0) ? 0.0 : -s[2] * s[3] / s[0];
        s[0] = s[4];
    }
}

static int blocks_for(int64_t C) {
    const int64_t b = (C + kT - 1) / kT;
    return (int)(b < kB ? b : kB);
}

int fcg_partials_count() { return 2 * kB; }

void launch_fcg_init(const Level &L, const float *p32, const float *b, double inv_h2, double *x, double *r, float *r32,
                     double *partials, double *s, hipStream_t st) {
    const int nb = blocks_for(L.C);
    hipLaunchKernelGGL(fcg_init_kernel, dim3(nb), dim3(kT), 0, st, L, p32, b, inv_h2, x, r, r32, partials);
    hipLaunchKernelGGL(fcg_scalars_kernel, dim3(1), dim3(kT), 0, st, partials, nb, 0, 0, s);
}

static int64_t work_of(const Level &L, const int32_t *rows, int64_t nrows) { return rows ? nrows * L.nz : L.C; }

void launch_fcg_dots(const Level &L, const uint8_t *op, const float *z, const double *qv, const double *r, double *w,
                     int first, double inv_h2, const int32_t *rows, int64_t nrows, double *partials, double *s,
                     hipStream_t st) {
    const int64_t work = work_of(L, rows, nrows);
    const int nb = blocks_for(work);
    hipLaunchKernelGGL(fcg_dots_kernel, dim3(nb), dim3(kT), 0, st, L, op, z, qv, r, w, inv_h2, rows, work, partials);
    hipLaunchKernelGGL(fcg_scalars_kernel, dim3(1), dim3(kT), 0, st, partials, nb, 3, first, s);
}

void launch_fcg_direction(const Level &L, const uint8_t *op, const float *z, double *d, const double *w, double *qv,
                          int first, const int32_t *rows, int64_t nrows, double *partials, double *s, hipStream_t st) {
    const int64_t work = work_of(L, rows, nrows);
    const int nb = blocks_for(work);
    hipLaunchKernelGGL(fcg_direction_kernel, dim3(nb), dim3(kT), 0, st, L, op, z, d, w, qv, s, first, rows, work,
                       partials);
    hipLaunchKernelGGL(fcg_scalars_kernel, dim3(1), dim3(kT), 0, st, partials, nb, 1, 0, s);
}

void launch_fcg_update(const Level &L, double *x, double *r, float *r32, const double *d, const double *qv,
                       const int32_t *rows, int64_t nrows, const uint8_t *op, float *e, float delta, double *partials,
                       double *s, hipStream_t st) {
    const int64_t work = work_of(L, rows, nrows);
    const int nb = blocks_for(work);
    hipLaunchKernelGGL(fcg_update_kernel, dim3(nb), dim3(kT), 0, st, L, x, r, r32, d, qv, s, rows, work, op, e,
                       delta * delta, partials);
    hipLaunchKernelGGL(fcg_scalars_kernel, dim3(1), dim3(kT), 0, st, partials, nb, 2, 0, s);
}

}  // namespace clq
