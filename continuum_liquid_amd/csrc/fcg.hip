// Flexible conjugate gradients in FP64 on the level-0 pressure system, with a
// preconditioner applied by the caller (the mixed-precision W-cycle, or the
// learned network followed by it).  No reference counterpart: the reference
// runs a fixed number of stationary W-cycles (pressure_solve.py:61-83).
//
// Vectors over all C cells, FP64: x (iterate), r (residual), d (direction),
// q = A d.  The preconditioned residual z arrives as FP32 in the level-0
// correction array.  Two kernels per iteration after the preconditioner:
//   dots:   w = A z (not stored); z.q, z.r, z.w
//           -> beta = -(z.q) / (d.q)_old, and the new direction's
//              d.q = z.w + 2 beta z.q + beta^2 (d.q)_old (A symmetric, so
//              d_old . A z = z . q_old), alpha = z.r / d.q
//   update: w = A z again, d = z + beta d, q = w + beta q, x += alpha d,
//           r -= alpha q; r.r
// (the direction needs no reduction of its own, so it streams with the
// update).  Scalars live in device memory, so alpha and beta never
// round-trip through the host:
//   s[0] rz = z.r   s[1] d.q   s[2] alpha   s[3] z.q   s[4] z.w   s[5] beta
//   s[6] r.r        s[7] b.b
// beta = -(z_{k+1} . q_k) / (d_k . q_k), the flexible (Polak-Ribiere) form
// written with q_k = -(r_{k+1} - r_k) / alpha_k, so r_k need not be kept.
// The initialisation writes all cells (0 outside the interior fluid cells),
// or with the row lists the active rows (rows holding an interior fluid cell)
// and x = 0 on the inactive rows that may hold another value; the
// per-iteration kernels visit the active rows only (all cells when no row
// list is given): outside them x, r, z, d and q are zero or never read.
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

// A z at cell q = (i, j, k) (0 off the interior fluid cells), from the
// packed level-0 operator (PACKED) or the float arrays.  Every load is issued
// before any test: off the interior the neighbour indices fall back to q
// itself (in bounds), and the result is dropped.
template <bool PACKED>
__device__ __forceinline__ double apply_a_z(const Level &L, const uint8_t *__restrict__ op,
                                            const float *__restrict__ z, int64_t q, bool inner, double inv_h2,
                                            bool &fluid, double &dg) {
    const int64_t sx = inner ? (int64_t)L.ny * L.nz : 0, sy = inner ? L.nz : 0, sz = inner ? 1 : 0;
    const double z0 = z[q + sx], z1 = z[q - sx], z2 = z[q + sy], z3 = z[q - sy], z4 = z[q + sz], z5 = z[q - sz];
    const double zc = z[q];
    double s = 0.0;
    bool fl;
    if constexpr (PACKED) {
        const uint32_t o = op[q], ox = op[q - sx], oy = op[q - sy], oz = op[q - sz];
        fl = (o & 8) != 0;
        dg = (double)(o >> 4);
        if (o & 1) s = s + z0;
        if (ox & 1) s = s + z1;
        if (o & 2) s = s + z2;
        if (oy & 2) s = s + z3;
        if (o & 4) s = s + z4;
        if (oz & 4) s = s + z5;
    } else {
        const float w0 = L.wx[q], w1 = L.wx[q - sx], w2 = L.wy[q], w3 = L.wy[q - sy], w4 = L.wz[q], w5 = L.wz[q - sz];
        fl = L.mask[q] > 0.0f;
        dg = (double)L.diag[q];
        if (w0 != 0.0f) s = s + (double)w0 * z0;
        if (w1 != 0.0f) s = s + (double)w1 * z1;
        if (w2 != 0.0f) s = s + (double)w2 * z2;
        if (w3 != 0.0f) s = s + (double)w3 * z3;
        if (w4 != 0.0f) s = s + (double)w4 * z4;
        if (w5 != 0.0f) s = s + (double)w5 * z5;
    }
    fluid = inner && fl;
    return fluid ? (s - dg * zc) * inv_h2 : 0.0;
}

// The per-iteration kernels walk rows: block b takes a contiguous run of the
// work's rows (the listed active rows, or all nx * ny rows), runs laid out
// per XCD so the stencil's neighbour rows stay in that XCD's L2; a row's index
// and (i, j) are block-uniform, and thread t takes the cells k = t, t + kT, ...
template <typename F>
__device__ __forceinline__ void for_rows(const Level &L, const int32_t *__restrict__ rows, int64_t nrows, F cell) {
    const int64_t per = (nrows + gridDim.x - 1) / gridDim.x;
    const int64_t a0 = (int64_t)xcd_item(blockIdx.x, gridDim.x) * per, a1 = min(a0 + per, nrows);
    for (int64_t a = a0; a < a1; ++a) {
        const int64_t row = rows ? rows[a] : a;
        const int i = (int)(row / L.ny), j = (int)(row - (int64_t)i * L.ny);
        const bool inner_row = i >= 1 && i <= L.nx - 2 && j >= 1 && j <= L.ny - 2;
        for (int k = threadIdx.x; k < L.nz; k += kT) cell(row * L.nz + k, inner_row && k >= 1 && k <= L.nz - 2, i, j, k);
    }
}

// fcg_init_kernel over the listed active rows (rows holding an interior
// fluid cell): the same values, A x from the packed operator (PACKED) or the
// float arrays, every load issued before the tests.
template <bool PACKED>
__global__ __launch_bounds__(kT) void fcg_init_rows_kernel(Level L, const uint8_t *__restrict__ op,
                                                          const float *__restrict__ p32, const float *__restrict__ b,
                                                          double inv_h2, const int32_t *__restrict__ rows,
                                                          int64_t nrows, double *__restrict__ x,
                                                          double *__restrict__ r, float *__restrict__ r32,
                                                          double *__restrict__ partials) {
    double acc[2] = {0.0, 0.0};
    for_rows(L, rows, nrows, [&](int64_t q, bool inner, int, int, int) {
        const double bv = (double)b[q], pv = (double)p32[q];
        bool fluid;
        double dg;
        const double av = apply_a_z<PACKED>(L, op, p32, q, inner, inv_h2, fluid, dg);
        const double rv = fluid ? bv - av : 0.0;
        x[q] = fluid ? pv : 0.0;
        r[q] = rv;
        r32[q] = (float)rv;
        acc[0] += rv * rv;
        acc[1] += bv * bv;
    });
    block_sums<2>(acc, partials);
}

// x = 0 on the inactive level-0 rows that may hold a non-zero value (x_rows
// set, or every inactive row when !known); x_rows = flags afterwards.  On
// those rows the pressure correction reads x, while r, r32, b are read there
// only under zero weights or a fluid select.  Block b tests rows
// [kT b, kT b + kT), one per thread, lists the rows to clear in LDS, and
// clears them together (thread = k).
__global__ __launch_bounds__(kT) void fcg_clear_rows_kernel(const int32_t *__restrict__ flags,
                                                           int32_t *__restrict__ x_rows, int known, int64_t nrows,
                                                           int nz, double *__restrict__ x) {
    __shared__ int32_t list[kT];
    __shared__ int count;
    if (threadIdx.x == 0) count = 0;
    __syncthreads();
    const int64_t a = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (a < nrows) {
        const int32_t f = flags[a];
        if (f == 0 && (!known || x_rows[a] != 0)) list[atomicAdd(&count, 1)] = threadIdx.x;
        x_rows[a] = f;
    }
    __syncthreads();
    const int n = count;
    for (int t = 0; t < n; ++t) {
        const int64_t row = (int64_t)blockIdx.x * kT + list[t];
        for (int k = threadIdx.x; k < nz; k += kT) x[row * nz + k] = 0.0;
    }
}

// w = A z (not stored); partial z.q, z.r, z.w
template <bool PACKED>
__global__ __launch_bounds__(kT) void fcg_dots_kernel(Level L, const uint8_t *__restrict__ op,
                                                     const float *__restrict__ z, const double *__restrict__ qv,
                                                     const double *__restrict__ r, double inv_h2,
                                                     const int32_t *__restrict__ rows, int64_t nrows,
                                                     double *__restrict__ partials) {
    double acc[3] = {0.0, 0.0, 0.0};
    for_rows(L, rows, nrows, [&](int64_t q, bool inner, int, int, int) {
        const double qo = qv[q], ro = r[q], zq = (double)z[q];
        bool fluid;
        double dg;
        const double wv = apply_a_z<PACKED>(L, op, z, q, inner, inv_h2, fluid, dg);
        acc[0] += zq * qo;
        acc[1] += zq * ro;
        acc[2] += zq * wv;
    });
    block_sums<3>(acc, partials);
}

// d = z + beta d, q = A z + beta q on the interior fluid cells (0 elsewhere;
// beta = 0 on the first iteration), then x += alpha d, r -= alpha q,
// r32 = float(r); partial r.r.  FOLD (packed operator): also the correction
// array e of the next preconditioner cycle after its first red half-sweep
// from e = 0 on r32: e = fmaf(-dd, r32, +0) / diag on the red interior fluid
// cells with diag > 0 (every neighbour term w * 0 is +0), 0 elsewhere, as
// zero_rows and rbgs_rows_packed_kernel would leave it.  e must not be z (z
// is read at the neighbours).
template <bool PACKED, bool FOLD>
__global__ __launch_bounds__(kT) void fcg_update_kernel(Level L, const uint8_t *__restrict__ op,
                                                       const float *__restrict__ z, double *__restrict__ x,
                                                       double *__restrict__ r, float *__restrict__ r32,
                                                       double *__restrict__ d, double *__restrict__ qv,
                                                       const double *__restrict__ s, int first, double inv_h2,
                                                       const int32_t *__restrict__ rows, int64_t nrows,
                                                       float *__restrict__ e, float dd, double *__restrict__ partials) {
    const double alpha = s[2], beta = first ? 0.0 : s[5];
    double acc[1] = {0.0};
    for_rows(L, rows, nrows, [&](int64_t q, bool inner, int i, int j, int k) {
        const double xo = x[q], ro = r[q], dold = d[q], qold = qv[q], zv = (double)z[q];
        bool fluid;
        double dg;
        const double wv = apply_a_z<PACKED>(L, op, z, q, inner, inv_h2, fluid, dg);
        const double dv = fluid ? zv + beta * dold : 0.0;
        const double qn = fluid ? wv + beta * qold : 0.0;
        d[q] = dv;
        qv[q] = qn;
        const double rv = ro - alpha * qn;
        const float rf = (float)rv;
        x[q] = xo + alpha * dv;
        r[q] = rv;
        r32[q] = rf;
        acc[0] += rv * rv;
        if constexpr (FOLD) {  // dg: the packed diagonal, a small integer
            float ev = 0.0f;
            if (((i + j + k) & 1) == 0 && fluid && dg > 0.0)
                ev = fmaf(-dd, rf, 0.0f) / (float)dg;  // multigrid.cl:227 on e = 0
            e[q] = ev;
        }
    });
    block_sums<1>(acc, partials);
}

// Reduce `n` partial arrays of `blocks` entries and update the scalars.
// stage 0: init (s6 = r.r, s7 = b.b); 2: r.r; 3: z.q, z.r, z.w -> beta,
// d.q, alpha (first: beta = 0).
constexpr int kScalarThreads = 1024;

__global__ __launch_bounds__(kScalarThreads) void fcg_scalars_kernel(const double *__restrict__ partials, int blocks,
                                                                    int stage, int first, double *__restrict__ s) {
    // each lane sums its partials (all loads in flight together), a
    // butterfly sums each wave, and wave 0 adds the 16 wave sums
    __shared__ double wsum[3][kScalarThreads / 64];
    const int n = stage == 0 ? 2 : stage == 3 ? 3 : 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double v[3] = {0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < blocks; q += kScalarThreads)
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (a < n) v[a] += partials[a * kB + q];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[a] += __shfl_xor(v[a], off, 64);
        if (lane == 0) wsum[a][wave] = v[a];
    }
    __syncthreads();
    if (wave != 0) return;
    double sh[3][1];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        double t = lane < kScalarThreads / 64 ? wsum[a][lane] : 0.0;
#pragma unroll
        for (int off = 8; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        sh[a][0] = t;
    }
    if (threadIdx.x != 0) return;
    if (stage == 0) {
        s[6] = sh[0][0];
        s[7] = sh[1][0];
    } else if (stage == 2) {
        s[6] = sh[0][0];
    } else {
        const double zq = sh[0][0], zr = sh[1][0], zw = sh[2][0], dq_old = s[1];
        const double beta = (first || dq_old == 0.0) ? 0.0 : -zq / dq_old;
        const double dq = first ? zw : zw + 2.0 * beta * zq + beta * beta * dq_old;
        s[0] = zr;
        s[1] = dq;
        s[2] = dq != 0.0 ? zr / dq : 0.0;
        s[3] = zq;
        s[4] = zw;
        s[5] = beta;
    }
}

static int blocks_for(int64_t C) {
    const int64_t b = (C + kT - 1) / kT;
    return (int)(b < kB ? b : kB);
}

int fcg_partials_count() { return 3 * kB; }

// rows of the per-iteration kernels' work, and their block count
static int64_t rows_of(const Level &L, const int32_t *rows, int64_t nrows) {
    return rows ? nrows : (int64_t)L.nx * L.ny;
}
static int row_blocks_for(int64_t nr) { return (int)(nr < kB ? (nr > 0 ? nr : 1) : kB); }

void launch_fcg_init(const Level &L, const uint8_t *op, const float *p32, const float *b, double inv_h2, double *x,
                     double *r, float *r32, const int32_t *flags, const int32_t *rows, int64_t nrows,
                     int32_t *x_rows, int x_rows_known, double *partials, double *s, hipStream_t st) {
    int nb;
    if (flags) {
        const int64_t all = (int64_t)L.nx * L.ny;
        hipLaunchKernelGGL(fcg_clear_rows_kernel, dim3((unsigned)((all + kT - 1) / kT)), dim3(kT), 0, st, flags,
                           x_rows, x_rows_known, all, L.nz, x);
        nb = row_blocks_for(nrows);
        if (op)
            hipLaunchKernelGGL(fcg_init_rows_kernel<true>, dim3(nb), dim3(kT), 0, st, L, op, p32, b, inv_h2, rows,
                               nrows, x, r, r32, partials);
        else
            hipLaunchKernelGGL(fcg_init_rows_kernel<false>, dim3(nb), dim3(kT), 0, st, L, op, p32, b, inv_h2, rows,
                               nrows, x, r, r32, partials);
    } else {
        nb = blocks_for(L.C);
        hipLaunchKernelGGL(fcg_init_kernel, dim3(nb), dim3(kT), 0, st, L, p32, b, inv_h2, x, r, r32, partials);
    }
    hipLaunchKernelGGL(fcg_scalars_kernel, dim3(1), dim3(kScalarThreads), 0, st, partials, nb, 0, 0, s);
}


void launch_fcg_dots(const Level &L, const uint8_t *op, const float *z, const double *qv, const double *r, int first,
                     double inv_h2, const int32_t *rows, int64_t nrows, double *partials, double *s, hipStream_t st) {
    const int64_t nr = rows_of(L, rows, nrows);
    const int nb = row_blocks_for(nr);
    if (op)
        hipLaunchKernelGGL(fcg_dots_kernel<true>, dim3(nb), dim3(kT), 0, st, L, op, z, qv, r, inv_h2, rows, nr,
                           partials);
    else
        hipLaunchKernelGGL(fcg_dots_kernel<false>, dim3(nb), dim3(kT), 0, st, L, op, z, qv, r, inv_h2, rows, nr,
                           partials);
    hipLaunchKernelGGL(fcg_scalars_kernel, dim3(1), dim3(kScalarThreads), 0, st, partials, nb, 3, first, s);
}

void launch_fcg_update(const Level &L, const uint8_t *op, const float *z, double *x, double *r, float *r32, double *d,
                       double *qv, int first, double inv_h2, const int32_t *rows, int64_t nrows, float *e, float delta,
                       double *partials, double *s, hipStream_t st) {
    const int64_t nr = rows_of(L, rows, nrows);
    const int nb = row_blocks_for(nr);
    const float dd = delta * delta;
    if (op && e)
        hipLaunchKernelGGL((fcg_update_kernel<true, true>), dim3(nb), dim3(kT), 0, st, L, op, z, x, r, r32, d, qv, s,
                           first, inv_h2, rows, nr, e, dd, partials);
    else if (op)
        hipLaunchKernelGGL((fcg_update_kernel<true, false>), dim3(nb), dim3(kT), 0, st, L, op, z, x, r, r32, d, qv, s,
                           first, inv_h2, rows, nr, e, dd, partials);
    else
        hipLaunchKernelGGL((fcg_update_kernel<false, false>), dim3(nb), dim3(kT), 0, st, L, op, z, x, r, r32, d, qv,
                           s, first, inv_h2, rows, nr, e, dd, partials);
    hipLaunchKernelGGL(fcg_scalars_kernel, dim3(1), dim3(kScalarThreads), 0, st, partials, nb, 2, 0, s);
}

}  // namespace clq
