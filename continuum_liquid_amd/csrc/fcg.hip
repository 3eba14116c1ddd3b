// Flexible conjugate gradients in FP64 on the level-0 pressure system, with a
// preconditioner applied by the caller (the mixed-precision W-cycle, or the
// learned network followed by it).  No reference counterpart: the reference
// runs a fixed number of stationary W-cycles (pressure_solve.py:61-83).
//
// Vectors over all C cells: x (iterate) and r (residual) in FP64, the
// direction d in FP32 (two arrays, read one, write the other).  The
// preconditioned residual z arrives as FP32 in the level-0 correction array.
// Two kernels per iteration after the preconditioner:
//   dots:   w = A z (not stored); (A z).d_old = z.(A d_old), z.r, z.w
//           -> beta = -(z.A d_old) / (d.A d)_old, and the new direction's
//              d.A d = z.w + 2 beta z.A d_old + beta^2 (d.A d)_old (A
//              symmetric), alpha = z.r / d.A d
//   update: d = fmaf(float(beta), d_old, z) at the cell and its six
//           neighbours, q = A d in FP64 from those values, x += alpha d,
//           r -= alpha q; r.r
// q is formed from the stored (rounded) direction itself, never carried by a
// recurrence, so x and r move by alpha d and alpha A d of the same d and the
// recurrence residual tracks b - A x to FP64 rounding; rounding d to FP32
// only perturbs the direction, not that consistency.  (The direction needs
// no reduction of its own, so it streams with the update.)  Scalars live in
// device memory, so alpha and beta never round-trip through the host:
//   s[0] rz = z.r   s[1] d.A d   s[2] alpha   s[3] z.A d_old   s[4] z.w
//   s[5] beta       s[6] r.r     s[7] b.b
// beta = -(z_{k+1} . A d_k) / (d_k . A d_k), the flexible (Polak-Ribiere)
// form written with A d_k = -(r_{k+1} - r_k) / alpha_k, so r_k need not be
// kept.  The initialisation writes all cells (0 outside the interior fluid
// cells), or with the row lists the active rows (rows holding an interior
// fluid cell) and x = 0 on the inactive rows that may hold another value;
// the per-iteration kernels visit the active rows only (all cells when no
// row list is given): outside them x, r, z and d are zero or never read.
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

// A v at cell q = (i, j, k) (0 off the interior fluid cells) for the vector
// whose value at cell t is v(t), from the packed level-0 operator (PACKED) or
// the float arrays.  Every load is issued before any test: off the interior
// the neighbour indices fall back to q itself (in bounds), and the result is
// dropped.
// The neighbour values arrive as z0..z5 (+x, -x, +y, -y, +z, -z) and zc.
template <bool PACKED>
__device__ __forceinline__ double apply_a_vals(const Level &L, const uint8_t *__restrict__ op, int64_t q, bool inner,
                                               double inv_h2, bool &fluid, double &dg, double z0, double z1,
                                               double z2, double z3, double z4, double z5, double zc) {
    const int64_t sx = inner ? (int64_t)L.ny * L.nz : 0, sy = inner ? L.nz : 0, sz = inner ? 1 : 0;
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

template <bool PACKED, typename F>
__device__ __forceinline__ double apply_a_fn(const Level &L, const uint8_t *__restrict__ op, int64_t q, bool inner,
                                             double inv_h2, bool &fluid, double &dg, F v) {
    const int64_t sx = inner ? (int64_t)L.ny * L.nz : 0, sy = inner ? L.nz : 0, sz = inner ? 1 : 0;
    const double z0 = v(q + sx), z1 = v(q - sx), z2 = v(q + sy), z3 = v(q - sy), z4 = v(q + sz), z5 = v(q - sz);
    const double zc = v(q);
    return apply_a_vals<PACKED>(L, op, q, inner, inv_h2, fluid, dg, z0, z1, z2, z3, z4, z5, zc);
}

// apply_a_fn of an array
template <bool PACKED, typename T>
__device__ __forceinline__ double apply_a_z(const Level &L, const uint8_t *__restrict__ op,
                                            const T *__restrict__ z, int64_t q, bool inner, double inv_h2,
                                            bool &fluid, double &dg) {
    return apply_a_fn<PACKED>(L, op, q, inner, inv_h2, fluid, dg, [&](int64_t t) { return (double)z[t]; });
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
        const double av = apply_a_z<PACKED, float>(L, op, p32, q, inner, inv_h2, fluid, dg);
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

// w = A z (not stored); partial (A z).d_old (= z.(A d_old), A symmetric),
// z.r, z.w
template <bool PACKED>
__global__ __launch_bounds__(kT) void fcg_dots_kernel(Level L, const uint8_t *__restrict__ op,
                                                     const float *__restrict__ z, const float *__restrict__ dold,
                                                     const double *__restrict__ r, double inv_h2,
                                                     const int32_t *__restrict__ rows, int64_t nrows,
                                                     double *__restrict__ partials) {
    double acc[3] = {0.0, 0.0, 0.0};
    for_rows(L, rows, nrows, [&](int64_t q, bool inner, int, int, int) {
        const double dq = (double)dold[q], ro = r[q], zq = (double)z[q];
        bool fluid;
        double dg;
        const double wv = apply_a_z<PACKED, float>(L, op, z, q, inner, inv_h2, fluid, dg);
        acc[0] += wv * dq;
        acc[1] += zq * ro;
        acc[2] += zq * wv;
    });
    block_sums<3>(acc, partials);
}

// d = fmaf(float(beta), d_old, z) on the interior fluid cells (0 elsewhere;
// d = z on the first iteration), q = A d in FP64 from the same values at the
// neighbours (computed here from z and d_old, the z-neighbours' taken from
// the adjacent lanes, which hold cells k -/+ 1 of the same row; at a
// neighbour that is not fluid the operator weight is 0 and the value
// unused), then x += alpha d,
// r -= alpha q, r32 = float(r); partial r.r.  FOLD (packed operator): also
// the correction array e of the next preconditioner cycle after its first
// red half-sweep from e = 0 on r32: e = fmaf(-dd, r32, +0) / diag on the red
// interior fluid cells with diag > 0 (every neighbour term w * 0 is +0), 0
// elsewhere, as zero_rows and rbgs_rows_packed_kernel would leave it.  e must
// not be z (z is read at the neighbours); dnew must not be dold.
template <bool PACKED, bool FOLD>
__global__ __launch_bounds__(kT) void fcg_update_kernel(Level L, const uint8_t *__restrict__ op,
                                                       const float *__restrict__ z, double *__restrict__ x,
                                                       double *__restrict__ r, float *__restrict__ r32,
                                                       const float *__restrict__ dold, float *__restrict__ dnew,
                                                       const double *__restrict__ s, int first, double inv_h2,
                                                       const int32_t *__restrict__ rows, int64_t nrows,
                                                       float *__restrict__ e, float dd, double *__restrict__ partials) {
    const double alpha = s[2];
    const float beta = first ? 0.0f : (float)s[5];
    double acc[1] = {0.0};
    // the new direction at cell t, as the cell's own thread stores it
    auto dir = [&](int64_t t) -> float { return first ? z[t] : fmaf(beta, dold[t], z[t]); };
    const int lane = threadIdx.x & 63;
    for_rows(L, rows, nrows, [&](int64_t q, bool inner, int i, int j, int k) {
        const double xo = x[q], ro = r[q];
        const int64_t sx = inner ? (int64_t)L.ny * L.nz : 0, sy = inner ? L.nz : 0, sz = inner ? 1 : 0;
        const float dc = dir(q), dxp = dir(q + sx), dxm = dir(q - sx), dyp = dir(q + sy), dym = dir(q - sy);
        // cells k + 1 and k - 1: the next and previous lanes (a wave's first
        // and last lanes read theirs); lanes off the row only feed ghost
        // cells, whose result is dropped
        float dzp = __shfl_down(dc, 1), dzm = __shfl_up(dc, 1);
        if (lane == 63) dzp = dir(q + sz);
        if (lane == 0) dzm = dir(q - sz);
        bool fluid;
        double dg;
        const double qn = apply_a_vals<PACKED>(L, op, q, inner, inv_h2, fluid, dg, dxp, dxm, dyp, dym, dzp, dzm, dc);
        const float dv = fluid ? dc : 0.0f;
        dnew[q] = dv;
        const double rv = ro - alpha * qn;
        const float rf = (float)rv;
        x[q] = xo + alpha * (double)dv;
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
// stage 0: init (s6 = r.r, s7 = b.b); 2: r.r; 3: z.(A d_old), z.r, z.w ->
// beta, d.A d, alpha (first: beta = 0).
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


void launch_fcg_dots(const Level &L, const uint8_t *op, const float *z, const float *dold, const double *r, int first,
                     double inv_h2, const int32_t *rows, int64_t nrows, double *partials, double *s, hipStream_t st) {
    const int64_t nr = rows_of(L, rows, nrows);
    const int nb = row_blocks_for(nr);
    if (op)
        hipLaunchKernelGGL(fcg_dots_kernel<true>, dim3(nb), dim3(kT), 0, st, L, op, z, dold, r, inv_h2, rows, nr,
                           partials);
    else
        hipLaunchKernelGGL(fcg_dots_kernel<false>, dim3(nb), dim3(kT), 0, st, L, op, z, dold, r, inv_h2, rows, nr,
                           partials);
    hipLaunchKernelGGL(fcg_scalars_kernel, dim3(1), dim3(kScalarThreads), 0, st, partials, nb, 3, first, s);
}

void launch_fcg_update(const Level &L, const uint8_t *op, const float *z, double *x, double *r, float *r32,
                       const float *dold, float *dnew, int first, double inv_h2, const int32_t *rows, int64_t nrows,
                       float *e, float delta, double *partials, double *s, hipStream_t st) {
    const int64_t nr = rows_of(L, rows, nrows);
    const int nb = row_blocks_for(nr);
    const float dd = delta * delta;
    if (op && e)
        hipLaunchKernelGGL((fcg_update_kernel<true, true>), dim3(nb), dim3(kT), 0, st, L, op, z, x, r, r32, dold, dnew,
                           s, first, inv_h2, rows, nr, e, dd, partials);
    else if (op)
        hipLaunchKernelGGL((fcg_update_kernel<true, false>), dim3(nb), dim3(kT), 0, st, L, op, z, x, r, r32, dold,
                           dnew, s, first, inv_h2, rows, nr, e, dd, partials);
    else
        hipLaunchKernelGGL((fcg_update_kernel<false, false>), dim3(nb), dim3(kT), 0, st, L, op, z, x, r, r32, dold,
                           dnew, s, first, inv_h2, rows, nr, e, dd, partials);
    hipLaunchKernelGGL(fcg_scalars_kernel, dim3(1), dim3(kScalarThreads), 0, st, partials, nb, 2, 0, s);
}


// ---- the true residual of an FP64 iterate -----------------------------------------------
// r = b - A x on the interior fluid cells of the listed rows (all rows when
// rows is NULL), 0 elsewhere; r32 = float(r); partial r.r and b.b.  Run after
// the FCG loop stops, so the reported residual is recomputed from x rather
// than carried by the r -= alpha q recurrence, and as the restart of the
// iteration when the two disagree.
template <bool PACKED>
__global__ __launch_bounds__(kT) void fcg_residual_rows_kernel(Level L, const uint8_t *__restrict__ op,
                                                              const double *__restrict__ x,
                                                              const float *__restrict__ b, double inv_h2,
                                                              const int32_t *__restrict__ rows, int64_t nrows,
                                                              double *__restrict__ r, float *__restrict__ r32,
                                                              double *__restrict__ partials) {
    double acc[2] = {0.0, 0.0};
    for_rows(L, rows, nrows, [&](int64_t q, bool inner, int, int, int) {
        const double bv = (double)b[q];
        bool fluid;
        double dg;
        const double av = apply_a_z<PACKED, double>(L, op, x, q, inner, inv_h2, fluid, dg);
        const double rv = fluid ? bv - av : 0.0;
        r[q] = rv;
        r32[q] = (float)rv;
        acc[0] += rv * rv;
        acc[1] += bv * bv;
    });
    block_sums<2>(acc, partials);
}

void launch_fcg_residual(const Level &L, const uint8_t *op, const double *x, const float *b, double inv_h2,
                         const int32_t *rows, int64_t nrows, double *r, float *r32, double *partials, double *s,
                         hipStream_t st) {
    const int64_t nr = rows_of(L, rows, nrows);
    const int nb = row_blocks_for(nr);
    if (op)
        hipLaunchKernelGGL(fcg_residual_rows_kernel<true>, dim3(nb), dim3(kT), 0, st, L, op, x, b, inv_h2, rows, nr,
                           r, r32, partials);
    else
        hipLaunchKernelGGL(fcg_residual_rows_kernel<false>, dim3(nb), dim3(kT), 0, st, L, op, x, b, inv_h2, rows,
                           nr, r, r32, partials);
    hipLaunchKernelGGL(fcg_scalars_kernel, dim3(1), dim3(kScalarThreads), 0, st, partials, nb, 0, 0, s);
}

// ---- FP64 check solve: Jacobi-preconditioned CG -----------------------------------------
// An independent solve of the same level-0 system A x = b, sharing nothing
// with the multigrid path but the operator arrays: CG in FP64 preconditioned
// by the operator's diagonal a = -diag / h^2 (z = r / a; 0 on a fluid cell
// without open neighbours).  Iteration k:
//   apply:  p = z + beta p (beta = 0 first; p double-buffered, the
//           neighbours' new p formed from z and the old p), q = A p; p.q
//   update: alpha = r.z / p.q; x += alpha p, r -= alpha q, z = r / a; r.z, r.r
// Scalars on the device: c[0] r.z, c[1] p.q, c[2] alpha, c[3] beta, c[4] r.r,
// c[5] b.b.  The float operator arrays are read (not the packed bytes).
// from_x 0: x = 0, r = b; 1: restart from the current x, r = b - A x
__global__ __launch_bounds__(kT) void check_init_kernel(Level L, const float *__restrict__ b, double inv_h2,
                                                       const int32_t *__restrict__ rows, int64_t nrows, int from_x,
                                                       double *__restrict__ x, double *__restrict__ r,
                                                       double *__restrict__ z, double *__restrict__ p,
                                                       double *__restrict__ partials) {
    double acc[3] = {0.0, 0.0, 0.0};
    for_rows(L, rows, nrows, [&](int64_t q, bool inner, int, int, int) {
        const double bv = (double)b[q], dg = (double)L.diag[q];
        bool fluid;
        double dd;
        const double av = apply_a_z<false, double>(L, nullptr, x, q, inner, inv_h2, fluid, dd);
        const double rv = fluid ? bv - (from_x ? av : 0.0) : 0.0;
        const double zv = fluid && dg > 0.0 ? rv / (-dg * inv_h2) : 0.0;
        if (!from_x) x[q] = 0.0;
        r[q] = rv;
        z[q] = zv;
        p[q] = 0.0;
        acc[0] += rv * zv;
        acc[1] += rv * rv;
        acc[2] += bv * bv;
    });
    block_sums<3>(acc, partials);
}

__global__ __launch_bounds__(kT) void check_apply_kernel(Level L, const double *__restrict__ z,
                                                        const double *__restrict__ po, double *__restrict__ pn,
                                                        double *__restrict__ qv, const double *__restrict__ c,
                                                        double inv_h2, const int32_t *__restrict__ rows,
                                                        int64_t nrows, double *__restrict__ partials) {
    const double beta = c[3];
    double acc[1] = {0.0};
    for_rows(L, rows, nrows, [&](int64_t q, bool inner, int, int, int) {
        const int64_t sx = inner ? (int64_t)L.ny * L.nz : 0, sy = inner ? L.nz : 0, sz = inner ? 1 : 0;
        auto pnew = [&](int64_t t) { return z[t] + beta * po[t]; };
        const double v0 = pnew(q + sx), v1 = pnew(q - sx), v2 = pnew(q + sy), v3 = pnew(q - sy), v4 = pnew(q + sz),
                     v5 = pnew(q - sz), vc = pnew(q);
        const float w0 = L.wx[q], w1 = L.wx[q - sx], w2 = L.wy[q], w3 = L.wy[q - sy], w4 = L.wz[q], w5 = L.wz[q - sz];
        const bool fluid = inner && L.mask[q] > 0.0f;
        double sum = 0.0;
        if (w0 != 0.0f) sum = sum + (double)w0 * v0;
        if (w1 != 0.0f) sum = sum + (double)w1 * v1;
        if (w2 != 0.0f) sum = sum + (double)w2 * v2;
        if (w3 != 0.0f) sum = sum + (double)w3 * v3;
        if (w4 != 0.0f) sum = sum + (double)w4 * v4;
        if (w5 != 0.0f) sum = sum + (double)w5 * v5;
        const double pv = fluid ? vc : 0.0;
        const double av = fluid ? (sum - (double)L.diag[q] * vc) * inv_h2 : 0.0;
        pn[q] = pv;
        qv[q] = av;
        acc[0] += pv * av;
    });
    block_sums<1>(acc, partials);
}

__global__ __launch_bounds__(kT) void check_update_kernel(Level L, const double *__restrict__ p,
                                                         const double *__restrict__ qv, double *__restrict__ x,
                                                         double *__restrict__ r, double *__restrict__ z,
                                                         const double *__restrict__ c, double inv_h2,
                                                         const int32_t *__restrict__ rows, int64_t nrows,
                                                         double *__restrict__ partials) {
    const double alpha = c[2];
    double acc[2] = {0.0, 0.0};
    for_rows(L, rows, nrows, [&](int64_t q, bool inner, int, int, int) {
        const double dg = (double)L.diag[q];
        const bool fluid = inner && L.mask[q] > 0.0f;
        const double xv = x[q] + alpha * p[q];
        const double rv = r[q] - alpha * qv[q];
        const double zv = fluid && dg > 0.0 ? rv / (-dg * inv_h2) : 0.0;
        x[q] = xv;
        r[q] = rv;
        z[q] = zv;
        acc[0] += rv * zv;
        acc[1] += rv * rv;
    });
    block_sums<2>(acc, partials);
}

// Error of a pressure against the check solution over the interior fluid
// cells: sum (p - x)^2, sum x^2 (block sums) and max |p - x|, max |x| (block
// maxima).
template <typename T>
__global__ __launch_bounds__(kT) void check_error_kernel(Level L, const T *__restrict__ pr,
                                                        const double *__restrict__ x, const int32_t *__restrict__ rows,
                                                        int64_t nrows, double *__restrict__ partials) {
    double acc[2] = {0.0, 0.0}, mx[2] = {0.0, 0.0};
    for_rows(L, rows, nrows, [&](int64_t q, bool inner, int, int, int) {
        if (!(inner && L.mask[q] > 0.0f)) return;
        const double xv = x[q], e = (double)pr[q] - xv;
        acc[0] += e * e;
        acc[1] += xv * xv;
        mx[0] = fmax(mx[0], fabs(e));
        mx[1] = fmax(mx[1], fabs(xv));
    });
    block_sums<2>(acc, partials);
    __shared__ double sh[2][kT];
    sh[0][threadIdx.x] = mx[0];
    sh[1][threadIdx.x] = mx[1];
    __syncthreads();
    for (int off = kT / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            sh[0][threadIdx.x] = fmax(sh[0][threadIdx.x], sh[0][threadIdx.x + off]);
            sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + off]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * kB + blockIdx.x] = sh[0][0];
        partials[3 * kB + blockIdx.x] = sh[1][0];
    }
}

// stage 0 (init): r.z, r.r, b.b; 1 (apply): p.q -> alpha; 2 (update): r.z,
// r.r -> beta; 3 (error): two sums and two maxima into c[8..11].
__global__ __launch_bounds__(1024) void check_scalars_kernel(const double *__restrict__ partials, int blocks,
                                                            int stage, double *__restrict__ c) {
    __shared__ double sh[4][1024];
    const int n = stage == 0 ? 3 : stage == 1 ? 1 : stage == 2 ? 2 : 4;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = threadIdx.x; q < blocks; q += 1024)
        for (int a = 0; a < n; ++a) {
            const double t = partials[a * kB + q];
            v[a] = (stage == 3 && a >= 2) ? fmax(v[a], t) : v[a] + t;
        }
    for (int a = 0; a < 4; ++a) sh[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (threadIdx.x < off)
            for (int a = 0; a < n; ++a)
                sh[a][threadIdx.x] = (stage == 3 && a >= 2) ? fmax(sh[a][threadIdx.x], sh[a][threadIdx.x + off])
                                                            : sh[a][threadIdx.x] + sh[a][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if (stage == 0) {
        c[0] = sh[0][0];
        c[3] = 0.0;
        c[4] = sh[1][0];
        c[5] = sh[2][0];
    } else if (stage == 1) {
        c[1] = sh[0][0];
        c[2] = c[1] != 0.0 ? c[0] / c[1] : 0.0;
    } else if (stage == 2) {
        c[3] = c[0] != 0.0 ? sh[0][0] / c[0] : 0.0;
        c[0] = sh[0][0];
        c[4] = sh[1][0];
    } else {
        for (int a = 0; a < 4; ++a) c[8 + a] = sh[a][0];
    }
}

int check_partials_count() { return 4 * kB; }

void launch_check_init(const Level &L, const float *b, double inv_h2, const int32_t *rows, int64_t nrows, int from_x,
                       double *x, double *r, double *z, double *p, double *partials, double *c, hipStream_t st) {
    const int64_t nr = rows_of(L, rows, nrows);
    const int nb = row_blocks_for(nr);
    hipLaunchKernelGGL(check_init_kernel, dim3(nb), dim3(kT), 0, st, L, b, inv_h2, rows, nr, from_x, x, r, z, p,
                       partials);
    hipLaunchKernelGGL(check_scalars_kernel, dim3(1), dim3(1024), 0, st, partials, nb, 0, c);
}

void launch_check_iteration(const Level &L, double inv_h2, const int32_t *rows, int64_t nrows, double *x, double *r,
                            double *z, double *po, double *pn, double *qv, double *partials, double *c,
                            hipStream_t st) {
    const int64_t nr = rows_of(L, rows, nrows);
    const int nb = row_blocks_for(nr);
    hipLaunchKernelGGL(check_apply_kernel, dim3(nb), dim3(kT), 0, st, L, z, po, pn, qv, c, inv_h2, rows, nr,
                       partials);
    hipLaunchKernelGGL(check_scalars_kernel, dim3(1), dim3(1024), 0, st, partials, nb, 1, c);
    hipLaunchKernelGGL(check_update_kernel, dim3(nb), dim3(kT), 0, st, L, pn, qv, x, r, z, c, inv_h2, rows, nr,
                       partials);
    hipLaunchKernelGGL(check_scalars_kernel, dim3(1), dim3(1024), 0, st, partials, nb, 2, c);
}

void launch_check_error(const Level &L, const double *p64, const float *p32, const double *x, const int32_t *rows,
                        int64_t nrows, double *partials, double *c, hipStream_t st) {
    const int64_t nr = rows_of(L, rows, nrows);
    const int nb = row_blocks_for(nr);
    if (p64)
        hipLaunchKernelGGL(check_error_kernel<double>, dim3(nb), dim3(kT), 0, st, L, p64, x, rows, nr, partials);
    else
        hipLaunchKernelGGL(check_error_kernel<float>, dim3(nb), dim3(kT), 0, st, L, p32, x, rows, nr, partials);
    hipLaunchKernelGGL(check_scalars_kernel, dim3(1), dim3(1024), 0, st, partials, nb, 3, c);
}

}  // namespace clq
