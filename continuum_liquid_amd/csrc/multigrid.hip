// Multigrid kernels of the pressure solve (reference
// Solver/Kernel/OpenCL/multigrid.cl, driven by Solver/Kernel/multigrid.py).
//
//   rbgs            rbgs_step (multigrid.cl:203-230), one lane per cell of
//                   the colour being updated
//   coarse_operator Galerkin coarse diagonal/weights/mask of
//                   restrict_residual (multigrid.cl:122-179), built once per
//                   solve instead of on every restriction visit: the fine
//                   operator does not change inside a solve, so the values are
//                   the ones every visit of the reference recomputes
//   restrict        coarse b = sum of child residuals, coarse p = 0
//                   (multigrid.cl:90-180)
//   prolongate      multigrid.cl:182-201
//   coarse_smooth   coarse_smooth (multigrid.cl:232-264) in one work-group
//                   with the whole coarsest level staged in LDS
//   residual_norm   FP64 ||b - A p||^2 and ||b||^2, fixed reduction tree
#include <mutex>

#include "common.hpp"

namespace clq {

// weighted_neighbor_sum (multigrid.cl:7-23) as clang contracts it.
__device__ __forceinline__ float wsum(const float *__restrict__ p, const float *__restrict__ wx,
                                      const float *__restrict__ wy, const float *__restrict__ wz, int64_t idx,
                                      int64_t sx, int64_t sy) {
    float s = wx[idx - sx] * p[idx - sx];
    s = fmaf(wx[idx], p[idx + sx], s);
    s = fmaf(wy[idx], p[idx + sy], s);
    s = fmaf(wy[idx - sy], p[idx - sy], s);
    s = fmaf(wz[idx], p[idx + 1], s);
    s = fmaf(wz[idx - 1], p[idx - 1], s);
    return s;
}

// One colour of red-black Gauss-Seidel.  Lane t handles the t-th cell of the
// colour in row-major order: row (i, j), k = 2*kk + ((parity + i + j) & 1).
// Non-fluid cells of that colour are zeroed, as rbgs_step zeroes every
// non-fluid cell; after a full sweep both colours have been zeroed.
__device__ __forceinline__ void rbgs_cell(const Level &L, float dd, int parity, uint32_t row, int kk) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int i = (int)(row / (uint32_t)ny), j = (int)(row - (uint32_t)i * (uint32_t)ny);
    const int k = 2 * kk + ((parity + i + j) & 1);
    if (k >= nz) return;
    const int64_t idx = (int64_t)row * nz + k;
    float *__restrict__ p = L.p;
    if (L.mask[idx] <= 0.0f) {
        p[idx] = 0.0f;
        return;
    }
    if (!interior(i, j, k, nx, ny, nz)) return;
    const float d = L.diag[idx];
    if (d > 0.0f) {
        const int64_t sx = (int64_t)ny * nz, sy = nz;
        p[idx] = fmaf(-dd, load_b(L.b, L.b16, L.bkind, L.bscale, idx), wsum(p, L.wx, L.wy, L.wz, idx, sx, sy)) /
                 d;  // multigrid.cl:227 (fmuladd)
    }
}

// The same update with every load issued before the first branch on loaded
// data, b read without a branch on its storage form (one memory round trip
// instead of three); the extra loads at air cells are cheap on the coarse
// levels, whose sweeps are latency-bound.
__device__ __forceinline__ void rbgs_cell_eager(const Level &L, float dd, int parity, uint32_t row, int kk) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int i = (int)(row / (uint32_t)ny), j = (int)(row - (uint32_t)i * (uint32_t)ny);
    const int k = 2 * kk + ((parity + i + j) & 1);
    if (k >= nz) return;
    const int64_t idx = (int64_t)row * nz + k;
    const bool inner = interior(i, j, k, nx, ny, nz);
    const float m = L.mask[idx];
    float d = 0.0f, num = 0.0f;
    if (inner) {
        const int64_t sx = (int64_t)ny * nz, sy = nz;
        d = L.diag[idx];
        num = fmaf(-dd, load_b_flat(L.b, L.b16, L.bkind, L.bscale, idx), wsum(L.p, L.wx, L.wy, L.wz, idx, sx, sy));
    }
    if (m <= 0.0f) {
        L.p[idx] = 0.0f;
        return;
    }
    if (inner && d > 0.0f) L.p[idx] = num / d;  // multigrid.cl:227
}

__device__ __forceinline__ void rbgs_item(const Level &L, float dd, int parity, int64_t t) {
    const uint32_t half = (uint32_t)(L.nz + 1) >> 1;
    const uint32_t row = (uint32_t)t / half;
    rbgs_cell(L, dd, parity, row, (int)((uint32_t)t - row * half));
}

// Four same-colour cells per lane with every load issued before any store:
// cells of one colour never read each other, so the batch is exact and the
// lane keeps four memory requests in flight (used by the single-work-group
// bottom kernel, where latency, not bandwidth, bounds a half-sweep).
__device__ __forceinline__ void rbgs_batch4(const Level &L, float dd, int parity, int64_t t0, int64_t stride,
                                            int64_t work) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const uint32_t half = (uint32_t)(nz + 1) >> 1;
    const int64_t sx = (int64_t)ny * nz, sy = nz;
    int64_t idx[4];
    int mode[4];  // 0 skip, 1 zero, 2 update
    float num[4], den[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int64_t t = t0 + u * stride;
        mode[u] = 0;
        idx[u] = 0;
        num[u] = 0.0f;
        den[u] = 1.0f;
        if (t >= work) continue;
        const uint32_t row = (uint32_t)t / half;
        const int kk = (int)((uint32_t)t - row * half);
        const int i = (int)(row / (uint32_t)ny), j = (int)(row - (uint32_t)i * (uint32_t)ny);
        const int k = 2 * kk + ((parity + i + j) & 1);
        if (k >= nz) continue;
        const int64_t q = (int64_t)row * nz + k;
        idx[u] = q;
        if (L.mask[q] <= 0.0f) {
            mode[u] = 1;
            continue;
        }
        if (!interior(i, j, k, nx, ny, nz)) continue;
        const float d = L.diag[q];
        if (!(d > 0.0f)) continue;
        mode[u] = 2;
        num[u] = fmaf(-dd, load_b(L.b, L.b16, L.bkind, L.bscale, q), wsum(L.p, L.wx, L.wy, L.wz, q, sx, sy));
        den[u] = d;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (mode[u] == 1) L.p[idx[u]] = 0.0f;
        else if (mode[u] == 2) L.p[idx[u]] = num[u] / den[u];  // multigrid.cl:227
    }
}

template <bool EAGER>
__global__ __launch_bounds__(256) void rbgs_kernel(Level L, float dd, int parity) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (int64_t)L.nx * L.ny * ((L.nz + 1) >> 1)) return;
    if (EAGER) {
        const uint32_t half = (uint32_t)(L.nz + 1) >> 1;
        const uint32_t row = (uint32_t)t / half;
        rbgs_cell_eager(L, dd, parity, row, (int)((uint32_t)t - row * half));
    } else {
        rbgs_item(L, dd, parity, t);
    }
}

// Levels below this many cells take the eager (latency-bound) variants.
constexpr int64_t kEagerCells = (int64_t)1 << 21;

void launch_rbgs(const Level &L, float delta, int parity, hipStream_t s) {
    const int half = (L.nz + 1) >> 1;
    const int64_t work = (int64_t)L.nx * L.ny * half;
    if (L.C < kEagerCells)
        hipLaunchKernelGGL(rbgs_kernel<true>, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, L, delta * delta,
                           parity);
    else
        hipLaunchKernelGGL(rbgs_kernel<false>, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, L,
                           delta * delta, parity);
}

// ---- active rows ----------------------------------------------------------------------
// A row is the run of cells (i, j, 0..nz-1).  It is active when it holds an
// interior cell with mask > 0.  Every kernel below restricted to active rows
// computes exactly what its dense version computes on those rows; on the
// other rows the dense smoother only writes p = 0 (mask <= 0 there), which
// zero_inactive_rows does once per solve, and no fluid cell reads them with a
// non-zero weight (the operator's weight towards a cell with mask <= 0 is 0).

// flags[row] = 1 when active; one wave per row.
__global__ __launch_bounds__(256) void row_flags_kernel(Level L, int32_t *__restrict__ flags) {
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t rows = (int64_t)L.nx * L.ny;
    if (row >= rows) return;
    const int i = (int)(row / L.ny), j = (int)(row - (int64_t)i * L.ny);
    bool any = false;
    if (i > 0 && i < L.nx - 1 && j > 0 && j < L.ny - 1)
        for (int k = 1 + lane; k < L.nz - 1; k += 64) any |= L.mask[row * L.nz + k] > 0.0f;
    const bool active = __ballot(any) != 0;
    if (lane == 0) flags[row] = active ? 1 : 0;
}

// Coarse row (a, b) is active when one of the fine rows it restricts from is.
__global__ __launch_bounds__(256) void coarse_row_flags_kernel(const int32_t *__restrict__ fine, int fnx, int fny,
                                                               int cnx, int cny, int32_t *__restrict__ flags) {
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (row >= (int64_t)cnx * cny) return;
    const int a = (int)(row / cny), b = (int)(row - (int64_t)a * cny);
    int32_t f = 0;
    if (a > 0 && a < cnx - 1 && b > 0 && b < cny - 1) {
        const int f0 = 1 + 2 * (a - 1), f1 = 1 + 2 * (b - 1);
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj) {
                const int i = f0 + di, j = f1 + dj;
                if (i < fnx - 1 && j < fny - 1) f |= fine[(int64_t)i * fny + j];
            }
    }
    flags[row] = f;
}

__global__ __launch_bounds__(256) void compact_rows_kernel(const int32_t *__restrict__ flags,
                                                           const int32_t *__restrict__ pos, int64_t rows,
                                                           int32_t *__restrict__ list, int32_t *__restrict__ count) {
    const int64_t row = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (row >= rows) return;
    if (flags[row]) list[pos[row]] = (int32_t)row;
    if (row == rows - 1) *count = pos[row] + flags[row];
}

// On the inactive rows: level 0 (coarse = 0) p = 0 where mask <= 0, what the
// first dense sweep does there; a coarse level p = 0 and b = 0 everywhere,
// what every dense restriction writes there (no child is fluid).
__global__ __launch_bounds__(256) void zero_inactive_rows_kernel(Level L, const int32_t *__restrict__ flags,
                                                                 int coarse) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= L.C) return;
    if (flags[q / L.nz]) return;
    if (coarse) {
        L.p[q] = 0.0f;
        L.b[q] = 0.0f;
        if (L.b16) L.b16[q] = 0;
    } else if (L.mask[q] <= 0.0f) {
        L.p[q] = 0.0f;
    }
}

// The row kernels take one lane per same-colour cell of the active rows:
// item t is cell kk = t mod half of active row t / half (half = cells of one
// colour in a row), so no lane idles on rows shorter than a block.
__device__ __forceinline__ bool row_item(const int32_t *__restrict__ rows, int64_t nrows, int half, uint32_t &row,
                                        int &kk) {
    const uint32_t t = xcd_item(blockIdx.x, gridDim.x) * 256u + threadIdx.x;
    const uint32_t ar = t / (uint32_t)half;
    if (ar >= (uint32_t)nrows) return false;
    row = (uint32_t)rows[ar];
    kk = (int)(t - ar * (uint32_t)half);
    return true;
}

static unsigned row_blocks(int64_t nrows, int nz) { return (unsigned)((nrows * ((nz + 1) >> 1) + 255) / 256); }

// all loads issued first
__global__ __launch_bounds__(256) void rbgs_rows_kernel(Level L, float dd, int parity, const int32_t *__restrict__ rows,
                                                        int64_t nrows) {
    uint32_t row;
    int kk;
    if (row_item(rows, nrows, (L.nz + 1) >> 1, row, kk)) rbgs_cell_eager(L, dd, parity, row, kk);
}

__global__ __launch_bounds__(256) void zero_rows_kernel(float *__restrict__ p, const int32_t *__restrict__ rows, int nz) {
    const int64_t base = (int64_t)rows[blockIdx.x] * nz;
    for (int k = threadIdx.x; k < nz; k += 256) p[base + k] = 0.0f;
}

void launch_zero_rows(float *p, const int32_t *rows, int64_t nrows, int nz, hipStream_t s) {
    if (nrows == 0) return;
    hipLaunchKernelGGL(zero_rows_kernel, dim3((unsigned)nrows), dim3(256), 0, s, p, rows, nz);
}

void launch_row_flags(const Level &L, int32_t *flags, hipStream_t s) {
    const int64_t rows = (int64_t)L.nx * L.ny;
    hipLaunchKernelGGL(row_flags_kernel, dim3((unsigned)((rows + 3) / 4)), dim3(256), 0, s, L, flags);
}

void launch_coarse_row_flags(const int32_t *fine, const Level &F, const Level &K, int32_t *flags, hipStream_t s) {
    const int64_t rows = (int64_t)K.nx * K.ny;
    hipLaunchKernelGGL(coarse_row_flags_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, fine, F.nx,
                       F.ny, K.nx, K.ny, flags);
}

void launch_compact_rows(const int32_t *flags, const int32_t *pos, int64_t rows, int32_t *list, int32_t *count,
                         hipStream_t s) {
    hipLaunchKernelGGL(compact_rows_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, s, flags, pos, rows,
                       list, count);
}

void launch_zero_inactive_rows(const Level &L, const int32_t *flags, int coarse, hipStream_t s) {
    hipLaunchKernelGGL(zero_inactive_rows_kernel, dim3((unsigned)((L.C + 255) / 256)), dim3(256), 0, s, L, flags,
                       coarse);
}

// ---- packed finest-level operator -----------------------------------------------------
// The finest operator (build_finest_operator) has weights 0 or 1 and an
// integer diagonal 0..6, and the sweep only tests mask > 0, so one byte per
// cell holds it exactly: bit 0-2 wx, wy, wz != 0, bit 3 mask > 0, bits 4-7 the
// diagonal.  Decoding gives the same floats, so the packed sweep is
// bit-identical; pack_operator_kernel raises *bad for any cell whose values
// are not of that form (the caller then keeps the float arrays).
__global__ __launch_bounds__(256) void pack_operator_kernel(Level L, uint8_t *__restrict__ op, int *__restrict__ bad) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= L.C) return;
    const float x = L.wx[q], y = L.wy[q], z = L.wz[q], d = L.diag[q];
    const bool ok = (x == 0.0f || x == 1.0f) && (y == 0.0f || y == 1.0f) && (z == 0.0f || z == 1.0f) && d >= 0.0f &&
                    d <= 15.0f && d == (float)(int)d;
    if (!ok) atomicOr(bad, 1);
    op[q] = (uint8_t)((x != 0.0f ? 1 : 0) | (y != 0.0f ? 2 : 0) | (z != 0.0f ? 4 : 0) | (L.mask[q] > 0.0f ? 8 : 0) |
                      ((ok ? (int)d : 0) << 4));
}

void launch_pack_operator(const Level &L, uint8_t *op, int *bad, hipStream_t s) {
    hipLaunchKernelGGL(pack_operator_kernel, dim3((unsigned)((L.C + 255) / 256)), dim3(256), 0, s, L, op, bad);
}

// rbgs_cell on the packed operator: same loads of p and b, the operator from
// op (own byte and the -x, -y, -z neighbours' bytes), same arithmetic.
__global__ __launch_bounds__(256) void rbgs_rows_packed_kernel(Level L, const uint8_t *__restrict__ op, float dd,
                                                               int parity, const int32_t *__restrict__ rows,
                                                               int64_t nrows) {
    uint32_t row;
    int kk;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    if (!row_item(rows, nrows, (nz + 1) >> 1, row, kk)) return;
    const int i = (int)(row / (uint32_t)ny), j = (int)(row - (uint32_t)i * (uint32_t)ny);
    const int k = 2 * kk + ((parity + i + j) & 1);
    if (k >= nz) return;
    const int64_t sx = (int64_t)ny * nz, sy = nz;
    float *__restrict__ p = L.p;
    const int64_t idx = (int64_t)row * nz + k;
    const bool inner = interior(i, j, k, nx, ny, nz);
    const uint32_t o = op[idx];
    float num = 0.0f;
    if (inner) {
        const uint32_t ox = op[idx - sx], oy = op[idx - sy], oz = op[idx - 1];
        float t = bitw(ox, 1) * p[idx - sx];
        t = fmaf(bitw(o, 1), p[idx + sx], t);
        t = fmaf(bitw(o, 2), p[idx + sy], t);
        t = fmaf(bitw(oy, 2), p[idx - sy], t);
        t = fmaf(bitw(o, 4), p[idx + 1], t);
        t = fmaf(bitw(oz, 4), p[idx - 1], t);
        num = fmaf(-dd, load_b(L.b, L.b16, L.bkind, L.bscale, idx), t);
    }
    if (!(o & 8)) {
        p[idx] = 0.0f;
        return;
    }
    const float d = (float)(o >> 4);
    if (inner && d > 0.0f) p[idx] = num / d;  // multigrid.cl:227
}

void launch_rbgs_rows_packed(const Level &L, const uint8_t *op, float delta, int parity, const int32_t *rows,
                             int64_t nrows, hipStream_t s) {
    if (nrows == 0) return;
    hipLaunchKernelGGL(rbgs_rows_packed_kernel, dim3(row_blocks(nrows, L.nz)), dim3(256), 0, s, L, op, delta * delta,
                       parity, rows, nrows);
}

void launch_rbgs_rows(const Level &L, float delta, int parity, const int32_t *rows, int64_t nrows, hipStream_t s) {
    if (nrows == 0) return;
    hipLaunchKernelGGL(rbgs_rows_kernel, dim3(row_blocks(nrows, L.nz)), dim3(256), 0, s, L, delta * delta, parity, rows,
                       nrows);
}

__global__ __launch_bounds__(256) void coarse_operator_kernel(
    const float *__restrict__ fmask, const float *__restrict__ fdiag, const float *__restrict__ fwx,
    const float *__restrict__ fwy, const float *__restrict__ fwz, int fnx, int fny, int fnz, float *__restrict__ cmask,
    float *__restrict__ cdiag, float *__restrict__ cwx, float *__restrict__ cwy, float *__restrict__ cwz, int cnx,
    int cny, int cnz) {
    const int64_t ci = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (ci >= (int64_t)cnx * cny * cnz) return;
    int a, b, c;
    coords(ci, cny, cnz, a, b, c);
    float m = 0.0f, ds = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
    if (interior(a, b, c, cnx, cny, cnz)) {
        const int f0 = 1 + 2 * (a - 1), f1 = 1 + 2 * (b - 1), f2 = 1 + 2 * (c - 1);
        int active = 0;
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj)
                for (int dk = 0; dk < 2; ++dk) {
                    const int i = f0 + di, j = f1 + dj, k = f2 + dk;
                    if (!interior(i, j, k, fnx, fny, fnz)) continue;
                    const int64_t idx = gidx(i, j, k, fny, fnz);
                    if (fmask[idx] <= 0.0f) continue;
                    ++active;
                    ds = ds + fdiag[idx];
                    const float x = fwx[idx], y = fwy[idx], z = fwz[idx];
                    if (x > 0.0f) { if (di == 0) ds = fmaf(-2.0f, x, ds); else cx = cx + x; }
                    if (y > 0.0f) { if (dj == 0) ds = fmaf(-2.0f, y, ds); else cy = cy + y; }
                    if (z > 0.0f) { if (dk == 0) ds = fmaf(-2.0f, z, ds); else cz = cz + z; }
                }
        if (active > 0) m = 1.0f;
        else ds = cx = cy = cz = 0.0f;
    }
    cmask[ci] = m;
    cdiag[ci] = ds;
    cwx[ci] = cx;
    cwy[ci] = cy;
    cwz[ci] = cz;
}

void launch_coarse_operator(const Level &F, const Level &K, hipStream_t s) {
    hipLaunchKernelGGL(coarse_operator_kernel, dim3((unsigned)((K.C + 255) / 256)), dim3(256), 0, s, F.mask, F.diag,
                       F.wx, F.wy, F.wz, F.nx, F.ny, F.nz, K.mask, K.diag, K.wx, K.wy, K.wz, K.nx, K.ny, K.nz);
}

// residual (multigrid.cl:45-61) at an interior fluid cell.
__device__ __forceinline__ float residual_at(const float *__restrict__ p, float bv, const float *__restrict__ diag,
                                             const float *__restrict__ wx, const float *__restrict__ wy,
                                             const float *__restrict__ wz, int64_t idx, int64_t sx, int64_t sy,
                                             float inv_delta2) {
    const float ap = fmaf(-diag[idx], p[idx], wsum(p, wx, wy, wz, idx, sx, sy)) * inv_delta2;  // :41 (fmuladd)
    return bv - ap;
}

// Coarse cell ci's b and pressure after the restriction (p = 0) and, with
// red, after the first (red) half-sweep of the coarse pre-smoothing too: on
// p = 0 every neighbour term w * 0 is +0, so the red update is
// fmaf(-dd, b, +0) / d with b as the smoother would read it back from
// storage (stored_b).
__device__ __forceinline__ void restrict_store(const Level &K, int64_t ci, bool kin, float rhs, float dd, bool red) {
    store_b(K.b, K.b16, K.bkind, K.bscale, ci, rhs);
    float pv = 0.0f;
    if (red && kin) {
        int a, b, c;
        coords(ci, K.ny, K.nz, a, b, c);
        if (((a + b + c) & 1) == 0) {
            const float m = K.mask[ci], d = K.diag[ci];
            if (m > 0.0f && d > 0.0f) pv = fmaf(-dd, stored_b(K.bkind, K.bscale, rhs), 0.0f) / d;
        }
    }
    K.p[ci] = pv;
}

__device__ __forceinline__ void restrict_item(const Level &F, const Level &K, float inv_delta2, int64_t ci,
                                              float dd = 0.0f, bool red = false) {
    int a, b, c;
    coords(ci, K.ny, K.nz, a, b, c);
    float rhs = 0.0f;
    if (interior(a, b, c, K.nx, K.ny, K.nz)) {
        const int64_t sx = (int64_t)F.ny * F.nz, sy = F.nz;
        const int f0 = 1 + 2 * (a - 1), f1 = 1 + 2 * (b - 1), f2 = 1 + 2 * (c - 1);
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj)
                for (int dk = 0; dk < 2; ++dk) {
                    const int i = f0 + di, j = f1 + dj, k = f2 + dk;
                    if (!interior(i, j, k, F.nx, F.ny, F.nz)) continue;
                    const int64_t idx = gidx(i, j, k, F.ny, F.nz);
                    if (F.mask[idx] <= 0.0f) continue;
                    const float bv = load_b(F.b, F.b16, F.bkind, F.bscale, idx);
                    rhs = rhs + residual_at(F.p, bv, F.diag, F.wx, F.wy, F.wz, idx, sx, sy, inv_delta2);
                }
    }
    restrict_store(K, ci, interior(a, b, c, K.nx, K.ny, K.nz), rhs, dd, red);
}

// restrict_item with the eight children's loads issued together; the sum
// takes the fluid children in the same order.
__device__ __forceinline__ void restrict_item_eager(const Level &F, const Level &K, float inv_delta2, int64_t ci,
                                                    float dd = 0.0f, bool red = false) {
    int a, b, c;
    coords(ci, K.ny, K.nz, a, b, c);
    float rhs = 0.0f;
    if (interior(a, b, c, K.nx, K.ny, K.nz)) {
        const int64_t sx = (int64_t)F.ny * F.nz, sy = F.nz;
        const int f0 = 1 + 2 * (a - 1), f1 = 1 + 2 * (b - 1), f2 = 1 + 2 * (c - 1);
        float m[8], res[8];
        bool in[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = f0 + (q >> 2), j = f1 + ((q >> 1) & 1), k = f2 + (q & 1);
            in[q] = interior(i, j, k, F.nx, F.ny, F.nz);
            m[q] = 0.0f;
            res[q] = 0.0f;
            if (in[q]) {
                const int64_t idx = gidx(i, j, k, F.ny, F.nz);
                m[q] = F.mask[idx];
                res[q] = residual_at(F.p, load_b_flat(F.b, F.b16, F.bkind, F.bscale, idx), F.diag, F.wx, F.wy, F.wz, idx,
                                     sx, sy, inv_delta2);
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (in[q] && m[q] > 0.0f) rhs = rhs + res[q];
    }
    restrict_store(K, ci, interior(a, b, c, K.nx, K.ny, K.nz), rhs, dd, red);
}

// One lane per coarse cell (restrict_item): the finer levels, where the
// eight-lane form below measured slower (its lanes repeat the index work and
// the neighbour loads that one lane shares between its children).
__global__ __launch_bounds__(256) void restrict_kernel(Level F, Level K, float inv_delta2, float dd, int red) {
    const int64_t ci = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (ci >= K.C) return;
    restrict_item(F, K, inv_delta2, ci, dd, red != 0);
}

// G lanes per coarse cell, lane g computing the residuals of children
// g * 8/G .. (g + 1) * 8/G - 1 (0 where the child is outside the interior or
// not fluid); lane 0 of the G sums the eight in restrict_item's order.
// Adding the +0 of a skipped child leaves the sum unchanged: it starts at +0
// and a round-to-nearest sum is -0 only when both terms are.  More lanes per
// coarse cell shorten each lane's chain of loads (128 for one lane) at the
// cost of repeated index work: eight lanes measured 5.1 and 4.1 us instead of
// 9.8 and 9.3 us for the 66^3 -> 34^3 and 34^3 -> 18^3 restrictions on MI355X.
template <int G>
__device__ __forceinline__ float sum_children(const float (&res)[8 / G]) {
    const int base = (int)(threadIdx.x & 63u) & ~(G - 1);
    float rhs = 0.0f;
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int u = 0; u < 8 / G; ++u) rhs = rhs + __shfl(res[u], base + g, 64);
    return rhs;
}

// child q of coarse cell (a, b, c): its residual (residual_at) where it is an
// interior fluid cell, else 0; from the packed operator when op != NULL
__device__ __forceinline__ float child_residual(const Level &F, const uint8_t *__restrict__ op, float inv_delta2,
                                                int a, int b, int c, int q) {
    const int i = 1 + 2 * (a - 1) + (q >> 2), j = 1 + 2 * (b - 1) + ((q >> 1) & 1), k = 1 + 2 * (c - 1) + (q & 1);
    if (!interior(i, j, k, F.nx, F.ny, F.nz)) return 0.0f;
    const int64_t idx = gidx(i, j, k, F.ny, F.nz), sx = (int64_t)F.ny * F.nz, sy = F.nz;
    if (op) {
        const uint32_t o = op[idx], ox = op[idx - sx], oy = op[idx - sy], oz = op[idx - 1];
        const float *__restrict__ p = F.p;
        float t = bitw(ox, 1) * p[idx - sx];
        t = fmaf(bitw(o, 1), p[idx + sx], t);
        t = fmaf(bitw(o, 2), p[idx + sy], t);
        t = fmaf(bitw(oy, 2), p[idx - sy], t);
        t = fmaf(bitw(o, 4), p[idx + 1], t);
        t = fmaf(bitw(oz, 4), p[idx - 1], t);
        const float r = load_b(F.b, F.b16, F.bkind, F.bscale, idx) - fmaf(-(float)(o >> 4), p[idx], t) * inv_delta2;
        return (o & 8) ? r : 0.0f;  // residual_at
    }
    const float m = F.mask[idx];
    const float r = residual_at(F.p, load_b_flat(F.b, F.b16, F.bkind, F.bscale, idx), F.diag, F.wx, F.wy, F.wz, idx,
                                sx, sy, inv_delta2);
    return m > 0.0f ? r : 0.0f;
}

// coarse cell w of the work list: krows == NULL: all cells, else the active
// coarse rows' cells
template <int G>
__global__ __launch_bounds__(256) void restrict_group_kernel(Level F, const uint8_t *__restrict__ op, Level K,
                                                             float inv_delta2, const int32_t *__restrict__ krows,
                                                             int64_t work, float dd, int red) {
    const int64_t t = (int64_t)xcd_item(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    const int64_t w = t / G;
    const int g = (int)(t % G);
    const bool live = w < work;  // the same for the G lanes of a coarse cell
    int64_t ci = 0;
    bool kin = false;
    float res[8 / G];
#pragma unroll
    for (int u = 0; u < 8 / G; ++u) res[u] = 0.0f;
    if (live) {
        if (krows) {
            const int64_t ar = w / K.nz;
            ci = (int64_t)krows[ar] * K.nz + (w - ar * K.nz);
        } else {
            ci = w;
        }
        int a, b, c;
        coords(ci, K.ny, K.nz, a, b, c);
        kin = interior(a, b, c, K.nx, K.ny, K.nz);
        if (kin)
#pragma unroll
            for (int u = 0; u < 8 / G; ++u) res[u] = child_residual(F, op, inv_delta2, a, b, c, g * (8 / G) + u);
    }
    const float rhs = sum_children<G>(res);  // every lane takes part in the shuffles
    if (live && g == 0) restrict_store(K, ci, kin, kin ? rhs : 0.0f, dd, red != 0);
}

template <int G>
static void launch_restrict_group(const Level &F, const uint8_t *op, const Level &K, float inv_delta2,
                                  const int32_t *krows, int64_t work, float dd, int red, hipStream_t s) {
    if (work == 0) return;
    hipLaunchKernelGGL(restrict_group_kernel<G>, dim3((unsigned)((G * work + 255) / 256)), dim3(256), 0, s, F, op, K,
                       inv_delta2, krows, work, dd, red);
}

// lanes per coarse cell: levels below 2M fine cells are latency-bound
constexpr int kRestrictLanesCoarse = 8, kRestrictLanesFine = 1, kRestrictLanesRows = 1;

void launch_restrict_residual(const Level &F, const Level &K, float inv_delta2, float delta, int red, hipStream_t s) {
    const float dd = delta * delta;
    if (F.C >= kEagerCells) {
        if (kRestrictLanesFine == 1)
            hipLaunchKernelGGL(restrict_kernel, dim3((unsigned)((K.C + 255) / 256)), dim3(256), 0, s, F, K, inv_delta2,
                               dd, red);
        else
            launch_restrict_group<kRestrictLanesFine>(F, nullptr, K, inv_delta2, nullptr, K.C, dd, red, s);
        return;
    }
    launch_restrict_group<kRestrictLanesCoarse>(F, nullptr, K, inv_delta2, nullptr, K.C, dd, red, s);
}

// restrict_item_eager with the fine operator read from the packed bytes
// (same decoded values, same operations).
__device__ __forceinline__ void restrict_item_packed(const Level &F, const uint8_t *__restrict__ op, const Level &K,
                                                     float inv_delta2, int64_t ci, float dd, bool red) {
    int a, b, c;
    coords(ci, K.ny, K.nz, a, b, c);
    float rhs = 0.0f;
    if (interior(a, b, c, K.nx, K.ny, K.nz)) {
        const int64_t sx = (int64_t)F.ny * F.nz, sy = F.nz;
        const int f0 = 1 + 2 * (a - 1), f1 = 1 + 2 * (b - 1), f2 = 1 + 2 * (c - 1);
        float res[8];
        bool use[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int i = f0 + (q >> 2), j = f1 + ((q >> 1) & 1), k = f2 + (q & 1);
            use[q] = false;
            res[q] = 0.0f;
            if (interior(i, j, k, F.nx, F.ny, F.nz)) {
                const int64_t idx = gidx(i, j, k, F.ny, F.nz);
                const uint32_t o = op[idx], ox = op[idx - sx], oy = op[idx - sy], oz = op[idx - 1];
                const float *__restrict__ p = F.p;
                float t = bitw(ox, 1) * p[idx - sx];
                t = fmaf(bitw(o, 1), p[idx + sx], t);
                t = fmaf(bitw(o, 2), p[idx + sy], t);
                t = fmaf(bitw(oy, 2), p[idx - sy], t);
                t = fmaf(bitw(o, 4), p[idx + 1], t);
                t = fmaf(bitw(oz, 4), p[idx - 1], t);
                const float ap = fmaf(-(float)(o >> 4), p[idx], t) * inv_delta2;  // residual_at
                res[q] = load_b(F.b, F.b16, F.bkind, F.bscale, idx) - ap;
                use[q] = (o & 8) != 0;
            }
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (use[q]) rhs = rhs + res[q];
    }
    restrict_store(K, ci, interior(a, b, c, K.nx, K.ny, K.nz), rhs, dd, red);
}

// restriction onto the active coarse rows only (packed fine operator when op != NULL)
__global__ __launch_bounds__(256) void restrict_rows_kernel(Level F, const uint8_t *__restrict__ op, Level K,
                                                            float inv_delta2, const int32_t *__restrict__ krows,
                                                            int64_t work, float dd, int red) {
    const int64_t t = (int64_t)xcd_item(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (t >= work) return;
    const int64_t ar = t / K.nz;
    const int64_t ci = (int64_t)krows[ar] * K.nz + (t - ar * K.nz);
    if (op) restrict_item_packed(F, op, K, inv_delta2, ci, dd, red != 0);
    else restrict_item_eager(F, K, inv_delta2, ci, dd, red != 0);
}

void launch_restrict_rows(const Level &F, const uint8_t *op, const Level &K, float inv_delta2, const int32_t *krows,
                          int64_t nkrows, float delta, int red, hipStream_t s) {
    const int64_t work = nkrows * K.nz;
    if (work == 0) return;
    // one lane per coarse cell measured best for the restrictions onto the
    // active rows of levels 1 and 2 (onto level 2 at 256^3: 11.5 against
    // 12.2 us with eight lanes per coarse cell)
    if (kRestrictLanesRows == 1)
        hipLaunchKernelGGL(restrict_rows_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, F, op, K,
                           inv_delta2, krows, work, delta * delta, red);
    else
        launch_restrict_group<kRestrictLanesRows>(F, op, K, inv_delta2, krows, work, delta * delta, red, s);
}

__device__ __forceinline__ void prolong_item(const Level &F, const Level &K, int64_t fi) {
    int i, j, k;
    coords(fi, F.ny, F.nz, i, j, k);
    if (!interior(i, j, k, F.nx, F.ny, F.nz) || F.mask[fi] <= 0.0f) return;
    const int a = min((i - 1) / 2 + 1, K.nx - 2), b = min((j - 1) / 2 + 1, K.ny - 2), c = min((k - 1) / 2 + 1, K.nz - 2);
    F.p[fi] = fmaf(K.corr, K.p[gidx(a, b, c, K.ny, K.nz)], F.p[fi]);
}

// prolong_item with its three loads issued together
__device__ __forceinline__ void prolong_item_eager(const Level &F, const Level &K, int64_t fi) {
    int i, j, k;
    coords(fi, F.ny, F.nz, i, j, k);
    if (!interior(i, j, k, F.nx, F.ny, F.nz)) return;
    const int a = min((i - 1) / 2 + 1, K.nx - 2), b = min((j - 1) / 2 + 1, K.ny - 2), c = min((k - 1) / 2 + 1, K.nz - 2);
    const float m = F.mask[fi], pf = F.p[fi], pc = K.p[gidx(a, b, c, K.ny, K.nz)];
    if (m > 0.0f) F.p[fi] = fmaf(K.corr, pc, pf);
}

__global__ __launch_bounds__(256) void prolongate_kernel(Level F, Level K) {
    const int64_t fi = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (fi >= F.C) return;
    if (F.C < kEagerCells) prolong_item_eager(F, K, fi);
    else prolong_item(F, K, fi);
}

void launch_prolongate(const Level &F, const Level &K, hipStream_t s) {
    hipLaunchKernelGGL(prolongate_kernel, dim3((unsigned)((F.C + 255) / 256)), dim3(256), 0, s, F, K);
}

__global__ __launch_bounds__(256) void prolongate_rows_kernel(Level F, Level K, const int32_t *__restrict__ frows,
                                                              int64_t work) {
    const int64_t t = (int64_t)xcd_item(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
    if (t >= work) return;
    const int64_t ar = t / F.nz;
    prolong_item_eager(F, K, (int64_t)frows[ar] * F.nz + (t - ar * F.nz));
}

void launch_prolongate_rows(const Level &F, const Level &K, const int32_t *frows, int64_t nfrows, hipStream_t s) {
    const int64_t work = nfrows * F.nz;
    if (work == 0) return;
    hipLaunchKernelGGL(prolongate_rows_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, F, K, frows,
                       work);
}

// ---- transfers folded into the neighbouring half-sweeps --------------------------------
// A level visit runs restrict -> [coarse visit] -> prolong -> red sweep ->
// black sweep.  The prolongation (prolong_item: p += w p_coarse[parent] on the
// interior fluid cells) is folded into the two post-smoothing half-sweeps:
// the red one reads its black neighbours as they would be after the
// prolongation (p + w p_coarse[parent] where the neighbour is fluid), and each
// colour adds the prolongation to its own fluid cells that the sweep does not
// overwrite (d <= 0).  Every value is the one the separate kernels produce.
__device__ __forceinline__ int64_t parent_of(const Level &K, int i, int j, int k) {
    return gidx(min((i - 1) / 2 + 1, K.nx - 2), min((j - 1) / 2 + 1, K.ny - 2), min((k - 1) / 2 + 1, K.nz - 2), K.ny,
                K.nz);
}

// the value of cell q = (i, j, k) after the prolongation
__device__ __forceinline__ float prolonged(const Level &L, const Level &K, int64_t q, int i, int j, int k) {
    const float m = L.mask[q], pv = L.p[q], pc = K.p[parent_of(K, i, j, k)];
    return m > 0.0f ? fmaf(K.corr, pc, pv) : pv;
}

// rbgs_cell_eager with the prolongation from K folded in (first: the red
// half-sweep, whose neighbours still lack it)
__device__ __forceinline__ void rbgs_cell_prolong(const Level &L, const Level &K, float dd, int parity, uint32_t row,
                                                  int kk, bool first) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int i = (int)(row / (uint32_t)ny), j = (int)(row - (uint32_t)i * (uint32_t)ny);
    const int k = 2 * kk + ((parity + i + j) & 1);
    if (k >= nz) return;
    const int64_t idx = (int64_t)row * nz + k;
    const bool inner = interior(i, j, k, nx, ny, nz);
    const float m = L.mask[idx];
    float d = 0.0f, num = 0.0f, own = 0.0f;
    if (inner) {
        const int64_t sx = (int64_t)ny * nz, sy = nz;
        d = L.diag[idx];
        own = fmaf(K.corr, K.p[parent_of(K, i, j, k)], L.p[idx]);  // prolong_item
        float t;
        if (first) {  // wsum over the prolonged neighbours
            const float vxm = prolonged(L, K, idx - sx, i - 1, j, k), vxp = prolonged(L, K, idx + sx, i + 1, j, k);
            const float vyp = prolonged(L, K, idx + sy, i, j + 1, k), vym = prolonged(L, K, idx - sy, i, j - 1, k);
            const float vzp = prolonged(L, K, idx + 1, i, j, k + 1), vzm = prolonged(L, K, idx - 1, i, j, k - 1);
            t = L.wx[idx - sx] * vxm;
            t = fmaf(L.wx[idx], vxp, t);
            t = fmaf(L.wy[idx], vyp, t);
            t = fmaf(L.wy[idx - sy], vym, t);
            t = fmaf(L.wz[idx], vzp, t);
            t = fmaf(L.wz[idx - 1], vzm, t);
        } else {
            t = wsum(L.p, L.wx, L.wy, L.wz, idx, sx, sy);
        }
        num = fmaf(-dd, load_b_flat(L.b, L.b16, L.bkind, L.bscale, idx), t);
    }
    if (m <= 0.0f) {
        L.p[idx] = 0.0f;
        return;
    }
    if (inner) L.p[idx] = d > 0.0f ? num / d : own;  // multigrid.cl:227 or the prolongation
}

__global__ __launch_bounds__(256) void rbgs_prolong_kernel(Level L, Level K, float dd, int parity, int first) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const uint32_t half = (uint32_t)(L.nz + 1) >> 1;
    if (t >= (int64_t)L.nx * L.ny * half) return;
    const uint32_t row = (uint32_t)t / half;
    rbgs_cell_prolong(L, K, dd, parity, row, (int)((uint32_t)t - row * half), first != 0);
}

__global__ __launch_bounds__(256) void rbgs_rows_prolong_kernel(Level L, Level K, float dd, int parity, int first,
                                                                const int32_t *__restrict__ rows, int64_t nrows) {
    uint32_t row;
    int kk;
    if (row_item(rows, nrows, (L.nz + 1) >> 1, row, kk)) rbgs_cell_prolong(L, K, dd, parity, row, kk, first != 0);
}

// the packed level-0 form (rbgs_rows_packed_kernel): a neighbour is fluid when
// bit 3 of its operator byte is set
__global__ __launch_bounds__(256) void rbgs_rows_packed_prolong_kernel(Level L, const uint8_t *__restrict__ op,
                                                                       Level K, float dd, int parity, int first,
                                                                       const int32_t *__restrict__ rows,
                                                                       int64_t nrows) {
    uint32_t row;
    int kk;
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    if (!row_item(rows, nrows, (nz + 1) >> 1, row, kk)) return;
    const int i = (int)(row / (uint32_t)ny), j = (int)(row - (uint32_t)i * (uint32_t)ny);
    const int64_t sx = (int64_t)ny * nz, sy = nz;
    const float *__restrict__ pc = K.p;
    const float w = K.corr;
    float *__restrict__ p = L.p;
    {
        const int k = 2 * kk + ((parity + i + j) & 1);
        if (k >= nz) return;
        const int64_t idx = (int64_t)row * nz + k;
        const bool inner = interior(i, j, k, nx, ny, nz);
        const uint32_t o = op[idx];
        float num = 0.0f, own = 0.0f;
        if (inner) {
            const uint32_t ox = op[idx - sx], oy = op[idx - sy], oz = op[idx - 1];
            own = fmaf(w, pc[parent_of(K, i, j, k)], p[idx]);  // prolong_item
            float vxm, vxp, vyp, vym, vzp, vzm;
            if (first) {
                const uint32_t oxp = op[idx + sx], oyp = op[idx + sy], ozp = op[idx + 1];
                vxm = p[idx - sx];
                vxp = p[idx + sx];
                vyp = p[idx + sy];
                vym = p[idx - sy];
                vzp = p[idx + 1];
                vzm = p[idx - 1];
                const float cxm = pc[parent_of(K, i - 1, j, k)], cxp = pc[parent_of(K, i + 1, j, k)];
                const float cyp = pc[parent_of(K, i, j + 1, k)], cym = pc[parent_of(K, i, j - 1, k)];
                const float czp = pc[parent_of(K, i, j, k + 1)], czm = pc[parent_of(K, i, j, k - 1)];
                if (ox & 8) vxm = fmaf(w, cxm, vxm);
                if (oxp & 8) vxp = fmaf(w, cxp, vxp);
                if (oyp & 8) vyp = fmaf(w, cyp, vyp);
                if (oy & 8) vym = fmaf(w, cym, vym);
                if (ozp & 8) vzp = fmaf(w, czp, vzp);
                if (oz & 8) vzm = fmaf(w, czm, vzm);
            } else {
                vxm = p[idx - sx];
                vxp = p[idx + sx];
                vyp = p[idx + sy];
                vym = p[idx - sy];
                vzp = p[idx + 1];
                vzm = p[idx - 1];
            }
            float t = bitw(ox, 1) * vxm;
            t = fmaf(bitw(o, 1), vxp, t);
            t = fmaf(bitw(o, 2), vyp, t);
            t = fmaf(bitw(oy, 2), vym, t);
            t = fmaf(bitw(o, 4), vzp, t);
            t = fmaf(bitw(oz, 4), vzm, t);
            num = fmaf(-dd, load_b(L.b, L.b16, L.bkind, L.bscale, idx), t);
        }
        if (!(o & 8)) {
            p[idx] = 0.0f;
            return;
        }
        const float d = (float)(o >> 4);
        if (inner) p[idx] = d > 0.0f ? num / d : own;  // multigrid.cl:227 or the prolongation
    }
}

// Both post-smoothing half-sweeps of a dense level with the prolongation
// from K folded in, in one launch.  A block owns a tile of kTi x kTj x kTk
// cells: it computes the red half-sweep's values (rbgs_cell_prolong, first)
// for the red cells of the tile and a one-cell halo into LDS, then the black
// half-sweep's values (rbgs_cell_prolong, not first) of its black cells from
// them, and writes every cell of the tile to `out`.  The red values of a
// halo cell are the same numbers its own tile computes; L.p is only read
// (other tiles' blocks still read their halos' old values from it), so the
// caller swaps `out` in as the level's pressure.  Every value equals the two
// separate launches'.
constexpr int kTi = 4, kTj = 8, kTk = 32;
constexpr int kBi = kTi + 2, kBj = kTj + 2, kBk = kTk + 2;

// the value rbgs_cell_prolong leaves in cell (i, j, k) after its half-sweep:
// red (first) from the prolonged black neighbours in L.p, black from the red
// values in r (the LDS box, origin (i0, j0, k0) of the box)
template <bool RED>
__device__ __forceinline__ float post_value(const Level &L, const Level &K, float dd, int i, int j, int k,
                                            const float *r, int i0, int j0, int k0) {
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int64_t idx = gidx(i, j, k, ny, nz), sx = (int64_t)ny * nz, sy = nz;
    const float m = L.mask[idx], pold = L.p[idx];
    if (m <= 0.0f) return 0.0f;
    if (!interior(i, j, k, nx, ny, nz)) return pold;
    const float d = L.diag[idx];
    const float own = fmaf(K.corr, K.p[parent_of(K, i, j, k)], pold);  // prolong_item
    float vxm, vxp, vyp, vym, vzp, vzm;
    if constexpr (RED) {
        vxm = prolonged(L, K, idx - sx, i - 1, j, k);
        vxp = prolonged(L, K, idx + sx, i + 1, j, k);
        vyp = prolonged(L, K, idx + sy, i, j + 1, k);
        vym = prolonged(L, K, idx - sy, i, j - 1, k);
        vzp = prolonged(L, K, idx + 1, i, j, k + 1);
        vzm = prolonged(L, K, idx - 1, i, j, k - 1);
    } else {
        auto at = [&](int a, int b, int c) { return r[((a - i0) * kBj + (b - j0)) * kBk + (c - k0)]; };
        vxm = at(i - 1, j, k);
        vxp = at(i + 1, j, k);
        vyp = at(i, j + 1, k);
        vym = at(i, j - 1, k);
        vzp = at(i, j, k + 1);
        vzm = at(i, j, k - 1);
    }
    float t = L.wx[idx - sx] * vxm;
    t = fmaf(L.wx[idx], vxp, t);
    t = fmaf(L.wy[idx], vyp, t);
    t = fmaf(L.wy[idx - sy], vym, t);
    t = fmaf(L.wz[idx], vzp, t);
    t = fmaf(L.wz[idx - 1], vzm, t);
    const float num = fmaf(-dd, load_b_flat(L.b, L.b16, L.bkind, L.bscale, idx), t);
    return d > 0.0f ? num / d : own;  // multigrid.cl:227 or the prolongation
}

__global__ __launch_bounds__(1024) void rbgs_post_fused_kernel(Level L, Level K, float dd, float *__restrict__ out) {
    __shared__ float r[kBi * kBj * kBk];
    const int ti = (int)blockIdx.z * kTi, tj = (int)blockIdx.y * kTj, tk = (int)blockIdx.x * kTk;
    const int i0 = ti - 1, j0 = tj - 1, k0 = tk - 1;  // box origin
    // red cells of the box: (bi, bj) pairs x kBk / 2 cells each
    for (int e = threadIdx.x; e < kBi * kBj * (kBk / 2); e += 1024) {
        const int pair = e / (kBk / 2), kk = e - pair * (kBk / 2);
        const int bi = pair / kBj, bj = pair - bi * kBj;
        const int i = i0 + bi, j = j0 + bj, k = k0 + 2 * kk + ((i + j + k0) & 1);  // i + j + k even
        if (i < 0 || i >= L.nx || j < 0 || j >= L.ny || k < 0 || k >= L.nz) continue;
        r[(bi * kBj + bj) * kBk + (k - k0)] = post_value<true>(L, K, dd, i, j, k, r, i0, j0, k0);
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kTi * kTj * kTk; e += 1024) {
        const int li = e / (kTj * kTk), rest = e - li * (kTj * kTk), lj = rest / kTk, lk = rest - lj * kTk;
        const int i = ti + li, j = tj + lj, k = tk + lk;
        if (i >= L.nx || j >= L.ny || k >= L.nz) continue;
        const float v = ((i + j + k) & 1) == 0 ? r[((li + 1) * kBj + (lj + 1)) * kBk + (lk + 1)]
                                                 : post_value<false>(L, K, dd, i, j, k, r, i0, j0, k0);
        out[gidx(i, j, k, L.ny, L.nz)] = v;
    }
}

void launch_rbgs_post_fused(const Level &L, const Level &K, float delta, float *out, hipStream_t s) {
    const dim3 grid((unsigned)((L.nz + kTk - 1) / kTk), (unsigned)((L.ny + kTj - 1) / kTj),
                    (unsigned)((L.nx + kTi - 1) / kTi));
    hipLaunchKernelGGL(rbgs_post_fused_kernel, grid, dim3(1024), 0, s, L, K, delta * delta, out);
}

void launch_rbgs_prolong(const Level &L, const Level &K, const uint8_t *op, float delta, int parity, int first,
                         const int32_t *rows, int64_t nrows, hipStream_t s) {
    const float dd = delta * delta;
    if (rows) {
        if (nrows == 0) return;
        if (op)
            hipLaunchKernelGGL(rbgs_rows_packed_prolong_kernel, dim3(row_blocks(nrows, L.nz)), dim3(256), 0, s, L, op,
                               K, dd, parity, first, rows, nrows);
        else
            hipLaunchKernelGGL(rbgs_rows_prolong_kernel, dim3(row_blocks(nrows, L.nz)), dim3(256), 0, s, L, K, dd,
                               parity, first, rows, nrows);
        return;
    }
    const int64_t work = (int64_t)L.nx * L.ny * ((L.nz + 1) >> 1);
    hipLaunchKernelGGL(rbgs_prolong_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, L, K, dd, parity,
                       first);
}

// coarse_smooth: one work-group of 1024 lanes owns the whole coarsest level
// (up to 2048 cells, two per lane).  Only p lives in LDS; each lane keeps its
// cells' weights, b, diagonal and colour in registers, so a half-sweep is six
// LDS reads, one fused chain and one division per active cell, then a
// barrier.  T = float restates coarse_smooth (multigrid.cl:232-264); T =
// double is the FP64 coarsest smoother of the mixed-precision solve (the
// same operations widened: b and the operator are exact in float).
constexpr int kCoarseThreads = 1024;
constexpr int kCoarseCells = 2 * kCoarseThreads;

template <typename T>
__device__ void coarse_smooth_body(const Level &L, T dd, int iterations) {
    __shared__ T sp[kCoarseCells];
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int C = nx * ny * nz;
    const int sx = ny * nz, sy = nz;
    int q[2], par[2];
    bool act[2];
    T w[2][6], bv[2], dv[2];
#pragma unroll
    for (int slot = 0; slot < 2; ++slot) {
        const int cell = threadIdx.x + slot * kCoarseThreads;
        q[slot] = cell;
        act[slot] = false;
        par[slot] = 0;
        if (cell < C) {
            sp[cell] = (T)L.p[cell];
            int i, j, k;
            coords(cell, ny, nz, i, j, k);
            par[slot] = (i + j + k) & 1;
            if (interior(i, j, k, nx, ny, nz) && L.mask[cell] > 0.0f && L.diag[cell] > 0.0f) {
                act[slot] = true;
                w[slot][0] = (T)L.wx[cell - sx];
                w[slot][1] = (T)L.wx[cell];
                w[slot][2] = (T)L.wy[cell];
                w[slot][3] = (T)L.wy[cell - sy];
                w[slot][4] = (T)L.wz[cell];
                w[slot][5] = (T)L.wz[cell - 1];
                bv[slot] = (T)load_b(L.b, L.b16, L.bkind, L.bscale, cell);
                dv[slot] = (T)L.diag[cell];
            }
        }
    }
    __syncthreads();
    for (int it = 0; it < iterations; ++it)
        for (int parity = 0; parity < 2; ++parity) {
#pragma unroll
            for (int slot = 0; slot < 2; ++slot) {
                if (!act[slot] || par[slot] != parity) continue;
                const int c = q[slot];
                T t = w[slot][0] * sp[c - sx];
                t = fma(w[slot][1], sp[c + sx], t);
                t = fma(w[slot][2], sp[c + sy], t);
                t = fma(w[slot][3], sp[c - sy], t);
                t = fma(w[slot][4], sp[c + 1], t);
                t = fma(w[slot][5], sp[c - 1], t);
                sp[c] = fma(-dd, bv[slot], t) / dv[slot];  // multigrid.cl:257 (fmuladd)
            }
            __syncthreads();
        }
#pragma unroll
    for (int slot = 0; slot < 2; ++slot)
        if (q[slot] < C) L.p[q[slot]] = (float)sp[q[slot]];
}

template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void coarse_smooth_kernel(Level L, T dd, int iterations) {
    coarse_smooth_body<T>(L, dd, iterations);
}

void launch_coarse_smooth(const Level &L, float delta, int iterations, hipStream_t s) {
    if (L.C > kCoarseCells) {
        // Larger coarsest levels (only for elongated grids) fall back to the
        // colour kernels, which give the same result: fluid cells only are
        // updated, and non-fluid cells are not touched by coarse_smooth.
        for (int it = 0; it < iterations; ++it)
            for (int parity = 0; parity < 2; ++parity) launch_rbgs(L, delta, parity, s);
        return;
    }
    hipLaunchKernelGGL(coarse_smooth_kernel<float>, dim3(1), dim3(kCoarseThreads), 0, s, L, delta * delta, iterations);
}

void launch_coarse_smooth_f64(const Level &L, double delta, int iterations, hipStream_t s) {
    hipLaunchKernelGGL(coarse_smooth_kernel<double>, dim3(1), dim3(kCoarseThreads), 0, s, L, delta * delta, iterations);
}

// ---- bottom of the W-cycle in one work-group --------------------------------------
// The levels below ~64k cells cost a few microseconds of dispatch each and
// almost no work.  One work-group of 1024 lanes runs the whole recursion below
// a given level (op list built by the host), each phase a strided loop over
// the level's cells with the same per-cell code as the standalone kernels,
// separated by work-group barriers; the data stays in L2.
__global__ __launch_bounds__(kCoarseThreads) void wcycle_bottom_kernel(BottomArgs a) {
    for (int o = 0; o < a.nops; ++o) {
        const BottomOp op = a.ops[o];
        const Level &L = a.lv[op.level];
        if (op.code == kOpSmooth) {
            const int64_t work = (int64_t)L.nx * L.ny * ((L.nz + 1) >> 1);
            for (int it = 0; it < op.count; ++it)
                for (int parity = 0; parity < 2; ++parity) {
                    for (int64_t t = threadIdx.x; t < work; t += 4 * kCoarseThreads)
                        rbgs_batch4(L, a.dd, parity, t, kCoarseThreads, work);
                    __syncthreads();
                }
        } else if (op.code == kOpRestrict) {
            const Level &K = a.lv[op.level + 1];
            for (int64_t ci = threadIdx.x; ci < K.C; ci += kCoarseThreads) restrict_item_eager(L, K, a.inv_delta2, ci);
            __syncthreads();
        } else if (op.code == kOpProlong) {
            const Level &K = a.lv[op.level + 1];
            for (int64_t fi = threadIdx.x; fi < L.C; fi += kCoarseThreads) prolong_item_eager(L, K, fi);
            __syncthreads();
        } else if (op.code == kOpCoarse32 || op.code == kOpCoarse64) {
            if (L.C > kCoarseCells) {  // same result as coarse_smooth (see launch_coarse_smooth)
                const int64_t work = (int64_t)L.nx * L.ny * ((L.nz + 1) >> 1);
                for (int it = 0; it < op.count; ++it)
                    for (int parity = 0; parity < 2; ++parity) {
                        for (int64_t t = threadIdx.x; t < work; t += kCoarseThreads) rbgs_item(L, a.dd, parity, t);
                        __syncthreads();
                    }
            } else {
                if (op.code == kOpCoarse32) coarse_smooth_body<float>(L, a.dd, op.count);
                else coarse_smooth_body<double>(L, a.dd64, op.count);
                __syncthreads();
            }
        }
    }
}

void launch_wcycle_bottom(const BottomArgs &a, hipStream_t s) {
    hipLaunchKernelGGL(wcycle_bottom_kernel, dim3(1), dim3(kCoarseThreads), 0, s, a);
}

// ---- the last two levels, resident in one work-group ----------------------------------
// When the bottom is level B (at most kResB cells) above the coarsest level K
// (at most kResK cells), the cycle below B is: smooth B, restrict to K,
// coarse-smooth K once or twice, prolong, smooth B.  Here B's pressure lives
// in LDS and its operator and b in registers (six cells per lane), K's cell
// tid in lane tid; B's residuals reach the restriction through LDS.  Every
// value is computed by the same operations in the same order as the kernels
// above (rbgs_cell, restrict_item, coarse_smooth_body, prolong_item), so the
// result is bit-identical; between two coarse calls the FP64 iterate is
// rounded to FP32, as the level array stores it.
//
// With a row list (bottom_rows_kernel), only B's rows that hold an interior
// fluid cell are loaded, swept and stored.  A cell of any other row is
// either not interior (the sweeps, the restriction and the prolongation
// leave it as it is) or has mask <= 0 (the sweeps set it to +0), so when
// every such row holds p = +0 on entry, which is what the restriction onto
// B writes there (and what the row-list solve's zeroing leaves in the rows
// it does not restrict onto), the pressure the kernel leaves is the same
// with and without the list.  A single work-group reads all of B through
// one CU, so B's bytes set the kernel's time: at 256^3 the rows with fluid
// are about a quarter of B.
constexpr int kResB = 6 * kCoarseThreads, kResK = kCoarseThreads;

int resident_bottom_fits(const Level &B, const Level &K) { return B.C <= kResB && K.C <= kResK; }

// rows[0 .. *count) = the rows (i, j) of B with an interior cell whose mask
// is > 0, in increasing order (one work-group)
__global__ __launch_bounds__(kCoarseThreads) void bottom_rows_kernel(Level B, int32_t *__restrict__ rows,
                                                                     int32_t *__restrict__ count) {
    __shared__ int wave_count[kCoarseThreads / 64];
    __shared__ int base;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nrows = B.nx * B.ny;
    if (tid == 0) base = 0;
    __syncthreads();
    for (int r0 = 0; r0 < nrows; r0 += kCoarseThreads) {
        const int r = r0 + tid;
        bool active = false;
        if (r < nrows) {
            const int i = r / B.ny, j = r - i * B.ny;
            if (i > 0 && i < B.nx - 1 && j > 0 && j < B.ny - 1)
                for (int k = 1; k < B.nz - 1; ++k) active |= B.mask[(int64_t)r * B.nz + k] > 0.0f;
        }
        const uint64_t bal = __ballot(active);
        if (lane == 0) wave_count[wave] = (int)__popcll(bal);
        __syncthreads();
        int at = base + (int)__popcll(bal & ((1ull << lane) - 1ull));
        for (int w = 0; w < wave; ++w) at += wave_count[w];
        if (active) rows[at] = r;
        __syncthreads();
        if (tid == 0)
            for (int w = 0; w < kCoarseThreads / 64; ++w) base += wave_count[w];
        __syncthreads();
    }
    if (tid == 0) *count = base;
}

void launch_bottom_rows(const Level &B, int32_t *rows, int32_t *count, hipStream_t s) {
    hipLaunchKernelGGL(bottom_rows_kernel, dim3(1), dim3(kCoarseThreads), 0, s, B, rows, count);
}

template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void bottom_resident_kernel(Level B, Level K, float dd, T ddk,
                                                                        float inv_delta2, int pre, int post,
                                                                        int coarse_iters, int coarse_calls,
                                                                        const int32_t *__restrict__ brows,
                                                                        const int32_t *__restrict__ bcount) {
    __shared__ float pB[kResB];
    __shared__ float rB[kResB];
    __shared__ T pK[kResK];
    const int tid = threadIdx.x;
    const int bnx = B.nx, bny = B.ny, bnz = B.nz, BC = (int)B.C;
    const int bsx = bny * bnz, bsy = bnz;
    // level B: item t = tid + s * 1024 is cell t, or with the row list cell
    // t % bnz of the listed row t / bnz; cell BC marks an item past the end.
    // Slot s is live (block-uniform) when item s * 1024 is.  cf[s] holds the
    // item's cell << 4 and its flags: bit 0 parity, 1 zero (mask <= 0), 2
    // update (interior, fluid, d > 0), 3 fluid (interior, mask > 0)
    const int nact = brows ? *bcount * bnz : BC;
    int cf[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int t = tid + s * kCoarseThreads;
        if (brows) {
            const int ar = t / bnz;
            const int row = brows[min(ar, bnx * bny - 1)];  // in the list's bounds; used when t < nact
            cf[s] = (t < nact ? row * bnz + (t - ar * bnz) : BC) << 4;
        } else {
            cf[s] = (t < BC ? t : BC) << 4;
        }
        // rows off the list hold +0 (B's pressure on entry there), and so
        // do their residuals
        if (t < BC) {
            pB[t] = 0.0f;
            rB[t] = 0.0f;
        }
    }
    auto live = [&](int s) { return s * kCoarseThreads < nact; };
    __syncthreads();
    float w[6][6], bv[6], dv[6];
    // every load of a live slot unconditional (indices clamped into the
    // arrays; the values of cells that do not use them are dropped) and
    // issued for all slots before any is used: one memory round trip
    float lp[6], lm[6], ld[6], lb[6], lw[6][6];
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        lp[s] = lm[s] = ld[s] = lb[s] = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; ++q) lw[s][q] = 0.0f;
        if (!live(s)) continue;
        const int cc = (cf[s] >> 4) < BC ? (cf[s] >> 4) : BC - 1;
        int i, j, k;
        coords(cc, bny, bnz, i, j, k);
        const bool inner = interior(i, j, k, bnx, bny, bnz);
        const int cxm = inner ? cc - bsx : cc, cym = inner ? cc - bsy : cc, czm = inner ? cc - 1 : cc;
        lp[s] = B.p[cc];
        lm[s] = B.mask[cc];
        ld[s] = B.diag[cc];
        lb[s] = load_b(B.b, B.b16, B.bkind, B.bscale, cc);
        lw[s][0] = B.wx[cxm];
        lw[s][1] = B.wx[cc];
        lw[s][2] = B.wy[cc];
        lw[s][3] = B.wy[cym];
        lw[s][4] = B.wz[cc];
        lw[s][5] = B.wz[czm];
    }
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int c = cf[s] >> 4;
        int fl = 0;
        bv[s] = 0.0f;
        dv[s] = 1.0f;
#pragma unroll
        for (int q = 0; q < 6; ++q) w[s][q] = 0.0f;
        if (live(s) && c < BC) {
            int i, j, k;
            coords(c, bny, bnz, i, j, k);
            pB[c] = lp[s];
            fl = (i + j + k) & 1;
            if (lm[s] <= 0.0f) {
                fl |= 2;
            } else if (interior(i, j, k, bnx, bny, bnz)) {
                fl |= 8;
                dv[s] = ld[s];
                if (ld[s] > 0.0f) fl |= 4;
                bv[s] = lb[s];
#pragma unroll
                for (int q = 0; q < 6; ++q) w[s][q] = lw[s][q];
            }
        }
        cf[s] |= fl;
    }
    // level K: cell tid
    const int knx = K.nx, kny = K.ny, knz = K.nz, KC = (int)K.C;
    const int ksx = kny * knz, ksy = knz;
    int ka, kb, kc;
    const int kt = tid < KC ? tid : KC - 1;
    coords(kt, kny, knz, ka, kb, kc);
    const bool kin = tid < KC && interior(ka, kb, kc, knx, kny, knz);
    bool kact = false;
    T kw[6] = {T(0), T(0), T(0), T(0), T(0), T(0)}, kd = T(1);
    {
        const int kxm = kin ? kt - ksx : kt, kym = kin ? kt - ksy : kt, kzm = kin ? kt - 1 : kt;
        const float m = K.mask[kt], d = K.diag[kt];
        const float w0 = K.wx[kxm], w1 = K.wx[kt], w2 = K.wy[kt], w3 = K.wy[kym], w4 = K.wz[kt], w5 = K.wz[kzm];
        if (kin && m > 0.0f && d > 0.0f) {
            kact = true;
            kw[0] = (T)w0;
            kw[1] = (T)w1;
            kw[2] = (T)w2;
            kw[3] = (T)w3;
            kw[4] = (T)w4;
            kw[5] = (T)w5;
            kd = (T)d;
        }
    }
    const int kpar = (ka + kb + kc) & 1;
    __syncthreads();

    auto smooth_b = [&](int iterations) {
        for (int it = 0; it < iterations; ++it)
            for (int parity = 0; parity < 2; ++parity) {
#pragma unroll
                for (int s = 0; s < 6; ++s) {
                    int c = cf[s] >> 4;
                    // the neighbours' LDS addresses are formed here, not
                    // kept in registers across the sweeps
                    asm volatile("" : "+v"(c));
                    if (!live(s) || c >= BC || (cf[s] & 1) != parity) continue;
                    if (cf[s] & 2) {
                        pB[c] = 0.0f;
                    } else if (cf[s] & 4) {  // wsum (multigrid.cl:7-23) and :227
                        float t = w[s][0] * pB[c - bsx];
                        t = fmaf(w[s][1], pB[c + bsx], t);
                        t = fmaf(w[s][2], pB[c + bsy], t);
                        t = fmaf(w[s][3], pB[c - bsy], t);
                        t = fmaf(w[s][4], pB[c + 1], t);
                        t = fmaf(w[s][5], pB[c - 1], t);
                        pB[c] = fmaf(-dd, bv[s], t) / dv[s];
                    }
                }
                __syncthreads();
            }
    };

    smooth_b(pre);
    // restriction: residuals of B's interior fluid cells (+0 elsewhere, which
    // leaves the child sums unchanged), then the sum over the eight children
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int c = cf[s] >> 4;
        if (!live(s) || c >= BC) continue;
        float r = 0.0f;
        if (cf[s] & 8) {
            float t = w[s][0] * pB[c - bsx];
            t = fmaf(w[s][1], pB[c + bsx], t);
            t = fmaf(w[s][2], pB[c + bsy], t);
            t = fmaf(w[s][3], pB[c - bsy], t);
            t = fmaf(w[s][4], pB[c + 1], t);
            t = fmaf(w[s][5], pB[c - 1], t);
            r = bv[s] - fmaf(-dv[s], pB[c], t) * inv_delta2;
        }
        rB[c] = r;
    }
    __syncthreads();
    T kbv = T(0);
    if (tid < KC) {
        float rhs = 0.0f;
        if (kin) {
            const int f0 = 1 + 2 * (ka - 1), f1 = 1 + 2 * (kb - 1), f2 = 1 + 2 * (kc - 1);
            for (int q = 0; q < 8; ++q) {
                const int i = f0 + (q >> 2), j = f1 + ((q >> 1) & 1), k = f2 + (q & 1);
                if (interior(i, j, k, bnx, bny, bnz)) rhs = rhs + rB[gidx(i, j, k, bny, bnz)];
            }
        }
        store_b(K.b, K.b16, K.bkind, K.bscale, tid, rhs);
        kbv = (T)stored_b(K.bkind, K.bscale, rhs);
        pK[tid] = T(0);
    }
    __syncthreads();
    // coarsest level (coarse_smooth_body), coarse_calls calls of coarse_iters sweeps
    for (int call = 0; call < coarse_calls; ++call) {
        for (int it = 0; it < coarse_iters; ++it)
            for (int parity = 0; parity < 2; ++parity) {
                if (kact && kpar == parity) {
                    const int c = tid;
                    T t = kw[0] * pK[c - ksx];
                    t = fma(kw[1], pK[c + ksx], t);
                    t = fma(kw[2], pK[c + ksy], t);
                    t = fma(kw[3], pK[c - ksy], t);
                    t = fma(kw[4], pK[c + 1], t);
                    t = fma(kw[5], pK[c - 1], t);
                    pK[c] = fma(-ddk, kbv, t) / kd;  // multigrid.cl:257
                }
                __syncthreads();
            }
        if (tid < KC) pK[tid] = (T)(float)pK[tid];  // the level array holds FP32
        __syncthreads();
    }
    // prolongation (prolong_item)
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int c = cf[s] >> 4;
        if (!live(s) || c >= BC || !(cf[s] & 8)) continue;
        int i, j, k;
        coords(c, bny, bnz, i, j, k);
        const int a = min((i - 1) / 2 + 1, knx - 2), b = min((j - 1) / 2 + 1, kny - 2), cc = min((k - 1) / 2 + 1, knz - 2);
        pB[c] = fmaf(K.corr, (float)pK[gidx(a, b, cc, kny, knz)], pB[c]);
    }
    __syncthreads();
    smooth_b(post);
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        const int c = cf[s] >> 4;
        if (live(s) && c < BC) B.p[c] = pB[c];
    }
    if (tid < KC) K.p[tid] = (float)pK[tid];
}

void launch_bottom_resident(const Level &B, const Level &K, float delta, bool f64, int pre, int post, int coarse_iters,
                            int coarse_calls, const int32_t *rows, const int32_t *count, hipStream_t s) {
    const float dd = delta * delta;
    if (f64)
        hipLaunchKernelGGL(bottom_resident_kernel<double>, dim3(1), dim3(kCoarseThreads), 0, s, B, K, dd,
                           (double)delta * (double)delta, 1.0f / dd, pre, post, coarse_iters, coarse_calls, rows,
                           count);
    else
        hipLaunchKernelGGL(bottom_resident_kernel<float>, dim3(1), dim3(kCoarseThreads), 0, s, B, K, dd, dd, 1.0f / dd,
                           pre, post, coarse_iters, coarse_calls, rows, count);
}

int coarse_f64_capacity() { return kCoarseCells; }

// ---- FP64 iterative-refinement update -----------------------------------------------
// p_new = p_old + e on interior fluid cells (0 elsewhere), r = b - A p_new in
// FP64 with left-to-right sums over the non-zero weights, r32 = float(r), and
// per-block partial sums of r^2 and b^2.  With from32 the iterate is the FP32
// pressure (start of a solve).  Neighbour values are recomputed from p_old
// and e, so p_new goes to a separate buffer.
constexpr int kNormThreadsIR = 256;
constexpr int kNormBlocksIR = 2048;

__global__ __launch_bounds__(kNormThreadsIR) void ir_update_residual_kernel(
    int from32, const float *__restrict__ p32, const double *__restrict__ pin, double *__restrict__ pout,
    float *__restrict__ r32, const float *__restrict__ b, const float *__restrict__ mask,
    const float *__restrict__ diag, const float *__restrict__ wx, const float *__restrict__ wy,
    const float *__restrict__ wz, double inv_h2, int nx, int ny, int nz, double *__restrict__ partials) {
    const int64_t C = (int64_t)nx * ny * nz;
    const int64_t sx = (int64_t)ny * nz, sy = nz;
    double rr = 0.0, bb = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * kNormThreadsIR + threadIdx.x; q < C;
         q += (int64_t)gridDim.x * kNormThreadsIR) {
        int i, j, k;
        coords(q, ny, nz, i, j, k);
        const bool fluid = interior(i, j, k, nx, ny, nz) && mask[q] > 0.0f;
        double r = 0.0;
        const float bq = b[q];
        if (fluid) {
            auto pnew = [&](int64_t x) -> double { return from32 ? (double)p32[x] : pin[x] + (double)p32[x]; };
            const double pc = pnew(q);
            double s = 0.0;
            if (wx[q] != 0.0f) s = s + (double)wx[q] * pnew(q + sx);
            if (wx[q - sx] != 0.0f) s = s + (double)wx[q - sx] * pnew(q - sx);
            if (wy[q] != 0.0f) s = s + (double)wy[q] * pnew(q + sy);
            if (wy[q - sy] != 0.0f) s = s + (double)wy[q - sy] * pnew(q - sy);
            if (wz[q] != 0.0f) s = s + (double)wz[q] * pnew(q + 1);
            if (wz[q - 1] != 0.0f) s = s + (double)wz[q - 1] * pnew(q - 1);
            pout[q] = pc;
            r = (double)bq - (s - (double)diag[q] * pc) * inv_h2;
        } else {
            pout[q] = 0.0;
        }
        r32[q] = (float)r;
        rr += r * r;
        bb += (double)bq * (double)bq;
    }
    __shared__ double s_r[kNormThreadsIR], s_b[kNormThreadsIR];
    s_r[threadIdx.x] = rr;
    s_b[threadIdx.x] = bb;
    __syncthreads();
    for (int off = kNormThreadsIR / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s_r[threadIdx.x] += s_r[threadIdx.x + off];
            s_b[threadIdx.x] += s_b[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s_r[0];
        partials[kNormBlocksIR + blockIdx.x] = s_b[0];
    }
}

__global__ __launch_bounds__(256) void narrow_kernel(const double *__restrict__ p64, float *__restrict__ p32,
                                                     int64_t C) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q < C) p32[q] = (float)p64[q];
}

void launch_widen_narrow(const double *p64, float *p32, int64_t C, hipStream_t s) {
    hipLaunchKernelGGL(narrow_kernel, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, s, p64, p32, C);
}

// narrow_kernel on the cells of the listed rows (item t: cell t mod nz of
// row t / nz)
__global__ __launch_bounds__(256) void narrow_rows_kernel(const double *__restrict__ p64, float *__restrict__ p32,
                                                          const int32_t *__restrict__ rows, int64_t work, int nz) {
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= work) return;
    const int64_t a = t / nz, q = (int64_t)rows[a] * nz + (t - a * nz);
    p32[q] = (float)p64[q];
}

void launch_narrow_rows(const double *p64, float *p32, const int32_t *rows, int64_t nrows, int nz, hipStream_t s) {
    const int64_t work = nrows * nz;
    if (work == 0) return;
    hipLaunchKernelGGL(narrow_rows_kernel, dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, p64, p32, rows,
                       work, nz);
}

// ---- FP64 residual norm -------------------------------------------------------------
constexpr int kNormThreads = 256;
constexpr int kNormBlocks = 2048;

int residual_partials_count(int64_t) { return 2 * kNormBlocks; }

__global__ __launch_bounds__(kNormThreads) void residual_norm_kernel(
    const float *__restrict__ p, const float *__restrict__ b, const float *__restrict__ mask,
    const float *__restrict__ diag, const float *__restrict__ wx, const float *__restrict__ wy,
    const float *__restrict__ wz, double inv_h2, int nx, int ny, int nz, double *__restrict__ partials) {
    const int64_t C = (int64_t)nx * ny * nz;
    const int64_t sx = (int64_t)ny * nz, sy = nz;
    double rr = 0.0, bb = 0.0;
    for (int64_t idx = (int64_t)blockIdx.x * kNormThreads + threadIdx.x; idx < C;
         idx += (int64_t)gridDim.x * kNormThreads) {
        int i, j, k;
        coords(idx, ny, nz, i, j, k);
        const double bv = b[idx];
        double ap = 0.0;
        if (interior(i, j, k, nx, ny, nz) && mask[idx] > 0.0f) {
            const double s = (double)wx[idx] * p[idx + sx] + (double)wx[idx - sx] * p[idx - sx] +
                             (double)wy[idx] * p[idx + sy] + (double)wy[idx - sy] * p[idx - sy] +
                             (double)wz[idx] * p[idx + 1] + (double)wz[idx - 1] * p[idx - 1];
            ap = (s - (double)diag[idx] * p[idx]) * inv_h2;
        }
        const double r = bv - ap;
        rr += r * r;
        bb += bv * bv;
    }
    __shared__ double s_r[kNormThreads], s_b[kNormThreads];
    s_r[threadIdx.x] = rr;
    s_b[threadIdx.x] = bb;
    __syncthreads();
    for (int off = kNormThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s_r[threadIdx.x] += s_r[threadIdx.x + off];
            s_b[threadIdx.x] += s_b[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s_r[0];
        partials[kNormBlocks + blockIdx.x] = s_b[0];
    }
}

// out[0], out[1] = the sums of the partials of r.r and b.b (SQRT: their
// square roots, the norms)
template <bool SQRT = false>
__global__ __launch_bounds__(kNormThreads) void reduce_partials_kernel(const double *__restrict__ partials, int count,
                                                                       double *__restrict__ out) {
    __shared__ double s_r[kNormThreads], s_b[kNormThreads];
    double rr = 0.0, bb = 0.0;
    for (int q = threadIdx.x; q < count; q += kNormThreads) {
        rr += partials[q];
        bb += partials[kNormBlocks + q];
    }
    s_r[threadIdx.x] = rr;
    s_b[threadIdx.x] = bb;
    __syncthreads();
    for (int off = kNormThreads / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            s_r[threadIdx.x] += s_r[threadIdx.x + off];
            s_b[threadIdx.x] += s_b[threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = SQRT ? sqrt(s_r[0]) : s_r[0];
        out[1] = SQRT ? sqrt(s_b[0]) : s_b[0];
    }
}

void launch_ir_update_residual(const Level &L, int from32, const float *p32, const double *pin, double *pout,
                               float *r32, const float *b, double inv_h2, double *partials, double *out2,
                               hipStream_t s) {
    int blocks = (int)((L.C + kNormThreadsIR - 1) / kNormThreadsIR);
    if (blocks > kNormBlocksIR) blocks = kNormBlocksIR;
    hipLaunchKernelGGL(ir_update_residual_kernel, dim3(blocks), dim3(kNormThreadsIR), 0, s, from32, p32, pin, pout, r32,
                       b, L.mask, L.diag, L.wx, L.wy, L.wz, inv_h2, L.nx, L.ny, L.nz, partials);
    hipLaunchKernelGGL(reduce_partials_kernel<false>, dim3(1), dim3(kNormThreads), 0, s, partials, blocks, out2);
}

void launch_residual_norm(const Level &L, double inv_h2, double *partials, double *out2, hipStream_t s) {
    int blocks = (int)((L.C + kNormThreads - 1) / kNormThreads);
    if (blocks > kNormBlocks) blocks = kNormBlocks;
    hipLaunchKernelGGL(residual_norm_kernel, dim3(blocks), dim3(kNormThreads), 0, s, L.p, L.b, L.mask, L.diag, L.wx,
                       L.wy, L.wz, inv_h2, L.nx, L.ny, L.nz, partials);
    hipLaunchKernelGGL(reduce_partials_kernel<false>, dim3(1), dim3(kNormThreads), 0, s, partials, blocks, out2);
}

}  // namespace clq

// ---- FP64 residual norm without allocation ----------------------------------------
// clq_residual_norm_noalloc keeps its per-block partial sums (2 x kNormBlocks
// doubles, 32 KiB, enough for any grid: the norm kernel is grid-strided over
// at most kNormBlocks blocks) in this file-scope device array.  Each GPU holds
// its own copy until the process ends; its device address is looked up once
// per device (the array's name in host code is only a stand-in).  Every block
// overwrites its two partials before the reduction reads them, so nothing has
// to be cleared between calls; calls that overlap on one GPU race on it.
static __device__ __align__(256) double scratch[2 * clq::kNormBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
double *g_scratch[kMaxDevices] = {};

// scratch's device address on the current device (set up on first use)
int scratch_on_device(double **out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return clq::set_error(CLQ_ERR_DEVICE, "residual_norm_noalloc: no usable current device (%d)", dev);
    std::lock_guard<std::mutex> guard(g_scratch_lock);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return clq::set_error(CLQ_ERR_DEVICE, "residual_norm_noalloc: hipGetSymbolAddress: %s",
                                  hipGetErrorString(e));
        g_scratch[dev] = (double *)ptr;
    }
    *out = g_scratch[dev];
    return CLQ_OK;
}
}  // namespace

extern "C" int clq_residual_norm_noalloc(const float *p, const float *rhs, const float *mask, const float *diag,
                                         const float *wx, const float *wy, const float *wz, int32_t nx, int32_t ny,
                                         int32_t nz, double dx, double *norms, void *stream) {
    using namespace clq;
    if (!p || !rhs || !mask || !diag || !wx || !wy || !wz || !norms)
        return set_error(CLQ_ERR_INVALID, "residual_norm_noalloc: NULL array");
    if (nx < 3 || ny < 3 || nz < 3)
        return set_error(CLQ_ERR_INVALID, "residual_norm_noalloc: grid %dx%dx%d too small (need >= 3)", nx, ny, nz);
    if ((int64_t)nx * ny * nz >= (1LL << 31))
        return set_error(CLQ_ERR_INVALID,
                         "residual_norm_noalloc: grid of %lld cells above the maximum of 2^31 - 1 cells",
                         (long long)nx * ny * nz);
    if (!(dx > 0.0)) return set_error(CLQ_ERR_INVALID, "residual_norm_noalloc: dx must be positive");
    double *partials = nullptr;
    if (int rc = scratch_on_device(&partials)) return rc;
    const hipStream_t st = (hipStream_t)stream;
    const double h = (double)(float)dx;  // the operator's cell size, as in clq_residual_norm
    const int64_t C = (int64_t)nx * ny * nz;
    int blocks = (int)((C + kNormThreads - 1) / kNormThreads);
    if (blocks > kNormBlocks) blocks = kNormBlocks;
    hipLaunchKernelGGL(residual_norm_kernel, dim3(blocks), dim3(kNormThreads), 0, st, p, rhs, mask, diag, wx, wy, wz,
                       1.0 / (h * h), nx, ny, nz, partials);
    hipLaunchKernelGGL(reduce_partials_kernel<true>, dim3(1), dim3(kNormThreads), 0, st, partials, blocks, norms);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(CLQ_ERR_DEVICE, "residual_norm_noalloc: %s", hipGetErrorString(e));
    return CLQ_OK;
}
