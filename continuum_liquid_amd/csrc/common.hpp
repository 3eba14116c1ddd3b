// Shared declarations of the gfx950 HIP implementation of include/cliquid.h.
//
// Layout in HBM (one context = one simulation grid):
//   grid arrays      float32[C], C = nx*ny*nz, k fastest (reference helper.cl:4-6)
//   particles        AoS float32[n][3] for position, velocity, work copies and
//                    the previous positions (reference solver.py:161-217)
//   bins             int32 counts/offsets[C], int32 key/rank/perm[capacity]
//   multigrid        7 float32 arrays per coarse level (p, b, mask, diagonal,
//                    weight x/y/z), allocated once at context creation
//
// Every kernel is compiled with -ffp-contract=off: a fused multiply-add
// appears only where written as fmaf(), at exactly the sites where the
// reference OpenCL source contracts (llvm.fmuladd), so results can be compared
// bit for bit with the CPU oracle.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cliquid.h"

namespace clq {

// Work-groups are dealt round-robin to the 8 XCDs, each with its own L2.
// Block b of nb takes item xcd_item(b, nb), a bijection of [0, nb) that gives
// every XCD one contiguous range of items, so a stencil finds the neighbouring
// rows and planes it reads in the L2 of its own XCD.
// decoded weight bit of a packed level-0 operator byte
__device__ __forceinline__ float bitw(uint32_t o, uint32_t bit) { return (o & bit) ? 1.0f : 0.0f; }

__device__ __forceinline__ uint32_t xcd_item(uint32_t b, uint32_t nb) {
    const uint32_t q = nb >> 3, rem = nb & 7, x = b & 7, y = b >> 3;
    return x * q + (x < rem ? x : rem) + y;
}

constexpr int kMaxLevels = 16;

struct Level {
    int nx = 0, ny = 0, nz = 0;
    int64_t C = 0;
    float *p = nullptr, *b = nullptr, *mask = nullptr, *diag = nullptr;
    float *wx = nullptr, *wy = nullptr, *wz = nullptr;
    // 16-bit storage of b for the mixed-precision solve: bkind 0 = float b,
    // 1 = fp16, 2 = bf16, stored value b / bscale (bscale a power of two).
    uint16_t *b16 = nullptr;
    int bkind = 0;
    float bscale = 1.0f;
    // weight of this level's pressure when it is prolongated into the finer
    // level: p_fine = fmaf(corr, p[parent], p_fine), which for corr = 1 is
    // the plain sum p_fine + p[parent] bit for bit
    float corr = 1.0f;
};

// b of a level through its storage format (uniform branch per launch).
__device__ __forceinline__ float bits_to_f16f(uint16_t h) { return __half2float(__ushort_as_half(h)); }
__device__ __forceinline__ float bits_to_bf16f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ float load_b(const float *__restrict__ b, const uint16_t *__restrict__ b16, int kind,
                                        float scale, int64_t i) {
    if (kind == 0) return b[i];
    const float v = kind == 1 ? bits_to_f16f(b16[i]) : bits_to_bf16f(b16[i]);
    return v * scale;
}
// load_b without branches: both storage forms are read (the 16-bit one
// through the float array when there is none) and the value selected, so a
// batch of such loads can all be in flight together.
__device__ __forceinline__ float load_b_flat(const float *__restrict__ b, const uint16_t *__restrict__ b16, int kind,
                                             float scale, int64_t i) {
    const uint16_t *h = b16 ? b16 : reinterpret_cast<const uint16_t *>(b);
    const float f = b[i];
    const uint16_t raw = h[i];
    const float v = (kind == 1 ? bits_to_f16f(raw) : bits_to_bf16f(raw)) * scale;
    return kind == 0 ? f : v;
}
// The 16-bit storage of v (kind 1 fp16, 2 bf16): v / scale rounded to
// nearest even (NaN kept quiet).
__device__ __forceinline__ uint16_t encode_b16(int kind, float scale, float v) {
    const float x = v * (1.0f / scale);
    if (kind == 1) return __half_as_ushort(__float2half_rn(x));
    const uint32_t u = __float_as_uint(x);
    return (u & 0x7fffffffu) > 0x7f800000u ? (uint16_t)((u >> 16) | 0x40u)
                                           : (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ void store_b(float *__restrict__ b, uint16_t *__restrict__ b16, int kind, float scale,
                                        int64_t i, float v) {
    if (kind == 0) b[i] = v;
    else b16[i] = encode_b16(kind, scale, v);
}

// What load_b returns after store_b(v): the stored value, without the round
// trip through memory.
__device__ __forceinline__ float stored_b(int kind, float scale, float v) {
    if (kind == 0) return v;
    const uint16_t h = encode_b16(kind, scale, v);
    return (kind == 1 ? bits_to_f16f(h) : bits_to_bf16f(h)) * scale;
}

// ---- device helpers (reference helper.cl:4-23) -----------------------------
__device__ __forceinline__ int64_t gidx(int i, int j, int k, int ny, int nz) {
    return ((int64_t)i * ny + j) * nz + k;
}
__device__ __forceinline__ bool interior(int i, int j, int k, int nx, int ny, int nz) {
    return i > 0 && j > 0 && k > 0 && i < nx - 1 && j < ny - 1 && k < nz - 1;
}
// Grids have fewer than 2^31 cells (checked at context creation), so the
// index arithmetic runs in 32-bit unsigned integers.
__device__ __forceinline__ void coords(int64_t index, int ny, int nz, int &i, int &j, int &k) {
    const uint32_t q = (uint32_t)index, yz = (uint32_t)ny * (uint32_t)nz;
    const uint32_t a = q / yz, r = q - a * yz;
    const uint32_t b = r / (uint32_t)nz;
    i = (int)a;
    j = (int)b;
    k = (int)(r - b * (uint32_t)nz);
}
// convert_int_rtn + clamp; NaN and out-of-range inputs clamp (NaN to lo:
// fmaxf returns the other operand), without branches.
__device__ __forceinline__ int floor_clamp(float t, int lo, int hi) {
    return (int)fminf(fmaxf(floorf(t), (float)lo), (float)hi);
}
// convert_int_rtz + clamp.
// trunc(t) clamped to [lo, hi] (lo for NaN), without branches: fmaxf
// returns the other operand for a NaN
__device__ __forceinline__ int trunc_clamp(float t, int lo, int hi) {
    return (int)fminf(fmaxf(truncf(t), (float)lo), (float)hi);
}

// ---- host launchers ----------------------------------------------------------
// binning.hip
void launch_bin_count(const float *pos, int n, float inv_dx, int nx, int ny, int nz, int32_t *key,
                      int32_t *rank, int32_t *counts, hipStream_t s);
int64_t scan_scratch_elems(int64_t count);
void launch_exclusive_scan(const int32_t *in, int32_t *out, int64_t count, int32_t *scratch, hipStream_t s);
// also sets *big_count to 0 for launch_bin_order
void launch_bin_permute(const int32_t *key, const int32_t *rank, const int32_t *offsets, int n, int32_t *perm,
                        int32_t *big_count, hipStream_t s);
// stable order within bins and the move of positions and velocities; bins of
// more than 64 members go through big_list (capacity big_bin_capacity(n),
// *big_count zeroed by launch_bin_permute) and are merge-sorted using scratch
// (n entries, free at this point)
int big_bin_capacity(int64_t n);
void launch_bin_order(const int32_t *perm, const int32_t *offsets, const int32_t *counts, int64_t C, int n,
                      const float *pos, const float *vel, float *pos_out, float *vel_out, int32_t *big_list,
                      int32_t *big_count, int big_cap, int32_t *scratch, hipStream_t s);

// p2g.hip
void launch_p2g(const float *pos, const float *vel, const int32_t *offsets, const int32_t *counts, float *u,
                float *v, float *w, float *ff, float inv_dx, float inv_ppc, int nx, int ny, int nz, hipStream_t s);

// grid.hip
void launch_grid_obstacle_bc(float *u, float *v, float *w, float *ff, const uint8_t *ob, int nx, int ny, int nz,
                             hipStream_t s);
void launch_grid_domain_bc(float *u, float *v, float *w, int nx, int ny, int nz, hipStream_t s);
void launch_update_velocity(const float *u, const float *v, const float *w, float *ru, float *rv, float *rw,
                            const float *ff, float ax, float ay, float az, float dt, float nudt, float inv_dx2,
                            int nx, int ny, int nz, hipStream_t s);
void launch_pressure_rhs(const float *u, const float *v, const float *w, const float *ff, float *rhs,
                         float rho_over_dt, float inv_dx, int nx, int ny, int nz, hipStream_t s);
// op, flags (may be NULL): also the packed level-0 operator and row flags
void launch_build_finest_operator(const float *mask, float *diag, float *wx, float *wy, float *wz, uint8_t *op,
                                  int32_t *flags, int nx, int ny, int nz, hipStream_t s);
void launch_neumann_bc(float *p, int nx, int ny, int nz, hipStream_t s);
void launch_pressure_correction(float *u, float *v, float *w, const float *p, float dt_over_rho, float inv_dx,
                                int nx, int ny, int nz, hipStream_t s);
// flags (may be NULL): level-0 row flags of a pressure that is 0 off the
// interior fluid cells; rows that cannot change are skipped
void launch_pressure_correction64(float *u, float *v, float *w, const double *p, const int32_t *flags,
                                  double dt_over_rho, double inv_dx, int nx, int ny, int nz, hipStream_t s);

// multigrid.hip
void launch_rbgs(const Level &L, float delta, int parity, hipStream_t s);
void launch_coarse_operator(const Level &F, const Level &K, hipStream_t s);
// red: also make the first (red) half-sweep of the coarse level's pre-smoothing
void launch_restrict_residual(const Level &F, const Level &K, float inv_delta2, float delta, int red, hipStream_t s);
void launch_prolongate(const Level &F, const Level &K, hipStream_t s);
void launch_coarse_smooth(const Level &L, float delta, int iterations, hipStream_t s);
void launch_coarse_smooth_f64(const Level &L, double delta, int iterations, hipStream_t s);
int coarse_f64_capacity();
// One work-group runs a list of multigrid phases on small levels.
enum { kOpSmooth = 0, kOpRestrict = 1, kOpProlong = 2, kOpCoarse32 = 3, kOpCoarse64 = 4 };
struct BottomOp {
    int code, level, count;
};
constexpr int kMaxBottomOps = 96;
struct BottomArgs {
    Level lv[kMaxLevels];
    BottomOp ops[kMaxBottomOps];
    int nops;
    float dd, inv_delta2;
    double dd64;
};
void launch_wcycle_bottom(const BottomArgs &a, hipStream_t s);
int resident_bottom_fits(const Level &B, const Level &K);
// rows / count: B's rows holding an interior fluid cell (launch_bottom_rows),
// with every other row of B at p = +0; NULL: all of B
void launch_bottom_resident(const Level &B, const Level &K, float delta, bool f64, int pre, int post, int coarse_iters,
                            int coarse_calls, const int32_t *rows, const int32_t *count, hipStream_t s);
void launch_bottom_rows(const Level &B, int32_t *rows, int32_t *count, hipStream_t s);
// active rows (cells (i, j, *) holding an interior cell with mask > 0)
void launch_row_flags(const Level &L, int32_t *flags, hipStream_t s);
void launch_coarse_row_flags(const int32_t *fine, const Level &F, const Level &K, int32_t *flags, hipStream_t s);
void launch_compact_rows(const int32_t *flags, const int32_t *pos, int64_t rows, int32_t *list, int32_t *count,
                         hipStream_t s);
void launch_zero_inactive_rows(const Level &L, const int32_t *flags, int coarse, hipStream_t s);
void launch_rbgs_rows(const Level &L, float delta, int parity, const int32_t *rows, int64_t nrows, hipStream_t s);
void launch_zero_rows(float *p, const int32_t *rows, int64_t nrows, int nz, hipStream_t s);
void launch_pack_operator(const Level &L, uint8_t *op, int *bad, hipStream_t s);
void launch_rbgs_rows_packed(const Level &L, const uint8_t *op, float delta, int parity, const int32_t *rows,
                             int64_t nrows, hipStream_t s);
void launch_restrict_rows(const Level &F, const uint8_t *op, const Level &K, float inv_delta2, const int32_t *krows,
                          int64_t nkrows, float delta, int red, hipStream_t s);
// One half-sweep of L with the prolongation from K folded in (first: the red
// half-sweep, run before the black one); rows / op as for the row sweeps,
// rows == NULL for the dense form.
void launch_rbgs_prolong(const Level &L, const Level &K, const uint8_t *op, float delta, int parity, int first,
                         const int32_t *rows, int64_t nrows, hipStream_t s);
void launch_prolongate_rows(const Level &F, const Level &K, const int32_t *frows, int64_t nfrows, hipStream_t s);
// both post-smoothing half-sweeps of a dense level with the prolongation from K
// folded in; writes every cell of L into out (L.p is only read): the caller
// makes out the level's pressure array
void launch_rbgs_post_fused(const Level &L, const Level &K, float delta, float *out, hipStream_t s);
void launch_ir_update_residual(const Level &L, int from32, const float *p32, const double *pin, double *pout,
                               float *r32, const float *b, double inv_h2, double *partials, double *out2,
                               hipStream_t s);
void launch_widen_narrow(const double *p64, float *p32, int64_t C, hipStream_t s);
void launch_narrow_rows(const double *p64, float *p32, const int32_t *rows, int64_t nrows, int nz, hipStream_t s);
int residual_partials_count(int64_t C);
void launch_residual_norm(const Level &L, double inv_h2, double *partials, double *out2, hipStream_t s);

// fcg.hip
int fcg_partials_count();
// x = p32 on the interior fluid cells, r = b - A x, r32; r.r and b.b.  With
// the level-0 row flags (flags, rows, nrows; else NULL): the listed active
// rows get the full update (op: packed operator, may be NULL), an inactive
// row only x = 0, and only when x_rows marks it as possibly non-zero or
// x_rows_known is 0; x_rows then holds the row flags.
void launch_fcg_init(const Level &L, const uint8_t *op, const float *p32, const float *b, double inv_h2, double *x,
                     double *r, float *r32, const int32_t *flags, const int32_t *rows, int64_t nrows,
                     int32_t *x_rows, int x_rows_known, double *partials, double *s, hipStream_t st);
// the dots z.(A d_old), z.r, z.w (w = A z) -> beta, d.A d of the new
// direction, alpha
void launch_fcg_dots(const Level &L, const uint8_t *op, const float *z, const float *dold, const double *r, int first,
                     double inv_h2, const int32_t *rows, int64_t nrows, double *partials, double *s, hipStream_t st);
// dnew = float(z + beta dold), q = A dnew, x += alpha dnew, r -= alpha q and
// r.r (dnew != dold); e != NULL (with op, e != z): also the next cycle's e
// after its first red half-sweep
void launch_fcg_update(const Level &L, const uint8_t *op, const float *z, double *x, double *r, float *r32,
                       const float *dold, float *dnew, int first, double inv_h2, const int32_t *rows, int64_t nrows,
                       float *e, float delta, double *partials, double *s, hipStream_t st);

// r = b - A x from the FP64 iterate x over the listed rows, r32; r.r and
// b.b into s[6], s[7]
void launch_fcg_residual(const Level &L, const uint8_t *op, const double *x, const float *b, double inv_h2,
                         const int32_t *rows, int64_t nrows, double *r, float *r32, double *partials, double *s,
                         hipStream_t st);
// FP64 Jacobi-preconditioned CG check solve (scalars c[0..5], error c[8..11])
int check_partials_count();
void launch_check_init(const Level &L, const float *b, double inv_h2, const int32_t *rows, int64_t nrows, int from_x,
                       double *x, double *r, double *z, double *p, double *partials, double *c, hipStream_t st);
void launch_check_iteration(const Level &L, double inv_h2, const int32_t *rows, int64_t nrows, double *x, double *r,
                            double *z, double *po, double *pn, double *qv, double *partials, double *c,
                            hipStream_t st);
// p64 or p32 (one of them NULL) against x: sum (p-x)^2, sum x^2, max|p-x|, max|x|
void launch_check_error(const Level &L, const double *p64, const float *p32, const double *x, const int32_t *rows,
                        int64_t nrows, double *partials, double *c, hipStream_t st);

// sources.hip
void launch_outflow_flags(const float *pos, int n, const uint8_t *sink, float delta, int nx, int ny, int nz,
                          int32_t *keep, hipStream_t s);
void launch_compact(const float *pos, const float *vel, const int32_t *keep, const int32_t *slot, int n,
                    float *pos_out, float *vel_out, hipStream_t s);
void launch_count_open(const float *pos, int n, const uint8_t *open, float delta, int nx, int ny, int nz,
                       int32_t *present, hipStream_t s);
void launch_missing(int32_t *present, const uint8_t *open, int density, int64_t C, hipStream_t s);
// rng: numpy PCG64 state {state hi, lo, inc hi, lo}; the emission starts
// draws_before draws after it
void launch_emit(const int32_t *missing, const int32_t *first, int64_t C, int ny, int nz, double delta,
                 const uint64_t *rng, uint64_t draws_before, int64_t n0, const float *velocity, float *pos, float *vel,
                 hipStream_t s);

// context.hip: sets the message returned by clq_last_error(), returns code
int set_error(int code, const char *fmt, ...);

// unet.hip
int unet_check_grid(const clq_net *n, int nx, int ny, int nz, double dx);
int unet_device(const clq_net *n);
void unet_predict(clq_net *n, const float *rhs, const float *mask, const float *prev, float *out, hipStream_t s);

// voxelise.hip
int voxelise_max_scale();
void launch_voxelise(const float *pos, int n, const int32_t *offsets, const int32_t *counts, int s, float sinv,
                     double scale, int nx, int ny, int nz, uint8_t *occ, float *dens, hipStream_t st);

// g2p.hip
struct ParticleUpdate {
    int flags;  // bit 0: G2P, 1: save previous, 2: advect, 3: domain BC, 4: obstacle BC
    float inv_dx, a, b, dt, lower, upper[3];
    int pure_flip;
};
enum { kG2P = 1, kSavePrev = 2, kAdvect = 4, kDomainBC = 8, kObstacleBC = 16 };
// offsets/counts: the bins of the current particle order (the G2P samples
// are then staged per bin-row segment), or NULL
void launch_particle_update(float *pos, float *vel, float *prev, const float *u, const float *v, const float *w,
                            const float *uo, const float *vo, const float *wo, const int32_t *offsets,
                            const int32_t *counts, const uint8_t *ob, int n, int nx, int ny, int nz,
                            const ParticleUpdate &pu, hipStream_t s);

// mesh.hip: ray-parity mesh voxelisation (scratch: mesh_scratch_bytes(count))
size_t mesh_scratch_bytes(int64_t count);
void launch_voxelise_mesh(const float *tris, int64_t count, void *scratch, int nx, int ny, int nz, float delta,
                          const float origin[3], uint8_t *mask, hipStream_t s);

}  // namespace clq
