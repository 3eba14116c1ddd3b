/*
 * cliquid_oracle.c -- CPU restatement of the Continuum Liquid FLIP/PIC step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library
 * (continuum_liquid_amd/csrc): only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  No product code path calls it.
 *
 * It exports the C-ABI of include/cliquid.h over host arrays so that the tests
 * can drive the oracle and the HIP library with identical call sequences.
 * Every kernel restates one OpenCL kernel of the reference (cited per
 * function, paths relative to the reference root) with the same floating-point
 * operations in the same order.  Where the reference expression is marked
 * llvm.fmuladd by clang (fp-contract=on, the OpenCL default) the restatement
 * uses a fused fmaf(): every target the reference runs on (x86-64 with FMA3,
 * AMD GPUs) lowers fmuladd to a hardware FMA.  Everywhere else the file is
 * compiled with -ffp-contract=off so that no other operation is fused.
 *
 * Deviations from the reference, each deliberate:
 *   - Within-bin particle order is the stable (original index) order.  The
 *     reference scatters with atomic_inc (particles.cl:129) and its order
 *     within a bin is not reproducible from run to run.
 *   - Out-of-range float->int conversions (NaN, |x| > 2^31) are clamped;
 *     OpenCL leaves them undefined.
 *   - The solve modes with a residual test (A14) and the FLIP/PIC blend are
 *     additions; at flip_ratio 1 and mode CLQ_SOLVE_FIXED_CYCLES the step is
 *     the reference's.
 *
 * Parity status per function (how each is pinned) is listed in the header of
 * tests/test_oracle.py.
 */
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <immintrin.h>

#include "../include/cliquid.h"

#define MAX_LEVELS 16

typedef struct level_t {
    int nx, ny, nz;
    long long C;
    float *p, *b, *mask, *diag, *wx, *wy, *wz;
    int owns;
    /* 16-bit storage of b used by the mixed-precision solve on levels >= 1:
     * bkind 0 = float b, 1 = fp16, 2 = bf16; the stored value is
     * b / bscale with bscale a power of two, so scaling is exact. */
    uint16_t *b16;
    int bkind;
    float bscale;
} level_t;

/* IEEE round-to-nearest-even conversions (F16C for fp16). */
static inline uint16_t f32_to_f16(float v) { return (uint16_t)_cvtss_sh(v, _MM_FROUND_TO_NEAREST_INT); }
static inline float f16_to_f32(uint16_t h) { return _cvtsh_ss(h); }
static inline uint16_t f32_to_bf16(float v) {
    uint32_t x;
    memcpy(&x, &v, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u); /* quiet NaN */
    x += 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(x >> 16);
}
static inline float bf16_to_f32(uint16_t h) {
    uint32_t x = (uint32_t)h << 16;
    float v;
    memcpy(&v, &x, 4);
    return v;
}
static inline float load_b(const level_t *L, long long i) {
    if (L->bkind == 0) return L->b[i];
    const float v = L->bkind == 1 ? f16_to_f32(L->b16[i]) : bf16_to_f32(L->b16[i]);
    return v * L->bscale;
}
static inline void store_b(level_t *L, long long i, float v) {
    if (L->bkind == 0) { L->b[i] = v; return; }
    const float scaled = v * (1.0f / L->bscale);
    L->b16[i] = L->bkind == 1 ? f32_to_f16(scaled) : f32_to_bf16(scaled);
}

struct clq_ctx {
    int nx, ny, nz;
    long long C;
    double dx;
    long long cap, n;
    long long tickets; /* frame downloads started (clq_positions_to_host_async) */
    float *pos, *vel, *pos_work, *vel_work, *prev;
    int32_t *counts, *offsets;
    float *u, *v, *w, *uo, *vo, *wo, *ff, *p, *rhs, *diag, *wx, *wy, *wz;
    uint8_t *obstacle;
    int nlevels;
    level_t lv[MAX_LEVELS];
    float coarse_weight; /* weight of the cycles' prolongations: 1 outside a solve */
    double stage_ms[CLQ_STAGE_COUNT];
    /* mixed-precision solve state, allocated on first use */
    double *p64a, *p64b;
    float *r32;
    double *r64;                     /* flexible CG: residual */
    float *dir0, *dir1;              /* FP32 directions, old and new */
    int p64_valid;                    /* p64a holds the pressure of the last solve (modes 2-4) */
    int bins_current; /* counts/offsets describe the current particle order */
    /* particle sources (clq_set_sources) */
    int nsources;
    int32_t *src_type, *src_density;
    float *src_velocity;
    uint8_t *src_open; /* [nsources][C]: mask and not solid */
    uint8_t *sink;     /* union of the OUTFLOW masks */
};

static __thread char g_error[512];

static int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
    return code;
}

static int non_finite(double r) {
    return fail(CLQ_ERR_NOT_CONVERGED, "pressure solve: residual norm is not finite (%g); the velocity field "
                "or the operator holds NaN or Inf", r);
}

static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

const char *clq_last_error(void) { return g_error; }
int clq_abi_version(void) { return CLQ_ABI_VERSION; }
int clq_device_count(int32_t *count) {
    if (!count) return fail(CLQ_ERR_INVALID, "count is NULL");
    *count = 1;
    return CLQ_OK;
}

/* ---- index helpers (helper.cl:4-23) -------------------------------------- */
static inline long long gidx(int i, int j, int k, int ny, int nz) { return ((long long)i * ny + j) * nz + k; }
static inline int interior(int i, int j, int k, int nx, int ny, int nz) {
    return i > 0 && j > 0 && k > 0 && i < nx - 1 && j < ny - 1 && k < nz - 1;
}
static inline void coords(long long index, int ny, int nz, int *i, int *j, int *k) {
    long long yz = (long long)ny * nz;
    *i = (int)(index / yz);
    long long r = index - (long long)(*i) * yz;
    *j = (int)(r / nz);
    *k = (int)(r - (long long)(*j) * nz);
}
/* convert_int_rtn + clamp(0, n-2) with out-of-range inputs clamped. */
static inline int floor_clamp(float t, int lo, int hi) {
    t = floorf(t);
    if (!(t >= (float)lo)) return lo;
    if (t > (float)hi) return hi;
    return (int)t;
}
/* convert_int_rtz + clamp(lo, hi). */
static inline int trunc_clamp(float t, int lo, int hi) {
    t = truncf(t);
    if (!(t >= (float)lo)) return lo;
    if (t > (float)hi) return hi;
    return (int)t;
}

/* ---- context ------------------------------------------------------------- */
static int coarse_size(int n) { /* multigrid.py:12-15 */
    int interior_size = n - 2 > 0 ? n - 2 : 0;
    return (interior_size + 1) / 2 + 2;
}

static void *xcalloc(size_t count, size_t size) { return calloc(count ? count : 1, size); }

static int alloc_particles(clq_ctx *c, long long cap) {
    free(c->pos); free(c->vel); free(c->pos_work); free(c->vel_work); free(c->prev);
    c->pos = xcalloc(cap * 3, 4);
    c->vel = xcalloc(cap * 3, 4);
    c->pos_work = xcalloc(cap * 3, 4);
    c->vel_work = xcalloc(cap * 3, 4);
    c->prev = xcalloc(cap * 3, 4);
    c->cap = cap;
    if (!c->pos || !c->vel || !c->pos_work || !c->vel_work || !c->prev)
        return fail(CLQ_ERR_NOMEM, "particle allocation of %lld failed", cap);
    return CLQ_OK;
}

int clq_destroy(clq_ctx *c) {
    if (!c) return CLQ_OK;
    free(c->pos); free(c->vel); free(c->pos_work); free(c->vel_work); free(c->prev);
    free(c->counts); free(c->offsets);
    float *grids[] = {c->u, c->v, c->w, c->uo, c->vo, c->wo, c->ff, c->p, c->rhs, c->diag, c->wx, c->wy, c->wz};
    for (size_t i = 0; i < sizeof grids / sizeof grids[0]; ++i) free(grids[i]);
    free(c->obstacle);
    for (int l = 1; l < c->nlevels; ++l) {
        level_t *L = &c->lv[l];
        free(L->p); free(L->b); free(L->mask); free(L->diag); free(L->wx); free(L->wy); free(L->wz);
        free(L->b16);
    }
    free(c->p64a); free(c->p64b); free(c->r32);
    free(c->r64); free(c->dir0); free(c->dir1);
    free(c->src_type); free(c->src_density); free(c->src_velocity); free(c->src_open); free(c->sink);
    free(c);
    return CLQ_OK;
}

int clq_create(int32_t nx, int32_t ny, int32_t nz, double dx, int64_t capacity, clq_ctx **out) {
    if (!out) return fail(CLQ_ERR_INVALID, "out is NULL");
    *out = NULL;
    if (nx < 3 || ny < 3 || nz < 3) return fail(CLQ_ERR_INVALID, "grid %dx%dx%d too small (need >= 3)", nx, ny, nz);
    if ((long long)nx * ny * nz >= (1LL << 31)) return fail(CLQ_ERR_INVALID, "grid has >= 2^31 cells");
    if (!(dx > 0.0)) return fail(CLQ_ERR_INVALID, "dx must be positive");
    if (capacity < 1) capacity = 1;
    clq_ctx *c = calloc(1, sizeof *c);
    if (!c) return fail(CLQ_ERR_NOMEM, "context allocation failed");
    c->nx = nx; c->ny = ny; c->nz = nz; c->dx = dx;
    c->coarse_weight = 1.0f;
    c->C = (long long)nx * ny * nz;
    if (alloc_particles(c, capacity)) { clq_destroy(c); return CLQ_ERR_NOMEM; }
    c->counts = xcalloc(c->C, 4);
    c->offsets = xcalloc(c->C, 4);
    float **grids[] = {&c->u, &c->v, &c->w, &c->uo, &c->vo, &c->wo, &c->ff, &c->p, &c->rhs, &c->diag, &c->wx, &c->wy, &c->wz};
    for (size_t i = 0; i < sizeof grids / sizeof grids[0]; ++i) {
        *grids[i] = xcalloc(c->C, 4);
        if (!*grids[i]) { clq_destroy(c); return fail(CLQ_ERR_NOMEM, "grid allocation failed"); }
    }
    c->obstacle = xcalloc(c->C, 1);
    /* level 0 aliases the simulation grid (pressure_solve.py:62-76) */
    level_t *L0 = &c->lv[0];
    L0->nx = nx; L0->ny = ny; L0->nz = nz; L0->C = c->C;
    L0->p = c->p; L0->b = c->rhs; L0->mask = c->ff; L0->diag = c->diag; L0->wx = c->wx; L0->wy = c->wy; L0->wz = c->wz;
    /* create_multigrid_levels(min_size=8): multigrid.py:18-50, solver.py:212 */
    int n0 = coarse_size(nx), n1 = coarse_size(ny), n2 = coarse_size(nz);
    c->nlevels = 1;
    while (n0 >= 8 && n1 >= 8 && n2 >= 8 && c->nlevels < MAX_LEVELS) {
        level_t *L = &c->lv[c->nlevels++];
        L->nx = n0; L->ny = n1; L->nz = n2; L->C = (long long)n0 * n1 * n2;
        float **arrs[] = {&L->p, &L->b, &L->mask, &L->diag, &L->wx, &L->wy, &L->wz};
        for (int a = 0; a < 7; ++a) {
            *arrs[a] = xcalloc(L->C, 4);
            if (!*arrs[a]) { clq_destroy(c); return fail(CLQ_ERR_NOMEM, "level allocation failed"); }
        }
        n0 = coarse_size(n0); n1 = coarse_size(n1); n2 = coarse_size(n2);
    }
    *out = c;
    return CLQ_OK;
}

int clq_sync(clq_ctx *c) { return c ? CLQ_OK : fail(CLQ_ERR_INVALID, "ctx is NULL"); }

/* ---- data movement ------------------------------------------------------- */
int clq_set_particles(clq_ctx *c, const float *pos, const float *vel, int64_t n) {
    if (!c || (n > 0 && !pos) || n < 0) return fail(CLQ_ERR_INVALID, "bad particle upload");
    if (n > c->cap) {
        long long cap = c->cap * 3 / 2; /* solver.py:290-291 */
        if (cap < n) cap = n;
        int rc = alloc_particles(c, cap);
        if (rc) return rc;
    }
    memcpy(c->pos, pos, (size_t)n * 12);
    if (vel) memcpy(c->vel, vel, (size_t)n * 12);
    else memset(c->vel, 0, (size_t)n * 12);
    c->n = n;
    c->bins_current = 0;
    return CLQ_OK;
}

int clq_get_particles(clq_ctx *c, float *pos, float *vel, int64_t n) {
    if (!c || n != c->n) return fail(CLQ_ERR_INVALID, "particle count mismatch (%lld requested, %lld held)",
                                     (long long)n, c ? c->n : -1LL);
    if (pos) memcpy(pos, c->pos, (size_t)n * 12);
    if (vel) memcpy(vel, c->vel, (size_t)n * 12);
    return CLQ_OK;
}

int clq_get_previous_positions(clq_ctx *c, float *pos, int64_t n) {
    if (!c || n != c->n || !pos) return fail(CLQ_ERR_INVALID, "bad previous-position download");
    memcpy(pos, c->prev, (size_t)n * 12);
    return CLQ_OK;
}

int clq_get_particle_count(clq_ctx *c, int64_t *n) {
    if (!c || !n) return fail(CLQ_ERR_INVALID, "NULL argument");
    *n = c->n;
    return CLQ_OK;
}

/* Frame output (Solver/Kernel/output.py:56-75, 178-195): on the CPU the
 * "pinned" slot is ordinary memory and the copy completes before the call
 * returns; the ticket rules are the library's (-1 for an empty set, the last
 * 16 tickets valid). */
int clq_host_alloc(clq_ctx *c, int64_t bytes, void **host) {
    if (!c || !host || bytes < 0) return fail(CLQ_ERR_INVALID, "bad host allocation");
    *host = malloc((size_t)(bytes > 16 ? bytes : 16));
    return *host ? CLQ_OK : fail(CLQ_ERR_NOMEM, "malloc(%lld) failed", (long long)bytes);
}

int clq_host_free(clq_ctx *c, void *host) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    free(host);
    return CLQ_OK;
}

int clq_positions_to_host_async(clq_ctx *c, float *host, int64_t capacity, int64_t *n, int64_t *ticket) {
    if (!c || !n || !ticket) return fail(CLQ_ERR_INVALID, "NULL argument");
    *n = c->n;
    *ticket = -1;
    if (c->n == 0) return CLQ_OK;
    if (!host || capacity < c->n)
        return fail(CLQ_ERR_INVALID, "host buffer holds %lld particles, %lld needed", (long long)capacity, c->n);
    memcpy(host, c->pos, (size_t)c->n * 12);
    *ticket = c->tickets++;
    return CLQ_OK;
}

int clq_host_copy_wait(clq_ctx *c, int64_t ticket) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    if (ticket < 0) return CLQ_OK;
    if (ticket >= c->tickets || ticket < c->tickets - 16)
        return fail(CLQ_ERR_INVALID, "unknown download ticket %lld", (long long)ticket);
    return CLQ_OK;
}

static void *field_ptr(clq_ctx *c, int f, long long *bytes) {
    long long fb = c->C * 4;
    void *p = NULL;
    switch (f) {
    case CLQ_FIELD_U: p = c->u; break;
    case CLQ_FIELD_V: p = c->v; break;
    case CLQ_FIELD_W: p = c->w; break;
    case CLQ_FIELD_U_OLD: p = c->uo; break;
    case CLQ_FIELD_V_OLD: p = c->vo; break;
    case CLQ_FIELD_W_OLD: p = c->wo; break;
    case CLQ_FIELD_FLUID_FRACTION: p = c->ff; break;
    case CLQ_FIELD_PRESSURE: p = c->p; break;
    case CLQ_FIELD_RHS: p = c->rhs; break;
    case CLQ_FIELD_DIAGONAL: p = c->diag; break;
    case CLQ_FIELD_WEIGHT_X: p = c->wx; break;
    case CLQ_FIELD_WEIGHT_Y: p = c->wy; break;
    case CLQ_FIELD_WEIGHT_Z: p = c->wz; break;
    case CLQ_FIELD_BIN_COUNTS: p = c->counts; break;
    case CLQ_FIELD_BIN_OFFSETS: p = c->offsets; break;
    case CLQ_FIELD_OBSTACLE: p = c->obstacle; fb = c->C; break;
    case CLQ_FIELD_PRESSURE_FP64: p = c->p64a; fb = c->C * 8; break;
    default: return NULL;
    }
    *bytes = fb;
    return p;
}

int clq_set_field(clq_ctx *c, int32_t f, const void *host, int64_t bytes) {
    long long fb;
    void *p = c ? field_ptr(c, f, &fb) : NULL;
    if (!p || !host || bytes != fb) return fail(CLQ_ERR_INVALID, "bad field upload (field %d, %lld bytes)", f, (long long)bytes);
    memcpy(p, host, (size_t)bytes);
    if (f == CLQ_FIELD_PRESSURE) c->p64_valid = 0;
    return CLQ_OK;
}

int clq_get_field(clq_ctx *c, int32_t f, void *host, int64_t bytes) {
    long long fb;
    void *p = c ? field_ptr(c, f, &fb) : NULL;
    if (!p || !host || bytes != fb) return fail(CLQ_ERR_INVALID, "bad field download (field %d, %lld bytes)", f, (long long)bytes);
    memcpy(host, p, (size_t)bytes);
    return CLQ_OK;
}

int clq_level_count(clq_ctx *c, int32_t *levels) {
    if (!c || !levels) return fail(CLQ_ERR_INVALID, "NULL argument");
    *levels = c->nlevels;
    return CLQ_OK;
}

int clq_level_shape(clq_ctx *c, int32_t l, int32_t *nx, int32_t *ny, int32_t *nz) {
    if (!c || l < 0 || l >= c->nlevels) return fail(CLQ_ERR_INVALID, "bad level %d", l);
    *nx = c->lv[l].nx; *ny = c->lv[l].ny; *nz = c->lv[l].nz;
    return CLQ_OK;
}

static float *level_array(level_t *L, int f) {
    switch (f) {
    case CLQ_LEVEL_P: return L->p;
    case CLQ_LEVEL_B: return L->b;
    case CLQ_LEVEL_MASK: return L->mask;
    case CLQ_LEVEL_DIAGONAL: return L->diag;
    case CLQ_LEVEL_WEIGHT_X: return L->wx;
    case CLQ_LEVEL_WEIGHT_Y: return L->wy;
    case CLQ_LEVEL_WEIGHT_Z: return L->wz;
    default: return NULL;
    }
}

int clq_set_level_field(clq_ctx *c, int32_t l, int32_t f, const float *host, int64_t bytes) {
    if (!c || l < 0 || l >= c->nlevels || !host) return fail(CLQ_ERR_INVALID, "bad level upload");
    float *a = level_array(&c->lv[l], f);
    if (!a || bytes != c->lv[l].C * 4) return fail(CLQ_ERR_INVALID, "bad level field %d / size", f);
    memcpy(a, host, (size_t)bytes);
    return CLQ_OK;
}

int clq_get_level_field(clq_ctx *c, int32_t l, int32_t f, float *host, int64_t bytes) {
    if (!c || l < 0 || l >= c->nlevels || !host) return fail(CLQ_ERR_INVALID, "bad level download");
    float *a = level_array(&c->lv[l], f);
    if (!a || bytes != c->lv[l].C * 4) return fail(CLQ_ERR_INVALID, "bad level field %d / size", f);
    memcpy(host, a, (size_t)bytes);
    return CLQ_OK;
}

/* ---- binning: particles.cl:3-6, 44-133; particles.py:39-140 -------------- */
/* interpolation_base (particles.cl:3-6): floor(fma(x, inv_dx, -0.5)) clamped
 * to [0, n-2]; the subtraction is contracted (llvm.fmuladd at particles.cl:4). */
static inline long long bin_of(const float *x, float inv_dx, int nx, int ny, int nz) {
    int bx = floor_clamp(fmaf(x[0], inv_dx, -0.5f), 0, nx - 2);
    int by = floor_clamp(fmaf(x[1], inv_dx, -0.5f), 0, ny - 2);
    int bz = floor_clamp(fmaf(x[2], inv_dx, -0.5f), 0, nz - 2);
    return gidx(bx, by, bz, ny, nz);
}

int clq_bin_particles(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    double t0 = now_ms();
    const float inv_dx = (float)(1.0 / c->dx); /* FIELD_DTYPE(1.0 / dx), particles.py:103 */
    const long long n = c->n;
    int32_t *key = malloc((size_t)(n ? n : 1) * 4);
    if (!key) return fail(CLQ_ERR_NOMEM, "key allocation failed");
    /* count_particles_per_cell (particles.cl:44-59) after the zero fill of particles.py:94 */
    memset(c->counts, 0, (size_t)c->C * 4);
    for (long long i = 0; i < n; ++i) {
        key[i] = (int32_t)bin_of(c->pos + 3 * i, inv_dx, c->nx, c->ny, c->nz);
        c->counts[key[i]]++;
    }
    /* exclusive_prefix_sum (particles.py:39-70, particles.cl:61-108): an
     * integer exclusive scan, identical for every summation order. */
    int32_t run = 0;
    for (long long b = 0; b < c->C; ++b) {
        c->offsets[b] = run;
        run += c->counts[b];
    }
    /* sort_particles (particles.cl:110-133) with cursors zeroed
     * (particles.py:119); stable order: cursor[b] advances in index order. */
    int32_t *cursor = xcalloc(c->C, 4);
    if (!cursor) { free(key); return fail(CLQ_ERR_NOMEM, "cursor allocation failed"); }
    for (long long i = 0; i < n; ++i) {
        long long dst = c->offsets[key[i]] + cursor[key[i]]++;
        memcpy(c->pos_work + 3 * dst, c->pos + 3 * i, 12);
        memcpy(c->vel_work + 3 * dst, c->vel + 3 * i, 12);
    }
    free(cursor);
    free(key);
    /* swap (particles.py:136-140) */
    float *t = c->pos; c->pos = c->pos_work; c->pos_work = t;
    t = c->vel; c->vel = c->vel_work; c->vel_work = t;
    c->stage_ms[CLQ_STAGE_BIN] = now_ms() - t0;
    c->bins_current = 1;
    return CLQ_OK;
}

/* particle_to_grid_gather (particles.cl:135-227), restated with the
 * reference's work-group of 8 lanes (kernel_config.py:28-29): lane l visits
 * neighbour bins l, l+8, l+16, l+24 in that order, accumulates its own
 * partial sums, and the 8 partials are combined by the LDS tree of
 * particles.cl:204-211. */
int clq_particle_to_grid(clq_ctx *c, double particles_per_cell) {
    if (!c || !(particles_per_cell > 0)) return fail(CLQ_ERR_INVALID, "bad P2G arguments");
    double t0 = now_ms();
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const float inv_dx = (float)(1.0 / c->dx);
    const float inv_ppc = (float)(1.0 / particles_per_cell);
#pragma omp parallel for schedule(dynamic, 256)
    for (long long index = 0; index < c->C; ++index) {
        int ci, cj, ck;
        coords(index, ny, nz, &ci, &cj, &ck);
        float vs[8][3], ws[8][3];
        int cnt[8];
        memset(vs, 0, sizeof vs);
        memset(ws, 0, sizeof ws);
        memset(cnt, 0, sizeof cnt);
        const float gx = (float)ci, gy = (float)cj, gz = (float)ck;
        const float upos[3] = {gx + 1.0f, gy + 0.5f, gz + 0.5f};
        const float vpos[3] = {gx + 0.5f, gy + 1.0f, gz + 0.5f};
        const float wpos[3] = {gx + 0.5f, gy + 0.5f, gz + 1.0f};
        for (int lane = 0; lane < 8; ++lane) {
            for (int nb = lane; nb < 27; nb += 8) {
                const int ox = nb / 9 - 1, oy = (nb / 3) % 3 - 1, oz = nb % 3 - 1;
                const int bx = ci + ox, by = cj + oy, bz = ck + oz;
                if (bx < 0 || bx >= nx - 1) continue;
                if (by < 0 || by >= ny - 1) continue;
                if (bz < 0 || bz >= nz - 1) continue;
                const long long bin = gidx(bx, by, bz, ny, nz);
                const int begin = c->offsets[bin], end = begin + c->counts[bin];
                for (int pi = begin; pi < end; ++pi) {
                    const float *x = c->pos + 3LL * pi;
                    const float *vel = c->vel + 3LL * pi;
                    const int pcx = trunc_clamp(x[0] * inv_dx, 1, nx - 2);
                    const int pcy = trunc_clamp(x[1] * inv_dx, 1, ny - 2);
                    const int pcz = trunc_clamp(x[2] * inv_dx, 1, nz - 2);
                    cnt[lane] += (pcx == ci && pcy == cj && pcz == ck);
                    float p[3] = {x[0] * inv_dx, x[1] * inv_dx, x[2] * inv_dx};
                    float wu[3], wv[3], ww[3];
                    for (int a = 0; a < 3; ++a) {
                        wu[a] = fmaxf(0.0f, 1.0f - fabsf(p[a] - upos[a]));
                        wv[a] = fmaxf(0.0f, 1.0f - fabsf(p[a] - vpos[a]));
                        ww[a] = fmaxf(0.0f, 1.0f - fabsf(p[a] - wpos[a]));
                    }
                    const float weight[3] = {wu[0] * wu[1] * wu[2], wv[0] * wv[1] * wv[2], ww[0] * ww[1] * ww[2]};
                    for (int a = 0; a < 3; ++a) {
                        vs[lane][a] = fmaf(weight[a], vel[a], vs[lane][a]); /* particles.cl:195 fmuladd */
                        ws[lane][a] = ws[lane][a] + weight[a];
                    }
                }
            }
        }
        /* LDS tree, offsets 4, 2, 1 (particles.cl:204-211) */
        for (int off = 4; off > 0; off >>= 1)
            for (int l = 0; l < off; ++l) {
                for (int a = 0; a < 3; ++a) {
                    vs[l][a] = vs[l][a] + vs[l + off][a];
                    ws[l][a] = ws[l][a] + ws[l + off][a];
                }
                cnt[l] += cnt[l + off];
            }
        c->u[index] = ws[0][0] > 0.0f ? vs[0][0] / ws[0][0] : 0.0f;
        c->v[index] = ws[0][1] > 0.0f ? vs[0][1] / ws[0][1] : 0.0f;
        c->w[index] = ws[0][2] > 0.0f ? vs[0][2] / ws[0][2] : 0.0f;
        c->ff[index] = (float)cnt[0] * inv_ppc;
    }
    c->stage_ms[CLQ_STAGE_P2G] = now_ms() - t0;
    return CLQ_OK;
}

/* ---- grid boundary conditions: boundary_conditions.cl ------------------- */
int clq_grid_obstacle_bc(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const uint8_t *ob = c->obstacle;
    /* grid_obstacle_bc (boundary_conditions.cl:99-119) */
#pragma omp parallel for schedule(static)
    for (long long index = 0; index < c->C; ++index) {
        int i, j, k;
        coords(index, ny, nz, &i, &j, &k);
        if (ob[index]) c->ff[index] = 0.0f;
        const int xs = ob[index] || (i + 1 < nx && ob[gidx(i + 1, j, k, ny, nz)]);
        const int ys = ob[index] || (j + 1 < ny && ob[gidx(i, j + 1, k, ny, nz)]);
        const int zs = ob[index] || (k + 1 < nz && ob[gidx(i, j, k + 1, ny, nz)]);
        if (xs) c->u[index] = 0.0f;
        if (ys) c->v[index] = 0.0f;
        if (zs) c->w[index] = 0.0f;
    }
    /* extrapolate_obstacle_velocity (boundary_conditions.cl:121-173): solid
     * cells read only fluid neighbours, so the in-place update is race free. */
#pragma omp parallel for schedule(static)
    for (long long index = 0; index < c->C; ++index) {
        if (!ob[index]) continue;
        int i, j, k;
        coords(index, ny, nz, &i, &j, &k);
        float t[3] = {0.0f, 0.0f, 0.0f};
        int nf = 0, ax[3] = {0, 0, 0};
        const int ni[6][4] = {{i - 1, j, k, 0}, {i + 1, j, k, 0}, {i, j - 1, k, 1},
                              {i, j + 1, k, 1}, {i, j, k - 1, 2}, {i, j, k + 1, 2}};
        const int ok[6] = {i > 0, i + 1 < nx, j > 0, j + 1 < ny, k > 0, k + 1 < nz};
        for (int q = 0; q < 6; ++q) {
            if (!ok[q]) continue;
            const long long nb = gidx(ni[q][0], ni[q][1], ni[q][2], ny, nz);
            if (!ob[nb]) {
                t[0] = t[0] + c->u[nb];
                t[1] = t[1] + c->v[nb];
                t[2] = t[2] + c->w[nb];
                nf += 1;
                ax[ni[q][3]] = 1;
            }
        }
        if (nf == 0) continue;
        const float fn = (float)nf;
        t[0] = t[0] / fn; t[1] = t[1] / fn; t[2] = t[2] / fn;
        c->u[index] = ax[0] ? 0.0f : t[0];
        c->v[index] = ax[1] ? 0.0f : t[1];
        c->w[index] = ax[2] ? 0.0f : t[2];
    }
    return CLQ_OK;
}

/* grid_domain_bc (boundary_conditions.cl:33-67).  A ghost cell copies from
 * the clamped interior cell only components that are not walls there, and
 * the interior thread writes only wall components, so there is no race. */
int clq_grid_domain_bc(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    const int nx = c->nx, ny = c->ny, nz = c->nz;
#pragma omp parallel for schedule(static)
    for (long long index = 0; index < c->C; ++index) {
        int i, j, k;
        coords(index, ny, nz, &i, &j, &k);
        const int ghost = !interior(i, j, k, nx, ny, nz);
        const int wu = i == 0 || i >= nx - 2;
        const int wv = j == 0 || j >= ny - 2;
        const int ww = k == 0 || k >= nz - 2;
        if (!ghost && !wu && !wv && !ww) continue;
        if (ghost) {
            const int ii = i < 1 ? 1 : (i > nx - 2 ? nx - 2 : i);
            const int jj = j < 1 ? 1 : (j > ny - 2 ? ny - 2 : j);
            const int kk = k < 1 ? 1 : (k > nz - 2 ? nz - 2 : k);
            const long long in = gidx(ii, jj, kk, ny, nz);
            c->u[index] = wu ? 0.0f : c->u[in];
            c->v[index] = wv ? 0.0f : c->v[in];
            c->w[index] = ww ? 0.0f : c->w[in];
            continue;
        }
        if (wu) c->u[index] = 0.0f;
        if (wv) c->v[index] = 0.0f;
        if (ww) c->w[index] = 0.0f;
    }
    return CLQ_OK;
}

/* laplacian (helper.cl:54-66): fma(-6, centre, neighbour_sum) * inv_dx2. */
static inline float laplacian(const float *f, int i, int j, int k, float inv_dx2, int ny, int nz) {
    const float centre = f[gidx(i, j, k, ny, nz)];
    float s = f[gidx(i + 1, j, k, ny, nz)] + f[gidx(i - 1, j, k, ny, nz)];
    s = s + f[gidx(i, j + 1, k, ny, nz)];
    s = s + f[gidx(i, j - 1, k, ny, nz)];
    s = s + f[gidx(i, j, k + 1, ny, nz)];
    s = s + f[gidx(i, j, k - 1, ny, nz)];
    return fmaf(-6.0f, centre, s) * inv_dx2;
}

/* update_velocity (velocity_update.cl:3-43) into the work buffers, then the
 * swap of solver.py:374-376. */
int clq_update_velocity(clq_ctx *c, const double *gravity, const double *force, double rho, double nu, double dt) {
    if (!c || !gravity || !force) return fail(CLQ_ERR_INVALID, "bad update_velocity arguments");
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const float g[3] = {(float)gravity[0], (float)gravity[1], (float)gravity[2]};
    const float f[3] = {(float)force[0], (float)force[1], (float)force[2]};
    const float frho = (float)rho, fnu = (float)nu, fdt = (float)dt;
    const float inv_dx2 = (float)(1.0 / (c->dx * c->dx)); /* velocity_update.py:46 */
    const float acc[3] = {g[0] + f[0] / frho, g[1] + f[1] / frho, g[2] + f[2] / frho};
    const float nudt = fnu * fdt;
#pragma omp parallel for schedule(static)
    for (long long index = 0; index < c->C; ++index) {
        int i, j, k;
        coords(index, ny, nz, &i, &j, &k);
        float r0 = c->u[index], r1 = c->v[index], r2 = c->w[index];
        if (c->ff[index] > 0.0f) { /* velocity_update.cl:31 fmuladd */
            r0 = fmaf(fdt, acc[0], r0);
            r1 = fmaf(fdt, acc[1], r1);
            r2 = fmaf(fdt, acc[2], r2);
        }
        if (interior(i, j, k, nx, ny, nz)) { /* velocity_update.cl:35-37 fmuladd */
            r0 = fmaf(nudt, laplacian(c->u, i, j, k, inv_dx2, ny, nz), r0);
            r1 = fmaf(nudt, laplacian(c->v, i, j, k, inv_dx2, ny, nz), r1);
            r2 = fmaf(nudt, laplacian(c->w, i, j, k, inv_dx2, ny, nz), r2);
        }
        c->uo[index] = r0;
        c->vo[index] = r1;
        c->wo[index] = r2;
    }
    float *t;
    t = c->u; c->u = c->uo; c->uo = t;
    t = c->v; c->v = c->vo; c->vo = t;
    t = c->w; c->w = c->wo; c->wo = t;
    return CLQ_OK;
}

/* ---- pressure ------------------------------------------------------------- */
/* pressure_rhs (pressure_solve.cl:3-24, divergence helper.cl:39-52). */
int clq_pressure_rhs(clq_ctx *c, double rho, double dt) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const float rho_over_dt = (float)(rho / dt), inv_dx = (float)(1.0 / c->dx);
#pragma omp parallel for schedule(static)
    for (long long index = 0; index < c->C; ++index) {
        int i, j, k;
        coords(index, ny, nz, &i, &j, &k);
        if (!interior(i, j, k, nx, ny, nz) || c->ff[index] <= 0.0f) {
            c->rhs[index] = 0.0f;
            continue;
        }
        const float du = c->u[index] - c->u[gidx(i - 1, j, k, ny, nz)];
        const float dv = c->v[index] - c->v[gidx(i, j - 1, k, ny, nz)];
        const float dw = c->w[index] - c->w[gidx(i, j, k - 1, ny, nz)];
        c->rhs[index] = rho_over_dt * ((du + dv + dw) * inv_dx);
    }
    return CLQ_OK;
}

/* build_finest_operator (multigrid.cl:63-88, domain_diagonal :3-5). */
int clq_build_operator(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const float *mask = c->ff;
#pragma omp parallel for schedule(static)
    for (long long index = 0; index < c->C; ++index) {
        int i, j, k;
        coords(index, ny, nz, &i, &j, &k);
        c->diag[index] = 0.0f; c->wx[index] = 0.0f; c->wy[index] = 0.0f; c->wz[index] = 0.0f;
        if (!interior(i, j, k, nx, ny, nz) || mask[index] <= 0.0f) continue;
        c->diag[index] = 6.0f - (float)(i == 1) - (float)(i == nx - 2) - (float)(j == 1) - (float)(j == ny - 2) -
                         (float)(k == 1) - (float)(k == nz - 2);
        if (i < nx - 2) c->wx[index] = mask[gidx(i + 1, j, k, ny, nz)] > 0.0f;
        if (j < ny - 2) c->wy[index] = mask[gidx(i, j + 1, k, ny, nz)] > 0.0f;
        if (k < nz - 2) c->wz[index] = mask[gidx(i, j, k + 1, ny, nz)] > 0.0f;
    }
    return CLQ_OK;
}

/* weighted_neighbor_sum (multigrid.cl:7-23): clang contracts the chain as
 * fma(wz-,p, fma(wz,p, fma(wy-,p, fma(wy,p, fma(wx,p[i+1], wx[i-1]*p[i-1]))))). */
static inline float wsum(const level_t *L, const float *p, int i, int j, int k) {
    const int ny = L->ny, nz = L->nz;
    const long long idx = gidx(i, j, k, ny, nz);
    float s = L->wx[gidx(i - 1, j, k, ny, nz)] * p[gidx(i - 1, j, k, ny, nz)];
    s = fmaf(L->wx[idx], p[gidx(i + 1, j, k, ny, nz)], s);
    s = fmaf(L->wy[idx], p[gidx(i, j + 1, k, ny, nz)], s);
    s = fmaf(L->wy[gidx(i, j - 1, k, ny, nz)], p[gidx(i, j - 1, k, ny, nz)], s);
    s = fmaf(L->wz[idx], p[gidx(i, j, k + 1, ny, nz)], s);
    s = fmaf(L->wz[gidx(i, j, k - 1, ny, nz)], p[gidx(i, j, k - 1, ny, nz)], s);
    return s;
}

/* rbgs_step (multigrid.cl:203-230) for both parities in turn (multigrid.py:190-220). */
static void smooth_level(clq_ctx *c, level_t *L, int iterations) {
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    const float delta = (float)c->dx;
    const float dd = delta * delta;
    for (int it = 0; it < iterations; ++it)
        for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for schedule(static)
            for (long long index = 0; index < L->C; ++index) {
                int i, j, k;
                coords(index, ny, nz, &i, &j, &k);
                if (!interior(i, j, k, nx, ny, nz) || L->mask[index] <= 0.0f || ((i + j + k) & 1) != parity) {
                    if (L->mask[index] <= 0.0f) L->p[index] = 0.0f;
                    continue;
                }
                const float d = L->diag[index];
                if (d > 0.0f) L->p[index] = fmaf(-dd, load_b(L, index), wsum(L, L->p, i, j, k)) / d; /* :227 fmuladd */
            }
        }
}

int clq_smooth(clq_ctx *c, int32_t level, int32_t iterations) {
    if (!c || level < 0 || level >= c->nlevels || iterations < 0) return fail(CLQ_ERR_INVALID, "bad smooth arguments");
    smooth_level(c, &c->lv[level], iterations);
    return CLQ_OK;
}

/* apply_operator / residual (multigrid.cl:25-61). */
static inline float residual_at(const level_t *L, int i, int j, int k, float inv_delta2) {
    const long long idx = gidx(i, j, k, L->ny, L->nz);
    float ap = 0.0f;
    if (interior(i, j, k, L->nx, L->ny, L->nz) && L->mask[idx] > 0.0f)
        ap = fmaf(-L->diag[idx], L->p[idx], wsum(L, L->p, i, j, k)) * inv_delta2; /* :41 fmuladd */
    return load_b(L, idx) - ap;
}

/* restrict_residual (multigrid.cl:90-180). */
int clq_restrict(clq_ctx *c, int32_t level) {
    if (!c || level < 0 || level + 1 >= c->nlevels) return fail(CLQ_ERR_INVALID, "bad restrict level %d", level);
    const level_t *F = &c->lv[level];
    level_t *K = &c->lv[level + 1];
    const float delta = (float)c->dx;
    const float inv_delta2 = 1.0f / (delta * delta);
#pragma omp parallel for schedule(static)
    for (long long ci = 0; ci < K->C; ++ci) {
        K->p[ci] = 0.0f; store_b(K, ci, 0.0f); K->mask[ci] = 0.0f; K->diag[ci] = 0.0f;
        K->wx[ci] = 0.0f; K->wy[ci] = 0.0f; K->wz[ci] = 0.0f;
        int a, b, d;
        coords(ci, K->ny, K->nz, &a, &b, &d);
        if (!interior(a, b, d, K->nx, K->ny, K->nz)) continue;
        const int fs[3] = {1 + 2 * (a - 1), 1 + 2 * (b - 1), 1 + 2 * (d - 1)};
        float rhs_sum = 0.0f, diag_sum = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
        int active = 0;
        for (int di = 0; di < 2; ++di)
            for (int dj = 0; dj < 2; ++dj)
                for (int dk = 0; dk < 2; ++dk) {
                    const int i = fs[0] + di, j = fs[1] + dj, k = fs[2] + dk;
                    if (!interior(i, j, k, F->nx, F->ny, F->nz)) continue;
                    const long long idx = gidx(i, j, k, F->ny, F->nz);
                    if (F->mask[idx] <= 0.0f) continue;
                    ++active;
                    rhs_sum = rhs_sum + residual_at(F, i, j, k, inv_delta2);
                    diag_sum = diag_sum + F->diag[idx];
                    const float wx = F->wx[idx], wy = F->wy[idx], wz = F->wz[idx];
                    if (wx > 0.0f) { if (di == 0) diag_sum = fmaf(-2.0f, wx, diag_sum); else cx = cx + wx; }
                    if (wy > 0.0f) { if (dj == 0) diag_sum = fmaf(-2.0f, wy, diag_sum); else cy = cy + wy; }
                    if (wz > 0.0f) { if (dk == 0) diag_sum = fmaf(-2.0f, wz, diag_sum); else cz = cz + wz; }
                }
        if (active > 0) {
            store_b(K, ci, rhs_sum); K->mask[ci] = 1.0f; K->diag[ci] = diag_sum;
            K->wx[ci] = cx; K->wy[ci] = cy; K->wz[ci] = cz;
        }
    }
    return CLQ_OK;
}

/* prolongate (multigrid.cl:182-201): p += w * p_coarse[parent], w = 1 in the
 * reference; fmaf(1, x, y) is the plain sum x + y, so weight 1 is the
 * reference's arithmetic exactly. */
static void prolongate_weighted(clq_ctx *c, int level, float w) {
    level_t *F = &c->lv[level];
    const level_t *K = &c->lv[level + 1];
#pragma omp parallel for schedule(static)
    for (long long fi = 0; fi < F->C; ++fi) {
        int i, j, k;
        coords(fi, F->ny, F->nz, &i, &j, &k);
        if (!interior(i, j, k, F->nx, F->ny, F->nz) || F->mask[fi] <= 0.0f) continue;
        int a = (i - 1) / 2 + 1, b = (j - 1) / 2 + 1, d = (k - 1) / 2 + 1;
        if (a > K->nx - 2) a = K->nx - 2;
        if (b > K->ny - 2) b = K->ny - 2;
        if (d > K->nz - 2) d = K->nz - 2;
        F->p[fi] = fmaf(w, K->p[gidx(a, b, d, K->ny, K->nz)], F->p[fi]);
    }
}

int clq_prolongate(clq_ctx *c, int32_t level) {
    if (!c || level < 0 || level + 1 >= c->nlevels) return fail(CLQ_ERR_INVALID, "bad prolongate level %d", level);
    prolongate_weighted(c, level, 1.0f);
    return CLQ_OK;
}

/* coarse_smooth (multigrid.cl:232-264): colours in turn, fluid cells only,
 * no zeroing of non-fluid cells (unlike rbgs_step). */
int clq_coarse_smooth(clq_ctx *c, int32_t iterations) {
    if (!c || iterations < 0) return fail(CLQ_ERR_INVALID, "bad coarse_smooth arguments");
    level_t *L = &c->lv[c->nlevels - 1];
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    const float delta = (float)c->dx;
    const float delta2 = delta * delta;
    for (int it = 0; it < iterations; ++it)
        for (int parity = 0; parity < 2; ++parity)
            for (long long index = 0; index < L->C; ++index) {
                int i, j, k;
                coords(index, ny, nz, &i, &j, &k);
                if (!interior(i, j, k, nx, ny, nz) || L->mask[index] <= 0.0f || ((i + j + k) & 1) != parity) continue;
                const float d = L->diag[index];
                if (d > 0.0f) L->p[index] = fmaf(-delta2, load_b(L, index), wsum(L, L->p, i, j, k)) / d; /* :257 */
            }
    return CLQ_OK;
}

/* Whether a visit of `level` runs its coarse-grid correction twice (the W
 * shape, multigrid.py:53-187) or once (V): vcycle_from 0 at every level,
 * k > 0 on the levels below k, -m < 0 on the levels whose bit is set in m.
 * Same rule as the library's cycle. */
static int twice_at(int vfrom, int level) {
    if (vfrom < 0) return level < 31 && ((((uint32_t)(-(int64_t)vfrom)) >> level) & 1u) != 0;
    return vfrom == 0 || level < vfrom;
}

/* w_cycle (multigrid.py:53-187): the operator of every coarse level is rebuilt
 * by each restriction, as in the reference. */
static void w_cycle(clq_ctx *c, int level, int pre, int post, int extra, int vfrom) {
    if (level == c->nlevels - 1) {
        clq_coarse_smooth(c, pre + extra);
        return;
    }
    smooth_level(c, &c->lv[level], pre);
    clq_restrict(c, level);
    w_cycle(c, level + 1, pre, post, extra, vfrom);
    if (twice_at(vfrom, level)) w_cycle(c, level + 1, pre, post, extra, vfrom);
    prolongate_weighted(c, level, c->coarse_weight);
    smooth_level(c, &c->lv[level], post);
}

int clq_w_cycle(clq_ctx *c, int32_t pre, int32_t post, int32_t extra) {
    if (!c || pre < 0 || post < 0 || extra < 0) return fail(CLQ_ERR_INVALID, "bad w_cycle arguments");
    if (c->nlevels < 2) { /* a grid too small to coarsen is solved by the coarse smoother alone */
        clq_coarse_smooth(c, pre + extra);
        return CLQ_OK;
    }
    w_cycle(c, 0, pre, post, extra, 0);
    return CLQ_OK;
}

/* neumann_domain_bc (boundary_conditions.cl:18-31). */
int clq_neumann_bc(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    const int nx = c->nx, ny = c->ny, nz = c->nz;
#pragma omp parallel for schedule(static)
    for (long long index = 0; index < c->C; ++index) {
        int i, j, k;
        coords(index, ny, nz, &i, &j, &k);
        if (interior(i, j, k, nx, ny, nz)) continue;
        const int ii = i < 1 ? 1 : (i > nx - 2 ? nx - 2 : i);
        const int jj = j < 1 ? 1 : (j > ny - 2 ? ny - 2 : j);
        const int kk = k < 1 ? 1 : (k > nz - 2 ? nz - 2 : k);
        c->p[index] = c->p[gidx(ii, jj, kk, ny, nz)];
    }
    return CLQ_OK;
}

/* ||b - A p||_2 and ||b||_2 on level 0 with A applied and accumulated in
 * FP64 from the FP32 arrays: A p = (sum w p - d p) / h^2, h = float(dx), on
 * interior cells with mask > 0 and 0 elsewhere (multigrid.cl:25-43). */
int clq_residual_norm(clq_ctx *c, double *residual, double *rhs_norm) {
    if (!c || !residual || !rhs_norm) return fail(CLQ_ERR_INVALID, "NULL argument");
    const level_t *L = &c->lv[0];
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    const double h = (double)(float)c->dx;
    const double inv_h2 = 1.0 / (h * h);
    double rr = 0.0, bb = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr, bb)
    for (long long index = 0; index < L->C; ++index) {
        int i, j, k;
        coords(index, ny, nz, &i, &j, &k);
        const double b = L->b[index];
        double ap = 0.0;
        if (interior(i, j, k, nx, ny, nz) && L->mask[index] > 0.0f) {
            const float *p = L->p;
            double s = (double)L->wx[index] * p[gidx(i + 1, j, k, ny, nz)] +
                       (double)L->wx[gidx(i - 1, j, k, ny, nz)] * p[gidx(i - 1, j, k, ny, nz)] +
                       (double)L->wy[index] * p[gidx(i, j + 1, k, ny, nz)] +
                       (double)L->wy[gidx(i, j - 1, k, ny, nz)] * p[gidx(i, j - 1, k, ny, nz)] +
                       (double)L->wz[index] * p[gidx(i, j, k + 1, ny, nz)] +
                       (double)L->wz[gidx(i, j, k - 1, ny, nz)] * p[gidx(i, j, k - 1, ny, nz)];
            ap = (s - (double)L->diag[index] * p[index]) * inv_h2;
        }
        const double r = b - ap;
        rr += r * r;
        bb += b * b;
    }
    *residual = sqrt(rr);
    *rhs_norm = sqrt(bb);
    return CLQ_OK;
}

/* ---- mixed-precision iterative refinement (no reference counterpart) ------- */
/* coarse_smooth in FP64: the coarsest level's p, b and operator are widened
 * to double, swept with the same red-black order, and p is rounded back. */
static void coarse_smooth_f64(clq_ctx *c, int iterations) {
    level_t *L = &c->lv[c->nlevels - 1];
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    const long long sx = (long long)ny * nz, sy = nz;
    const double delta = (double)(float)c->dx;
    const double dd = delta * delta;
    double *p = malloc((size_t)L->C * sizeof(double));
    for (long long q = 0; q < L->C; ++q) p[q] = L->p[q];
    for (int it = 0; it < iterations; ++it)
        for (int parity = 0; parity < 2; ++parity)
            for (long long q = 0; q < L->C; ++q) {
                int i, j, k;
                coords(q, ny, nz, &i, &j, &k);
                if (!interior(i, j, k, nx, ny, nz) || L->mask[q] <= 0.0f || ((i + j + k) & 1) != parity) continue;
                const double d = L->diag[q];
                if (!(d > 0.0)) continue;
                double s = (double)L->wx[q - sx] * p[q - sx];
                s = fma((double)L->wx[q], p[q + sx], s);
                s = fma((double)L->wy[q], p[q + sy], s);
                s = fma((double)L->wy[q - sy], p[q - sy], s);
                s = fma((double)L->wz[q], p[q + 1], s);
                s = fma((double)L->wz[q - 1], p[q - 1], s);
                p[q] = fma(-dd, (double)load_b(L, q), s) / d;
            }
    for (long long q = 0; q < L->C; ++q) L->p[q] = (float)p[q];
    free(p);
}

/* f64: the coarsest sweeps in FP64 (mode 2's refinement); otherwise the
 * reference's FP32 coarse_smooth (the FCG preconditioner). */
static void w_cycle_mixed(clq_ctx *c, int level, int pre, int post, int extra, int vfrom, int f64) {
    if (level == c->nlevels - 1) {
        if (f64) coarse_smooth_f64(c, pre + extra);
        else clq_coarse_smooth(c, pre + extra);
        return;
    }
    smooth_level(c, &c->lv[level], pre);
    clq_restrict(c, level);
    w_cycle_mixed(c, level + 1, pre, post, extra, vfrom, f64);
    if (twice_at(vfrom, level)) w_cycle_mixed(c, level + 1, pre, post, extra, vfrom, f64);
    prolongate_weighted(c, level, c->coarse_weight);
    smooth_level(c, &c->lv[level], post);
}

/* One refinement update on level 0: p_new = p_old + e on interior fluid
 * cells (0 elsewhere), r = b - A p_new in FP64 (left-to-right sums, no fused
 * operations), r32 = float(r); returns ||r||^2 and ||b||^2.  With from32 the
 * iterate is taken from the FP32 pressure instead (start of the solve). */
static void ir_update_residual(clq_ctx *c, int from32, const double *pin, double *pout, double *rr_out,
                               double *bb_out) {
    const level_t *L = &c->lv[0];
    const int nx = L->nx, ny = L->ny, nz = L->nz;
    const long long sx = (long long)ny * nz, sy = nz;
    const double h = (double)(float)c->dx;
    const double inv_h2 = 1.0 / (h * h);
    const float *e = c->p;
    const float *b = c->rhs;
    double rr = 0.0, bb = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr, bb)
    for (long long q = 0; q < L->C; ++q) {
        int i, j, k;
        coords(q, ny, nz, &i, &j, &k);
        const int fluid = interior(i, j, k, nx, ny, nz) && L->mask[q] > 0.0f;
        double r = 0.0;
        if (fluid) {
#define PNEW(x) (from32 ? (double)c->p[x] : pin[x] + (double)e[x])
            const double pc = PNEW(q);
            double s = 0.0;
            if (L->wx[q] != 0.0f) s = s + (double)L->wx[q] * PNEW(q + sx);
            if (L->wx[q - sx] != 0.0f) s = s + (double)L->wx[q - sx] * PNEW(q - sx);
            if (L->wy[q] != 0.0f) s = s + (double)L->wy[q] * PNEW(q + sy);
            if (L->wy[q - sy] != 0.0f) s = s + (double)L->wy[q - sy] * PNEW(q - sy);
            if (L->wz[q] != 0.0f) s = s + (double)L->wz[q] * PNEW(q + 1);
            if (L->wz[q - 1] != 0.0f) s = s + (double)L->wz[q - 1] * PNEW(q - 1);
#undef PNEW
            pout[q] = pc;
            r = (double)b[q] - (s - (double)L->diag[q] * pc) * inv_h2;
        } else {
            pout[q] = 0.0;
        }
        c->r32[q] = (float)r;
        rr += r * r;
        bb += (double)b[q] * (double)b[q];
    }
    *rr_out = rr;
    *bb_out = bb;
}

static int solve_mixed(clq_ctx *c, const clq_solve_params *sp, clq_solve_stats *s) {
    if (!c->p64a) {
        c->p64a = xcalloc(c->C, sizeof(double));
        c->p64b = xcalloc(c->C, sizeof(double));
        c->r32 = xcalloc(c->C, sizeof(float));
        for (int l = 1; l < c->nlevels; ++l) c->lv[l].b16 = xcalloc(c->lv[l].C, sizeof(uint16_t));
        if (!c->p64a || !c->p64b || !c->r32) return fail(CLQ_ERR_NOMEM, "mixed-precision buffers");
    }
    double rr, bb;
    ir_update_residual(c, 1, NULL, c->p64a, &rr, &bb);
    double r = sqrt(rr);
    const double bn = sqrt(bb);
    if (!isfinite(r) || !isfinite(bn)) return non_finite(r + bn);
    s->initial_residual = r;
    s->rhs_norm = bn;
    const int kind = sp->coarse_dtype == 1 ? 2 : 1;
    level_t *L0 = &c->lv[0];
    for (int cyc = 0; cyc < sp->max_cycles && r > sp->tolerance * bn; ++cyc) {
        /* per-level power-of-two scale: coarse b grows ~8x per level from the
         * RMS of the fine residual */
        const int es = (int)lround(log2(r / sqrt((double)c->C) + 1e-300));
        for (int l = 1; l < c->nlevels; ++l) {
            c->lv[l].bkind = kind;
            c->lv[l].bscale = ldexpf(1.0f, 3 * l + es);
        }
        L0->b = c->r32;
        memset(c->p, 0, (size_t)c->C * sizeof(float));
        w_cycle_mixed(c, 0, sp->pre_smooth, sp->post_smooth, sp->coarse_extra, sp->vcycle_from, 1);
        L0->b = c->rhs;
        ir_update_residual(c, 0, c->p64a, c->p64b, &rr, &bb);
        double *t = c->p64a; c->p64a = c->p64b; c->p64b = t;
        s->cycles++;
        const double before = r;
        r = sqrt(rr);
        if (!isfinite(r)) {
            for (int l = 1; l < c->nlevels; ++l) c->lv[l].bkind = 0;
            L0->b = c->rhs;
            return non_finite(r);
        }
        if (r > 0.95 * before) break;
    }
    for (int l = 1; l < c->nlevels; ++l) c->lv[l].bkind = 0;
    for (long long q = 0; q < c->C; ++q) c->p[q] = (float)c->p64a[q];
    s->final_residual = r;
    s->relative_residual = bn > 0 ? r / bn : 0.0;
    return CLQ_OK;
}


/* ---- flexible CG (no reference counterpart) --------------------------------- */
/* A v at a fluid cell in FP64, left-to-right over the non-zero weights. */
static inline double apply_a64(const level_t *L, long long q, const double *v, double inv_h2) {
    const long long sx = (long long)L->ny * L->nz, sy = L->nz;
    double s = 0.0;
    if (L->wx[q] != 0.0f) s = s + (double)L->wx[q] * v[q + sx];
    if (L->wx[q - sx] != 0.0f) s = s + (double)L->wx[q - sx] * v[q - sx];
    if (L->wy[q] != 0.0f) s = s + (double)L->wy[q] * v[q + sy];
    if (L->wy[q - sy] != 0.0f) s = s + (double)L->wy[q - sy] * v[q - sy];
    if (L->wz[q] != 0.0f) s = s + (double)L->wz[q] * v[q + 1];
    if (L->wz[q - 1] != 0.0f) s = s + (double)L->wz[q - 1] * v[q - 1];
    return (s - (double)L->diag[q] * v[q]) * inv_h2;
}

/* apply_a64 of an FP32 vector (the preconditioned residual z) */
static inline double apply_a64_f32(const level_t *L, long long q, const float *v, double inv_h2) {
    const long long sx = (long long)L->ny * L->nz, sy = L->nz;
    double s = 0.0;
    if (L->wx[q] != 0.0f) s = s + (double)L->wx[q] * v[q + sx];
    if (L->wx[q - sx] != 0.0f) s = s + (double)L->wx[q - sx] * v[q - sx];
    if (L->wy[q] != 0.0f) s = s + (double)L->wy[q] * v[q + sy];
    if (L->wy[q - sy] != 0.0f) s = s + (double)L->wy[q - sy] * v[q - sy];
    if (L->wz[q] != 0.0f) s = s + (double)L->wz[q] * v[q + 1];
    if (L->wz[q - 1] != 0.0f) s = s + (double)L->wz[q - 1] * v[q - 1];
    return (s - (double)L->diag[q] * v[q]) * inv_h2;
}

static inline int fluid_at(const level_t *L, long long q) {
    int i, j, k;
    coords(q, L->ny, L->nz, &i, &j, &k);
    return interior(i, j, k, L->nx, L->ny, L->nz) && L->mask[q] > 0.0f;
}

/* z = M r: the mixed cycle on A e = r32 from e = 0, 16-bit coarse b and the
 * FP64 coarse smoother on the coarsest level. */
static void precondition_wcycle(clq_ctx *c, const clq_solve_params *sp, double rnorm) {
    const int kind = sp->coarse_dtype == 1 ? 2 : 1;
    const int es = (int)lround(log2(rnorm / sqrt((double)c->C) + 1e-300));
    for (int l = 1; l < c->nlevels; ++l) {
        c->lv[l].bkind = kind;
        c->lv[l].bscale = ldexpf(1.0f, 3 * l + es);
    }
    c->lv[0].b = c->r32;
    memset(c->p, 0, (size_t)c->C * sizeof(float));
    w_cycle_mixed(c, 0, sp->pre_smooth, sp->post_smooth, sp->coarse_extra, sp->vcycle_from, 1);
    c->lv[0].b = c->rhs;
    for (int l = 1; l < c->nlevels; ++l) c->lv[l].bkind = 0;
}

static int solve_fcg(clq_ctx *c, const clq_solve_params *sp, clq_solve_stats *s) {
    if (!c->p64a) {
        c->p64a = xcalloc(c->C, sizeof(double));
        c->p64b = xcalloc(c->C, sizeof(double));
        c->r32 = xcalloc(c->C, sizeof(float));
        for (int l = 1; l < c->nlevels; ++l) c->lv[l].b16 = xcalloc(c->lv[l].C, sizeof(uint16_t));
    }
    if (!c->r64) {
        c->r64 = xcalloc(c->C, sizeof(double));
        c->dir0 = xcalloc(c->C, sizeof(float));
        c->dir1 = xcalloc(c->C, sizeof(float));
    }
    if (!c->p64a || !c->r64 || !c->dir0 || !c->dir1) return fail(CLQ_ERR_NOMEM, "FCG buffers");
    const level_t *L = &c->lv[0];
    const double h = (double)(float)c->dx;
    const double inv_h2 = 1.0 / (h * h);
    double *x = c->p64a, *r = c->r64;
    float *dold = c->dir0, *dnew = c->dir1;
    /* x = p (fluid), r = b - A x */
    double rr = 0.0, bb = 0.0;
    for (long long t = 0; t < c->C; ++t) x[t] = fluid_at(L, t) ? (double)c->p[t] : 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr, bb)
    for (long long t = 0; t < c->C; ++t) {
        double rv = 0.0;
        if (fluid_at(L, t)) rv = (double)c->rhs[t] - apply_a64(L, t, x, inv_h2);
        r[t] = rv;
        c->r32[t] = (float)rv;
        rr += rv * rv;
        bb += (double)c->rhs[t] * (double)c->rhs[t];
    }
    double rn = sqrt(rr);
    const double bn = sqrt(bb);
    if (!isfinite(rn) || !isfinite(bn)) return non_finite(rn + bn);
    s->initial_residual = rn;
    s->rhs_norm = bn;
    s->recurrence_residual = rn;
    int plain = 0;
    /* Each pass starts the direction recurrence from r = b - A x: the
     * initialisation, or the residual recomputed from x after a pass whose
     * recurrence met the tolerance (or ran out of cycles). */
    while (rn > sp->tolerance * bn && s->cycles < sp->max_cycles) {
        /* The direction streams with the update: the dots of z give
         * beta = -(z.A d_old)/(d.A d)_old and the new direction's
         * d.A d = z.w + 2 beta z.A d_old + beta^2 (d.A d)_old with w = A z
         * (A symmetric, so z.A d_old = w.d_old), hence alpha = z.r / d.A d
         * before d = fmaf(float(beta), d_old, z) is formed; q = A d is then
         * applied to that stored FP32 direction, so x and r move by alpha d
         * and alpha A d of the same d. */
        double alpha = 0.0, beta = 0.0, dq = 0.0;
        int first = 1;
        precondition_wcycle(c, sp, rn);
        s->cycles++;
        s->fallback_cycles += plain;
        for (;;) {
            double zq = 0.0, zr = 0.0, zw = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : zq, zr, zw)
            for (long long t = 0; t < c->C; ++t) {
                const double wv = fluid_at(L, t) ? apply_a64_f32(L, t, c->p, inv_h2) : 0.0;
                zq += wv * (double)dold[t];
                zr += (double)c->p[t] * r[t];
                zw += (double)c->p[t] * wv;
            }
            beta = (first || dq == 0.0) ? 0.0 : -zq / dq;
            dq = first ? zw : zw + 2.0 * beta * zq + beta * beta * dq;
            alpha = dq != 0.0 ? zr / dq : 0.0;
            const float beta32 = (float)beta;
#pragma omp parallel for schedule(static)
            for (long long t = 0; t < c->C; ++t)
                dnew[t] = !fluid_at(L, t) ? 0.0f : first ? c->p[t] : fmaf(beta32, dold[t], c->p[t]);
            rr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr)
            for (long long t = 0; t < c->C; ++t) {
                const double qn = fluid_at(L, t) ? apply_a64_f32(L, t, dnew, inv_h2) : 0.0;
                x[t] = x[t] + alpha * (double)dnew[t];
                r[t] = r[t] - alpha * qn;
                c->r32[t] = (float)r[t];
                rr += r[t] * r[t];
            }
            float *tmp = dold;
            dold = dnew;
            dnew = tmp;
            rn = sqrt(rr);
            if (!isfinite(rn)) return non_finite(rn);
            if (rn <= sp->tolerance * bn || s->cycles >= sp->max_cycles) break;
            /* A weighted cycle that is not definite for this residual gives a
             * step length alpha = z.r / d.Ad <= 0 (for a definite one both
             * dots carry the sign of the negative definite operator): go on
             * with weight 1 and restart the direction recurrence. */
            first = c->coarse_weight != 1.0f && !(alpha > 0.0);
            if (first) c->coarse_weight = 1.0f;
            plain = plain || first;
            precondition_wcycle(c, sp, rn);
            s->cycles++;
            s->fallback_cycles += plain;
        }
        /* the residual recomputed from x: the reported value, and the start
         * of another pass when it misses the tolerance */
        s->recurrence_residual = rn;
        rr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rr)
        for (long long t = 0; t < c->C; ++t) {
            double rv = 0.0;
            if (fluid_at(L, t)) rv = (double)c->rhs[t] - apply_a64(L, t, x, inv_h2);
            r[t] = rv;
            c->r32[t] = (float)rv;
            rr += rv * rv;
        }
        rn = sqrt(rr);
        if (!isfinite(rn)) return non_finite(rn);
        if (rn > sp->tolerance * bn && s->cycles < sp->max_cycles) s->restarts++;
    }
    for (long long t = 0; t < c->C; ++t) c->p[t] = (float)x[t];
    s->final_residual = rn;
    s->relative_residual = bn > 0 ? rn / bn : 0.0;
    return CLQ_OK;
}

/* Cells of the level-0 rows (i, j, *) holding an interior cell with mask > 0
 * (the statistic the device reports for its row-restricted sweeps). */
static long long active_cells(const clq_ctx *c) {
    if (c->nlevels < 2) return c->C;
    const level_t *L = &c->lv[0];
    long long rows = 0;
    for (int i = 1; i < L->nx - 1; ++i)
        for (int j = 1; j < L->ny - 1; ++j) {
            const long long base = ((long long)i * L->ny + j) * L->nz;
            for (int k = 1; k < L->nz - 1; ++k)
                if (L->mask[base + k] > 0.0f) {
                    ++rows;
                    break;
                }
        }
    return rows * L->nz;
}

static int pressure_solve(clq_ctx *c, const clq_solve_params *sp, clq_solve_stats *st) {
    if (sp->pre_smooth < 0 || sp->post_smooth < 0 || sp->coarse_extra < 0 || sp->max_cycles < 0)
        return fail(CLQ_ERR_INVALID, "negative sweep or cycle count");
    if (sp->mode != CLQ_SOLVE_FIXED_CYCLES && sp->mode != CLQ_SOLVE_FP32_TOLERANCE &&
        sp->mode != CLQ_SOLVE_MIXED_REFINEMENT && sp->mode != CLQ_SOLVE_FCG)
        return fail(CLQ_ERR_UNSUPPORTED, "solve mode %d is not available in the oracle", sp->mode);
    double t0 = now_ms();
    clq_solve_stats s;
    memset(&s, 0, sizeof s);
    c->p64_valid = 0;
    if (sp->mode == CLQ_SOLVE_MIXED_REFINEMENT || sp->mode == CLQ_SOLVE_FCG) {
        int rc = sp->mode == CLQ_SOLVE_FCG ? solve_fcg(c, sp, &s) : solve_mixed(c, sp, &s);
        if (rc) return rc;
        c->p64_valid = 1;
        clq_neumann_bc(c);
        s.solve_ms = now_ms() - t0;
        s.active_cells = active_cells(c);
        c->stage_ms[CLQ_STAGE_PRESSURE_SOLVE] = s.solve_ms;
        if (st) *st = s;
        return CLQ_OK;
    }
    double r0, bn;
    clq_residual_norm(c, &r0, &bn);
    s.initial_residual = r0;
    s.rhs_norm = bn;
    double r = r0;
    const double target = sp->tolerance * bn;
    for (int cyc = 0; cyc < sp->max_cycles; ++cyc) {
        if (sp->mode == CLQ_SOLVE_FP32_TOLERANCE && r <= target) break;
        w_cycle(c, 0, sp->pre_smooth, sp->post_smooth, sp->coarse_extra, sp->vcycle_from);
        s.cycles++;
        if (sp->mode == CLQ_SOLVE_FP32_TOLERANCE) {
            const double before = r;
            clq_residual_norm(c, &r, &bn);
            if (r > 0.95 * before) break; /* stagnation at the FP32 floor */
        }
    }
    if (sp->mode == CLQ_SOLVE_FIXED_CYCLES) clq_residual_norm(c, &r, &bn);
    if (!isfinite(r) || !isfinite(bn)) return non_finite(r + bn);
    s.final_residual = r;
    s.relative_residual = bn > 0 ? r / bn : 0.0;
    clq_neumann_bc(c); /* pressure_solve.py:84 */
    s.solve_ms = now_ms() - t0;
    s.active_cells = active_cells(c);
    c->stage_ms[CLQ_STAGE_PRESSURE_SOLVE] = s.solve_ms;
    if (st) *st = s;
    return CLQ_OK;
}

/* The solve with the cycles' coarse-grid corrections weighted by
 * params->coarse_weight (0 selects 1, the reference's plain sum). */
int clq_pressure_solve(clq_ctx *c, const clq_solve_params *sp, clq_solve_stats *st) {
    if (!c || !sp) return fail(CLQ_ERR_INVALID, "NULL argument");
    const float w = sp->coarse_weight == 0.0f ? 1.0f : sp->coarse_weight;
    if (!(w > 0.0f && w < 2.0f)) return fail(CLQ_ERR_INVALID, "coarse_weight %g outside (0, 2)", (double)w);
    c->coarse_weight = w;
    const int rc = pressure_solve(c, sp, st);
    c->coarse_weight = 1.0f;
    return rc;
}

/* FP64 check solve (no reference counterpart): Jacobi-preconditioned CG on
 * the level-0 system from x = 0, the preconditioner the operator's diagonal
 * a = -diag / h^2 (z = 0 on a fluid cell without open neighbours); then the
 * residual recomputed from x and the error of the last solve's pressure
 * (the FP64 iterate after modes 2-4, else the FP32 pressure). */
int clq_fp64_check_solve(clq_ctx *c, double tolerance, int32_t max_iterations, clq_check_stats *st) {
    if (!c || !st || !(tolerance > 0.0) || max_iterations < 0) return fail(CLQ_ERR_INVALID, "bad check-solve arguments");
    const double t0 = now_ms();
    const level_t *L = &c->lv[0];
    const long long C = c->C;
    const double h = (double)(float)c->dx;
    const double inv_h2 = 1.0 / (h * h);
    double *x = xcalloc(C, 8), *r = xcalloc(C, 8), *z = xcalloc(C, 8), *p = xcalloc(C, 8), *q = xcalloc(C, 8);
    if (!x || !r || !z || !p || !q) {
        free(x); free(r); free(z); free(p); free(q);
        return fail(CLQ_ERR_NOMEM, "check-solve buffers");
    }
    double bn = 0.0, rel = 0.0;
    int it = 0;
    /* passes of CG: from x = 0, then restarts from the current x while the
     * residual recomputed at the end of a pass misses the tolerance */
    for (int pass = 0; pass < 4; ++pass) {
        double rz = 0.0, rr = 0.0, bb = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rz, rr, bb)
        for (long long t = 0; t < C; ++t) {
            const double bv = (double)c->rhs[t], dg = (double)L->diag[t];
            const int fl = fluid_at(L, t);
            r[t] = fl ? bv - (pass > 0 ? apply_a64(L, t, x, inv_h2) : 0.0) : 0.0;
            z[t] = fl && dg > 0.0 ? r[t] / (-dg * inv_h2) : 0.0;
            p[t] = z[t];
            rz += r[t] * z[t];
            rr += r[t] * r[t];
            bb += bv * bv;
        }
        bn = sqrt(bb);
        /* the recurrence is driven to half the tolerance, so that the
         * residual recomputed from x at the end meets it */
        while (sqrt(rr) > 0.5 * tolerance * bn && it < max_iterations) {
            double pq = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : pq)
            for (long long t = 0; t < C; ++t) {
                q[t] = fluid_at(L, t) ? apply_a64(L, t, p, inv_h2) : 0.0;
                pq += p[t] * q[t];
            }
            const double alpha = pq != 0.0 ? rz / pq : 0.0;
            double rz_new = 0.0;
            rr = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rz_new, rr)
            for (long long t = 0; t < C; ++t) {
                x[t] += alpha * p[t];
                r[t] -= alpha * q[t];
                const double dg = (double)L->diag[t];
                z[t] = fluid_at(L, t) && dg > 0.0 ? r[t] / (-dg * inv_h2) : 0.0;
                rz_new += r[t] * z[t];
                rr += r[t] * r[t];
            }
            const double beta = rz != 0.0 ? rz_new / rz : 0.0;
            rz = rz_new;
#pragma omp parallel for schedule(static)
            for (long long t = 0; t < C; ++t) p[t] = z[t] + beta * p[t];
            ++it;
            if (!isfinite(rr)) break;
        }
        double rt = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : rt)
        for (long long t = 0; t < C; ++t) {
            if (!fluid_at(L, t)) continue;
            const double rv = (double)c->rhs[t] - apply_a64(L, t, x, inv_h2);
            rt += rv * rv;
        }
        if (!isfinite(rt)) {
            free(x); free(r); free(z); free(p); free(q);
            return fail(CLQ_ERR_NOT_CONVERGED, "check solve: residual norm is not finite");
        }
        rel = bn > 0 ? sqrt(rt) / bn : 0.0;
        if (rel <= tolerance || it >= max_iterations) break;
    }
    double e2 = 0.0, x2 = 0.0, emax = 0.0, xmax = 0.0;
    for (long long t = 0; t < C; ++t) {
        if (!fluid_at(L, t)) continue;
        const double pv = c->p64_valid ? c->p64a[t] : (double)c->p[t];
        const double e = pv - x[t];
        e2 += e * e;
        x2 += x[t] * x[t];
        if (fabs(e) > emax) emax = fabs(e);
        if (fabs(x[t]) > xmax) xmax = fabs(x[t]);
    }
    free(x); free(r); free(z); free(p); free(q);
    st->iterations = it;
    st->relative_residual = rel;
    st->converged = rel <= tolerance;
    st->error_l2 = x2 > 0 ? sqrt(e2 / x2) : sqrt(e2);
    st->error_max = xmax > 0 ? emax / xmax : emax;
    st->solve_ms = now_ms() - t0;
    return CLQ_OK;
}

/* pressure_correction (pressure_solve.cl:26-47, gradient helper.cl:25-37). */
int clq_pressure_correction(clq_ctx *c, double rho, double dt) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const float dt_over_rho = (float)(dt / rho), inv_dx = (float)(1.0 / c->dx);
    if (c->p64_valid) { /* FP64 pressure of modes 2-4; +1 neighbour clamped to the interior */
        const double *p = c->p64a;
        /* rounded through float explicitly: gcc's OpenMP outlining of a
         * const float widened to double has been seen to keep the unrounded
         * double initialiser */
        volatile float k32 = (float)(dt / rho), id32 = (float)(1.0 / c->dx);
        const double k = (double)k32, id = (double)id32;
#pragma omp parallel for schedule(static)
        for (long long index = 0; index < c->C; ++index) {
            int i, j, kk;
            coords(index, ny, nz, &i, &j, &kk);
            if (!interior(i, j, kk, nx, ny, nz)) continue;
            const double centre = p[index];
            const double gx = (p[gidx(i + 1 > nx - 2 ? nx - 2 : i + 1, j, kk, ny, nz)] - centre) * id;
            const double gy = (p[gidx(i, j + 1 > ny - 2 ? ny - 2 : j + 1, kk, ny, nz)] - centre) * id;
            const double gz = (p[gidx(i, j, kk + 1 > nz - 2 ? nz - 2 : kk + 1, ny, nz)] - centre) * id;
            c->u[index] = (float)((double)c->u[index] - k * gx);
            c->v[index] = (float)((double)c->v[index] - k * gy);
            c->w[index] = (float)((double)c->w[index] - k * gz);
        }
        return CLQ_OK;
    }
#pragma omp parallel for schedule(static)
    for (long long index = 0; index < c->C; ++index) {
        int i, j, k;
        coords(index, ny, nz, &i, &j, &k);
        if (!interior(i, j, k, nx, ny, nz)) continue;
        const float centre = c->p[index];
        const float gx = (c->p[gidx(i + 1, j, k, ny, nz)] - centre) * inv_dx;
        const float gy = (c->p[gidx(i, j + 1, k, ny, nz)] - centre) * inv_dx;
        const float gz = (c->p[gidx(i, j, k + 1, ny, nz)] - centre) * inv_dx;
        c->u[index] = fmaf(-dt_over_rho, gx, c->u[index]); /* :44-46 fmuladd */
        c->v[index] = fmaf(-dt_over_rho, gy, c->v[index]);
        c->w[index] = fmaf(-dt_over_rho, gz, c->w[index]);
    }
    return CLQ_OK;
}

/* ---- particles ------------------------------------------------------------ */
/* sample_field (particles.cl:30-42) at the staggered base/fraction of
 * particles.cl:12-21: base = floor(fma(x, inv_dx, -off)) clamped, fraction =
 * clamp(fma(x, inv_dx, -off) - base, 0, 1). */
static inline float sample_staggered(const float *field, const float *x, float inv_dx, const float *off, int nx, int ny,
                                     int nz) {
    const int n[3] = {nx, ny, nz};
    int base[3];
    float fr[3];
    for (int a = 0; a < 3; ++a) {
        const float t = fmaf(x[a], inv_dx, -off[a]);
        base[a] = floor_clamp(t, 0, n[a] - 2);
        float f = t - (float)base[a];
        fr[a] = fminf(fmaxf(f, 0.0f), 1.0f);
    }
    float value = 0.0f;
    for (int ox = 0; ox <= 1; ++ox)
        for (int oy = 0; oy <= 1; ++oy)
            for (int oz = 0; oz <= 1; ++oz) {
                const float wx = ox ? fr[0] : 1.0f - fr[0];
                const float wy = oy ? fr[1] : 1.0f - fr[1];
                const float wz = oz ? fr[2] : 1.0f - fr[2];
                const float wgt = wx * wy * wz;
                value = fmaf(wgt, field[gidx(base[0] + ox, base[1] + oy, base[2] + oz, ny, nz)], value); /* :37 */
            }
    return value;
}

static const float U_OFF[3] = {1.0f, 0.5f, 0.5f};
static const float V_OFF[3] = {0.5f, 1.0f, 0.5f};
static const float W_OFF[3] = {0.5f, 0.5f, 1.0f};

/* grid_to_particle (particles.cl:229-262) with the FLIP/PIC blend. */
static inline void g2p_one(const clq_ctx *c, long long i, float inv_dx, int pure_flip, float a, float b) {
    const float *x = c->pos + 3 * i;
    float *vel = c->vel + 3 * i;
    const float nu = sample_staggered(c->u, x, inv_dx, U_OFF, c->nx, c->ny, c->nz);
    const float nv = sample_staggered(c->v, x, inv_dx, V_OFF, c->nx, c->ny, c->nz);
    const float nw = sample_staggered(c->w, x, inv_dx, W_OFF, c->nx, c->ny, c->nz);
    const float du = nu - sample_staggered(c->uo, x, inv_dx, U_OFF, c->nx, c->ny, c->nz);
    const float dv = nv - sample_staggered(c->vo, x, inv_dx, V_OFF, c->nx, c->ny, c->nz);
    const float dw = nw - sample_staggered(c->wo, x, inv_dx, W_OFF, c->nx, c->ny, c->nz);
    const float f0 = vel[0] + du, f1 = vel[1] + dv, f2 = vel[2] + dw;
    if (pure_flip) {
        vel[0] = f0; vel[1] = f1; vel[2] = f2;
    } else {
        vel[0] = fmaf(a, f0, b * nu);
        vel[1] = fmaf(a, f1, b * nv);
        vel[2] = fmaf(a, f2, b * nw);
    }
}

int clq_grid_to_particle(clq_ctx *c, double flip_ratio) {
    if (!c || !(flip_ratio >= 0.0 && flip_ratio <= 1.0)) return fail(CLQ_ERR_INVALID, "flip_ratio must be in [0, 1]");
    const float inv_dx = (float)(1.0 / c->dx);
    const int pure = flip_ratio == 1.0;
    const float a = (float)flip_ratio, b = (float)(1.0 - flip_ratio);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < c->n; ++i) g2p_one(c, i, inv_dx, pure, a, b);
    return CLQ_OK;
}

int clq_save_previous_positions(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    memcpy(c->prev, c->pos, (size_t)c->n * 12);
    return CLQ_OK;
}

int clq_advect_particles(clq_ctx *c, double dt) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    c->bins_current = 0;
    const float fdt = (float)dt;
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < 3 * c->n; ++i) c->pos[i] = fmaf(fdt, c->vel[i], c->pos[i]); /* particles.cl:275 */
    return CLQ_OK;
}

/* particle_domain_bc (boundary_conditions.cl:69-97). */
static inline void domain_bc_one(float *x, float *v, float lower, const float *upper) {
    int hit[3];
    for (int a = 0; a < 3; ++a) {
        const int lo = x[a] < lower, hi = x[a] > upper[a];
        hit[a] = lo | hi;
        float y = x[a];
        if (lo) y = fmaf(2.0f, lower, -y);
        if (hi) y = fmaf(2.0f, upper[a], -y);
        y = fminf(fmaxf(y, lower), upper[a]);
        x[a] = y;
    }
    for (int a = 0; a < 3; ++a)
        if (hit[a]) v[a] = 0.0f;
}

static void domain_bounds(const clq_ctx *c, float *lower, float *upper) {
    *lower = (float)(1.001 * c->dx); /* boundary_conditions.py:98-101 */
    upper[0] = (float)((c->nx - 1.001) * c->dx);
    upper[1] = (float)((c->ny - 1.001) * c->dx);
    upper[2] = (float)((c->nz - 1.001) * c->dx);
}

int clq_particle_domain_bc(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    c->bins_current = 0;
    float lower, upper[3];
    domain_bounds(c, &lower, upper);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < c->n; ++i) domain_bc_one(c->pos + 3 * i, c->vel + 3 * i, lower, upper);
    return CLQ_OK;
}

/* particle_obstacle_bc (boundary_conditions.cl:175-264).  length() is
 * sqrtf(x*x + y*y + z*z) and dot() is a left-to-right sum of products; both
 * are OpenCL builtins whose rounding the reference does not fix. */
static inline void obstacle_bc_one(const clq_ctx *c, long long pi, float inv_dx) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const uint8_t *ob = c->obstacle;
    float *x = c->pos + 3 * pi, *v = c->vel + 3 * pi;
    const float *pv = c->prev + 3 * pi;
    const int cx = trunc_clamp(x[0] * inv_dx, 0, nx - 1), cy = trunc_clamp(x[1] * inv_dx, 0, ny - 1),
              cz = trunc_clamp(x[2] * inv_dx, 0, nz - 1);
    if (!ob[gidx(cx, cy, cz, ny, nz)]) return;
    const int px = trunc_clamp(pv[0] * inv_dx, 0, nx - 1), py = trunc_clamp(pv[1] * inv_dx, 0, ny - 1),
              pz = trunc_clamp(pv[2] * inv_dx, 0, nz - 1);
    float n[3] = {0.0f, 0.0f, 0.0f};
    if ((px != cx || py != cy || pz != cz) && !ob[gidx(px, py, pz, ny, nz)]) {
        n[0] = (float)(px - cx); n[1] = (float)(py - cy); n[2] = (float)(pz - cz);
    } else {
        if (cx > 0 && !ob[gidx(cx - 1, cy, cz, ny, nz)]) n[0] -= 1.0f;
        if (cx + 1 < nx && !ob[gidx(cx + 1, cy, cz, ny, nz)]) n[0] += 1.0f;
        if (cy > 0 && !ob[gidx(cx, cy - 1, cz, ny, nz)]) n[1] -= 1.0f;
        if (cy + 1 < ny && !ob[gidx(cx, cy + 1, cz, ny, nz)]) n[1] += 1.0f;
        if (cz > 0 && !ob[gidx(cx, cy, cz - 1, ny, nz)]) n[2] -= 1.0f;
        if (cz + 1 < nz && !ob[gidx(cx, cy, cz + 1, ny, nz)]) n[2] += 1.0f;
    }
    const float len = sqrtf(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (len > 0.0f) {
        n[0] = n[0] / len; n[1] = n[1] / len; n[2] = n[2] / len;
    } else {
        const float f[3] = {pv[0] - x[0], pv[1] - x[1], pv[2] - x[2]};
        const float fl = sqrtf(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
        if (fl > 0.0f) { n[0] = f[0] / fl; n[1] = f[1] / fl; n[2] = f[2] / fl; }
        else { n[0] = 0.0f; n[1] = 0.0f; n[2] = 1.0f; }
    }
    const float vn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
    if (vn < 0.0f) { /* :237 fmuladd */
        v[0] = fmaf(-vn, n[0], v[0]); v[1] = fmaf(-vn, n[1], v[1]); v[2] = fmaf(-vn, n[2], v[2]);
    }
    const float dx = 1.0f / inv_dx;
    const int cell[3] = {cx, cy, cz};
    float corr[3] = {x[0], x[1], x[2]};
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 0.0f) corr[a] = fmaf(2.0f, (float)cell[a] * dx, -x[a]);
        else if (n[a] > 0.0f) corr[a] = fmaf(2.0f, (float)(cell[a] + 1) * dx, -x[a]);
    }
    const int qx = trunc_clamp(corr[0] * inv_dx, 0, nx - 1), qy = trunc_clamp(corr[1] * inv_dx, 0, ny - 1),
              qz = trunc_clamp(corr[2] * inv_dx, 0, nz - 1);
    if (ob[gidx(qx, qy, qz, ny, nz)]) { corr[0] = pv[0]; corr[1] = pv[1]; corr[2] = pv[2]; }
    x[0] = corr[0]; x[1] = corr[1]; x[2] = corr[2];
}

int clq_particle_obstacle_bc(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    c->bins_current = 0;
    const float inv_dx = (float)(1.0 / c->dx);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < c->n; ++i) obstacle_bc_one(c, i, inv_dx);
    return CLQ_OK;
}

int clq_grid_to_particle_advect(clq_ctx *c, double flip_ratio, double dt) {
    double t0 = now_ms();
    int rc = clq_grid_to_particle(c, flip_ratio);
    if (!rc) rc = clq_save_previous_positions(c);
    if (!rc) rc = clq_advect_particles(c, dt);
    if (!rc) rc = clq_particle_domain_bc(c);
    if (!rc) rc = clq_particle_obstacle_bc(c);
    if (!rc) c->stage_ms[CLQ_STAGE_G2P_ADVECT] = now_ms() - t0;
    return rc;
}

/* ---- voxelisation ---------------------------------------------------------- */
/* Occupancy at scale s: a particle marks the refined cell
 * clamp(trunc(x * (s * inv_dx)), s, s*(n-1)-1); at s = 1 this is the cell of
 * particles.cl:186-188, so occupancy equals fluid_fraction > 0.  Smoothed
 * density: each particle splats unit mass with trilinear weights onto the
 * refined cell centres, scaled by s^3 / particles_per_cell so a uniformly
 * seeded region reads 1.  Summation runs particle by particle in the current
 * (binned) order, in FP64, and is rounded once. */
/* ---- particle sources (source_particles.py:42-92, applied at solver.py:253-306) ---- */
int clq_set_sources(clq_ctx *c, int32_t count, const int32_t *types, const uint8_t *masks, const int32_t *densities,
                    const float *velocities) {
    if (!c || count < 0 || (count > 0 && (!types || !masks || !densities || !velocities)))
        return fail(CLQ_ERR_INVALID, "bad source arguments");
    for (int32_t q = 0; q < count; ++q)
        if (types[q] < CLQ_SOURCE_GEOMETRY || types[q] > CLQ_SOURCE_OUTFLOW)
            return fail(CLQ_ERR_INVALID, "unknown source type %d", types[q]);
    free(c->src_type); free(c->src_density); free(c->src_velocity); free(c->src_open); free(c->sink);
    c->nsources = count;
    c->src_type = xcalloc(count, 4);
    c->src_density = xcalloc(count, 4);
    c->src_velocity = xcalloc(3 * (size_t)count, 4);
    c->src_open = xcalloc((size_t)count * c->C, 1);
    c->sink = xcalloc(c->C, 1);
    if (!c->src_type || !c->src_density || !c->src_velocity || !c->src_open || !c->sink)
        return fail(CLQ_ERR_NOMEM, "source allocation failed");
    for (int32_t q = 0; q < count; ++q) {
        c->src_type[q] = types[q];
        c->src_density[q] = densities[q] > 1 ? densities[q] : 1;
        for (int a = 0; a < 3; ++a) c->src_velocity[3 * q + a] = velocities[3 * q + a];
        const uint8_t *m = masks + (size_t)q * c->C;
        for (long long i = 0; i < c->C; ++i) {
            c->src_open[(size_t)q * c->C + i] = m[i] && !c->obstacle[i];
            if (types[q] == CLQ_SOURCE_OUTFLOW && m[i]) c->sink[i] = 1;
        }
    }
    return CLQ_OK;
}

typedef unsigned __int128 u128;
static const u128 PCG_MULT = ((u128)0x2360ED051FC65DA4ull << 64) | (u128)0x4385DF649FCCF645ull;

/* numpy's Generator.random() on PCG64 (XSL-RR 128/64): the state steps, then
 * its output, >> 11, times 2^-53 */
static double pcg_next_double(u128 *state, u128 inc) {
    *state = *state * PCG_MULT + inc;
    const uint64_t x = (uint64_t)(*state >> 64) ^ (uint64_t)*state;
    const unsigned rot = (unsigned)(*state >> 122);
    const uint64_t r = (x >> rot) | (x << ((64u - rot) & 63u));
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

static long long source_cell(const clq_ctx *c, const float *x) {
    const float delta = (float)c->dx;
    return gidx(floor_clamp(x[0] / delta, 0, c->nx - 1), floor_clamp(x[1] / delta, 0, c->ny - 1),
                floor_clamp(x[2] / delta, 0, c->nz - 1), c->ny, c->nz);
}

int clq_apply_sources(clq_ctx *c, int32_t initial, const uint64_t *rng, int64_t *draws) {
    if (!c || !rng || !draws) return fail(CLQ_ERR_INVALID, "NULL argument");
    u128 state = ((u128)rng[0] << 64) | rng[1];
    const u128 inc = ((u128)rng[2] << 64) | rng[3];
    long long used = 0;
    /* sinks: the particles in any OUTFLOW cell go, the rest keep their order */
    int any_sink = 0;
    for (int q = 0; q < c->nsources; ++q) any_sink |= c->src_type[q] == CLQ_SOURCE_OUTFLOW;
    if (any_sink) {
        long long m = 0;
        for (long long i = 0; i < c->n; ++i) {
            if (c->sink[source_cell(c, c->pos + 3 * i)]) continue;
            memmove(c->pos + 3 * m, c->pos + 3 * i, 12);
            memmove(c->vel + 3 * m, c->vel + 3 * i, 12);
            ++m;
        }
        c->n = m;
    }
    int32_t *present = xcalloc(c->C, 4);
    if (!present) return fail(CLQ_ERR_NOMEM, "source counts");
    for (int q = 0; q < c->nsources; ++q) {
        const int t = c->src_type[q];
        if (!(t == CLQ_SOURCE_INFLOW || (initial && t == CLQ_SOURCE_GEOMETRY))) continue;
        const uint8_t *open = c->src_open + (size_t)q * c->C;
        memset(present, 0, (size_t)c->C * 4);
        for (long long i = 0; i < c->n; ++i) {
            const long long cell = source_cell(c, c->pos + 3 * i);
            if (open[cell]) present[cell]++;
        }
        long long total = 0;
        for (long long cell = 0; cell < c->C; ++cell)
            if (open[cell] && present[cell] < c->src_density[q]) total += c->src_density[q] - present[cell];
        if (total == 0) continue; /* no draw, as on the host */
        if (c->n + total > c->cap) { /* solver.py:290-291, keeping the particles */
            long long cap = c->cap * 3 / 2;
            if (cap < c->n + total) cap = c->n + total;
            float *np_ = xcalloc(cap * 3, 4), *nv = xcalloc(cap * 3, 4);
            if (!np_ || !nv) { free(np_); free(nv); free(present); return fail(CLQ_ERR_NOMEM, "source growth"); }
            memcpy(np_, c->pos, (size_t)c->n * 12);
            memcpy(nv, c->vel, (size_t)c->n * 12);
            free(c->pos_work); free(c->vel_work); free(c->prev);
            c->pos_work = xcalloc(cap * 3, 4);
            c->vel_work = xcalloc(cap * 3, 4);
            c->prev = xcalloc(cap * 3, 4);
            free(c->pos); free(c->vel);
            c->pos = np_; c->vel = nv; c->cap = cap;
            if (!c->pos_work || !c->vel_work || !c->prev) { free(present); return fail(CLQ_ERR_NOMEM, "source growth"); }
        }
        long long slot = c->n;
        for (long long cell = 0; cell < c->C; ++cell) {
            if (!open[cell] || present[cell] >= c->src_density[q]) continue;
            int i, j, k;
            coords(cell, c->ny, c->nz, &i, &j, &k);
            const double corner[3] = {(double)i, (double)j, (double)k};
            for (int m = present[cell]; m < c->src_density[q]; ++m, ++slot) {
                for (int a = 0; a < 3; ++a)
                    c->pos[3 * slot + a] = (float)((corner[a] + pcg_next_double(&state, inc)) * c->dx);
                for (int a = 0; a < 3; ++a) c->vel[3 * slot + a] = c->src_velocity[3 * q + a];
            }
        }
        used += 3 * total;
        c->n += total;
    }
    free(present);
    *draws = used;
    c->bins_current = 0;
    return CLQ_OK;
}

/* ---- mesh voxelisation (Solver/Kernel/OpenCL/voxelise_mesh.cl) ---------------
 * The OpenCL builtins are restated by their definitions with separately
 * rounded operations: cross(a, b) = (a.y b.z - a.z b.y, a.z b.x - a.x b.z,
 * a.x b.y - a.y b.x), dot as a left-to-right sum of products, normalize(v) =
 * v / sqrt(dot(v, v)).  The cell centre o + (i + 0.5) * delta is the one
 * expression of the kernel that fp-contract=on fuses (fmuladd). */
static void mesh_cross(const float a[3], const float b[3], float out[3]) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

static float mesh_dot(const float a[3], const float b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* normalize((float3)(1.0f, 0.3713907f, 0.2178311f)), voxelise_mesh.cl:42 */
static void mesh_direction(float d[3]) {
    const float v[3] = {1.0f, 0.3713907f, 0.2178311f};
    const float len = sqrtf(mesh_dot(v, v));
    for (int a = 0; a < 3; ++a) d[a] = v[a] / len;
}

/* ray_hit, voxelise_mesh.cl:5-22 (Moller-Trumbore with the kernel's
 * thresholds); tri = a, b, c */
static int mesh_ray_hit(const float o[3], const float d[3], const float *tri) {
    float e1[3], e2[3], p[3], t[3], q[3];
    for (int a = 0; a < 3; ++a) {
        e1[a] = tri[3 + a] - tri[a];
        e2[a] = tri[6 + a] - tri[a];
    }
    mesh_cross(d, e2, p);
    const float det = mesh_dot(e1, p);
    if (fabsf(det) < 1.0e-7f) return 0;
    const float inv = 1.0f / det;
    for (int a = 0; a < 3; ++a) t[a] = o[a] - tri[a];
    const float u = mesh_dot(t, p) * inv;
    if (u < 0.0f || u > 1.0f) return 0;
    mesh_cross(t, e1, q);
    const float v = mesh_dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0;
    return mesh_dot(e2, q) * inv > 1.0e-6f;
}

/* voxelise_mesh, voxelise_mesh.cl:24-48: every interior cell tests its
 * centre's ray against every triangle and is marked when the count is odd */
int clq_voxelise_mesh(clq_ctx *c, const float *triangles, int64_t count, double ox, double oy, double oz,
                      uint8_t *mask) {
    if (!c || count < 0 || (count > 0 && !triangles) || !mask) return fail(CLQ_ERR_INVALID, "bad mesh arguments");
    const float delta = (float)c->dx, o[3] = {(float)ox, (float)oy, (float)oz};
    float d[3];
    mesh_direction(d);
    const int nx = c->nx, ny = c->ny, nz = c->nz;
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int i = 1; i < nx - 1; ++i)
        for (int j = 1; j < ny - 1; ++j)
            for (int k = 1; k < nz - 1; ++k) {
                const float point[3] = {fmaf((float)i + 0.5f, delta, o[0]), fmaf((float)j + 0.5f, delta, o[1]),
                                        fmaf((float)k + 0.5f, delta, o[2])};
                int hits = 0;
                for (int64_t q = 0; q < count; ++q) hits += mesh_ray_hit(point, d, triangles + 9 * q);
                if (hits & 1) mask[((long long)i * ny + j) * nz + k] = 1;
            }
    return CLQ_OK;
}

int clq_voxelise(clq_ctx *c, int32_t s, double ppc, uint8_t *occ, float *density) {
    if (!c || s < 1 || !(ppc > 0)) return fail(CLQ_ERR_INVALID, "bad voxelise arguments");
    if (s > 2) return fail(CLQ_ERR_UNSUPPORTED, "voxelise scale %d > 2", s);
    /* the density is accumulated in binned particle order, as the GPU gather walks bins */
    if (density && !c->bins_current) clq_bin_particles(c);
    const long long sx = (long long)s * c->nx, sy = (long long)s * c->ny, sz = (long long)s * c->nz;
    const long long S = sx * sy * sz;
    const float sinv = (float)s * (float)(1.0 / c->dx);
    if (occ) {
        memset(occ, 0, (size_t)S);
        for (long long i = 0; i < c->n; ++i) {
            const float *x = c->pos + 3 * i;
            const int a = trunc_clamp(x[0] * sinv, s, (int)(s * (c->nx - 1) - 1));
            const int b = trunc_clamp(x[1] * sinv, s, (int)(s * (c->ny - 1) - 1));
            const int d = trunc_clamp(x[2] * sinv, s, (int)(s * (c->nz - 1) - 1));
            occ[(a * sy + b) * sz + d] = 1;
        }
    }
    if (density) {
        double *acc = xcalloc(S, 8);
        if (!acc) return fail(CLQ_ERR_NOMEM, "density accumulator allocation failed");
        const double scale = (double)s * s * s / ppc;
        for (long long i = 0; i < c->n; ++i) {
            const float *x = c->pos + 3 * i;
            int base[3];
            double fr[3];
            const long long dims[3] = {sx, sy, sz};
            for (int q = 0; q < 3; ++q) {
                const double t = (double)x[q] * (double)sinv - 0.5;
                double fl = floor(t);
                if (fl < 0) fl = 0;
                if (fl > dims[q] - 2) fl = (double)(dims[q] - 2);
                base[q] = (int)fl;
                double f = t - fl;
                fr[q] = f < 0 ? 0 : (f > 1 ? 1 : f);
            }
            for (int ox = 0; ox < 2; ++ox)
                for (int oy = 0; oy < 2; ++oy)
                    for (int oz = 0; oz < 2; ++oz) {
                        const double wgt = (ox ? fr[0] : 1 - fr[0]) * (oy ? fr[1] : 1 - fr[1]) * (oz ? fr[2] : 1 - fr[2]);
                        acc[((base[0] + ox) * sy + base[1] + oy) * sz + base[2] + oz] += wgt;
                    }
        }
        for (long long q = 0; q < S; ++q) density[q] = (float)(acc[q] * scale);
        free(acc);
    }
    return CLQ_OK;
}

int clq_kernel_timing(clq_ctx *c, int32_t enable) {
    return c ? CLQ_OK : fail(CLQ_ERR_INVALID, "ctx is NULL");
}

int clq_kernel_times(clq_ctx *c, double *total_ms, int64_t *launches, int32_t capacity) {
    if (!c || !total_ms || !launches) return fail(CLQ_ERR_INVALID, "NULL argument");
    int n = capacity < CLQ_KERNEL_CLASS_COUNT ? capacity : CLQ_KERNEL_CLASS_COUNT;
    for (int i = 0; i < n; ++i) { total_ms[i] = 0.0; launches[i] = 0; }
    return n;
}

int clq_stage_times(clq_ctx *c, double *ms, int32_t capacity) {
    if (!c || !ms) return fail(CLQ_ERR_INVALID, "NULL argument");
    int n = capacity < CLQ_STAGE_COUNT ? capacity : CLQ_STAGE_COUNT;
    for (int i = 0; i < n; ++i) ms[i] = c->stage_ms[i];
    return n;
}
