// C-ABI of include/cliquid.h on gfx950: context, buffers, data movement and
// the stage entry points.  Every stage enqueues on the context's stream; the
// solve loop and the W-cycle recursion run here, natively, not in Python.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "common.hpp"

using namespace clq;

namespace {

thread_local char g_error[512];

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return fail(CLQ_ERR_DEVICE, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                      \
    } while (0)

}  // namespace

int clq::set_error(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof g_error, fmt, ap);
    va_end(ap);
    return code;
}

namespace {

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(CLQ_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
    return CLQ_OK;
}

int non_finite(double r) {
    return fail(CLQ_ERR_NOT_CONVERGED, "pressure solve: residual norm is not finite (%g); the velocity field "
                "or the operator holds NaN or Inf", r);
}

int coarse_size(int n) {  // multigrid.py:12-15
    const int interior_size = std::max(0, n - 2);
    return (interior_size + 1) / 2 + 2;
}

}  // namespace

struct clq_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int nx = 0, ny = 0, nz = 0;
    int64_t C = 0;
    double dx = 0.0;
    int64_t cap = 0, n = 0;
    float *pos = nullptr, *vel = nullptr, *pos_work = nullptr, *vel_work = nullptr, *prev = nullptr;
    int32_t *key = nullptr, *rank = nullptr, *perm = nullptr;
    // bins too crowded for the in-LDS ranking (allocated with the particles)
    int32_t *big_list = nullptr, *big_count = nullptr;
    int big_cap = 0;
    int32_t *pscan = nullptr;  // scan scratch for particle-length scans
    // particle sources on the device (clq_set_sources): per source its type,
    // density, velocity and open cells (mask and not solid), and the union of
    // the OUTFLOW masks
    struct Source {
        int32_t type, density;
        float velocity[3];
        uint8_t *open;
    };
    std::vector<Source> sources;
    uint8_t *sink = nullptr;
    int32_t *src_scalar = nullptr, *src_scalar_host = nullptr;
    int32_t *counts = nullptr, *offsets = nullptr, *scan_scratch = nullptr;
    float *u = nullptr, *v = nullptr, *w = nullptr, *uo = nullptr, *vo = nullptr, *wo = nullptr;
    float *ff = nullptr, *p = nullptr, *rhs = nullptr, *diag = nullptr, *wx = nullptr, *wy = nullptr, *wz = nullptr;
    uint8_t *obstacle = nullptr;
    bool has_obstacle = false;
    bool bins_current = false;  // counts/offsets describe the current particle order
    int nlevels = 0;
    Level lv[kMaxLevels];
    bool coarse_ops_valid = false;  // coarse operators match the current level-0 operator
    double *norm_partials = nullptr, *norm_out = nullptr, *norm_host = nullptr;
    hipEvent_t ev_scal = nullptr;  // behind each FCG scalar read
    // mixed-precision solve state, allocated on first use
    double *p64a = nullptr, *p64b = nullptr;
    float *r32 = nullptr;
    bool p64_valid = false;  // p64a holds the pressure of the last solve (modes 2-4)
    // flexible-CG state, allocated on first use
    // FCG: the residual, the FP32 direction (two arrays: the update reads
    // the old direction at the neighbours while it writes the new one), and
    // the second level-0 correction array (the update writes the next
    // cycle's correction while it reads z)
    double *r64 = nullptr, *fcg_s = nullptr, *fcg_partials = nullptr;
    float *dir[2] = {nullptr, nullptr};
    float *p_alt = nullptr;
    // level-0 rows on which x (p64a) may be non-zero: the row flags of the
    // last FCG initialisation by rows; x_rows_known is false after any other
    // write of p64a (mixed refinement, a dense FCG solve, a field upload)
    int32_t *x_rows = nullptr;
    bool x_rows_known = false;
    // second pressure array of each dense level (>= 2): the fused
    // post-smoothing writes it, and it becomes the level's pressure
    float *p_swap[kMaxLevels] = {};
    // active rows of levels 0 and 1 (rows holding an interior fluid cell),
    // used by the smoother, transfers and FCG vectors inside clq_pressure_solve
    // active rows of levels 0-2 (level 2's only for the restriction onto it)
    int32_t *row_flags[3] = {nullptr, nullptr, nullptr}, *rows[3] = {nullptr, nullptr, nullptr}, *row_pos = nullptr;
    int32_t *row_count = nullptr, *row_count_host = nullptr;
    int64_t nrows[3] = {0, 0, 0};
    bool rows_valid = false, use_rows = false;
    // rows of the resident bottom level B (the second coarsest) holding an
    // interior fluid cell, and their number, on the device (bottom_listed
    // when they follow the current operator); used inside
    // clq_pressure_solve, where B's other rows hold p = +0
    int32_t *bottom_rows = nullptr, *bottom_count = nullptr;
    bool bottom_listed = false;
    uint8_t *op0 = nullptr;  // packed level-0 operator (valid when op0_ok)
    bool op0_ok = false;
    // op0 and the level-0 row flags were written by the operator build from
    // the current level-0 operator (cleared by every write of its inputs)
    bool level0_packed = false;
    // FP64 check solve (clq_fp64_check_solve), allocated on first use:
    // x, r, z, two direction buffers, q; a float residual copy; partials and
    // scalars
    double *chk[6] = {};
    float *chk_r32 = nullptr;
    double *chk_partials = nullptr, *chk_c = nullptr;
    // learned preconditioner (not owned) and its FP32 initial-guess buffer
    clq_net *net = nullptr;
    double stall_ratio = 0.5;
    float *guess32 = nullptr;
    hipEvent_t ev_start[CLQ_STAGE_COUNT] = {}, ev_stop[CLQ_STAGE_COUNT] = {};
    bool ev_used[CLQ_STAGE_COUNT] = {};
    // per-kernel-class event pairs while kernel timing is enabled
    bool ktiming = false;
    std::vector<hipEvent_t> kevents[CLQ_KERNEL_CLASS_COUNT];  // start, stop, start, stop, ...
    std::vector<hipEvent_t> event_pool;
    std::vector<void *> allocations;
    // frame downloads (clq_positions_to_host_async): a device snapshot of the
    // positions, the copy stream, the snapshot-ready event and one completion
    // event per ticket modulo kDownloadEvents
    static constexpr int kDownloadEvents = 16;
    float *snap = nullptr;
    int64_t snap_cap = 0;
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_ready = nullptr, copied[kDownloadEvents] = {};
    int64_t tickets = 0;

    int alloc(void **ptr, size_t bytes) {
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, bytes ? bytes : 16);
        if (e != hipSuccess) return fail(CLQ_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        allocations.push_back(q);
        *ptr = q;
        return CLQ_OK;
    }
    void release(void *q) {
        if (!q) return;
        auto it = std::find(allocations.begin(), allocations.end(), q);
        if (it != allocations.end()) allocations.erase(it);
        (void)hipFree(q);
    }
};

namespace {

// (Re)allocates the particle arrays for `cap` particles; the first `keep`
// positions and velocities are carried over.
int alloc_particles(clq_ctx *c, int64_t cap, int64_t keep = 0) {
    void *arrs[10];
    const size_t fb = (size_t)cap * 12, ib = (size_t)cap * 4;
    const int big_cap = big_bin_capacity(cap);
    const size_t sizes[10] = {fb, fb, fb, fb, fb, ib, ib, ib, (size_t)big_cap * 4,
                              (size_t)scan_scratch_elems(cap) * 4};
    for (int q = 0; q < 10; ++q) {
        if (int rc = c->alloc(&arrs[q], sizes[q])) {
            for (int r = 0; r < q; ++r) c->release(arrs[r]);
            return rc;
        }
    }
    if (keep > 0) {
        if (hipMemcpyAsync(arrs[0], c->pos, (size_t)keep * 12, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
            hipMemcpyAsync(arrs[1], c->vel, (size_t)keep * 12, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            for (void *q : arrs) c->release(q);
            return fail(CLQ_ERR_DEVICE, "particle copy on capacity growth failed");
        }
    }
    c->release(c->pos); c->release(c->vel); c->release(c->pos_work); c->release(c->vel_work); c->release(c->prev);
    c->release(c->key); c->release(c->rank); c->release(c->perm); c->release(c->big_list); c->release(c->pscan);
    c->pos = (float *)arrs[0]; c->vel = (float *)arrs[1]; c->pos_work = (float *)arrs[2];
    c->vel_work = (float *)arrs[3]; c->prev = (float *)arrs[4];
    c->key = (int32_t *)arrs[5]; c->rank = (int32_t *)arrs[6]; c->perm = (int32_t *)arrs[7];
    c->big_list = (int32_t *)arrs[8];
    c->pscan = (int32_t *)arrs[9];
    c->big_cap = big_cap;
    c->cap = cap;
    return CLQ_OK;
}

struct DeviceGuard {
    explicit DeviceGuard(int dev) { (void)hipSetDevice(dev); }
};

// Kernel-class timing: a pair of events around one launch unit.
void kernel_mark(clq_ctx *c, int cls) {
    if (!c->ktiming) return;
    hipEvent_t e = nullptr;
    if (!c->event_pool.empty()) {
        e = c->event_pool.back();
        c->event_pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
        c->ktiming = false;  // out of events: stop timing rather than fail the step
        return;
    }
    (void)hipEventRecord(e, c->stream);
    c->kevents[cls].push_back(e);
}

void stage_begin(clq_ctx *c, int stage) {
    (void)hipEventRecord(c->ev_start[stage], c->stream);
}
void stage_end(clq_ctx *c, int stage) {
    (void)hipEventRecord(c->ev_stop[stage], c->stream);
    c->ev_used[stage] = true;
}

}  // namespace

extern "C" {

const char *clq_last_error(void) { return g_error; }
int clq_abi_version(void) { return CLQ_ABI_VERSION; }

int clq_device_count(int32_t *count) {
    if (!count) return fail(CLQ_ERR_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return fail(CLQ_ERR_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    }
    *count = n;
    return CLQ_OK;
}

int clq_destroy(clq_ctx *c) {
    if (!c) return CLQ_OK;
    DeviceGuard g(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    for (void *q : c->allocations) (void)hipFree(q);
    if (c->snap_ready) (void)hipEventDestroy(c->snap_ready);
    for (hipEvent_t e : c->copied)
        if (e) (void)hipEventDestroy(e);
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    if (c->norm_host) (void)hipHostFree(c->norm_host);
    if (c->src_scalar_host) (void)hipHostFree(c->src_scalar_host);
    if (c->ev_scal) (void)hipEventDestroy(c->ev_scal);
    for (int s = 0; s < CLQ_STAGE_COUNT; ++s) {
        if (c->ev_start[s]) (void)hipEventDestroy(c->ev_start[s]);
        if (c->ev_stop[s]) (void)hipEventDestroy(c->ev_stop[s]);
    }
    for (auto &v : c->kevents)
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return CLQ_OK;
}

int clq_create(int32_t nx, int32_t ny, int32_t nz, double dx, int64_t capacity, clq_ctx **out) {
    if (!out) return fail(CLQ_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (nx < 3 || ny < 3 || nz < 3) return fail(CLQ_ERR_INVALID, "grid %dx%dx%d too small (need >= 3)", nx, ny, nz);
    if ((int64_t)nx * ny * nz >= (1LL << 31)) return fail(CLQ_ERR_INVALID, "grid has >= 2^31 cells");
    if (!(dx > 0.0)) return fail(CLQ_ERR_INVALID, "dx must be positive");
    if (capacity >= (1LL << 31)) return fail(CLQ_ERR_INVALID, "capacity must be < 2^31 particles");
    if (capacity < 1) capacity = 1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return fail(CLQ_ERR_DEVICE, "no HIP device available (the HIP library needs a GPU)");
    clq_ctx *c = new clq_ctx();
    (void)hipGetDevice(&c->device);
    c->nx = nx; c->ny = ny; c->nz = nz; c->dx = dx;
    c->C = (int64_t)nx * ny * nz;
    int rc = CLQ_OK;
    auto bail = [&](int code) { clq_destroy(c); return code; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(CLQ_ERR_DEVICE, "hipStreamCreate failed"));
    for (int s = 0; s < CLQ_STAGE_COUNT; ++s) {
        if (hipEventCreate(&c->ev_start[s]) != hipSuccess || hipEventCreate(&c->ev_stop[s]) != hipSuccess)
            return bail(fail(CLQ_ERR_DEVICE, "hipEventCreate failed"));
    }
    if ((rc = alloc_particles(c, capacity))) return bail(rc);
    const size_t gb = (size_t)c->C * 4;
    void *q;
    float **grids[] = {&c->u, &c->v, &c->w, &c->uo, &c->vo, &c->wo, &c->ff, &c->p, &c->rhs, &c->diag, &c->wx,
                       &c->wy, &c->wz};
    for (float **g : grids) {
        if ((rc = c->alloc(&q, gb))) return bail(rc);
        *g = (float *)q;
        if (hipMemsetAsync(q, 0, gb, c->stream) != hipSuccess) return bail(fail(CLQ_ERR_DEVICE, "memset failed"));
    }
    if ((rc = c->alloc(&q, gb))) return bail(rc);
    c->counts = (int32_t *)q;
    if ((rc = c->alloc(&q, 16))) return bail(rc);
    c->big_count = (int32_t *)q;
    if ((rc = c->alloc(&q, gb))) return bail(rc);
    c->offsets = (int32_t *)q;
    if ((rc = c->alloc(&q, (size_t)scan_scratch_elems(c->C) * 4))) return bail(rc);
    c->scan_scratch = (int32_t *)q;
    if ((rc = c->alloc(&q, (size_t)c->C))) return bail(rc);
    c->obstacle = (uint8_t *)q;
    (void)hipMemsetAsync(c->counts, 0, gb, c->stream);
    (void)hipMemsetAsync(c->offsets, 0, gb, c->stream);
    (void)hipMemsetAsync(c->obstacle, 0, (size_t)c->C, c->stream);
    // level 0 aliases the simulation grid (pressure_solve.py:62-76)
    Level &L0 = c->lv[0];
    L0.nx = nx; L0.ny = ny; L0.nz = nz; L0.C = c->C;
    L0.p = c->p; L0.b = c->rhs; L0.mask = c->ff; L0.diag = c->diag; L0.wx = c->wx; L0.wy = c->wy; L0.wz = c->wz;
    // create_multigrid_levels(min_size=8): multigrid.py:18-50, solver.py:212
    int a = coarse_size(nx), b = coarse_size(ny), d = coarse_size(nz);
    c->nlevels = 1;
    while (a >= 8 && b >= 8 && d >= 8 && c->nlevels < kMaxLevels) {
        Level &L = c->lv[c->nlevels++];
        L.nx = a; L.ny = b; L.nz = d; L.C = (int64_t)a * b * d;
        float **arrs[] = {&L.p, &L.b, &L.mask, &L.diag, &L.wx, &L.wy, &L.wz};
        for (float **arr : arrs) {
            if ((rc = c->alloc(&q, (size_t)L.C * 4))) return bail(rc);
            *arr = (float *)q;
            (void)hipMemsetAsync(q, 0, (size_t)L.C * 4, c->stream);
        }
        a = coarse_size(a); b = coarse_size(b); d = coarse_size(d);
    }
    for (int l = 2; l < c->nlevels; ++l) {
        if ((rc = c->alloc(&q, (size_t)c->lv[l].C * 4))) return bail(rc);
        c->p_swap[l] = (float *)q;
        (void)hipMemsetAsync(q, 0, (size_t)c->lv[l].C * 4, c->stream);
    }
    if ((rc = c->alloc(&q, (size_t)residual_partials_count(c->C) * 8))) return bail(rc);
    c->norm_partials = (double *)q;
    if ((rc = c->alloc(&q, 16))) return bail(rc);
    c->norm_out = (double *)q;
    if (hipHostMalloc((void **)&c->norm_host, 10 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return bail(fail(CLQ_ERR_NOMEM, "hipHostMalloc failed"));
    if (hipEventCreateWithFlags(&c->ev_scal, hipEventDisableTiming) != hipSuccess)
        return bail(fail(CLQ_ERR_DEVICE, "hipEventCreate failed"));
    if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(fail(CLQ_ERR_DEVICE, "initial clear failed"));
    *out = c;
    return CLQ_OK;
}

int clq_sync(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CLQ_OK;
}

// ---- data movement ---------------------------------------------------------------------
int clq_set_particles(clq_ctx *c, const float *pos, const float *vel, int64_t n) {
    if (!c || n < 0 || (n > 0 && !pos)) return fail(CLQ_ERR_INVALID, "bad particle upload");
    if (n >= (1LL << 31)) return fail(CLQ_ERR_INVALID, "particle count must be < 2^31");
    DeviceGuard g(c->device);
    if (n > c->cap) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        int64_t cap = std::max<int64_t>(n, c->cap * 3 / 2);  // solver.py:290-291
        if (int rc = alloc_particles(c, cap)) return rc;
    }
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(c->pos, pos, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
        if (vel) HIP_TRY(hipMemcpyAsync(c->vel, vel, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
        else HIP_TRY(hipMemsetAsync(c->vel, 0, (size_t)n * 12, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->n = n;
    c->bins_current = false;
    return CLQ_OK;
}

int clq_get_particles(clq_ctx *c, float *pos, float *vel, int64_t n) {
    if (!c || n != c->n)
        return fail(CLQ_ERR_INVALID, "particle count mismatch (%lld requested, %lld held)", (long long)n,
                    c ? (long long)c->n : -1LL);
    DeviceGuard g(c->device);
    if (n > 0 && pos) HIP_TRY(hipMemcpyAsync(pos, c->pos, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream));
    if (n > 0 && vel) HIP_TRY(hipMemcpyAsync(vel, c->vel, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CLQ_OK;
}

int clq_get_previous_positions(clq_ctx *c, float *pos, int64_t n) {
    if (!c || n != c->n || !pos) return fail(CLQ_ERR_INVALID, "bad previous-position download");
    DeviceGuard g(c->device);
    if (n > 0) HIP_TRY(hipMemcpyAsync(pos, c->prev, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CLQ_OK;
}

int clq_get_particle_count(clq_ctx *c, int64_t *n) {
    if (!c || !n) return fail(CLQ_ERR_INVALID, "NULL argument");
    *n = c->n;
    return CLQ_OK;
}

// ---- frame output ------------------------------------------------------------------------
int clq_host_alloc(clq_ctx *c, int64_t bytes, void **host) {
    if (!c || !host || bytes < 0) return fail(CLQ_ERR_INVALID, "bad host allocation");
    DeviceGuard g(c->device);
    *host = nullptr;
    if (hipHostMalloc(host, (size_t)std::max<int64_t>(bytes, 16), hipHostMallocDefault) != hipSuccess)
        return fail(CLQ_ERR_NOMEM, "hipHostMalloc(%lld) failed", (long long)bytes);
    return CLQ_OK;
}

int clq_host_free(clq_ctx *c, void *host) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    if (!host) return CLQ_OK;
    DeviceGuard g(c->device);
    // a copy into this memory may still be in flight
    HIP_TRY(c->copy_stream ? hipStreamSynchronize(c->copy_stream) : hipSuccess);
    HIP_TRY(hipHostFree(host));
    return CLQ_OK;
}

int clq_positions_to_host_async(clq_ctx *c, float *host, int64_t capacity, int64_t *n, int64_t *ticket) {
    if (!c || !n || !ticket) return fail(CLQ_ERR_INVALID, "NULL argument");
    *n = c->n;
    *ticket = -1;
    if (c->n == 0) return CLQ_OK;
    if (!host || capacity < c->n)
        return fail(CLQ_ERR_INVALID, "host buffer holds %lld particles, %lld needed", (long long)capacity,
                    (long long)c->n);
    DeviceGuard g(c->device);
    if (!c->copy_stream) {
        HIP_TRY(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreateWithFlags(&c->snap_ready, hipEventDisableTiming));
        for (hipEvent_t &e : c->copied) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    if (c->n > c->snap_cap) {
        HIP_TRY(hipStreamSynchronize(c->copy_stream));  // the old snapshot may still be read
        c->release(c->snap);
        c->snap = nullptr;
        c->snap_cap = 0;
        void *q = nullptr;
        if (int rc = c->alloc(&q, (size_t)c->cap * 12)) return rc;
        c->snap = (float *)q;
        c->snap_cap = c->cap;
    }
    // the snapshot is overwritten only after the previous copy has read it
    if (c->tickets > 0)
        HIP_TRY(hipStreamWaitEvent(c->stream, c->copied[(c->tickets - 1) % clq_ctx::kDownloadEvents], 0));
    const size_t bytes = (size_t)c->n * 12;
    HIP_TRY(hipMemcpyAsync(c->snap, c->pos, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIP_TRY(hipEventRecord(c->snap_ready, c->stream));
    HIP_TRY(hipStreamWaitEvent(c->copy_stream, c->snap_ready, 0));
    HIP_TRY(hipMemcpyAsync(host, c->snap, bytes, hipMemcpyDeviceToHost, c->copy_stream));
    HIP_TRY(hipEventRecord(c->copied[c->tickets % clq_ctx::kDownloadEvents], c->copy_stream));
    *ticket = c->tickets++;
    return CLQ_OK;
}

int clq_host_copy_wait(clq_ctx *c, int64_t ticket) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    if (ticket < 0) return CLQ_OK;
    if (ticket >= c->tickets || ticket < c->tickets - clq_ctx::kDownloadEvents)
        return fail(CLQ_ERR_INVALID, "unknown download ticket %lld", (long long)ticket);
    DeviceGuard g(c->device);
    // copies complete in order on the copy stream: the event of a later ticket
    // recorded into the same slot completes after this one
    HIP_TRY(hipEventSynchronize(c->copied[ticket % clq_ctx::kDownloadEvents]));
    return CLQ_OK;
}

static void *field_ptr(clq_ctx *c, int f, int64_t *bytes) {
    int64_t fb = c->C * 4;
    void *p = nullptr;
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
    default: return nullptr;
    }
    *bytes = fb;
    return p;
}

int clq_set_field(clq_ctx *c, int32_t f, const void *host, int64_t bytes) {
    int64_t fb = 0;
    void *p = c ? field_ptr(c, f, &fb) : nullptr;
    if (!p || !host || bytes != fb)
        return fail(CLQ_ERR_INVALID, "bad field upload (field %d, %lld bytes)", f, (long long)bytes);
    DeviceGuard g(c->device);
    HIP_TRY(hipMemcpyAsync(p, host, (size_t)bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (f == CLQ_FIELD_PRESSURE) c->p64_valid = false;
    if (f == CLQ_FIELD_PRESSURE_FP64) c->x_rows_known = false;
    if (f == CLQ_FIELD_OBSTACLE) {
        const uint8_t *o = (const uint8_t *)host;
        c->has_obstacle = std::any_of(o, o + bytes, [](uint8_t x) { return x != 0; });
    }
    if (f == CLQ_FIELD_FLUID_FRACTION || f == CLQ_FIELD_DIAGONAL || f == CLQ_FIELD_WEIGHT_X ||
        f == CLQ_FIELD_WEIGHT_Y || f == CLQ_FIELD_WEIGHT_Z)
        c->coarse_ops_valid = c->level0_packed = false;
    return CLQ_OK;
}

int clq_get_field(clq_ctx *c, int32_t f, void *host, int64_t bytes) {
    int64_t fb = 0;
    void *p = c ? field_ptr(c, f, &fb) : nullptr;
    if (!p || !host || bytes != fb)
        return fail(CLQ_ERR_INVALID, "bad field download (field %d, %lld bytes)", f, (long long)bytes);
    DeviceGuard g(c->device);
    HIP_TRY(hipMemcpyAsync(host, p, (size_t)bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CLQ_OK;
}

int clq_level_count(clq_ctx *c, int32_t *levels) {
    if (!c || !levels) return fail(CLQ_ERR_INVALID, "NULL argument");
    *levels = c->nlevels;
    return CLQ_OK;
}

int clq_level_shape(clq_ctx *c, int32_t l, int32_t *nx, int32_t *ny, int32_t *nz) {
    if (!c || l < 0 || l >= c->nlevels || !nx || !ny || !nz) return fail(CLQ_ERR_INVALID, "bad level %d", l);
    *nx = c->lv[l].nx; *ny = c->lv[l].ny; *nz = c->lv[l].nz;
    return CLQ_OK;
}

static float *level_array(Level &L, int f) {
    switch (f) {
    case CLQ_LEVEL_P: return L.p;
    case CLQ_LEVEL_B: return L.b;
    case CLQ_LEVEL_MASK: return L.mask;
    case CLQ_LEVEL_DIAGONAL: return L.diag;
    case CLQ_LEVEL_WEIGHT_X: return L.wx;
    case CLQ_LEVEL_WEIGHT_Y: return L.wy;
    case CLQ_LEVEL_WEIGHT_Z: return L.wz;
    default: return nullptr;
    }
}

int clq_set_level_field(clq_ctx *c, int32_t l, int32_t f, const float *host, int64_t bytes) {
    if (!c || l < 0 || l >= c->nlevels || !host) return fail(CLQ_ERR_INVALID, "bad level upload");
    float *a = level_array(c->lv[l], f);
    if (!a || bytes != c->lv[l].C * 4) return fail(CLQ_ERR_INVALID, "bad level field %d / size", f);
    DeviceGuard g(c->device);
    HIP_TRY(hipMemcpyAsync(a, host, (size_t)bytes, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (f >= CLQ_LEVEL_MASK) c->coarse_ops_valid = c->level0_packed = false;
    return CLQ_OK;
}

int clq_get_level_field(clq_ctx *c, int32_t l, int32_t f, float *host, int64_t bytes) {
    if (!c || l < 0 || l >= c->nlevels || !host) return fail(CLQ_ERR_INVALID, "bad level download");
    float *a = level_array(c->lv[l], f);
    if (!a || bytes != c->lv[l].C * 4) return fail(CLQ_ERR_INVALID, "bad level field %d / size", f);
    DeviceGuard g(c->device);
    HIP_TRY(hipMemcpyAsync(host, a, (size_t)bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return CLQ_OK;
}

// ---- particle_to_grid ---------------------------------------------------------------
int clq_bin_particles(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    DeviceGuard g(c->device);
    const float inv_dx = (float)(1.0 / c->dx);  // FIELD_DTYPE(1.0 / dx), particles.py:103
    const int n = (int)c->n;
    stage_begin(c, CLQ_STAGE_BIN);
    kernel_mark(c, CLQ_KERNEL_BIN);
    HIP_TRY(hipMemsetAsync(c->counts, 0, (size_t)c->C * 4, c->stream));
    launch_bin_count(c->pos, n, inv_dx, c->nx, c->ny, c->nz, c->key, c->rank, c->counts, c->stream);
    launch_exclusive_scan(c->counts, c->offsets, c->C, c->scan_scratch, c->stream);
    launch_bin_permute(c->key, c->rank, c->offsets, n, c->perm, c->big_count, c->stream);
    // key is free after the permutation: the big bins' merge sort uses it
    launch_bin_order(c->perm, c->offsets, c->counts, c->C, n, c->pos, c->vel, c->pos_work, c->vel_work, c->big_list,
                     c->big_count, c->big_cap, c->key, c->stream);
    kernel_mark(c, CLQ_KERNEL_BIN);
    stage_end(c, CLQ_STAGE_BIN);
    std::swap(c->pos, c->pos_work);  // particles.py:136-140
    std::swap(c->vel, c->vel_work);
    c->bins_current = true;
    return check_launch("bin_particles");
}

int clq_particle_to_grid(clq_ctx *c, double particles_per_cell) {
    if (!c || !(particles_per_cell > 0)) return fail(CLQ_ERR_INVALID, "bad P2G arguments");
    DeviceGuard g(c->device);
    stage_begin(c, CLQ_STAGE_P2G);
    kernel_mark(c, CLQ_KERNEL_P2G);
    launch_p2g(c->pos, c->vel, c->offsets, c->counts, c->u, c->v, c->w, c->ff, (float)(1.0 / c->dx),
               (float)(1.0 / particles_per_cell), c->nx, c->ny, c->nz, c->stream);
    kernel_mark(c, CLQ_KERNEL_P2G);
    stage_end(c, CLQ_STAGE_P2G);
    c->coarse_ops_valid = c->level0_packed = false;
    return check_launch("particle_to_grid");
}

// ---- grid glue ------------------------------------------------------------------------
int clq_grid_obstacle_bc(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    if (!c->has_obstacle) return CLQ_OK;  // every write of the two kernels needs a solid voxel
    DeviceGuard g(c->device);
    launch_grid_obstacle_bc(c->u, c->v, c->w, c->ff, c->obstacle, c->nx, c->ny, c->nz, c->stream);
    c->coarse_ops_valid = c->level0_packed = false;
    return check_launch("grid_obstacle_bc");
}

int clq_grid_domain_bc(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    DeviceGuard g(c->device);
    launch_grid_domain_bc(c->u, c->v, c->w, c->nx, c->ny, c->nz, c->stream);
    return check_launch("grid_domain_bc");
}

int clq_update_velocity(clq_ctx *c, const double *gravity, const double *force, double rho, double nu, double dt) {
    if (!c || !gravity || !force) return fail(CLQ_ERR_INVALID, "bad update_velocity arguments");
    DeviceGuard g(c->device);
    const float frho = (float)rho, fdt = (float)dt, fnu = (float)nu;
    const float ax = (float)gravity[0] + (float)force[0] / frho;
    const float ay = (float)gravity[1] + (float)force[1] / frho;
    const float az = (float)gravity[2] + (float)force[2] / frho;
    const float inv_dx2 = (float)(1.0 / (c->dx * c->dx));  // velocity_update.py:46
    launch_update_velocity(c->u, c->v, c->w, c->uo, c->vo, c->wo, c->ff, ax, ay, az, fdt, fnu * fdt, inv_dx2, c->nx,
                           c->ny, c->nz, c->stream);
    std::swap(c->u, c->uo);  // solver.py:374-376
    std::swap(c->v, c->vo);
    std::swap(c->w, c->wo);
    return check_launch("update_velocity");
}

// ---- pressure ----------------------------------------------------------------------------
int clq_pressure_rhs(clq_ctx *c, double rho, double dt) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    DeviceGuard g(c->device);
    launch_pressure_rhs(c->u, c->v, c->w, c->ff, c->rhs, (float)(rho / dt), (float)(1.0 / c->dx), c->nx, c->ny, c->nz,
                        c->stream);
    return check_launch("pressure_rhs");
}

static int ensure_row_buffers(clq_ctx *c);

int clq_build_operator(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    DeviceGuard g(c->device);
    // with the row lists in use (two levels or more), the build also packs
    // the level-0 operator and flags the active rows
    const bool pack = c->nlevels >= 2;
    if (pack)
        if (int rc = ensure_row_buffers(c)) return rc;
    launch_build_finest_operator(c->ff, c->diag, c->wx, c->wy, c->wz, pack ? c->op0 : nullptr,
                                 pack ? c->row_flags[0] : nullptr, c->nx, c->ny, c->nz, c->stream);
    c->coarse_ops_valid = false;
    c->rows_valid = c->bottom_listed = false;  // the row flags now follow the new operator
    c->level0_packed = pack;
    return check_launch("build_operator");
}

static void ensure_coarse_operators(clq_ctx *c) {
    if (c->coarse_ops_valid) return;
    for (int l = 0; l + 1 < c->nlevels; ++l) launch_coarse_operator(c->lv[l], c->lv[l + 1], c->stream);
    c->coarse_ops_valid = true;
    c->rows_valid = c->bottom_listed = false;
}

// Active-row lists of levels 0 and 1 for the current operator (one host
// read-back of the two counts per operator build).
// levels with active-row lists: 0 and 1 are swept by rows, level 2's list
// serves the restriction onto it (its sweeps are dense)
static int row_levels(const clq_ctx *c) { return c->nlevels < 3 ? c->nlevels : 3; }

static int ensure_row_buffers(clq_ctx *c) {
    if (!c->row_pos) {
        const int64_t r0 = (int64_t)c->nx * c->ny;
        void *q;
        for (int l = 0; l < row_levels(c); ++l) {
            const int64_t r = (int64_t)c->lv[l].nx * c->lv[l].ny;
            if (int rc = c->alloc(&q, (size_t)r * 4)) return rc;
            c->row_flags[l] = (int32_t *)q;
            if (int rc = c->alloc(&q, (size_t)r * 4)) return rc;
            c->rows[l] = (int32_t *)q;
        }
        if (int rc = c->alloc(&q, (size_t)r0 * 4)) return rc;
        c->row_pos = (int32_t *)q;
        if (int rc = c->alloc(&q, 16)) return rc;
        c->row_count = (int32_t *)q;
        HIP_TRY(hipHostMalloc((void **)&c->row_count_host, 16, hipHostMallocDefault));
        if (int rc = c->alloc(&q, (size_t)c->C)) return rc;
        c->op0 = (uint8_t *)q;
        const int b = c->nlevels - 2;
        if (b >= 1 && resident_bottom_fits(c->lv[b], c->lv[b + 1])) {
            if (int rc = c->alloc(&q, (size_t)c->lv[b].nx * c->lv[b].ny * 4)) return rc;
            c->bottom_rows = (int32_t *)q;
            if (int rc = c->alloc(&q, 16)) return rc;
            c->bottom_count = (int32_t *)q;
        }
    }
    return CLQ_OK;
}

static int ensure_active_rows(clq_ctx *c) {
    if (c->rows_valid) return CLQ_OK;
    if (c->nlevels < 2) return CLQ_OK;
    if (int rc = ensure_row_buffers(c)) return rc;
    const bool pack = !c->level0_packed;  // else the operator build's bytes are exact
    if (pack) {
        HIP_TRY(hipMemsetAsync(c->row_count + 3, 0, 4, c->stream));
        launch_pack_operator(c->lv[0], c->op0, c->row_count + 3, c->stream);
        launch_row_flags(c->lv[0], c->row_flags[0], c->stream);
    }
    for (int l = 1; l < row_levels(c); ++l)
        launch_coarse_row_flags(c->row_flags[l - 1], c->lv[l - 1], c->lv[l], c->row_flags[l], c->stream);
    for (int l = 0; l < row_levels(c); ++l) {
        const int64_t r = (int64_t)c->lv[l].nx * c->lv[l].ny;
        launch_exclusive_scan(c->row_flags[l], c->row_pos, r, c->scan_scratch, c->stream);
        launch_compact_rows(c->row_flags[l], c->row_pos, r, c->rows[l], c->row_count + l, c->stream);
    }
    if (c->bottom_rows) launch_bottom_rows(c->lv[c->nlevels - 2], c->bottom_rows, c->bottom_count, c->stream);
    c->bottom_listed = c->bottom_rows != nullptr;
    HIP_TRY(hipMemcpyAsync(c->row_count_host, c->row_count, 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int l = 0; l < row_levels(c); ++l) c->nrows[l] = c->row_count_host[l];
    c->op0_ok = !pack || c->row_count_host[3] == 0;
    c->rows_valid = true;
    return check_launch("active rows");
}

// Inside a pressure solve: the row lists, and the writes the dense kernels
// would make on the inactive rows.
static int begin_sparse_solve(clq_ctx *c) {
    if (int rc = ensure_active_rows(c)) return rc;
    if (c->nlevels < 2) return CLQ_OK;
    for (int l = 0; l < row_levels(c); ++l) launch_zero_inactive_rows(c->lv[l], c->row_flags[l], l > 0, c->stream);
    c->use_rows = true;
    return CLQ_OK;
}

// done: half-sweeps of the first sweep already made: 1 when the restriction
// onto this level made the red one (restrict_level with red) or the FCG
// update did on level 0, 2 when the FCG loop has also run the black one.
static void smooth_level(clq_ctx *c, int level, int iterations, int done = 0) {
    const float delta = (float)c->dx;
    const bool sparse = c->use_rows && level < 2;
    for (int it = 0; it < iterations; ++it) {
        const int skip = it == 0 ? done : 0;
        const bool whole = skip == 0;
        if (level == 0 && whole) kernel_mark(c, CLQ_KERNEL_RBGS_FINE);
        for (int parity = skip; parity < 2; ++parity) {
            if (sparse && level == 0 && c->op0_ok)
                launch_rbgs_rows_packed(c->lv[0], c->op0, delta, parity, c->rows[0], c->nrows[0], c->stream);
            else if (sparse)
                launch_rbgs_rows(c->lv[level], delta, parity, c->rows[level], c->nrows[level], c->stream);
            else
                launch_rbgs(c->lv[level], delta, parity, c->stream);
        }
        if (level == 0 && whole) kernel_mark(c, CLQ_KERNEL_RBGS_FINE);
    }
}

// Inside a solve (row lists set up) the transfers are folded into the
// neighbouring half-sweeps: the restriction also makes the coarse level's
// first red half-sweep, and the prolongation rides on the first
// post-smoothing sweep.  Same values as the separate kernels.
static bool fold_transfers(const clq_ctx *c) { return c->use_rows; }

// the first post-smoothing sweep of `level` with the prolongation from level + 1
static void smooth_prolong(clq_ctx *c, int level) {
    const bool sparse = level < 2;
    if (!sparse && c->p_swap[level]) {  // both half-sweeps in one launch
        launch_rbgs_post_fused(c->lv[level], c->lv[level + 1], (float)c->dx, c->p_swap[level], c->stream);
        std::swap(c->lv[level].p, c->p_swap[level]);
        return;
    }
    const uint8_t *op = level == 0 && c->op0_ok ? c->op0 : nullptr;
    for (int parity = 0; parity < 2; ++parity)
        launch_rbgs_prolong(c->lv[level], c->lv[level + 1], op, (float)c->dx, parity, parity == 0,
                            sparse ? c->rows[level] : nullptr, sparse ? c->nrows[level] : 0, c->stream);
}

static void restrict_level(clq_ctx *c, int level, bool red = false) {
    const float delta = (float)c->dx;
    if (c->use_rows && level + 1 < row_levels(c))  // onto the coarse level's active rows
        launch_restrict_rows(c->lv[level], level == 0 && c->op0_ok ? c->op0 : nullptr, c->lv[level + 1],
                             1.0f / (delta * delta), c->rows[level + 1], c->nrows[level + 1], delta, red, c->stream);
    else
        launch_restrict_residual(c->lv[level], c->lv[level + 1], 1.0f / (delta * delta), delta, red, c->stream);
}

static void prolongate_level(clq_ctx *c, int level) {
    if (c->use_rows && level == 0)
        launch_prolongate_rows(c->lv[0], c->lv[1], c->rows[0], c->nrows[0], c->stream);
    else
        launch_prolongate(c->lv[level], c->lv[level + 1], c->stream);
}

// Levels with at most this many cells run inside one work-group launch
// (wcycle_bottom_kernel) for the rest of the recursion below them.
constexpr int64_t kBottomCells = 8192;

// Whether a visit of `level` runs its coarse-grid correction twice (the W
// shape) or once (V): vcycle_from 0 at every level, k > 0 on the levels below
// k, -m < 0 on the levels whose bit is set in m.
static bool twice_at(int vfrom, int level) {
    if (vfrom < 0) return level < 31 && ((((uint32_t)(-(int64_t)vfrom)) >> level) & 1u) != 0;
    return vfrom == 0 || level < vfrom;
}

static void bottom_ops(const clq_ctx *c, int level, int pre, int post, int extra, int vfrom, bool f64,
                       BottomArgs &a) {
    if (level == c->nlevels - 1) {
        a.ops[a.nops++] = BottomOp{f64 ? kOpCoarse64 : kOpCoarse32, level, pre + extra};
        return;
    }
    a.ops[a.nops++] = BottomOp{kOpSmooth, level, pre};
    a.ops[a.nops++] = BottomOp{kOpRestrict, level, 0};
    bottom_ops(c, level + 1, pre, post, extra, vfrom, f64, a);
    if (twice_at(vfrom, level)) bottom_ops(c, level + 1, pre, post, extra, vfrom, f64, a);
    a.ops[a.nops++] = BottomOp{kOpProlong, level, 0};
    a.ops[a.nops++] = BottomOp{kOpSmooth, level, post};
}

static int bottom_ops_count(int depth) { return depth == 0 ? 1 : 4 + 2 * bottom_ops_count(depth - 1); }

// First level >= 1 whose remaining recursion runs in one work-group.
static int bottom_level(const clq_ctx *c) {
    for (int l = 1; l < c->nlevels; ++l)
        if (c->lv[l].C <= kBottomCells && bottom_ops_count(c->nlevels - 1 - l) <= kMaxBottomOps) return l;
    return c->nlevels;  // none
}

static void launch_bottom(clq_ctx *c, int level, int pre, int post, int extra, int vfrom, bool f64) {
    if (level == c->nlevels - 2 && resident_bottom_fits(c->lv[level], c->lv[level + 1])) {
        const int calls = twice_at(vfrom, level) ? 2 : 1;
        const bool listed = c->use_rows && c->bottom_listed;
        launch_bottom_resident(c->lv[level], c->lv[level + 1], (float)c->dx, f64, pre, post, pre + extra, calls,
                               listed ? c->bottom_rows : nullptr, listed ? c->bottom_count : nullptr, c->stream);
        return;
    }
    BottomArgs a;
    for (int l = 0; l < c->nlevels; ++l) a.lv[l] = c->lv[l];
    a.nops = 0;
    bottom_ops(c, level, pre, post, extra, vfrom, f64, a);
    const float delta = (float)c->dx;
    a.dd = delta * delta;
    a.inv_delta2 = 1.0f / (delta * delta);
    a.dd64 = (double)delta * (double)delta;
    launch_wcycle_bottom(a, c->stream);
}

// multigrid.py:53-187
// Whether a visit of `level` smooths it with the colour kernels (and not in
// the bottom kernel or the coarsest smoother).
static bool swept_level(const clq_ctx *c, int level) {
    return level < c->nlevels - 1 && !(level > 0 && level == bottom_level(c));
}

// The steps of one level visit shared by both cycles: pre-smoothing, the
// restriction, then (after the caller's coarse visits) the prolongation and
// post-smoothing.  Returns whether the coarse level's first red half-sweep
// was made by the restriction.
static bool visit_down(clq_ctx *c, int level, int pre, int done) {
    smooth_level(c, level, pre, done);
    const bool red = fold_transfers(c) && pre > 0 && swept_level(c, level + 1);
    restrict_level(c, level, red);
    return red;
}

static void visit_up(clq_ctx *c, int level, int post) {
    if (fold_transfers(c) && post > 0) {
        smooth_prolong(c, level);
        smooth_level(c, level, post - 1);
    } else {
        prolongate_level(c, level);
        smooth_level(c, level, post);
    }
}

static void w_cycle(clq_ctx *c, int level, int pre, int post, int extra, int vfrom, int done = 0) {
    if (level > 0 && level == bottom_level(c)) {
        launch_bottom(c, level, pre, post, extra, vfrom, false);
        return;
    }
    if (level == c->nlevels - 1) {
        launch_coarse_smooth(c->lv[level], (float)c->dx, pre + extra, c->stream);
        return;
    }
    const bool red = visit_down(c, level, pre, done);
    w_cycle(c, level + 1, pre, post, extra, vfrom, red);
    if (twice_at(vfrom, level)) w_cycle(c, level + 1, pre, post, extra, vfrom);
    visit_up(c, level, post);
}

int clq_smooth(clq_ctx *c, int32_t level, int32_t iterations) {
    if (!c || level < 0 || level >= c->nlevels || iterations < 0) return fail(CLQ_ERR_INVALID, "bad smooth arguments");
    DeviceGuard g(c->device);
    smooth_level(c, level, iterations);
    return check_launch("smooth");
}

int clq_restrict(clq_ctx *c, int32_t level) {
    if (!c || level < 0 || level + 1 >= c->nlevels) return fail(CLQ_ERR_INVALID, "bad restrict level");
    DeviceGuard g(c->device);
    // The standalone call restates restrict_residual whole: operator and b.
    launch_coarse_operator(c->lv[level], c->lv[level + 1], c->stream);
    restrict_level(c, level);
    if (level + 2 < c->nlevels) c->coarse_ops_valid = false;
    return check_launch("restrict");
}

int clq_prolongate(clq_ctx *c, int32_t level) {
    if (!c || level < 0 || level + 1 >= c->nlevels) return fail(CLQ_ERR_INVALID, "bad prolongate level");
    DeviceGuard g(c->device);
    launch_prolongate(c->lv[level], c->lv[level + 1], c->stream);
    return check_launch("prolongate");
}

int clq_coarse_smooth(clq_ctx *c, int32_t iterations) {
    if (!c || iterations < 0) return fail(CLQ_ERR_INVALID, "bad coarse_smooth arguments");
    DeviceGuard g(c->device);
    launch_coarse_smooth(c->lv[c->nlevels - 1], (float)c->dx, iterations, c->stream);
    return check_launch("coarse_smooth");
}

int clq_w_cycle(clq_ctx *c, int32_t pre, int32_t post, int32_t extra) {
    if (!c || pre < 0 || post < 0 || extra < 0) return fail(CLQ_ERR_INVALID, "bad w_cycle arguments");
    DeviceGuard g(c->device);
    ensure_coarse_operators(c);
    w_cycle(c, 0, pre, post, extra, 0);
    return check_launch("w_cycle");
}

int clq_neumann_bc(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    DeviceGuard g(c->device);
    launch_neumann_bc(c->p, c->nx, c->ny, c->nz, c->stream);
    return check_launch("neumann_bc");
}

static int residual_norm(clq_ctx *c, double *r, double *bn) {
    const double h = (double)(float)c->dx;
    launch_residual_norm(c->lv[0], 1.0 / (h * h), c->norm_partials, c->norm_out, c->stream);
    HIP_TRY(hipMemcpyAsync(c->norm_host, c->norm_out, 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *r = sqrt(c->norm_host[0]);
    *bn = sqrt(c->norm_host[1]);
    return CLQ_OK;
}

int clq_residual_norm(clq_ctx *c, double *residual, double *rhs_norm) {
    if (!c || !residual || !rhs_norm) return fail(CLQ_ERR_INVALID, "NULL argument");
    DeviceGuard g(c->device);
    return residual_norm(c, residual, rhs_norm);
}


// ---- mixed-precision iterative refinement -------------------------------------------
// f64: the coarsest level's sweeps in FP64 (the refinement of mode 2);
// otherwise the reference's FP32 coarse smoother
static void w_cycle_mixed(clq_ctx *c, int level, int pre, int post, int extra, int vfrom, bool f64,
                          int done = 0) {
    if (level > 0 && level == bottom_level(c)) {
        launch_bottom(c, level, pre, post, extra, vfrom, f64);
        return;
    }
    if (level == c->nlevels - 1) {
        const Level &L = c->lv[level];
        if (f64 && L.C <= coarse_f64_capacity())
            launch_coarse_smooth_f64(L, (double)(float)c->dx, pre + extra, c->stream);
        else launch_coarse_smooth(L, (float)c->dx, pre + extra, c->stream);
        return;
    }
    const bool red = visit_down(c, level, pre, done);
    w_cycle_mixed(c, level + 1, pre, post, extra, vfrom, f64, red);
    if (twice_at(vfrom, level)) w_cycle_mixed(c, level + 1, pre, post, extra, vfrom, f64);
    visit_up(c, level, post);
}

static int ir_residual(clq_ctx *c, int from32, const double *pin, double *pout, double *r, double *bn) {
    const double h = (double)(float)c->dx;
    launch_ir_update_residual(c->lv[0], from32, c->p, pin, pout, c->r32, c->rhs, 1.0 / (h * h), c->norm_partials,
                              c->norm_out, c->stream);
    HIP_TRY(hipMemcpyAsync(c->norm_host, c->norm_out, 16, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *r = sqrt(c->norm_host[0]);
    *bn = sqrt(c->norm_host[1]);
    return CLQ_OK;
}

// FP64 iterate p64 and residual, W-cycle on A e = r with FP32 smoothing and
// 16-bit coarse b (power-of-two scale 2^(3l + es) per level from the RMS of
// the fine residual), FP64 coarsest smoother, p64 += e.
static int ensure_mixed_buffers(clq_ctx *c);

static int solve_mixed(clq_ctx *c, const clq_solve_params *sp, clq_solve_stats *s) {
    if (int rc = ensure_mixed_buffers(c)) return rc;
    c->x_rows_known = false;
    double r = 0.0, bn = 0.0;
    if (int rc = ir_residual(c, 1, nullptr, c->p64a, &r, &bn)) return rc;
    if (!std::isfinite(r) || !std::isfinite(bn)) return non_finite(r + bn);
    s->initial_residual = r;
    s->rhs_norm = bn;
    const int kind = sp->coarse_dtype == 1 ? 2 : 1;
    Level &L0 = c->lv[0];
    int rc = CLQ_OK;
    for (int cyc = 0; cyc < sp->max_cycles && r > sp->tolerance * bn; ++cyc) {
        const int es = (int)lround(log2(r / sqrt((double)c->C) + 1e-300));
        for (int l = 1; l < c->nlevels; ++l) {
            c->lv[l].bkind = kind;
            c->lv[l].bscale = ldexpf(1.0f, 3 * l + es);
        }
        L0.b = c->r32;
        HIP_TRY(hipMemsetAsync(c->p, 0, (size_t)c->C * 4, c->stream));
        w_cycle_mixed(c, 0, sp->pre_smooth, sp->post_smooth, sp->coarse_extra, sp->vcycle_from, true);
        L0.b = c->rhs;
        const double before = r;
        if ((rc = ir_residual(c, 0, c->p64a, c->p64b, &r, &bn))) break;
        std::swap(c->p64a, c->p64b);
        s->cycles++;
        if (!std::isfinite(r)) {
            rc = non_finite(r);
            break;
        }
        if (r > 0.95 * before) break;
    }
    L0.b = c->rhs;
    for (int l = 1; l < c->nlevels; ++l) c->lv[l].bkind = 0;
    if (rc) return rc;
    launch_widen_narrow(c->p64a, c->p, c->C, c->stream);
    s->final_residual = r;
    s->relative_residual = bn > 0 ? r / bn : 0.0;
    return check_launch("solve_mixed");
}

// ---- flexible CG ---------------------------------------------------------------------------
static int alloc_zeroed(clq_ctx *c, void **q, size_t bytes) {
    if (int rc = c->alloc(q, bytes)) return rc;
    HIP_TRY(hipMemsetAsync(*q, 0, bytes, c->stream));
    return CLQ_OK;
}

static int ensure_mixed_buffers(clq_ctx *c) {
    if (c->p64a) return CLQ_OK;
    void *q;
    if (int rc = alloc_zeroed(c, &q, (size_t)c->C * 8)) return rc;
    c->p64a = (double *)q;
    if (int rc = alloc_zeroed(c, &q, (size_t)c->C * 8)) return rc;
    c->p64b = (double *)q;
    if (int rc = alloc_zeroed(c, &q, (size_t)c->C * 4)) return rc;
    c->r32 = (float *)q;
    for (int l = 1; l < c->nlevels; ++l) {
        if (int rc = alloc_zeroed(c, &q, (size_t)c->lv[l].C * 2)) return rc;
        c->lv[l].b16 = (uint16_t *)q;
    }
    return CLQ_OK;
}

// z = M r: the mixed-precision W-cycle on A e = r32 from e = 0 (e is the
// level-0 pressure array), FP32 smoothing, 16-bit coarse b scaled from the
// residual norm and FP64 coarsest sweeps (the coarsest-grid solve of the
// mixed-precision pattern; FP32 ones give the same iteration counts at
// 256^3 and save about 1 % of the step).
// done: the FCG loop already left e after the cycle's first red half-sweep
// on level 0 (1: fcg_update with e) or after its whole first sweep (2: and
// the black half-sweep queued during the scalar read), in place of zeroing it.
static int precondition_wcycle(clq_ctx *c, const clq_solve_params *sp, double rnorm, int done = 0) {
    const int kind = sp->coarse_dtype == 1 ? 2 : 1;
    const int es = (int)lround(log2(rnorm / sqrt((double)c->C) + 1e-300));
    for (int l = 1; l < c->nlevels; ++l) {
        c->lv[l].bkind = kind;
        c->lv[l].bscale = ldexpf(1.0f, 3 * l + es);
    }
    Level &L0 = c->lv[0];
    L0.b = c->r32;
    // e = 0 where the cycle reads it with a non-zero weight: the active rows
    if (!done) {
        if (c->use_rows) launch_zero_rows(L0.p, c->rows[0], c->nrows[0], c->nz, c->stream);
        else HIP_TRY(hipMemsetAsync(L0.p, 0, (size_t)c->C * 4, c->stream));
    }
    w_cycle_mixed(c, 0, sp->pre_smooth, sp->post_smooth, sp->coarse_extra, sp->vcycle_from, true, done);
    L0.b = c->rhs;
    for (int l = 1; l < c->nlevels; ++l) c->lv[l].bkind = 0;
    return CLQ_OK;
}

// the FCG scalars on the host: norm_host[0] = r.r, [1] = b.b, and the whole
// block s[0..8) from norm_host[2] (alpha of the last update at [4])
// The copy is waited for through an event recorded behind it, so that work
// queued after it keeps the GPU busy while the host reads the scalars.
static int queue_scalars(clq_ctx *c) {
    HIP_TRY(hipMemcpyAsync(c->norm_host + 2, c->fcg_s, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipEventRecord(c->ev_scal, c->stream));
    return CLQ_OK;
}
static int wait_scalars(clq_ctx *c) {
    HIP_TRY(hipEventSynchronize(c->ev_scal));
    c->norm_host[0] = c->norm_host[2 + 6];
    c->norm_host[1] = c->norm_host[2 + 7];
    return CLQ_OK;
}
static int read_scalars(clq_ctx *c) {
    if (int rc = queue_scalars(c)) return rc;
    return wait_scalars(c);
}

// A weighted cycle (coarse_weight != 1) that is not definite for the current
// residual shows as a step length alpha = z.r / d.Ad <= 0: for a definite
// preconditioner both dots carry the sign of the (negative definite)
// operator.  The solve then goes on with weight 1, restarting the direction
// recurrence; returns whether it did.
static bool drop_weight_on_breakdown(clq_ctx *c) {
    if (c->nlevels < 2 || c->lv[1].corr == 1.0f || c->norm_host[2 + 2] > 0.0) return false;
    for (int l = 1; l < c->nlevels; ++l) c->lv[l].corr = 1.0f;
    return true;
}

static int ensure_fcg_buffers(clq_ctx *c) {
    if (int rc = ensure_mixed_buffers(c)) return rc;
    if (!c->r64) {
        void *q;
        if (int rc = alloc_zeroed(c, &q, (size_t)c->C * 8)) return rc;
        c->r64 = (double *)q;
        for (float *&d : c->dir) {
            if (int rc = alloc_zeroed(c, &q, (size_t)c->C * 4)) return rc;
            d = (float *)q;
        }
        if (int rc = alloc_zeroed(c, &q, (size_t)c->C * 4)) return rc;
        c->p_alt = (float *)q;
        if (int rc = alloc_zeroed(c, &q, 16 * 8)) return rc;
        c->fcg_s = (double *)q;
        if (int rc = alloc_zeroed(c, &q, (size_t)fcg_partials_count() * 8)) return rc;
        c->fcg_partials = (double *)q;
        if (int rc = alloc_zeroed(c, &q, (size_t)c->nx * c->ny * 4)) return rc;
        c->x_rows = (int32_t *)q;
    }
    return CLQ_OK;
}

// level-0 active-row list of the FCG vector kernels (all cells without one)
#define ROWS0 (c->use_rows ? c->rows[0] : nullptr), c->nrows[0]
// level-0 row flags when the row lists are in use
#define FLAGS0 (c->use_rows ? c->row_flags[0] : nullptr)
// packed level-0 operator when the row lists are in use and it is exact
#define OP0 (c->use_rows && c->op0_ok ? c->op0 : nullptr)

// the FCG initialisation from the level-0 pressure guess p32 (x, r, r32,
// r.r and b.b)
static void fcg_init(clq_ctx *c, const float *p32) {
    const double h = (double)(float)c->dx;
    launch_fcg_init(c->lv[0], OP0, p32, c->rhs, 1.0 / (h * h), c->p64a, c->r64, c->r32, FLAGS0, ROWS0, c->x_rows,
                    c->x_rows_known, c->fcg_partials, c->fcg_s, c->stream);
    c->x_rows_known = c->use_rows;
}

// ||b - A x|| from the FP64 iterate (also r, r32 and r.r for a restart);
// returns through *r
static int true_residual(clq_ctx *c, double *r) {
    const double h = (double)(float)c->dx;
    launch_fcg_residual(c->lv[0], OP0, c->p64a, c->rhs, 1.0 / (h * h), ROWS0, c->r64, c->r32, c->fcg_partials,
                        c->fcg_s, c->stream);
    if (int rc = read_scalars(c)) return rc;
    *r = sqrt(c->norm_host[0]);
    return CLQ_OK;
}

static int solve_fcg(clq_ctx *c, const clq_solve_params *sp, clq_solve_stats *s) {
    if (int rc = ensure_fcg_buffers(c)) return rc;
    const Level &L0 = c->lv[0];
    const double h = (double)(float)c->dx;
    const double inv_h2 = 1.0 / (h * h);
    fcg_init(c, c->p);
    if (int rc = read_scalars(c)) return rc;
    double r = sqrt(c->norm_host[0]);
    const double bn = sqrt(c->norm_host[1]);
    if (!std::isfinite(r) || !std::isfinite(bn)) return non_finite(r + bn);
    s->initial_residual = r;
    s->rhs_norm = bn;
    s->recurrence_residual = r;
    // the update makes the next cycle's first level-0 red half-sweep
    const bool fold = fold_transfers(c) && c->op0_ok && sp->pre_smooth > 0 && swept_level(c, 0);
    bool plain = false;
    float *dold = c->dir[0], *dnew = c->dir[1];  // the direction, read and written by each update
    // each pass of the outer loop starts the direction recurrence from the
    // residual r = b - A x held in r64 / r32 (the initialisation, or the
    // recomputation after a pass whose recurrence met the tolerance)
    while (r > sp->tolerance * bn && s->cycles < sp->max_cycles) {
        // z: this iteration's preconditioned residual, the level-0
        // correction array of the cycle that made it; with the fold, the
        // update writes the next cycle's correction into the other array
        float *z = c->p, *zn = c->p_alt;
        c->lv[0].p = z;
        if (int rc = precondition_wcycle(c, sp, r)) return rc;
        s->cycles++;
        s->fallback_cycles += plain;
        launch_fcg_dots(L0, OP0, z, dold, c->r64, 1, inv_h2, ROWS0, c->fcg_partials, c->fcg_s, c->stream);
        bool first = true;
        for (;;) {
            launch_fcg_update(L0, OP0, z, c->p64a, c->r64, c->r32, dold, dnew, first, inv_h2, ROWS0,
                              fold ? zn : nullptr, (float)c->dx, c->fcg_partials, c->fcg_s, c->stream);
            std::swap(dold, dnew);
            if (int rc = queue_scalars(c)) return rc;
            if (fold) {
                // the next cycle's level-0 black half-sweep (on e = zn with
                // b = r32) runs while the host reads the scalars; it needs
                // neither the residual norm (the coarse b scale) nor the
                // weight, and when the solve stops here it is simply unused
                Level Lz = c->lv[0];
                Lz.p = zn;
                Lz.b = c->r32;
                launch_rbgs_rows_packed(Lz, c->op0, (float)c->dx, 1, c->rows[0], c->nrows[0], c->stream);
            }
            if (int rc = wait_scalars(c)) return rc;
            r = sqrt(c->norm_host[0]);
            if (!std::isfinite(r)) {
                c->lv[0].p = c->p;
                return non_finite(r);
            }
            if (r <= sp->tolerance * bn || s->cycles >= sp->max_cycles) break;
            first = drop_weight_on_breakdown(c);
            plain = plain || first;
            if (fold) {
                std::swap(z, zn);
                c->lv[0].p = z;
            }
            if (int rc = precondition_wcycle(c, sp, r, fold ? 2 : 0)) return rc;
            s->cycles++;
            s->fallback_cycles += plain;
            launch_fcg_dots(L0, OP0, z, dold, c->r64, first, inv_h2, ROWS0, c->fcg_partials, c->fcg_s, c->stream);
        }
        c->lv[0].p = c->p;
        // the recurrence says converged (or the cycles ran out): recompute
        // the residual from x; it also seeds the restart if one follows
        s->recurrence_residual = r;
        if (int rc = true_residual(c, &r)) return rc;
        if (!std::isfinite(r)) return non_finite(r);
        if (r > sp->tolerance * bn && s->cycles < sp->max_cycles) s->restarts++;
    }
    c->lv[0].p = c->p;
    // the pressure: float(x) on the active rows; the other rows' interior
    // cells are 0 already (cleared at the solve's start, never written by
    // the row kernels), and their ghost cells get the Neumann copy next
    if (c->use_rows) launch_narrow_rows(c->p64a, c->p, c->rows[0], c->nrows[0], c->nz, c->stream);
    else launch_widen_narrow(c->p64a, c->p, c->C, c->stream);
    s->final_residual = r;
    s->relative_residual = bn > 0 ? r / bn : 0.0;
    return check_launch("solve_fcg");
}

// Mode 3: FCG with the network as initial guess and preconditioner, the
// W-cycle taking over for good once an iteration reduces the residual by less
// than stall_ratio.  z (network or W-cycle output) lands in the level-0
// pressure array, as in mode 4.
static int solve_learned(clq_ctx *c, const clq_solve_params *sp, clq_solve_stats *s) {
    if (!c->net) return fail(CLQ_ERR_INVALID, "CLQ_SOLVE_LEARNED_FCG needs a network: call clq_attach_net first");
    if (int rc = ensure_fcg_buffers(c)) return rc;
    if (!c->guess32) {
        void *q;
        if (int rc = alloc_zeroed(c, &q, (size_t)c->C * 4)) return rc;
        c->guess32 = (float *)q;
    }
    const Level &L0 = c->lv[0];
    const double h = (double)(float)c->dx;
    const double inv_h2 = 1.0 / (h * h);
    // warm start (the previous pressure) against the network's guess
    double r_warm = 0.0, bn = 0.0;
    if (int rc = residual_norm(c, &r_warm, &bn)) return rc;
    unet_predict(c->net, c->rhs, L0.mask, c->p, c->guess32, c->stream);
    fcg_init(c, c->guess32);
    if (int rc = read_scalars(c)) return rc;
    double r = sqrt(c->norm_host[0]);
    bn = sqrt(c->norm_host[1]);
    if (!(r < r_warm)) {
        fcg_init(c, c->p);
        if (int rc = read_scalars(c)) return rc;
        r = sqrt(c->norm_host[0]);
    }
    s->initial_residual = r;
    s->rhs_norm = bn;
    bool use_net = true;
    auto precondition = [&](double rnorm) -> int {
        if (use_net) {
            unet_predict(c->net, c->r32, L0.mask, nullptr, c->p, c->stream);
        } else {
            if (int rc = precondition_wcycle(c, sp, rnorm)) return rc;
            s->fallback_cycles++;
        }
        s->cycles++;
        return CLQ_OK;
    };
    if (!std::isfinite(r) || !std::isfinite(bn)) return non_finite(r + bn);
    s->recurrence_residual = r;
    float *dold = c->dir[0], *dnew = c->dir[1];
    // passes of the direction recurrence as in mode 4: a pass that meets the
    // tolerance by its recurrence is checked against b - A x, which restarts
    // the recurrence (with the W-cycle) when it falls short
    while (r > sp->tolerance * bn && s->cycles < sp->max_cycles) {
        if (int rc = precondition(r)) return rc;
        launch_fcg_dots(L0, OP0, c->p, dold, c->r64, 1, inv_h2, ROWS0, c->fcg_partials, c->fcg_s, c->stream);
        bool first = true;
        for (;;) {
            launch_fcg_update(L0, OP0, c->p, c->p64a, c->r64, c->r32, dold, dnew, first, inv_h2, ROWS0, nullptr,
                              (float)c->dx, c->fcg_partials, c->fcg_s, c->stream);
            std::swap(dold, dnew);
            if (int rc = read_scalars(c)) return rc;
            const double before = r;
            r = sqrt(c->norm_host[0]);
            if (!std::isfinite(r)) return non_finite(r);
            if (r <= sp->tolerance * bn || s->cycles >= sp->max_cycles) break;
            // the weighted W-cycle's breakdown guard (the network's steps keep
            // the stall test only)
            const bool restart = !use_net && drop_weight_on_breakdown(c);
            if (use_net && r > c->stall_ratio * before) use_net = false;
            if (int rc = precondition(r)) return rc;
            launch_fcg_dots(L0, OP0, c->p, dold, c->r64, restart, inv_h2, ROWS0, c->fcg_partials, c->fcg_s,
                            c->stream);
            first = restart;
        }
        s->recurrence_residual = r;
        if (int rc = true_residual(c, &r)) return rc;
        if (!std::isfinite(r)) return non_finite(r);
        if (r > sp->tolerance * bn && s->cycles < sp->max_cycles) {
            s->restarts++;
            use_net = false;
        }
    }
    launch_widen_narrow(c->p64a, c->p, c->C, c->stream);
    s->final_residual = r;
    s->relative_residual = bn > 0 ? r / bn : 0.0;
    return check_launch("solve_learned");
}

// ---- FP64 check solve -----------------------------------------------------------------------
int clq_fp64_check_solve(clq_ctx *c, double tolerance, int32_t max_iterations, clq_check_stats *st) {
    if (!c || !st || !(tolerance > 0.0) || max_iterations < 0) return fail(CLQ_ERR_INVALID, "bad check-solve arguments");
    DeviceGuard g(c->device);
    const auto t0 = std::chrono::steady_clock::now();
    if (!c->chk_c) {
        void *q;
        for (double *&a : c->chk) {
            if (int rc = alloc_zeroed(c, &q, (size_t)c->C * 8)) return rc;
            a = (double *)q;
        }
        if (int rc = alloc_zeroed(c, &q, (size_t)c->C * 4)) return rc;
        c->chk_r32 = (float *)q;
        if (int rc = alloc_zeroed(c, &q, (size_t)check_partials_count() * 8)) return rc;
        c->chk_partials = (double *)q;
        if (int rc = alloc_zeroed(c, &q, 16 * 8)) return rc;
        c->chk_c = (double *)q;
    }
    // the active rows of the current operator hold every interior fluid cell
    if (int rc = ensure_active_rows(c)) return rc;
    const int32_t *rows = c->nlevels >= 2 ? c->rows[0] : nullptr;
    const int64_t nrows = c->nlevels >= 2 ? c->nrows[0] : 0;
    const Level &L0 = c->lv[0];
    const double h = (double)(float)c->dx;
    const double inv_h2 = 1.0 / (h * h);
    double *x = c->chk[0], *r = c->chk[1], *z = c->chk[2], *pa = c->chk[3], *pb = c->chk[4], *qv = c->chk[5];
    double host[16];
    auto read = [&]() -> int {
        HIP_TRY(hipMemcpyAsync(host, c->chk_c, sizeof host, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        return CLQ_OK;
    };
    double bn = 0.0, rel = 0.0;
    int it = 0;
    // passes of CG: from x = 0, then restarts from the current x while the
    // residual recomputed at the end of a pass misses the tolerance (the FP64
    // recurrence drifts from b - A x near 1e-12 on large grids)
    for (int pass = 0; pass < 4; ++pass) {
        launch_check_init(L0, c->rhs, inv_h2, rows, nrows, pass > 0, x, r, z, pa, c->chk_partials, c->chk_c,
                          c->stream);
        if (int rc = read()) return rc;
        bn = sqrt(host[5]);
        double rn = sqrt(host[4]);
        // the recurrence is driven to half the tolerance, so that the
        // residual recomputed from x meets it
        while (rn > 0.5 * tolerance * bn && it < max_iterations) {
            const int batch = std::min(32, max_iterations - it);
            for (int b = 0; b < batch; ++b, ++it) {
                launch_check_iteration(L0, inv_h2, rows, nrows, x, r, z, pa, pb, qv, c->chk_partials, c->chk_c,
                                       c->stream);
                std::swap(pa, pb);
            }
            if (int rc = read()) return rc;
            rn = sqrt(host[4]);
            if (!std::isfinite(rn)) return fail(CLQ_ERR_NOT_CONVERGED, "check solve: residual norm is not finite");
        }
        launch_fcg_residual(L0, nullptr, x, c->rhs, inv_h2, rows, nrows, qv, c->chk_r32, c->chk_partials, c->chk_c,
                            c->stream);
        if (int rc = read()) return rc;
        rel = bn > 0 ? sqrt(host[6]) / bn : 0.0;
        if (rel <= tolerance || it >= max_iterations) break;
    }
    // the error of the solver's pressure against x
    launch_check_error(L0, c->p64_valid ? c->p64a : nullptr, c->p64_valid ? nullptr : c->p, x, rows, nrows,
                       c->chk_partials, c->chk_c, c->stream);
    if (int rc = read()) return rc;
    st->iterations = it;
    st->relative_residual = rel;
    st->converged = st->relative_residual <= tolerance;
    st->error_l2 = host[9] > 0 ? sqrt(host[8] / host[9]) : sqrt(host[8]);
    st->error_max = host[11] > 0 ? host[10] / host[11] : host[10];
    st->solve_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return check_launch("fp64_check_solve");
}

int clq_attach_net(clq_ctx *c, clq_net *net, double stall_ratio) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    if (net) {
        if (int rc = unet_check_grid(net, c->nx, c->ny, c->nz, c->dx)) return rc;
        if (unet_device(net) != c->device) return fail(CLQ_ERR_INVALID, "network and context are on different devices");
        if (!(stall_ratio > 0.0 && stall_ratio <= 1.0)) return fail(CLQ_ERR_INVALID, "stall_ratio must be in (0, 1]");
    }
    c->net = net;
    c->stall_ratio = stall_ratio;
    return CLQ_OK;
}

static int pressure_solve(clq_ctx *c, const clq_solve_params *sp, clq_solve_stats *st);

int clq_pressure_solve(clq_ctx *c, const clq_solve_params *sp, clq_solve_stats *st) {
    if (!c || !sp) return fail(CLQ_ERR_INVALID, "NULL argument");
    const float w = sp->coarse_weight == 0.0f ? 1.0f : sp->coarse_weight;
    if (!(w > 0.0f && w < 2.0f)) return fail(CLQ_ERR_INVALID, "coarse_weight %g outside (0, 2)", (double)w);
    // every prolongation of the solve's cycles weighs the coarse pressure by
    // w; the stage entry points (clq_prolongate, clq_w_cycle) keep weight 1
    for (int l = 1; l < c->nlevels; ++l) c->lv[l].corr = w;
    const int rc = pressure_solve(c, sp, st);
    for (int l = 1; l < c->nlevels; ++l) c->lv[l].corr = 1.0f;
    if (!rc && st) st->active_cells = (c->nlevels < 2 ? c->C : c->nrows[0] * c->nz);
    c->use_rows = false;
    return rc;
}

static int pressure_solve(clq_ctx *c, const clq_solve_params *sp, clq_solve_stats *st) {
    if (sp->pre_smooth < 0 || sp->post_smooth < 0 || sp->coarse_extra < 0 || sp->max_cycles < 0)
        return fail(CLQ_ERR_INVALID, "negative sweep or cycle count");
    if (sp->mode < CLQ_SOLVE_FIXED_CYCLES || sp->mode > CLQ_SOLVE_FCG)
        return fail(CLQ_ERR_INVALID, "unknown solve mode %d", sp->mode);
    if (sp->mode == CLQ_SOLVE_LEARNED_FCG && !c->net)
        return fail(CLQ_ERR_INVALID, "CLQ_SOLVE_LEARNED_FCG needs a network: call clq_attach_net first");
    DeviceGuard g(c->device);
    clq_solve_stats s;
    memset(&s, 0, sizeof s);
    c->p64_valid = false;
    stage_begin(c, CLQ_STAGE_PRESSURE_SOLVE);
    ensure_coarse_operators(c);
    if (int rc = begin_sparse_solve(c)) return rc;
    if (sp->mode == CLQ_SOLVE_MIXED_REFINEMENT || sp->mode == CLQ_SOLVE_FCG || sp->mode == CLQ_SOLVE_LEARNED_FCG) {
        const int rc = sp->mode == CLQ_SOLVE_FCG           ? solve_fcg(c, sp, &s)
                       : sp->mode == CLQ_SOLVE_LEARNED_FCG ? solve_learned(c, sp, &s)
                                                           : solve_mixed(c, sp, &s);
        if (rc) return rc;
        c->p64_valid = true;
        launch_neumann_bc(c->p, c->nx, c->ny, c->nz, c->stream);
        stage_end(c, CLQ_STAGE_PRESSURE_SOLVE);
        HIP_TRY(hipEventSynchronize(c->ev_stop[CLQ_STAGE_PRESSURE_SOLVE]));
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, c->ev_start[CLQ_STAGE_PRESSURE_SOLVE], c->ev_stop[CLQ_STAGE_PRESSURE_SOLVE]);
        s.solve_ms = ms;
        if (st) *st = s;
        return check_launch("pressure_solve");
    }
    double r = 0.0, bn = 0.0;
    if (int rc = residual_norm(c, &r, &bn)) return rc;
    s.initial_residual = r;
    s.rhs_norm = bn;
    const double target = sp->tolerance * bn;
    for (int cyc = 0; cyc < sp->max_cycles; ++cyc) {
        if (sp->mode == CLQ_SOLVE_FP32_TOLERANCE && r <= target) break;
        w_cycle(c, 0, sp->pre_smooth, sp->post_smooth, sp->coarse_extra, sp->vcycle_from);
        s.cycles++;
        if (sp->mode == CLQ_SOLVE_FP32_TOLERANCE) {
            const double before = r;
            if (int rc = residual_norm(c, &r, &bn)) return rc;
            if (r > 0.95 * before) break;  // stagnation at the FP32 floor
        }
    }
    if (sp->mode == CLQ_SOLVE_FIXED_CYCLES)
        if (int rc = residual_norm(c, &r, &bn)) return rc;
    if (!std::isfinite(r) || !std::isfinite(bn)) return non_finite(r + bn);
    s.final_residual = r;
    s.relative_residual = bn > 0 ? r / bn : 0.0;
    launch_neumann_bc(c->p, c->nx, c->ny, c->nz, c->stream);  // pressure_solve.py:84
    stage_end(c, CLQ_STAGE_PRESSURE_SOLVE);
    HIP_TRY(hipEventSynchronize(c->ev_stop[CLQ_STAGE_PRESSURE_SOLVE]));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev_start[CLQ_STAGE_PRESSURE_SOLVE], c->ev_stop[CLQ_STAGE_PRESSURE_SOLVE]);
    s.solve_ms = ms;
    if (st) *st = s;
    return check_launch("pressure_solve");
}

int clq_pressure_correction(clq_ctx *c, double rho, double dt) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    DeviceGuard g(c->device);
    // the FP64 pressure of modes 2-4 is 0 off the interior fluid cells: with
    // the operator's row flags, rows that cannot change are skipped
    if (c->p64_valid)
        launch_pressure_correction64(c->u, c->v, c->w, c->p64a, c->rows_valid ? c->row_flags[0] : nullptr,
                                     (double)(float)(dt / rho), (double)(float)(1.0 / c->dx), c->nx, c->ny, c->nz,
                                     c->stream);
    else
        launch_pressure_correction(c->u, c->v, c->w, c->p, (float)(dt / rho), (float)(1.0 / c->dx), c->nx, c->ny,
                                   c->nz, c->stream);
    return check_launch("pressure_correction");
}

// ---- particles -------------------------------------------------------------------------------
static ParticleUpdate make_update(clq_ctx *c, int flags, double flip_ratio, double dt) {
    ParticleUpdate pu;
    pu.flags = flags;
    pu.inv_dx = (float)(1.0 / c->dx);
    pu.pure_flip = flip_ratio == 1.0;
    pu.a = (float)flip_ratio;
    pu.b = (float)(1.0 - flip_ratio);
    pu.dt = (float)dt;
    pu.lower = (float)(1.001 * c->dx);  // boundary_conditions.py:98-101
    pu.upper[0] = (float)((c->nx - 1.001) * c->dx);
    pu.upper[1] = (float)((c->ny - 1.001) * c->dx);
    pu.upper[2] = (float)((c->nz - 1.001) * c->dx);
    return pu;
}

static int particle_update(clq_ctx *c, int flags, double flip_ratio, double dt, const char *what) {
    DeviceGuard g(c->device);
    // with the particles in bin order, the G2P samples are staged per bin row
    const bool binned = c->bins_current;
    if (flags & (kAdvect | kDomainBC | kObstacleBC)) c->bins_current = false;
    const ParticleUpdate pu = make_update(c, flags, flip_ratio, dt);
    launch_particle_update(c->pos, c->vel, c->prev, c->u, c->v, c->w, c->uo, c->vo, c->wo,
                           binned ? c->offsets : nullptr, c->counts, c->obstacle, (int)c->n, c->nx, c->ny, c->nz, pu,
                           c->stream);
    return check_launch(what);
}

int clq_grid_to_particle(clq_ctx *c, double flip_ratio) {
    if (!c || !(flip_ratio >= 0.0 && flip_ratio <= 1.0)) return fail(CLQ_ERR_INVALID, "flip_ratio must be in [0, 1]");
    return particle_update(c, kG2P, flip_ratio, 0.0, "grid_to_particle");
}

int clq_save_previous_positions(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    return particle_update(c, kSavePrev, 1.0, 0.0, "save_previous_positions");
}

int clq_advect_particles(clq_ctx *c, double dt) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    return particle_update(c, kAdvect, 1.0, dt, "advect_particles");
}

int clq_particle_domain_bc(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    return particle_update(c, kDomainBC, 1.0, 0.0, "particle_domain_bc");
}

int clq_particle_obstacle_bc(clq_ctx *c) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    if (!c->has_obstacle) return CLQ_OK;  // no solid voxel: the kernel returns before any write
    return particle_update(c, kObstacleBC, 1.0, 0.0, "particle_obstacle_bc");
}

int clq_grid_to_particle_advect(clq_ctx *c, double flip_ratio, double dt) {
    if (!c || !(flip_ratio >= 0.0 && flip_ratio <= 1.0)) return fail(CLQ_ERR_INVALID, "flip_ratio must be in [0, 1]");
    // Without solids the previous positions are never read, so the copy is
    // skipped; clq_get_previous_positions then reports the last copy made.
    const int flags = c->has_obstacle ? (kG2P | kSavePrev | kAdvect | kDomainBC | kObstacleBC)
                                      : (kG2P | kAdvect | kDomainBC);
    stage_begin(c, CLQ_STAGE_G2P_ADVECT);
    kernel_mark(c, CLQ_KERNEL_PARTICLE_UPDATE);
    int rc = particle_update(c, flags, flip_ratio, dt, "grid_to_particle_advect");
    kernel_mark(c, CLQ_KERNEL_PARTICLE_UPDATE);
    stage_end(c, CLQ_STAGE_G2P_ADVECT);
    return rc;
}

// ---- particle sources --------------------------------------------------------------------
int clq_set_sources(clq_ctx *c, int32_t count, const int32_t *types, const uint8_t *masks, const int32_t *densities,
                    const float *velocities) {
    if (!c || count < 0 || (count > 0 && (!types || !masks || !densities || !velocities)))
        return fail(CLQ_ERR_INVALID, "bad source arguments");
    for (int32_t q = 0; q < count; ++q)
        if (types[q] < CLQ_SOURCE_GEOMETRY || types[q] > CLQ_SOURCE_OUTFLOW)
            return fail(CLQ_ERR_INVALID, "unknown source type %d", types[q]);
    DeviceGuard g(c->device);
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (auto &src : c->sources) c->release(src.open);
    c->sources.clear();
    void *q;
    if (!c->sink) {
        if (int rc = c->alloc(&q, (size_t)c->C)) return rc;
        c->sink = (uint8_t *)q;
        if (int rc = c->alloc(&q, 16)) return rc;
        c->src_scalar = (int32_t *)q;
        HIP_TRY(hipHostMalloc((void **)&c->src_scalar_host, 16, hipHostMallocDefault));
    }
    // open cells per source and the union of the sinks, on the host
    std::vector<uint8_t> solid((size_t)c->C), cells((size_t)c->C), sink((size_t)c->C, 0);
    HIP_TRY(hipMemcpy(solid.data(), c->obstacle, (size_t)c->C, hipMemcpyDeviceToHost));
    for (int32_t s = 0; s < count; ++s) {
        const uint8_t *m = masks + (size_t)s * c->C;
        clq_ctx::Source src;
        src.type = types[s];
        src.density = std::max(1, densities[s]);
        for (int a = 0; a < 3; ++a) src.velocity[a] = velocities[3 * s + a];
        for (int64_t i = 0; i < c->C; ++i) {
            cells[i] = m[i] && !solid[i];
            if (src.type == CLQ_SOURCE_OUTFLOW && m[i]) sink[i] = 1;
        }
        if (int rc = c->alloc(&q, (size_t)c->C)) return rc;
        src.open = (uint8_t *)q;
        HIP_TRY(hipMemcpy(src.open, cells.data(), (size_t)c->C, hipMemcpyHostToDevice));
        c->sources.push_back(src);
    }
    HIP_TRY(hipMemcpy(c->sink, sink.data(), (size_t)c->C, hipMemcpyHostToDevice));
    return CLQ_OK;
}

int clq_apply_sources(clq_ctx *c, int32_t initial, const uint64_t *rng, int64_t *draws) {
    if (!c || !rng || !draws) return fail(CLQ_ERR_INVALID, "NULL argument");
    DeviceGuard g(c->device);
    const float delta = (float)c->dx;
    *draws = 0;
    bool any_sink = false;
    for (auto &src : c->sources) any_sink = any_sink || src.type == CLQ_SOURCE_OUTFLOW;
    // sinks: stable compaction of the particles outside every OUTFLOW cell
    if (any_sink && c->n > 0) {
        const int n = (int)c->n;
        launch_outflow_flags(c->pos, n, c->sink, delta, c->nx, c->ny, c->nz, c->key, c->stream);
        launch_exclusive_scan(c->key, c->rank, n, c->pscan, c->stream);
        launch_compact(c->pos, c->vel, c->key, c->rank, n, c->pos_work, c->vel_work, c->stream);
        HIP_TRY(hipMemcpyAsync(c->src_scalar_host, c->rank + (n - 1), 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(c->src_scalar_host + 1, c->key + (n - 1), 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        std::swap(c->pos, c->pos_work);
        std::swap(c->vel, c->vel_work);
        c->n = (int64_t)c->src_scalar_host[0] + c->src_scalar_host[1];
    }
    // top-ups, each counting the particles present at that moment
    uint64_t used = 0;
    for (auto &src : c->sources) {
        if (!(src.type == CLQ_SOURCE_INFLOW || (initial && src.type == CLQ_SOURCE_GEOMETRY))) continue;
        HIP_TRY(hipMemsetAsync(c->counts, 0, (size_t)c->C * 4, c->stream));
        launch_count_open(c->pos, (int)c->n, src.open, delta, c->nx, c->ny, c->nz, c->counts, c->stream);
        launch_missing(c->counts, src.open, src.density, c->C, c->stream);
        launch_exclusive_scan(c->counts, c->offsets, c->C, c->scan_scratch, c->stream);
        HIP_TRY(hipMemcpyAsync(c->src_scalar_host, c->offsets + (c->C - 1), 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipMemcpyAsync(c->src_scalar_host + 1, c->counts + (c->C - 1), 4, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        const int64_t total = (int64_t)c->src_scalar_host[0] + c->src_scalar_host[1];
        if (total == 0) continue;  // no draw, as on the host
        if (c->n + total >= (1LL << 31)) return fail(CLQ_ERR_INVALID, "particle count would reach 2^31");
        if (c->n + total > c->cap) {  // solver.py:290-291
            const int64_t cap = std::max<int64_t>(c->n + total, c->cap * 3 / 2);
            if (int rc = alloc_particles(c, cap, c->n)) return rc;
        }
        launch_emit(c->counts, c->offsets, c->C, c->ny, c->nz, c->dx, rng, used, c->n, src.velocity, c->pos,
                    c->vel, c->stream);
        used += 3ull * (uint64_t)total;
        c->n += total;
    }
    *draws = (int64_t)used;
    c->bins_current = false;
    return check_launch("apply_sources");
}

int clq_voxelise_mesh(clq_ctx *c, const float *triangles, int64_t count, double ox, double oy, double oz,
                      uint8_t *mask) {
    if (!c || count < 0 || (count > 0 && !triangles) || !mask) return fail(CLQ_ERR_INVALID, "bad mesh arguments");
    if (count == 0) return CLQ_OK;  // no triangle: no cell changes
    if (count > (1LL << 31)) return fail(CLQ_ERR_INVALID, "more than 2^31 triangles");
    DeviceGuard g(c->device);
    // one-time set-up: the triangles, their prepared terms and the mask live
    // only for this call
    void *tri = nullptr, *scratch = nullptr, *dmask = nullptr;
    if (int rc = c->alloc(&tri, (size_t)count * 36)) return rc;
    if (int rc = c->alloc(&scratch, mesh_scratch_bytes(count))) { c->release(tri); return rc; }
    if (int rc = c->alloc(&dmask, (size_t)c->C)) { c->release(tri); c->release(scratch); return rc; }
    const float origin[3] = {(float)ox, (float)oy, (float)oz};
    int rc = CLQ_OK;
    if (hipMemcpyAsync(tri, triangles, (size_t)count * 36, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(dmask, mask, (size_t)c->C, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = fail(CLQ_ERR_DEVICE, "mesh upload failed");
    if (!rc) {
        launch_voxelise_mesh((const float *)tri, count, scratch, c->nx, c->ny, c->nz, (float)c->dx, origin,
                             (uint8_t *)dmask, c->stream);
        rc = check_launch("voxelise_mesh");
    }
    if (!rc && hipMemcpyAsync(mask, dmask, (size_t)c->C, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = fail(CLQ_ERR_DEVICE, "mask download failed");
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = fail(CLQ_ERR_DEVICE, "voxelise_mesh sync failed");
    c->release(tri);
    c->release(scratch);
    c->release(dmask);
    return rc;
}

int clq_voxelise(clq_ctx *c, int32_t scale, double particles_per_cell, uint8_t *occupancy, float *density) {
    if (!c || scale < 1 || !(particles_per_cell > 0)) return fail(CLQ_ERR_INVALID, "bad voxelise arguments");
    if (scale > voxelise_max_scale()) return fail(CLQ_ERR_UNSUPPORTED, "voxelise scale %d > %d", scale,
                                                  voxelise_max_scale());
    DeviceGuard g(c->device);
    if (density && !c->bins_current)
        if (int rc = clq_bin_particles(c)) return rc;  // the density gather walks the bins
    const int64_t S = (int64_t)scale * scale * scale * c->C;
    void *q = nullptr;
    uint8_t *docc = nullptr;
    float *ddens = nullptr;
    // output staging lives only for this call (the outputs are host arrays)
    if (occupancy) {
        if (int rc = c->alloc(&q, (size_t)S)) return rc;
        docc = (uint8_t *)q;
        HIP_TRY(hipMemsetAsync(docc, 0, (size_t)S, c->stream));
    }
    if (density) {
        if (int rc = c->alloc(&q, (size_t)S * 4)) { c->release(docc); return rc; }
        ddens = (float *)q;
    }
    const float sinv = (float)scale * (float)(1.0 / c->dx);
    launch_voxelise(c->pos, (int)c->n, c->offsets, c->counts, scale, sinv,
                    (double)scale * scale * scale / particles_per_cell, c->nx, c->ny, c->nz, docc, ddens, c->stream);
    int rc = check_launch("voxelise");
    if (!rc && docc) rc = hipMemcpyAsync(occupancy, docc, (size_t)S, hipMemcpyDeviceToHost, c->stream) == hipSuccess
                              ? CLQ_OK : fail(CLQ_ERR_DEVICE, "occupancy download failed");
    if (!rc && ddens) rc = hipMemcpyAsync(density, ddens, (size_t)S * 4, hipMemcpyDeviceToHost, c->stream) == hipSuccess
                               ? CLQ_OK : fail(CLQ_ERR_DEVICE, "density download failed");
    if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = fail(CLQ_ERR_DEVICE, "voxelise sync failed");
    c->release(docc);
    c->release(ddens);
    return rc;
}

int clq_kernel_timing(clq_ctx *c, int32_t enable) {
    if (!c) return fail(CLQ_ERR_INVALID, "ctx is NULL");
    DeviceGuard g(c->device);
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (auto &v : c->kevents) {
        c->event_pool.insert(c->event_pool.end(), v.begin(), v.end());
        v.clear();
    }
    c->ktiming = enable != 0;
    return CLQ_OK;
}

int clq_kernel_times(clq_ctx *c, double *total_ms, int64_t *launches, int32_t capacity) {
    if (!c || !total_ms || !launches) return fail(CLQ_ERR_INVALID, "NULL argument");
    DeviceGuard g(c->device);
    HIP_TRY(hipStreamSynchronize(c->stream));
    const int n = std::min<int>(capacity, CLQ_KERNEL_CLASS_COUNT);
    for (int k = 0; k < n; ++k) {
        double total = 0.0;
        const auto &v = c->kevents[k];
        for (size_t q = 0; q + 1 < v.size(); q += 2) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, v[q], v[q + 1]));
            total += ms;
        }
        total_ms[k] = total;
        launches[k] = (int64_t)(v.size() / 2);
    }
    return n;
}

int clq_stage_times(clq_ctx *c, double *ms, int32_t capacity) {
    if (!c || !ms) return fail(CLQ_ERR_INVALID, "NULL argument");
    DeviceGuard g(c->device);
    const int n = std::min<int>(capacity, CLQ_STAGE_COUNT);
    for (int s = 0; s < n; ++s) {
        ms[s] = 0.0;
        if (!c->ev_used[s]) continue;
        if (hipEventSynchronize(c->ev_stop[s]) != hipSuccess) continue;
        float t = 0.0f;
        if (hipEventElapsedTime(&t, c->ev_start[s], c->ev_stop[s]) == hipSuccess) ms[s] = t;
    }
    return n;
}

}  // extern "C"
