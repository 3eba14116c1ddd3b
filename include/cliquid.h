/*
 * cliquid.h -- C-ABI of the Continuum Liquid FLIP/PIC time step.
 *
 * One shared library exports this ABI: libcliquid_hip.so (hand-written gfx950
 * HIP kernels, the product).  The CPU oracle used by the tests
 * (oracle/libcliquid_oracle.so) exports the same symbols so that both can be
 * driven by identical call sequences and their downloads compared.
 *
 * What this replaces.  The reference drives 23 OpenCL kernels from Python
 * stage wrappers of the form (kernels: dict, queue, cl.Buffer..., numpy
 * scalars) -> buffers (Solver/Kernel/particles.py:73-226,
 * Solver/Kernel/multigrid.py:53-249, Solver/Kernel/pressure_solve.py:10-114,
 * Solver/Kernel/velocity_update.py:8-52,
 * Solver/Kernel/boundary_conditions.py:8-132).  solver.py allocates every
 * buffer (Solver/Kernel/helper.py:81-83, solver.py:161-217) and stages never
 * allocate.  Here the library owns an opaque context that holds every device
 * buffer of one simulation grid plus one HIP stream; each stage wrapper becomes
 * one entry point below that enqueues on that stream.
 *
 * Conventions (all entry points):
 *   - return 0 (CLQ_OK) on success, a CLQ_ERR_* code otherwise; the message of
 *     the last failure on the calling thread is clq_last_error().  No C++
 *     exception crosses the ABI.
 *   - grid arrays are C-ordered (nx, ny, nz), k fastest, one ghost layer per
 *     face (Solver/Kernel/OpenCL/helper.cl:4-6, :21-23).
 *   - particle arrays are AoS float32 [n][3] (Solver/Kernel/solver.py:164-173).
 *   - physical scalars are passed as double and rounded to float32 exactly as
 *     the reference host does before each launch (e.g. FIELD_DTYPE(1.0 / dx),
 *     Solver/Kernel/particles.py:103).
 *   - stage calls are asynchronous on the context stream; clq_get_* downloads
 *     and clq_sync() wait.  Calls into one context must be serialised by the
 *     caller; distinct contexts are independent.
 */
#ifndef CLIQUID_H
#define CLIQUID_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLQ_ABI_VERSION 4

enum clq_status {
    CLQ_OK = 0,
    CLQ_ERR_INVALID = 1,     /* bad argument, size or state              */
    CLQ_ERR_DEVICE = 2,      /* HIP runtime / kernel launch failure      */
    CLQ_ERR_NOMEM = 3,       /* allocation failed                        */
    CLQ_ERR_UNSUPPORTED = 4, /* mode not implemented by this library     */
    CLQ_ERR_NOT_CONVERGED = 5
};

/* Grid field identifiers for clq_set_field / clq_get_field. */
enum clq_field {
    CLQ_FIELD_U = 0,              /* float32[C]  MAC +x face velocity            */
    CLQ_FIELD_V = 1,              /* float32[C]  MAC +y face velocity            */
    CLQ_FIELD_W = 2,              /* float32[C]  MAC +z face velocity            */
    CLQ_FIELD_U_OLD = 3,          /* float32[C]  work buffers: after
                                     clq_update_velocity they hold the
                                     post-BC P2G velocity (solver.py:374-376) */
    CLQ_FIELD_V_OLD = 4,
    CLQ_FIELD_W_OLD = 5,
    CLQ_FIELD_FLUID_FRACTION = 6, /* float32[C]  occupancy, count/8            */
    CLQ_FIELD_PRESSURE = 7,       /* float32[C]                                */
    CLQ_FIELD_RHS = 8,            /* float32[C]  rho/dt * div u                */
    CLQ_FIELD_DIAGONAL = 9,       /* float32[C]  finest operator               */
    CLQ_FIELD_WEIGHT_X = 10,
    CLQ_FIELD_WEIGHT_Y = 11,
    CLQ_FIELD_WEIGHT_Z = 12,
    CLQ_FIELD_BIN_COUNTS = 13,    /* int32[C]    particles per bin              */
    CLQ_FIELD_BIN_OFFSETS = 14,   /* int32[C]    exclusive scan of counts       */
    CLQ_FIELD_OBSTACLE = 15,      /* uint8[C]    static solid voxels           */
    CLQ_FIELD_PRESSURE_FP64 = 16, /* float64[C]  FP64 iterate of solve modes 2-4;
                                     while valid, clq_pressure_correction uses
                                     it instead of the FP32 pressure            */
    CLQ_FIELD_COUNT = 17
};

/* Per-level multigrid arrays for clq_get_level_field / clq_set_level_field.
 * Level 0 is the simulation grid (p = pressure, b = rhs, mask = fluid
 * fraction), levels 1..L-1 are the coarse grids of
 * Solver/Kernel/multigrid.py:18-50. */
enum clq_level_field {
    CLQ_LEVEL_P = 0,
    CLQ_LEVEL_B = 1,
    CLQ_LEVEL_MASK = 2,
    CLQ_LEVEL_DIAGONAL = 3,
    CLQ_LEVEL_WEIGHT_X = 4,
    CLQ_LEVEL_WEIGHT_Y = 5,
    CLQ_LEVEL_WEIGHT_Z = 6
};

/* Pressure-solve modes. */
enum clq_solve_mode {
    /* Exactly `max_cycles` FP32 W-cycles, no residual test: the reference
     * driver (Solver/Kernel/pressure_solve.py:61-84, iterations from
     * settings.iterations, solver.py:125). */
    CLQ_SOLVE_FIXED_CYCLES = 0,
    /* FP32 W-cycles until the FP64-accumulated relative residual
     * ||b - A p||_2 / ||b||_2 <= tolerance, or max_cycles, or stagnation: a
     * cycle that leaves the residual above 0.95x its previous value ends the
     * solve (FP32 storage of p bounds the attainable residual). */
    CLQ_SOLVE_FP32_TOLERANCE = 1,
    /* FP64 iterative refinement around a W-cycle with FP32 smoothing and
     * 16-bit coarse-level residual storage; FP64 residual, FP64 correction
     * accumulation and FP64 coarsest solve. */
    CLQ_SOLVE_MIXED_REFINEMENT = 2,
    /* Flexible CG in FP64 with the learned network (clq_attach_net) as the
     * initial guess and as the preconditioner; once an iteration leaves the
     * residual above stall_ratio x its previous value, the mixed-precision
     * W-cycle of mode 4 preconditions every later iteration.  The network's
     * initial guess is kept only when its FP64 residual is below the warm
     * start's (the previous pressure, as the reference starts). */
    CLQ_SOLVE_LEARNED_FCG = 3,
    /* Flexible CG in FP64 preconditioned by the mixed-precision W-cycle of
     * mode 2 (one W-cycle per iteration). */
    CLQ_SOLVE_FCG = 4
};

typedef struct clq_solve_params {
    int32_t mode;          /* enum clq_solve_mode                               */
    int32_t pre_smooth;    /* reference: 2 (pressure_solve.py:77)               */
    int32_t post_smooth;   /* reference: 4 (pressure_solve.py:78)               */
    int32_t coarse_extra;  /* reference: 64 -> coarsest runs pre+64 sweeps      */
    int32_t max_cycles;    /* fixed count (mode 0) or cap                        */
    int32_t coarse_dtype;  /* modes 2-4: 0 = fp16, 1 = bf16 coarse b storage    */
    int32_t vcycle_from;   /* 0: W-cycle at every level (multigrid.py:53-187);
                              k > 0: the coarse-grid correction runs twice
                              below levels < k and once from level k down;
                              -m < 0: twice below the levels whose bit is set
                              in m (-2: level 1 only), once below the rest */
    float coarse_weight;   /* weight w of every coarse-grid correction of the
                              cycle, p += w * p_coarse[parent]; 0 selects 1,
                              the reference's plain sum (multigrid.cl:182-201);
                              otherwise 0 < w < 2.  Piecewise-constant
                              prolongation under-corrects smooth errors, and
                              w near 1.6 roughly halves the FCG iterations    */
    double tolerance;      /* relative residual target (modes 1-4)              */
} clq_solve_params;

typedef struct clq_solve_stats {
    int32_t cycles;            /* cycles (modes 0-2) or FCG iterations (3, 4)   */
    int32_t fallback_cycles;   /* mode 3: iterations preconditioned by the
                                  W-cycle after the network stalled; mode 4:
                                  iterations run with coarse_weight 1 after
                                  the weighted cycle broke down (a step
                                  length alpha <= 0)                         */
    double rhs_norm;           /* ||b||_2 (FP64 accumulation)                    */
    double initial_residual;   /* ||b - A p0||_2 before the solve               */
    double final_residual;     /* ||b - A p||_2 after the solve                 */
    double relative_residual;  /* final_residual / rhs_norm                     */
    double solve_ms;           /* device time of the solve (HIP events)         */
    int64_t active_cells;      /* cells of the level-0 rows (i, j, *) holding an
                                  interior fluid cell: what the finest-level
                                  smoother and FCG vectors visit             */
    double recurrence_residual; /* modes 3, 4: ||r|| of the r -= alpha q
                                  recurrence when the iteration stopped; the
                                  final_residual above is recomputed as
                                  ||b - A x|| from the FP64 iterate, and when it
                                  misses the tolerance while cycles remain the
                                  iteration restarts from it              */
    int32_t restarts;          /* modes 3, 4: such restarts                      */
} clq_solve_stats;

typedef struct clq_ctx clq_ctx;

/* ---- library / context ---------------------------------------------------- */
const char *clq_last_error(void);
int clq_abi_version(void);
/* Number of usable devices (the CPU oracle reports 1). */
int clq_device_count(int32_t *count);
/* Creates a context for an (nx, ny, nz) grid of spacing dx holding up to
 * `capacity` particles.  Allocates every buffer of solver.py:161-217,
 * including the multigrid hierarchy with min_size 8 (solver.py:203-213). */
int clq_create(int32_t nx, int32_t ny, int32_t nz, double dx, int64_t capacity, clq_ctx **out);
int clq_destroy(clq_ctx *ctx);
int clq_sync(clq_ctx *ctx);

/* ---- data movement -------------------------------------------------------- */
/* Uploads n particles (AoS float32[n][3] each); grows capacity to
 * max(n, 3*cap/2) when needed (solver.py:290-303).  vel may be NULL (zeros). */
int clq_set_particles(clq_ctx *ctx, const float *pos, const float *vel, int64_t n);
/* Downloads the current n particles; either pointer may be NULL. */
int clq_get_particles(clq_ctx *ctx, float *pos, float *vel, int64_t n);
int clq_get_previous_positions(clq_ctx *ctx, float *pos, int64_t n);
int clq_get_particle_count(clq_ctx *ctx, int64_t *n);

/* ---- frame output (Solver/Kernel/output.py:56-75, 178-195) -------------------
 * The reference maps ALLOC_HOST_PTR buffers as pinned frame slots and copies
 * the positions into one per output frame.  Here:
 *   clq_host_alloc / clq_host_free: pinned (page-locked) host memory of the
 *     context's device, for the frame slots;
 *   clq_positions_to_host_async: snapshots the current positions on the
 *     device (stream order: after every step enqueued so far) and copies the
 *     snapshot into `host` (float32[n][3], n = particle count, at most
 *     `capacity` rows) on a separate copy stream; returns at once with the
 *     count and a ticket.  Later steps run while the copy is in flight;
 *   clq_host_copy_wait: blocks until the copy of `ticket` has landed (a
 *     ticket of -1, returned for an empty set, needs no wait).
 * Tickets stay valid while fewer than 16 newer copies have been started. */
int clq_host_alloc(clq_ctx *ctx, int64_t bytes, void **host);
int clq_host_free(clq_ctx *ctx, void *host);
int clq_positions_to_host_async(clq_ctx *ctx, float *host, int64_t capacity, int64_t *n, int64_t *ticket);
int clq_host_copy_wait(clq_ctx *ctx, int64_t ticket);
int clq_set_field(clq_ctx *ctx, int32_t field, const void *host, int64_t bytes);
int clq_get_field(clq_ctx *ctx, int32_t field, void *host, int64_t bytes);
int clq_level_count(clq_ctx *ctx, int32_t *levels);
int clq_level_shape(clq_ctx *ctx, int32_t level, int32_t *nx, int32_t *ny, int32_t *nz);
int clq_set_level_field(clq_ctx *ctx, int32_t level, int32_t field, const float *host, int64_t bytes);
int clq_get_level_field(clq_ctx *ctx, int32_t level, int32_t field, float *host, int64_t bytes);

/* ---- particle_to_grid section (Solver/Kernel/particles.py:73-171) ---------- */
/* Histogram, exclusive scan and counting-sort scatter of positions and
 * velocities into bins (kernels count_particles_per_cell, scan_blocks_exclusive,
 * add_scanned_block_offsets, sort_particles: particles.cl:44-133), then the
 * double-buffer swap (particles.py:136-140).  The order within a bin is the
 * stable (original-index) order. */
int clq_bin_particles(clq_ctx *ctx);
/* P2G gather onto the MAC grid plus occupancy voxelisation
 * (particle_to_grid_gather, particles.cl:135-227).  Requires binned particles. */
int clq_particle_to_grid(clq_ctx *ctx, double particles_per_cell);

/* ---- grid glue ------------------------------------------------------------- */
/* grid_obstacle_bc + extrapolate_obstacle_velocity
 * (boundary_conditions.cl:99-173, boundary_conditions.py:53-79). */
int clq_grid_obstacle_bc(clq_ctx *ctx);
/* grid_domain_bc (boundary_conditions.cl:33-67). */
int clq_grid_domain_bc(clq_ctx *ctx);
/* update_velocity into the work buffers followed by the swap of
 * solver.py:374-376 (velocity_update.cl:3-43). gravity/force are float[3]. */
int clq_update_velocity(clq_ctx *ctx, const double *gravity, const double *force, double rho, double nu,
                        double dt);

/* ---- pressure_solve section (Solver/Kernel/pressure_solve.py:10-85) -------- */
/* pressure_rhs (pressure_solve.cl:3-24). */
int clq_pressure_rhs(clq_ctx *ctx, double rho, double dt);
/* build_finest_operator (multigrid.cl:63-88). */
int clq_build_operator(clq_ctx *ctx);
/* The solve proper followed by neumann_domain_bc (pressure_solve.py:61-84).
 * A residual that is not finite (a diverged or NaN-fed solve) returns
 * CLQ_ERR_NOT_CONVERGED: later stages would otherwise advect NaN velocities. */
int clq_pressure_solve(clq_ctx *ctx, const clq_solve_params *params, clq_solve_stats *stats);

/* FP64 check of a pressure solve.  No counterpart in the reference.  Solves
 * the current level-0 system (the rhs and the finest operator arrays) once
 * more, independently of the multigrid path, by Jacobi-preconditioned CG in
 * FP64 from x = 0 until its residual recurrence is below tolerance / 2 (or
 * max_iterations), then recomputes ||b - A x|| / ||b|| from x and reports
 * the error of the last clq_pressure_solve's pressure (the FP64 iterate
 * after modes 2-4, else the FP32 pressure) against that solution over the
 * interior fluid cells. */
typedef struct clq_check_stats {
    int32_t iterations;        /* CG iterations of the check solve              */
    int32_t converged;         /* 1 when the tolerance was met                  */
    double relative_residual;  /* ||b - A x_check|| / ||b||, recomputed          */
    double error_l2;           /* ||p - x_check||_2 / ||x_check||_2              */
    double error_max;          /* max |p - x_check| / max |x_check|              */
    double solve_ms;           /* wall time of the check solve                  */
} clq_check_stats;
int clq_fp64_check_solve(clq_ctx *ctx, double tolerance, int32_t max_iterations, clq_check_stats *stats);
/* One W-cycle on the current level-0 p/b/mask/operator (multigrid.py:53-187). */
int clq_w_cycle(clq_ctx *ctx, int32_t pre_smooth, int32_t post_smooth, int32_t coarse_extra);
/* Red-black Gauss-Seidel sweeps on one level (multigrid.py:190-220). */
int clq_smooth(clq_ctx *ctx, int32_t level, int32_t iterations);
/* restrict_residual from `level` to level+1 (multigrid.cl:90-180). */
int clq_restrict(clq_ctx *ctx, int32_t level);
/* prolongate from level+1 into `level` (multigrid.cl:182-201). */
int clq_prolongate(clq_ctx *ctx, int32_t level);
/* coarse_smooth on the coarsest level (multigrid.cl:232-264). */
int clq_coarse_smooth(clq_ctx *ctx, int32_t iterations);
/* neumann_domain_bc on the pressure (boundary_conditions.cl:18-31). */
int clq_neumann_bc(clq_ctx *ctx);
/* FP64-accumulated ||b - A p||_2 and ||b||_2 on level 0. */
int clq_residual_norm(clq_ctx *ctx, double *residual, double *rhs_norm);
/* clq_residual_norm without a context and without allocating.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops).  The arrays are
 * device pointers of the level-0 system of an (nx, ny, nz) grid of cell size
 * dx, laid out as the fields of the same names (p = pressure, rhs, mask =
 * fluid fraction, the finest operator); norms is a device double[2] that
 * receives ||b - A p||_2 and ||b||_2 (the same values as clq_residual_norm).
 * Enqueued on `stream` (a hipStream_t, NULL for the default stream) without
 * waiting.  The per-block partial sums live in a 32 KiB array of the
 * library's device memory, one per GPU; calls that overlap on one GPU race
 * on it.  Grids must have at least 3 and fewer than 2^31 cells (otherwise
 * CLQ_ERR_INVALID). */
int clq_residual_norm_noalloc(const float *p, const float *rhs, const float *mask, const float *diag,
                              const float *wx, const float *wy, const float *wz, int32_t nx, int32_t ny, int32_t nz,
                              double dx, double *norms, void *stream);
/* pressure_correction (pressure_solve.cl:26-47).  After a solve in modes 2-4
 * the FP64 pressure is used (ghost neighbours read the clamped interior value,
 * which is what neumann_domain_bc stores): the gradient and the update run in
 * FP64 and u is rounded once. */
int clq_pressure_correction(clq_ctx *ctx, double rho, double dt);

/* ---- grid_to_particle / advection section --------------------------------- */
/* G2P with FLIP/PIC blend (grid_to_particle, particles.cl:229-262).
 * flip_ratio 1.0 is the reference's pure FLIP update v += u(x) - u_old(x);
 * otherwise v <- a*(v + du) + (1-a)*u(x). */
int clq_grid_to_particle(clq_ctx *ctx, double flip_ratio);
/* Copy of positions into the previous-position buffer (solver.py:469-474). */
int clq_save_previous_positions(clq_ctx *ctx);
/* advect_particles (particles.cl:264-276). */
int clq_advect_particles(clq_ctx *ctx, double dt);
/* particle_domain_bc (boundary_conditions.cl:69-97, bounds from
 * boundary_conditions.py:98-101). */
int clq_particle_domain_bc(clq_ctx *ctx);
/* particle_obstacle_bc (boundary_conditions.cl:175-264). */
int clq_particle_obstacle_bc(clq_ctx *ctx);
/* The five calls above in order, as one pass over the particles. */
int clq_grid_to_particle_advect(clq_ctx *ctx, double flip_ratio, double dt);

/* ---- particle sources (Solver/Kernel/solver.py:253-306) ------------------------
 * The per-step OUTFLOW removal and INFLOW top-up of
 * Solver/Kernel/source_particles.py:42-92 on the device, replacing the host
 * round trip of every particle.  Semantics: a particle's cell is
 * floor(x / dx) (float32) clamped into the grid; OUTFLOW sources first remove
 * the particles in their cells (the rest keep their order); then each INFLOW
 * source (and each GEOMETRY source when `initial`), in order, tops every
 * open cell (mask and not solid) up to its density counting the particles
 * present, appending the new ones in C order of their cells at
 * float32(((double)cell + u) * dx) with the source's velocity.  The uniforms
 * u continue numpy's Generator.random() stream of the PCG64 state `rng`
 * ({state >> 64, state & (2^64-1), inc >> 64, inc & (2^64-1)}), three per
 * new particle as rng.random((count, 3)) draws them; *draws receives the
 * number of 64-bit draws consumed, by which the caller advances its own
 * generator.  The result equals the host seeding bit for bit.  Masks are
 * uint8[count][C] in the grid layout; the obstacle field must be set before
 * clq_set_sources. */
enum clq_source_type { CLQ_SOURCE_GEOMETRY = 0, CLQ_SOURCE_INFLOW = 1, CLQ_SOURCE_OUTFLOW = 2 };
int clq_set_sources(clq_ctx *ctx, int32_t count, const int32_t *types, const uint8_t *masks, const int32_t *densities,
                    const float *velocities);
int clq_apply_sources(clq_ctx *ctx, int32_t initial, const uint64_t *rng, int64_t *draws);

/* ---- mesh voxelisation (Solver/Kernel/voxelise_mesh.py:111-135) --------------
 * Kernel voxelise_mesh (Solver/Kernel/OpenCL/voxelise_mesh.cl:24-48) for one
 * geometry entry: `triangles` is float32[count][3][3] in world coordinates;
 * every interior cell (no index 0 or n-1) whose centre
 * (o + (index + 0.5) * dx, float32, o = (ox, oy, oz) rounded to float32)
 * sends a ray along normalize(1, 0.3713907, 0.2178311) that hits an odd
 * number of triangles (ray_hit, voxelise_mesh.cl:5-22) gets mask[cell] = 1;
 * other cells keep their value, so calls for several entries give the union.
 * `mask` is a host uint8[C] array in the grid layout, updated in place. */
int clq_voxelise_mesh(clq_ctx *ctx, const float *triangles, int64_t count, double ox, double oy, double oz,
                      uint8_t *mask);

/* ---- particle voxelisation ------------------------------------------------ */
/* Occupancy at `scale` x grid resolution (uint8, 1 where a particle's cell is
 * occupied) and, when density != NULL, smoothed density (float32, trilinear
 * splat of unit mass per particle at the scaled resolution, divided by the
 * particles-per-cell target).  Output arrays are (scale*nx, scale*ny, scale*nz). */
int clq_voxelise(clq_ctx *ctx, int32_t scale, double particles_per_cell, uint8_t *occupancy, float *density);

/* ---- timing --------------------------------------------------------------- */
/* Device time of the most recent call of each stage, in ms, indexed by
 * enum clq_stage; returns the number of stages written. */
enum clq_stage {
    CLQ_STAGE_BIN = 0,
    CLQ_STAGE_P2G = 1,
    CLQ_STAGE_PRESSURE_SOLVE = 2,
    CLQ_STAGE_G2P_ADVECT = 3,
    CLQ_STAGE_COUNT = 4
};
int clq_stage_times(clq_ctx *ctx, double *ms, int32_t capacity);

/* Accumulated device time per kernel class, for roofline reports.  While
 * enabled, HIP events bracket every launch unit of each class on the context
 * stream: one binning pass (all its kernels), one P2G gather, one full
 * red-black sweep (both colours) of the finest level, one fused particle
 * update.  Enabling resets the totals.  No counterpart in the reference
 * (its debug_timing sections time whole stages on the host,
 * Solver/Kernel/timing.py:12-29). */
enum clq_kernel_class {
    CLQ_KERNEL_BIN = 0,
    CLQ_KERNEL_P2G = 1,
    CLQ_KERNEL_RBGS_FINE = 2,
    CLQ_KERNEL_PARTICLE_UPDATE = 3,
    CLQ_KERNEL_CLASS_COUNT = 4
};
int clq_kernel_timing(clq_ctx *ctx, int32_t enable);
/* Writes total ms and launch-unit count per class; returns classes written. */
int clq_kernel_times(clq_ctx *ctx, double *total_ms, int64_t *launches, int32_t capacity);

/* ---- learned preconditioner --------------------------------------------------
 * No counterpart in the reference.  A 3-level 3D U-Net (16/32/64 channels,
 * 3x3x3 convolutions, ReLU, BF16 weights and activations with FP32
 * accumulation on the matrix cores) mapping (rhs, fluid mask, previous
 * pressure) on a pressure grid of cell size h to the pressure.  With s the RMS
 * of rhs over the interior cells with mask > 0 and k = s (N h)^2 / 1024 (N the
 * largest grid dimension), its inputs are rhs / s and prev / k and its output
 * is prev + k times the network's head (prev = 0 when absent): the head
 * predicts the change from prev.
 * Parameter vector (FP32), in order: for each of the 10 conv layers
 * (16->16, 16->16, 16->32, 32->32, 32->64, 64->64, 96->32, 32->32, 48->16,
 * 16->16) the weights W[tap][c_in][c_out] with tap = 9(dx+1) + 3(dy+1) + (dz+1)
 * then the c_out biases; then the 1x1 head's 16 weights and its bias.
 * Host arrays are float32 [nx][ny][nz] of the pressure grid. */
typedef struct clq_net clq_net;
int64_t clq_net_param_count(void);
/* A network for an (nx, ny, nz) pressure grid of cell size dx on the current
 * device (parameters are independent of the grid: copy them between nets). */
int clq_net_create(int32_t nx, int32_t ny, int32_t nz, double dx, clq_net **out);
int clq_net_destroy(clq_net *net);
/* He-normal weights from a fixed generator seeded with `seed`, zero biases. */
int clq_net_init(clq_net *net, uint64_t seed);
int clq_net_set_params(clq_net *net, const float *params, int64_t count);
int clq_net_get_params(clq_net *net, float *params, int64_t count);
/* out = network(rhs, mask, prev); prev may be NULL (zero channel). */
int clq_net_forward(clq_net *net, const float *rhs, const float *mask, const float *prev, float *out);
/* Parameter gradients of sum(grad_out * out) for the last clq_net_forward. */
int clq_net_backward(clq_net *net, const float *grad_out, float *grads, int64_t count);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on the mean over `batch` samples of
 * the mean squared error between (out - prev) / k and (target - prev) / k
 * over the interior cells with mask > 0 (k as above).  Arrays are [batch][grid];
 * prev may be NULL.  *loss receives the mean loss before the step. */
int clq_net_train_step(clq_net *net, int32_t batch, const float *rhs, const float *mask, const float *prev,
                       const float *target, double lr, double *loss);
/* Use `net` (created for the context's grid and cell size, on its device) in
 * CLQ_SOLVE_LEARNED_FCG; NULL detaches.  The context does not own the net:
 * detach before destroying it. */
int clq_attach_net(clq_ctx *ctx, clq_net *net, double stall_ratio);

#ifdef __cplusplus
}
#endif

#endif /* CLIQUID_H */
