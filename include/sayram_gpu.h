/*
 * sayram_gpu.h -- C ABI of the MI355X (gfx950) implicit PPFV time step.
 *
 * Drop-in for the reference's Solver (source/Solver.h:18-82). The reference has
 * no FFI: its only boundary is the C++ class Solver, constructed in
 * source/main.cc:51, stepped in source/main.cc:80 (Solver::update,
 * source/Solver.cc:270-290) and read in source/main.cc:74,83 (Solver::f(),
 * source/Solver.h:24). The C++ shim GpuSolver (sayram-2d_amd/host/GpuSolver.h)
 * reproduces that class surface on top of the functions below.
 *
 * Conventions
 *  - Every function returns SG_OK (0) or a negative sg_status; sg_last_error()
 *    returns the message of the last failure on the calling thread.
 *  - Cell arrays crossing this boundary use the reference's Xtensor2d layout:
 *    shape (nx, ny), element (i, j) at i*ny + j (source/common.h:28).
 *  - Boundary sides are ordered XMIN, XMAX, YMIN, YMAX (source/BCTypes.h:13).
 *    Dirichlet vertex values for XMIN/XMAX hold ny+1 values (vertex j), for
 *    YMIN/YMAX nx+1 values (vertex i), as Equation::dirichlet_vertex_value
 *    supplies them (source/Equation.h:59-65, source/Solver.cc:385-422).
 *  - The library owns all device memory; host buffers are copied on call.
 *  - A handle is not thread-safe: use one handle per host thread.
 */
#ifndef SAYRAM_GPU_H
#define SAYRAM_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SG_ABI_VERSION 4

typedef enum {
  SG_OK = 0,
  SG_ERR_ARG = -1,       /* invalid argument (shape, null pointer, bad enum) */
  SG_ERR_HIP = -2,       /* HIP runtime failure, or no GPU */
  SG_ERR_NOMEM = -3,     /* device allocation failed or HBM budget exceeded */
  SG_ERR_NOCONV = -4,    /* Krylov solve did not reach its tolerance */
  SG_ERR_BC = -5,        /* missing Dirichlet values (source/Solver.cc:392-394) */
  SG_ERR_STATE = -6      /* call not valid in the handle's current state */
} sg_status;

typedef enum { SG_DIRICHLET = 0, SG_ZEROFLUX = 1 } sg_bc_type;   /* source/BCTypes.h:16 */
typedef enum { SG_XMIN = 0, SG_XMAX = 1, SG_YMIN = 2, SG_YMAX = 3 } sg_side;
/* Precision of the implicit solve. Every mode assembles the operator, forms
 * the true residual R - M f and advances f in FP64 (source/Solver.cc:270-290).
 *  SG_FP64  FP64 BiCGSTAB preconditioned by an FP64 multigrid V-cycle.
 *  SG_MIXED FP32 BiCGSTAB with an FP32 V-cycle inside FP64 iterative
 *           refinement (the FP64 residual is corrected to rtol).
 *  SG_PC32  FP64 BiCGSTAB preconditioned by an FP32 V-cycle: the Krylov
 *           vectors, the operator and the convergence test stay FP64; only the
 *           preconditioner, most of the traffic, runs in FP32. */
typedef enum { SG_FP64 = 0, SG_MIXED = 1, SG_PC32 = 2 } sg_precision;

/* Grid2D (source/Grid2D.h:16-68): strictly increasing cell edges. */
typedef struct {
  int nx, ny;
  const double *x_edges;   /* nx+1 */
  const double *y_edges;   /* ny+1 */
} sg_grid;

/* Equation's per-cell arrays (source/Equation.h:44-50, 67-74), (nx,ny) layout.
 * inv_tau may be NULL (1/tau = 0, source/Equation.h:39-40). */
typedef struct {
  const double *G, *Dxx, *Dxy, *Dyy, *inv_tau;
} sg_coeffs;

/* Boundary types and Dirichlet vertex values; values[s] may be NULL for a
 * zero-flux side. */
typedef struct {
  int type[4];
  const double *values[4];
} sg_bc;

/* Solver options. Zero-initialised options select the defaults. */
typedef struct {
  int precision;        /* sg_precision; default SG_FP64 */
  double rtol;          /* stop when ||R - M f||_2 <= rtol ||R||_2; default 1e-14 */
  int max_iters;        /* BiCGSTAB iteration cap per step (FP64, PC32) or
                           refinement sweeps (MIXED); default 200 / 12 */
  double inner_rtol;    /* MIXED: FP32 inner solve tolerance; default 3e-3
                           (each refinement then gains at least that factor;
                           the outer FP64 residual still stops at rtol) */
  int max_inner;        /* MIXED: FP32 inner iteration cap; default 40 */
  int check_every;      /* iterations between host convergence checks; default 1
                           (the checks are pipelined one chunk ahead) */
  int smoother;         /* schedule of the finest level: bit0 pre x-lines,
                           bit1 pre y-lines, bit2 post x-lines, bit3 post
                           y-lines; 0 (default): 12, post-smoothing with
                           x-lines then y-lines */
  int coarse_smoother;  /* schedule of the levels below the finest: 4 post
                           x-lines only, 8 post y-lines only, 12 both, each
                           optionally plus 1 (an x-line pre-sweep first).
                           0: 13 when smoother is 0 too (the default: on the
                           batched workload 7% faster than 13 on every level,
                           which needs 5% fewer iterations), otherwise as
                           smoother. */
  int use_graph;        /* 0 (default) or 1: each chunk of check_every Krylov
                           iterations is captured once into a hipGraph and
                           replayed; -1: launch every kernel eagerly */
  int line_tiles[4];    /* line-solve tile shapes by line length: rows of at most
                           [0] cells use 64x16 tiles, of at most [1] cells 16x64,
                           longer rows 4x256; columns of at most [2] cells 64x16,
                           of at most [3] cells 32x32, longer columns 16x256. 0 selects the defaults 16, 160,
                           16, 768. They change the rounding of the line solves,
                           so results are reproducible for the same values. */
  int coarse_tiles[4];  /* as line_tiles, for the line solves of the coarser
                           levels (9-point operators); 0 selects the defaults
                           16, 160, 16, 768. */
  int fuse_cells;       /* the V-cycle levels of an instance from its first level
                           of at most fuse_cells cells down run in one fused
                           launch per instance (post-smoothing schedules);
                           0 selects 1024, the maximum (a level is kept in
                           LDS); -1 disables. Like line_tiles it changes the
                           rounding of those line solves. */
  int streams;          /* 2: the tile shapes of one finest-level line sweep
                           run on two HIP streams (the longest lines beside the
                           rest; a coarser level's shapes share one launch in
                           either mode), and the line factorisations beside
                           the Galerkin chain; 1:
                           one stream. 0 selects 1: with the large instances in
                           a concurrent group of their own and a step captured
                           as one graph, the second stream no longer pays
                           (its fork and join edges slow the graph down).
                           Results are identical. */
  int group_cells;      /* a batch steps its instances of at least group_cells
                           cells as a second group, on its own stream and host
                           thread, concurrently with the rest (their long line
                           solves are latency-bound, the others' work
                           bandwidth-bound). 0 selects 2^21; -1 keeps one group.
                           Results are identical. */
  int head_iters;       /* FP64 and PC32 Krylov forms: BiCGSTAB iterations the
                           step's first graph launch holds after the set-up
                           (further iterations run in pipelined chunks of
                           check_every). 0 (default): the previous step's
                           largest count, at most 16; -1: none; n > 0: n. */
  int no_clamp;         /* 0 (default): the step ends with f = max(x, 0), so
                           min f >= 0 holds exactly; the reference copies the
                           solution unclamped (source/Solver.cc:279-284), whose
                           negative entries can only be rounding-sized while M
                           is an M-matrix. Cells clamped are counted
                           (sg_step_stats.last_clamps, clamp_count): a count
                           above 0 marks a step that departs from the
                           reference. 1: f = x, as the reference. */
  int warm_start;       /* 0 or 1 (default): the solve of a step starts at
                           f + theta df, df the increment of the instance's
                           last accepted step and theta the ratio of its last
                           two increments, (df_n, df_n-1) / (df_n-1, df_n-1)
                           clamped to [0, 1] (an extrapolation in time; at f
                           until two steps have been accepted since creation or
                           since sg_set_f / sg_batch_set_f). A smoothly relaxing
                           trajectory then needs fewer iterations, and the
                           solution is the same to the solve tolerance. -1:
                           the solve starts at f, as a direct solve's result
                           does not depend on a start, so a resumed run
                           continues bit for bit. */
} sg_opts;

/* Acceptance of a step. The solve stops when ||R - M x||_2 <= rtol ||R||_2 (the
 * BiCGSTAB recurrence residual, or the FP64 residual of each refinement in
 * MIXED), or when that residual has stalled at the FP64 rounding floor, about
 * cond(M) eps, which a direct solve reaches too: no halving for 4 iterations
 * while below 1e-12 relative, or for 8 while below 1e-10 (BiCGSTAB can
 * plateau for several iterations above the floor before it converges on), or
 * in MIXED one refinement that does not halve it below 1e-10. The step is
 * then accepted only if the true
 * residual ||R - M x|| of the solution is within max(rtol, 1e-10) ||R||, every
 * x is finite and the solve did not break down or hit max_iters. A rejected
 * step leaves f and the step count unchanged and returns SG_ERR_NOCONV (the
 * reference's direct solve never produces an unconverged step). */
typedef struct {
  int64_t steps;              /* steps taken since creation / last set_f */
  int last_iters;             /* Krylov iterations of the last step (sum of
                                 inner iterations for MIXED) */
  int last_refinements;       /* MIXED: FP64 refinement sweeps of the last step */
  double last_rel_residual;   /* ||R - M x||/||R|| of the last solution x */
  double min_f;               /* min f over the grid after the last step (>= 0
                                 unless no_clamp) */
  int64_t clamp_count;        /* cumulative count of cells clamped to 0 */
  int64_t total_iters;        /* cumulative Krylov iterations */
  int status;                 /* sg_status of the last step */
  int last_nonfinite;         /* cells of the last solution that were not
                                 finite (the step was then rejected) */
  int64_t last_clamps;        /* cells clamped to 0 by the last step */
  double last_min_x;          /* min of the last solution before clamping */
} sg_step_stats;

typedef struct sg_solver sg_solver;   /* one instance (Solver equivalent) */
typedef struct sg_batch sg_batch;     /* many independent instances */

/* --- library ------------------------------------------------------------ */
int sg_abi_version(void);
const char *sg_last_error(void);
/* Number of visible GPUs (0 when none); never fails. */
int sg_device_count(void);

/* --- single instance: Solver surface ------------------------------------ */
/* Solver::Solver + Solver::init (source/Solver.cc:14-55): uploads the grid,
 * G and D (Lambda = G*D, source/Solver.cc:57-65), 1/tau, boundary data and f0
 * (Equation::init_f values). dt as Mesh::dt (source/Mesh.h:55). */
int sg_create(const sg_grid *grid, const sg_coeffs *coeffs, const sg_bc *bc, const double *f0,
              double dt, const sg_opts *opts, sg_solver **out);
/* Dirichlet vertex values of one side at the next step's time t
 * (fill_vertex_from_bcs, source/Solver.cc:385-422). */
int sg_set_bc_values(sg_solver *h, int side, const double *values);
/* New G, D and/or 1/tau after Equation::update (source/Solver.cc:287-288,
 * update_Lambda :57-65), (nx,ny) layout. A NULL array keeps its current values,
 * so a refresh moves only what changed; Lambda = G*D is formed on the device.
 * Non-finite values are rejected (SG_ERR_ARG) before anything is replaced. */
int sg_set_coeffs(sg_solver *h, const sg_coeffs *coeffs);
/* nsteps x Solver::update (source/Solver.cc:270-290); synchronous. */
int sg_step(sg_solver *h, int nsteps);
/* Solver::f() (source/Solver.h:24): f in (nx,ny) layout. */
int sg_get_f(sg_solver *h, double *out);
/* Resume: replace f and the step counter (the whole solver state). */
int sg_set_f(sg_solver *h, const double *f, int64_t istep);
/* Solver::t() = istep*dt (source/Solver.h:23). */
double sg_time(const sg_solver *h);
int sg_stats(const sg_solver *h, sg_step_stats *out);
void sg_destroy(sg_solver *h);

/* --- batched: many instances of different sizes per step ---------------- */
/* hbm_budget_bytes = 0 means no budget beyond the device's free memory. */
int sg_batch_create(const sg_opts *opts, size_t hbm_budget_bytes, sg_batch **out);
/* Device bytes one instance of this shape would occupy (admission control). */
size_t sg_instance_bytes(int nx, int ny, int precision);
/* Adds an instance; returns its slot id in *slot. SG_ERR_NOMEM when the
 * budget would be exceeded (nothing is allocated then). */
int sg_batch_add(sg_batch *b, const sg_grid *grid, const sg_coeffs *coeffs, const sg_bc *bc,
                 const double *f0, double dt, int *slot);
int sg_batch_remove(sg_batch *b, int slot);
/* One implicit step for every resident instance, nsteps times. */
int sg_batch_step(sg_batch *b, int nsteps);
/* As sg_batch_step, but instance k only advances while its step count is
 * below target_steps[k] (indexed by slot; entries of free slots ignored). */
int sg_batch_step_until(sg_batch *b, const int64_t *target_steps, int nslots, int nsteps);
int sg_batch_get_f(sg_batch *b, int slot, double *out);
/* sg_batch_get_f that allocates nothing: no hipMalloc, hipFree or other
 * allocating call, for callers that cannot allocate device memory (code
 * between graph captures, tight serving loops). It has no counterpart in the
 * reference. The layout transpose is staged in a static device array of
 * 4 Mi doubles (32 MiB) per GPU, held until the process ends; instances of
 * more than 4 Mi cells are rejected with SG_ERR_ARG. Every call on one GPU
 * shares that array, so calls that overlap on one GPU (from several host
 * threads) race; calls from one thread are ordered. */
int sg_batch_get_f_static(sg_batch *b, int slot, double *out);
int sg_batch_set_f(sg_batch *b, int slot, const double *f, int64_t istep);
int sg_batch_set_bc_values(sg_batch *b, int slot, int side, const double *values);
/* sg_set_coeffs for one instance of a batch. */
int sg_batch_set_coeffs(sg_batch *b, int slot, const sg_coeffs *coeffs);
int sg_batch_stats(const sg_batch *b, int slot, sg_step_stats *out);
int sg_batch_size(const sg_batch *b);
size_t sg_batch_bytes_in_use(const sg_batch *b);
void sg_batch_destroy(sg_batch *b);

/* --- instrumentation (benchmarks and tests) ----------------------------- */
/* Assembly only: writes the 5 FP64 bands (d|w|e|s|n, K = j*nx+i order) and R
 * of the next step into host buffers, for parity checks of the assembly. */
int sg_debug_assemble(sg_solver *h, double *bands5, double *R);
/* Kernel timing with HIP events around every dispatch of a kernel family, on
 * the stream it runs on (the steps then launch eagerly): kernel_id selects the
 * family (see sg_kernel_name), sg_kernel_count() selects every family, -1
 * disables. Selecting resets the totals. */
int sg_batch_time_kernel(sg_batch *b, int kernel_id);
/* Totals over the selected families: time, dispatches and algorithmic bytes
 * (each array a kernel reads or writes counted once per launch). */
int sg_batch_kernel_time(sg_batch *b, double *total_ms, int64_t *launches, double *bytes);
/* Totals of one family (bytes 0 where no byte model applies). */
int sg_batch_kernel_time_id(sg_batch *b, int kernel_id, double *total_ms, int64_t *launches,
                            double *bytes);
/* Mean HIP-event interval around a one-workgroup kernel on the batch's stream:
 * the bracket overhead included in every timed dispatch. */
int sg_batch_event_overhead(sg_batch *b, double *ms);
const char *sg_kernel_name(int kernel_id);
int sg_kernel_count(void);
int sg_solver_batch(sg_solver *h, sg_batch **out);

#ifdef __cplusplus
}
#endif
#endif
