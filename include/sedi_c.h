/*
 * sedi_c.h -- C ABI of libsedi_amd, the MI355X (gfx950) implementation of
 * intel/sedi's stochastic diag(A^-1) estimator path.
 *
 * Every entry point replaces one reference interface, cited as file:line of
 * intel/sedi.  Arrays are dense n1*n2*n3 doubles, i1 fastest (Slice3d.hpp:30-32),
 * without ghost cells; cells outside the grid count as 0, which is what the
 * reference's zero ghosts give (Cg.hpp:57-62).  An array argument may be a
 * host pointer or a HIP device pointer (detected with hipPointerGetAttributes);
 * host arrays are staged through library-owned device memory.
 *
 * Return values: >= 0 is success (for a solve, the reference's result code
 * 0..3); < 0 is a runtime error (SEDI_ERR_*) and sedi_last_error() describes
 * it.  No C++ exception crosses this boundary.
 *
 * Threading: a sedi_solver / sedi_estimator handle belongs to one host thread
 * at a time; different handles may be used from different threads.
 */
#ifndef SEDI_C_H
#define SEDI_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- codes ---------------------------------------------------------------- */
#define SEDI_OK               0
#define SEDI_ERR_ARG        (-1)   /* bad shape, pointer or option                */
#define SEDI_ERR_HIP        (-2)   /* HIP runtime error                            */
#define SEDI_ERR_NOMEM      (-3)   /* device allocation failed / over the budget   */
#define SEDI_ERR_STATE      (-4)   /* handle used out of order                     */
#define SEDI_ERR_UNSUPPORTED (-5)  /* option not available in this build           */

/* Solver result codes: Cg.hpp:31-34, PipelinedCg.hpp:31-34, BiCgStab.hpp:32-37. */
#define SEDI_SOLUTION_FOUND   0
#define SEDI_REACHED_MAXITER  1
#define SEDI_BREAKDOWN_RHO    2   /* BiCGStab only */
#define SEDI_BREAKDOWN_OMEGA  3   /* BiCGStab only */

/* Solver kinds: the SOLVER numbering of stochastic-estimator.cc:134-142. */
#define SEDI_BICGSTAB 1
#define SEDI_CG       2
#define SEDI_PCG      3

/* Arithmetic of a solve.  FP64: the reference's (real_t = double,
 * stochastic-estimator.cc:29).  FP32_MIXED: FP32 vector storage (CG computes
 * its stencil and updates in FP32, pipelined CG and BiCGStab in FP64
 * registers), FP64 dots, scalars, residual replacement and diagonal
 * accumulator. */
#define SEDI_FP64        0
#define SEDI_FP32_MIXED  1

/* Global grid: DomainDecomposition's ng1, ng2, ng3 (DomainDecomposition.hpp:8-9). */
typedef struct sedi_grid { int64_t n1, n2, n3; } sedi_grid;
/* Stencil<double>(a, b, c) (Stencil.hpp:18-19); the estimator uses (6, 1, 1)
 * (stochastic-estimator.cc:74). */
typedef struct sedi_stencil { double a, b, c; } sedi_stencil;

typedef struct sedi_solver sedi_solver;
typedef struct sedi_estimator sedi_estimator;

const char *sedi_last_error(void);
int sedi_version(void);
/* Selects the HIP device for this host thread (hipSetDevice) and loads the
 * kernels.  Optional: every entry point initialises lazily. */
int sedi_init(int device);

/* ---- operator ------------------------------------------------------------- */
/* y = A x on the whole grid.  Replaces Stencil<double>::apply
 * (Stencil.hpp:21-36) together with DistrAsyncStencil's
 * initiate/compute_local/finalize (DistrAsyncStencil.hpp:52-86) and the halo
 * exchange of CommHalo (CommHalo.hpp:84-181): on one GPU the grid is one
 * domain and its boundary is predicated in the kernel.  x and y may be host
 * or device arrays of n1*n2*n3 doubles (i1 fastest); any 8-byte alignment
 * works (16-byte aligned device arrays with n1 even take the faster march
 * with two cells per lane).  y is bitwise the reference's. */
int sedi_stencil_apply(const sedi_grid *grid, const sedi_stencil *stencil,
                       const double *x, double *y);

/* ---- solvers -------------------------------------------------------------- */
/* Replaces the solver constructors Cg(dd, nghost, mode, timers) (Cg.hpp:48),
 * PipelinedCg(...) (PipelinedCg.hpp:48) and BiCgStab(...) (BiCgStab.hpp:51):
 * allocates the solver's device temporaries once for reuse by every solve.
 * mode: the -f bits (bit 0 synchronous stencil, bit 1 synchronous dots,
 * Cg.hpp:36-40); accepted for compatibility, they change nothing on one GPU. */
int sedi_solver_create(const sedi_grid *grid, const sedi_stencil *stencil,
                       int kind, int mode, sedi_solver **out);

/* Replaces solver_t::solve(A, comm_halo, x, b, iM, maxiter, tol, &final_err,
 * &final_iter) (Cg.hpp:70-82, PipelinedCg.hpp:72-84, BiCgStab.hpp:77-89).
 * x: initial guess in, solution out.  inv_diag: the Jacobi vector iM, or NULL
 * for the constant 1/a that the estimator uses (stochastic-estimator.cc:103).
 * Returns the result code 0..3 and always sets *final_err and *final_iter. */
int sedi_solve(sedi_solver *solver, const double *b, double *x,
               const double *inv_diag, int maxiter, double tol,
               double *final_err, int *final_iter);

/* No counterpart in the reference: an allocation-free sedi_solve for callers
 * that cannot allocate device memory (code under HIP graph capture set-up,
 * tight serving loops, harnesses that forbid hipMalloc).  b, x and inv_diag
 * (or NULL for 1/a) must be device arrays; every temporary -- the solver's
 * vectors, its control block and err history -- lives in one fixed device
 * array of SEDI_NOALLOC_SCRATCH_BYTES that the library holds on each device
 * it runs on until the process ends.  Same kernels, launch plan, result
 * codes and final_err/final_iter as sedi_solver_create + sedi_solve on the
 * same grid (so the same x, bitwise); no err history is kept.  A solve
 * needing more (vectors: 5 for CG, 11 for pipelined CG, 7 for BiCGStab, each
 * 8*n1*n2*n3 bytes, plus 8*maxiter bytes and ~50 KB) returns SEDI_ERR_ARG.
 * Launches on the null stream and returns when the solve is done.  Calls on
 * one device share the array: calls that overlap on one device race. */
#define SEDI_NOALLOC_SCRATCH_BYTES (64u << 20)
int sedi_solve_noalloc(const sedi_grid *grid, const sedi_stencil *stencil, int kind,
                       const double *b, double *x, const double *inv_diag, int maxiter,
                       double tol, double *final_err, int *final_iter);

/* The err values that the reference prints after each iteration
 * ("Cg: after iter=%d, err=%.10e", Cg.hpp:192; PipelinedCg.hpp:228;
 * BiCgStab.hpp:228) for the last solve.  Copies min(cap, count) values
 * (none if out is NULL) and returns count. */
int sedi_solver_err_history(sedi_solver *solver, double *out, int cap);

void sedi_solver_destroy(sedi_solver *solver);

/* ---- probes --------------------------------------------------------------- */
/* Replaces DistrRngStd::fill_plusminusone (DistrRngStd.hpp:22-42):
 * v[c] = +1 if the c-th std::uniform_real_distribution<double>(0,1) draw from
 * std::mt19937(seed), after skipping first_draw 32-bit outputs, is > 0.5,
 * else -1.  Each cell takes two outputs.  The reference re-seeds from
 * std::random_device on each call; this ABI takes the seed explicitly. */
int sedi_rademacher_fill(uint32_t seed, uint64_t first_draw, int64_t ncells,
                         double *v);

/* Host-only utility, no reference counterpart: the 624-word state of
 * std::mt19937(seed) after `draws` outputs, oldest word first, computed by
 * jump-ahead (x^draws mod the generator's characteristic polynomial) instead
 * of generating the draws.  The batched estimator uses it to start every
 * probe's stream segment in parallel. */
int sedi_mt_state_after(uint32_t seed, uint64_t draws, uint32_t *state624);

/* ---- estimator ------------------------------------------------------------ */
/* The sample loop of stochastic-estimator.cc:57-232 (CLI -g -s -e -m -i -f)
 * with seeded probes: sample k solves A x = v_k from x = 0 with v_k cells
 * drawn from outputs [2Nk, 2N(k+1)) of std::mt19937(seed), accumulates
 * t += x.*v, q += v.*v (:195,198) and returns D = t./q (:216). */
typedef struct sedi_estimate_args {
    sedi_grid grid;
    sedi_stencil stencil;   /* (6,1,1) in the reference */
    int kind;               /* SEDI_BICGSTAB / SEDI_CG / SEDI_PCG        (-S) */
    int mode;               /* -f bits, accepted and ignored              (-f) */
    int samples;            /* number of probes                           (-s) */
    int maxiter;            /* solver iteration cap                        (-m) */
    double tol;             /* solver tolerance                            (-e) */
    int ignore_errors;      /* keep going after a failed solve             (-i) */
    uint32_t seed;          /* mt19937 seed of the probe stream                 */
    int precision;          /* SEDI_FP64 or SEDI_FP32_MIXED                     */
    int batch_k;            /* probes solved together (1 = plain, one at a time);
                               above 1 (or with SEDI_FP32_MIXED): the batched
                               solvers, probes in blocks of 4 (padded with
                               zero right-hand sides); 0 = the batch size the
                               autotuner measured best for this solver,
                               precision and grid (tuning.txt beside the
                               library), within 80% of free device memory  */
    int rep_every;          /* SEDI_FP32_MIXED: CG's FP64 residual
                               replacement period (0 = 50); see the batched
                               solvers below for the other triggers           */
} sedi_estimate_args;

/* Allocates everything the run needs (excluded from the timed region, as the
 * reference's TIMER_GLOBAL excludes setup: stochastic-estimator.cc:151-153). */
int sedi_estimator_create(const sedi_estimate_args *args, sedi_estimator **out);

/* Runs the sample loop.  D (n1*n2*n3, may be NULL), iters/errs/codes (per
 * sample, may be NULL).  *seconds (may be NULL) is the wall time of the loop,
 * the reference's TIMER_GLOBAL scope (stochastic-estimator.cc:152-153,211-212).
 * Returns 0, or 100 + code of the first failed solve when ignore_errors is 0
 * (the MPI_Abort code of stochastic-estimator.cc:181-184); iters/errs/codes
 * then hold the samples done so far. */
int sedi_estimator_run(sedi_estimator *est, double *D, int *iters, double *errs,
                       int *codes, double *seconds);

/* Optional line sink for the reference's stdout lines, called from
 * sedi_estimator_run: "Solving for sample %d/%d..." (stochastic-estimator.cc:160),
 * "<Solver>: after iter=%d, err=%.10e" after each solve (Cg.hpp:192,
 * PipelinedCg.hpp:228, BiCgStab.hpp:228) and "Incomplete convergence in
 * solver, errcode=%d" (stochastic-estimator.cc:182).  Lines have no '\n'. */
typedef void (*sedi_log_fn)(void *user, const char *line);
int sedi_estimator_set_log(sedi_estimator *est, sedi_log_fn fn, void *user);

/* Solver validation mode (stochastic-estimator.cc:162-171, ValidateSolver::
 * fill_with_ref_rhs, ValidateSolver.hpp:93-96): every sample solves with this
 * right-hand side (n cells, i1 fastest, host or device; copied) instead of a
 * probe.  NULL returns to the probes. */
int sedi_estimator_set_rhs(sedi_estimator *est, const double *v);

/* Per-sample observer for the validation checks (ValidateSolver::
 * compare_to_ref_solution after each solve, stochastic-estimator.cc:189-191;
 * ValidateEstimator::compare_to_ref_inv_diag on D = t./q after each
 * accumulation, :200-205).  Called from sedi_estimator_run after sample k with
 * host copies (valid during the call) of the requested arrays, NULL for the
 * others; the copies cost a device synchronisation per sample. */
#define SEDI_HOOK_X 1 /* the sample's solution x_k */
#define SEDI_HOOK_V 2 /* its right-hand side v_k   */
#define SEDI_HOOK_D 4 /* D = t./q so far           */
typedef void (*sedi_sample_fn)(void *user, int k, const double *x, const double *v,
                               const double *D);
int sedi_estimator_set_sample_hook(sedi_estimator *est, sedi_sample_fn fn, int want, void *user);

void sedi_estimator_destroy(sedi_estimator *est);

/* create + run + destroy. */
int sedi_estimate(const sedi_estimate_args *args, double *D, int *iters,
                  double *errs, int *codes, double *seconds);

/* ---- batched multi-RHS solvers (no reference counterpart) ------------------ */
/* The reference solves one probe at a time (stochastic-estimator.cc:170-199);
 * a batch solves up to K right-hand sides together with the recurrences of
 * Cg.hpp:102-208, PipelinedCg.hpp:113-243 or BiCgStab.hpp:117-264 (kind)
 * kept per right-hand side: one stencil sweep per pass serves every
 * right-hand side, with per-RHS scalars, convergence and masks.
 *   precision SEDI_FP64: FP64 vectors, as the plain solver;
 *   SEDI_FP32_MIXED: FP32 vectors, FP64 dot products, scalars and solution
 *   accumulator (CG: FP32 stencil and updates; pipelined CG and BiCGStab:
 *   FP64 arithmetic on the FP32-stored vectors), with an FP64 residual
 *   replacement r = b - A x before a right-hand side is declared converged
 *   (its returned err is then the true residual's) and: CG every rep_every
 *   iterations (0 = 50) and at each tenfold decrease; pipelined CG the same
 *   but at least every 10 iterations (all of r, u, w, s, q, z are replaced);
 *   BiCGStab at each tenfold decrease (its solution accumulated in FP64),
 *   restarting r^ and the directions there.
 * Results of a right-hand side do not depend on the others in its batch or on
 * K.  K is rounded up to a multiple of 4. */
typedef struct sedi_batch sedi_batch;
int sedi_batch_create(const sedi_grid *grid, const sedi_stencil *stencil, int kind, int K,
                      int precision, int rep_every, sedi_batch **out);
/* Solves A x_k = b_k from x_k = 0 for k < nrhs <= K; b and x are [nrhs][n]
 * FP64 (host or device).  iters/errs/codes (may be NULL) per right-hand side,
 * with the meaning of sedi_solve's final_iter/final_err/return code. */
int sedi_batch_solve(sedi_batch *batch, const double *b, int nrhs, double *x, int maxiter,
                     double tol, int *iters, double *errs, int *codes);
/* The per-iteration error of right-hand side k in the last solve; returns the
 * number of entries (written up to cap; none if out is NULL), or a negative
 * error code. */
int sedi_batch_err_history(sedi_batch *batch, int k, double *out, int cap);
void sedi_batch_destroy(sedi_batch *batch);
/* Lane compaction: a block of 4 right-hand sides runs until its slowest one
 * converges, so between iterations the still-running right-hand sides of the
 * emptiest block trade lanes with finished ones of other blocks (all at the
 * same iteration) and the emptied block stops; every right-hand side's
 * operations, and so its results, stay the same.  On by default for batch
 * solves and the batched estimator (the server's blocks start at different
 * times and keep their lanes).  sedi_batch_lane_moves gives the lanes moved
 * in the batch's last solve.  No reference counterpart. */
int sedi_batch_compact_enable(int on);
int sedi_batch_lane_moves(sedi_batch *batch, int64_t *moves);

/* Probes k0 .. k0+count-1 of the seeded stream (the estimator's probes), as
 * [count][ncells] doubles (host or device), each generated from its own
 * jump-ahead start: equal to sedi_rademacher_fill(seed, 2*ncells*k, ...). */
int sedi_probe_batch(uint32_t seed, int64_t ncells, int k0, int count, double *v);

/* ---- serving (no reference counterpart) ----------------------------------- */
/* A server accepts estimation jobs at any time and runs them concurrently on
 * the GPU with continuous batching: each (grid, solver, stencil) has a pool of
 * right-hand-side slots in blocks of 4 (the batched solvers above), a block
 * of a new job starts as soon as a slot is free while the pool's other blocks
 * are mid-solve, and pools of different grids run on their own streams.  Jobs
 * are admitted in arrival order against a device memory budget.  A job's D
 * is accumulated in probe order and equals sedi_estimate's batched result for
 * the same job with ignore_errors = 1 bitwise, whatever else is in flight:
 * like the reference's -i, every probe is accumulated whether or not its
 * solve converged, and the per-probe codes tell which did not.  Thread-safe
 * submit and wait; one engine thread owns the device. */
typedef struct sedi_server sedi_server;
typedef struct sedi_server_config {
    double hbm_budget_bytes; /* device memory for pools and jobs             */
    int precision;           /* SEDI_FP32_MIXED or SEDI_FP64                 */
    int rep_every;           /* mixed: replacement period (0 = 50)           */
    int chunk;               /* iterations per pool step (0 = 8)             */
    int max_slots;           /* right-hand sides per pool at most (0 = 64)   */
    int max_iter;            /* largest maxiter a job may ask for (0 = 10000) */
} sedi_server_config;
typedef struct sedi_job {
    sedi_grid grid;
    sedi_stencil stencil;
    int kind;       /* SEDI_CG / SEDI_PCG / SEDI_BICGSTAB                    */
    int samples;    /* probes                                                */
    double tol;
    int maxiter;
    uint32_t seed;  /* the job's probe stream (the estimator's convention)   */
} sedi_job;
int sedi_server_create(const sedi_server_config *cfg, sedi_server **out);
/* Returns the job id (>= 0) or a negative error code. */
int64_t sedi_server_submit(sedi_server *srv, const sedi_job *job);
/* Waits for a job (timeout_s < 0: no limit).  Returns 0 when it finished
 * (then *status is 0, or a negative error code if the job failed, and its
 * result is released; a probe that reached maxiter or broke down does not
 * fail the job, its code does not change *status), 1 on timeout, < 0 on a
 * bad call.  D (n cells),
 * iters/errs/codes (per probe) and times4 = {submitted, admitted, first
 * block started, done} in seconds since the server started may be NULL. */
int sedi_server_wait(sedi_server *srv, int64_t job, double timeout_s, double *D, int *iters,
                     double *errs, int *codes, double *times4, int *status);
typedef struct sedi_server_info {
    int64_t submitted, done, probes;
    double bytes_in_use, peak_bytes;
    int pools;
} sedi_server_info;
int sedi_server_stats(sedi_server *srv, sedi_server_info *out);
/* Finishes every submitted job, then stops. */
void sedi_server_destroy(sedi_server *srv);

/* ---- validation archives (HDF5) ------------------------------------------- */
/* The reference's validation archives (ValidateSolver.hpp:30-85: datasets
 * "x_k" and "v_k"; ValidateEstimator.hpp:27-62: "diag_inv") hold FP64 arrays
 * of shape [n1][n2][n3] with i3 fastest (Slice3d::pack_into, Slice3d.hpp:42-51).
 * These calls take and return i1-fastest arrays (host memory) and transpose.
 * libhdf5 is loaded at run time; without it they return SEDI_ERR_UNSUPPORTED.
 * They do not touch the GPU. */
int sedi_h5_available(void);
int sedi_h5_read(const char *path, const char *dataset, const sedi_grid *grid, double *out);
/* Creates (truncates) path with `count` datasets; the reference never writes
 * archives, this lets its -V / -E modes read the GPU's results. */
int sedi_h5_write(const char *path, int count, const char *const *datasets,
                  const double *const *arrays, const sedi_grid *grid);

/* ---- measurement ---------------------------------------------------------- */
/* Kernel timing with HIP events recorded on the launching stream around every
 * launch of the named kernel ("*" = all kernels, NULL = off).  The reference's
 * Timers (Timers.hpp:14-55) time phases with MPI_Wtime; this is the device
 * counterpart. */
int sedi_profile_enable(const char *kernel_name);
/* Fetches (launches, total milliseconds) of a kernel timed since the last
 * sedi_profile_reset().  Returns 0, or SEDI_ERR_ARG if it was never timed. */
int sedi_profile_get(const char *kernel_name, int64_t *launches, double *total_ms);
int sedi_profile_reset(void);

/* The solvers launch their iterations in chunks; from the second chunk of a
 * solve on, a chunk is captured once into a HIP graph and replayed (the
 * kernels read every per-iteration value from device memory, so one graph
 * serves every chunk).  This turns it off (0) or on (1, the default); kernels
 * timed with sedi_profile_enable are never captured.  No reference
 * counterpart (the reference's iterations are host loops). */
int sedi_graphs_enable(int on);

/* The launch parameters an offline autotuner measured (python -m
 * sedi_amd.autotune writes tuning.txt beside the library; SEDI_TUNING_FILE
 * names another table, empty for none): the value stored for `key` and the
 * grid, or `fallback`.  Keys "zc:<kernel>" (planes per tile of a solver
 * kernel), "sched:<kernel>" (1 = tiles in launch order, 0 = XCD-contiguous
 * runs, 2 = one workgroup per tile, 3/4 = launch order dealt to the XCDs in
 * chunks of 4/8 tile rows), "pf:stencil_apply_kernel" (planes of loads in
 * flight), "x2:stencil_apply_kernel" (1 = the march with two cells per lane,
 * the default where n1 is even and x, y are 16-byte aligned),
 * "wpc:<kernel>" (workgroups per CU; 0 = as occupancy allows),
 * "k:<solver>:<precision>" (probes per batch; nearest grid by cell
 * count), "ty:batch_upd_kernel" and "ty:batch_dir_kernel" (rows per tile,
 * 8 or 16, of the FP32 batched CG update and direction passes).  No reference counterpart (its tiling is Platform.hpp's
 * compile-time constants). */
int sedi_tuning_lookup(const char *key, const sedi_grid *grid, int fallback);

#ifdef __cplusplus
}
#endif
#endif /* SEDI_C_H */
