/*
 * sovlp.h -- C ABI of libsovlp.so, the MI355X (gfx950) PDLP engine.
 *
 * Conventions
 *   - Every function returns 0 on success and a nonzero code otherwise;
 *     sovlp_last_error() returns the message of the last failure on the
 *     calling thread.  Codes: 1 invalid argument, 2 HIP runtime error,
 *     4 unsupported, 5 self-test mismatch.
 *   - Device pointers are plain pointers owned by the caller (PyTorch tensors
 *     in the Python host layer).  The library allocates no device memory except
 *     inside sovlp_selftest; engines carve the caller's workspace.
 *   - `stream` is a hipStream_t (NULL = default stream).  Launches are
 *     asynchronous unless a function says it synchronises.
 *   - Row/column pointer arrays of the engines are int32 offsets (an LP or a
 *     batch member holds fewer than 2^31 nonzeros); the Backend surface keeps
 *     the reference's int64 pointers.
 *
 * The reference is SOVOPT (src/sovopt); each entry names what it replaces.
 */
#ifndef SOVLP_H
#define SOVLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *sovlp_stream_t;

/* ---- library --------------------------------------------------------- */

/* Message of the last error on this thread ("" if none). */
const char *sovlp_last_error(void);

/* ABI version (major*100 + minor). */
int sovlp_version(void);

/* Known-answer test [[1,0,2],[0,3,0]] . [1,2,3] == [7,6] through the plain
 * and the tiled SpMV kernels, FP64 and FP32; synchronises.
 * Replaces gpu_selftest, src/sovopt/core/backend.py:499-535. */
int sovlp_selftest(void);

/* Device wall-clock rate in Hz (for converting engine tick counts). */
int sovlp_clock_rate(int device, double *hz);

/* ---- Backend surface: src/sovopt/core/backend.py:317-493 ------------- */

/* y = A x, A in CSR (int64 pointers, int32 indices).  lanes_per_row in
 * {1,2,4,8,16,32,64} lanes reduce one row (the reference uses one 32-lane
 * warp per row, backend.py:118-152; 64 = one wavefront per row).
 * Replaces Backend.spmv, backend.py:373-384. */
int sovlp_spmv_csr_f64(int rows, const int64_t *rp, const int32_t *ri, const double *rx,
                       const double *x, double *y, int lanes_per_row, sovlp_stream_t stream);
int sovlp_spmv_csr_f32(int rows, const int64_t *rp, const int32_t *ri, const float *rx,
                       const float *x, float *y, int lanes_per_row, sovlp_stream_t stream);

/* out = A X with X dense (cols, k) and out (rows, k), row-major with the
 * batch index fastest (A in CSR, int64 pointers): one lane per (row, column)
 * walking the row's nonzeros in order, products and sums unfused, so the
 * results are the reference's CPU _csr_spmm (src/sovopt/core/sparse.py:72-83)
 * bit for bit.  Replaces Backend.spmm, backend.py:386-400 (kernels
 * csr_spmm_warp_*, backend.py:158-200). */
int sovlp_spmm_csr_f64(int rows, const int64_t *rp, const int32_t *ri, const double *rx,
                       const double *X, double *out, int k, sovlp_stream_t stream);
int sovlp_spmm_csr_f32(int rows, const int64_t *rp, const int32_t *ri, const float *rx,
                       const float *X, float *out, int k, sovlp_stream_t stream);

/* out = b - A x to about double-double accuracy (A in CSR, int64 pointers):
 * each row's dot product with Dekker's two_product and Knuth's two_sum in
 * CSR order, bit for bit the reference's compensated_residual
 * (src/sovopt/numerics/refine.py:48-84, 103-110).  out may not alias x or b.
 * Rows of more than 128 nonzeros are listed in `workspace`
 * (sovlp_residual_workspace_bytes) and summed by one workgroup each. */
int sovlp_residual_workspace_bytes(int rows, int64_t *bytes);
int sovlp_residual_f64(int rows, const int64_t *rp, const int32_t *ri, const double *rx,
                       const double *x, const double *b, double *out, void *workspace,
                       int64_t workspace_bytes, sovlp_stream_t stream);

/* sovlp_residual_f64 without a workspace argument and without allocating:
 * no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Its scratch is a device array of the library holding
 * the workspace of up to sovlp_residual_noalloc_max_rows() rows (1048576, 4
 * MiB), one copy per GPU, kept until the process ends; larger inputs are
 * refused (code 1).  Every call on a GPU shares that GPU's copy, so calls
 * that overlap on one GPU race; calls ordered on one stream do not.  Same
 * results as sovlp_residual_f64, bit for bit. */
int sovlp_residual_noalloc_max_rows(void);
int sovlp_residual_f64_noalloc(int rows, const int64_t *rp, const int32_t *ri, const double *rx,
                               const double *x, const double *b, double *out,
                               sovlp_stream_t stream);

/* The certificate of a point x and dual vector y (textbook signs, d = c - A^T y)
 * on an LP in the model's own sense:
 *   bound          certified_bound (src/sovopt/mip/safebound.py:252-307):
 *                  a lower bound on min c'x (upper bound on max c'x) valid
 *                  for any y, offset included; -inf (+inf) when vacuous.
 *                  Duals pointing at infinite row bounds are taken as 0;
 *                  reduced costs pointing at infinite column bounds whose
 *                  largest magnitude is <= dual_tol (1 + ||c||_inf) are
 *                  absorbed (taken as 0), larger ones make the bound vacuous.
 *                  Strict: exact-to-a-few-ulps sums, less 4 eps sum |terms|
 *                  (safebound.py:125-143).
 *   perturbation   that largest magnitude (0 when no reduced cost points at
 *                  an infinite bound).
 *   objective      c'x + offset, compensated (bench/verify.py:227).
 *   col_violation  max(lo - x, x - hi, 0) (bench/verify.py:208-211).
 *   row_violation  max(rl - Ax, Ax - ru, 0) with A x compensated
 *                  (bench/verify.py:214-218).
 *   vacuous        1 when the bound is infinite. */
typedef struct {
  double bound, perturbation, objective, col_violation, row_violation;
  int32_t vacuous, reserved;
} sovlp_certificate;

int sovlp_certify_workspace_bytes(int m, int n, int64_t *bytes);

/* Batched node relaxation (src/sovopt/mip/bnr.py:176-301 BNREngine.solve_batch
 * and mip/safebound.py:194-235 safe_dual_bound_batch): K LPs sharing A, c and
 * the row bounds, each with its own column bounds L, U, advanced together by
 * `iters` PDHG iterations with a per-node adaptive step (growth 1 - t^-g,
 * damping 1 + t^-d), then bounded: bounds[k] = sum_i y_ik (rl_i | ru_i) +
 * sum_j d_jk (L_jk | U_jk), d = c - A^T Y, -inf when a term points at an
 * infinite bound.  Every (length, K) block is row-major, node index fastest.
 * X, Y hold the start (X is clamped into each box) and receive the iterates;
 * eta, omega (K) the step sizes (omega is not changed); empty[k] = 1 when
 * some L_jk > U_jk + 1e-12.  1 <= K <= 256.  Everything device memory.
 * dual_sign: 1 bounds with Y as the iteration leaves it, as the reference
 * does (its PDHG dual is the negative of the textbook dual, so the bound is
 * valid but weak, vacuous whenever a one-sided row's multiplier is nonzero);
 * -1 bounds with -Y, the textbook dual, which tightens to the optimum as
 * the iterates converge. */
typedef struct {
  int32_t m, n, K, iters;
  const int64_t *rp, *cp;
  const int32_t *ri, *ci;
  const double *rx, *cx, *c, *rl, *ru;
  const double *L, *U;
  double *X, *Y, *eta, *omega;
  double step_growth, step_damp;
  double *bounds;
  int32_t *empty;
  double dual_sign;
} sovlp_bnr_batch;

int sovlp_bnr_workspace_bytes(int m, int n, int K, int64_t *bytes);
int sovlp_bnr_solve_batch(const sovlp_bnr_batch *b, void *workspace, int64_t workspace_bytes,
                          sovlp_stream_t stream);
/* A: CSR (rp m+1, ri, rx) and CSC (cp n+1, ci, cx), int64 pointers; every
 * array a device pointer.  `d_out` (n, optional) receives c - A^T y in the
 * minimisation sense (before absorption), `act_out` (m, optional) the
 * compensated row activities.  Synchronises `stream`; `out` is host memory. */
int sovlp_certify_f64(int m, int n, const int64_t *rp, const int32_t *ri, const double *rx,
                      const int64_t *cp, const int32_t *ci, const double *cx, const double *c,
                      const double *rl, const double *ru, const double *lo, const double *hi,
                      double offset, int maximise, const double *x, const double *y,
                      double dual_tol, void *workspace, int64_t workspace_bytes, double *d_out,
                      double *act_out, sovlp_certificate *out, sovlp_stream_t stream);

/* xnew = clamp(x - tau (c + aty), lo, hi).  Replaces Backend.primal_step,
 * backend.py:404-415 (kernel 204-215). */
int sovlp_primal_step_f64(double *xnew, const double *x, const double *c, const double *aty,
                          const double *lo, const double *hi, double tau, int n,
                          sovlp_stream_t stream);
/* ynew = w - sigma clamp(w/sigma, rl, ru), w = y + sigma ax.  Replaces
 * Backend.dual_step, backend.py:417-434 (kernel 219-231). */
int sovlp_dual_step_f64(double *ynew, const double *y, const double *ax, const double *rl,
                        const double *ru, double sigma, int m, sovlp_stream_t stream);
/* In place x <- clamp(x - tau (c + aty)), dx = xnew - xold.  Replaces
 * Backend.pdhg_primal, backend.py:441-452 (kernel 241-255). */
int sovlp_pdhg_primal_f64(double *x, double *dx, const double *c, const double *aty,
                          const double *lo, const double *hi, double tau, int n,
                          sovlp_stream_t stream);
/* ax += adx; y <- prox(y + sigma (ax + adx)); dy = ynew - yold.  Replaces
 * Backend.pdhg_dual, backend.py:454-467 (kernel 260-279). */
int sovlp_pdhg_dual_f64(double *y, double *dy, double *ax, const double *adx, const double *rl,
                        const double *ru, double sigma, int m, sovlp_stream_t stream);
/* acc += inc; sum += eta v.  Replaces Backend.pdhg_accum, backend.py:469-479
 * (kernel 282-291). */
int sovlp_pdhg_accum_f64(double *acc, const double *inc, double *sum, const double *v,
                         double eta, int n, sovlp_stream_t stream);

/* FP64 reductions into a device scalar, fixed summation order (bitwise
 * repeatable).  Replace Backend.dot / nrm2 / nrm_inf, backend.py:483-490. */
int sovlp_dot_f64(const double *a, const double *b, int64_t n, double *out, sovlp_stream_t stream);
int sovlp_nrm2_f64(const double *a, int64_t n, double *out, sovlp_stream_t stream);
int sovlp_nrminf_f64(const double *a, int64_t n, double *out, sovlp_stream_t stream);
/* FP32 vectors reduced in FP64. */
int sovlp_dot_f32(const float *a, const float *b, int64_t n, double *out, sovlp_stream_t stream);
int sovlp_nrm2_f32(const float *a, int64_t n, double *out, sovlp_stream_t stream);

/* ---- tiled (CSR-Adaptive) SpMV: the engines' sparse product ---------- */

/* Host function.  Cuts rows into tiles of consecutive rows with at most
 * tile_nnz nonzeros and at most tile_rows rows; a row longer than tile_nnz is
 * a tile of its own.  With block > 0 (the lanes of the workgroup that will run
 * the plan) a tile also ends before a row that would leave one lane more than
 * 64 serial additions of the row reductions.  tiles_out needs rows+1 entries;
 * *ntiles_out receives the tile count (tiles_out[0..ntiles] are row starts).
 * No reference counterpart (backend.py:118-152 is row-per-warp only). */
int sovlp_plan_tiles(int rows, const int64_t *ptr, int tile_nnz, int tile_rows, int block,
                     int32_t *tiles_out, int32_t *ntiles_out);

/* Host function.  Splits ntiles tiles into `parts` contiguous ranges of equal
 * cost (nonzeros + rows); part_out needs parts+1 entries. */
int sovlp_partition_tiles(int ntiles, const int32_t *tiles, const int64_t *ptr, int parts,
                          int32_t *part_out);

/* Tile geometry compiled into the single-LP engine and the tiled SpMV: the
 * workgroup size, and the row cap and nonzero cap of a tile (plan tiles with
 * these). */
int sovlp_tile_geometry(int precision, int *block, int *tile_rows, int *tile_nnz);

/* y = A x through the engines' tile pass (ptr32 = int32 row pointers).
 * `part` has grid+1 entries (sovlp_partition_tiles with parts = grid). */
int sovlp_spmv_tiled_f64(int rows, const int32_t *ptr32, const int32_t *idx, const double *val,
                         const int32_t *tiles, int ntiles, const int32_t *part, int grid,
                         const double *x, double *y, sovlp_stream_t stream);
int sovlp_spmv_tiled_f32(int rows, const int32_t *ptr32, const int32_t *idx, const float *val,
                         const int32_t *tiles, int ntiles, const int32_t *part, int grid,
                         const float *x, float *y, sovlp_stream_t stream);

/* y = A x with A stored as `panels` column panels (XCD-aware): ptrs holds
 * panels x (rows+1) int32 pointers into the concatenated idx / val (panel p
 * holds the entries of the columns assigned to it), tiles the concatenated
 * tile plans of the panels (panel p's start at tile_off[p]), parts panels x
 * (groups+1) tile-range splits (sovlp_partition_tiles per panel).  Workgroup b
 * works on panel b % panels, so with 8 panels each XCD gathers from one slice
 * of x.  partial: panels x rows scratch; y[r] = the panels' row sums added in
 * panel order.  Two launches.  No reference counterpart (backend.py:118-152 is
 * row-per-warp only). */
int sovlp_spmv_panels_f64(int panels, int rows, const int32_t *ptrs, const int32_t *idx,
                          const double *val, const int32_t *tiles, const int32_t *tile_off,
                          const int32_t *parts, int groups, const double *x, double *partial,
                          double *y, sovlp_stream_t stream);
int sovlp_spmv_panels_f32(int panels, int rows, const int32_t *ptrs, const int32_t *idx,
                          const float *val, const int32_t *tiles, const int32_t *tile_off,
                          const int32_t *parts, int groups, const float *x, float *partial,
                          float *y, sovlp_stream_t stream);

/* ---- PDLP engine: src/sovopt/lp/pdlp.py:238-494 ---------------------- */

/* PDLPParams, pdlp.py:79-116 (same fields and defaults), then the
 * refinement settings of the FP32 engines (no reference counterpart):
 * refine_rounds  0 = none (required at precision 64), else the most solves
 *                (the plain one included) of an outer iterative-refinement
 *                loop, at most 16.  Round r > 0 solves min c'x^ subject to
 *                rl <= A x^ <= ru and lo <= x^ <= hi shifted by the FP64
 *                accumulated X (and A X) and scaled by a power of two Delta;
 *                X += x^ / Delta.  The dual carries over between rounds.
 * refine_stall   a round run after an evaluated X has met _converged (one
 *                that only chases refine_tol) ends when its KKT ratio has not
 *                halved for this many iterations (0: 1024), or after 8x that
 * refine_stall_open  the same for the rounds before that (0: 4096); shorter
 *                finishes fewer LPs in less time
 * refine_eps     relative tolerance (eps_abs = eps_rel) of every round
 * refine_tol     the loop also stops once the primal residual of X is
 *                <= refine_tol (1 + ||finite row bounds||_inf) and the duality
 *                gap of X, Y is <= refine_tol (1 + |pobj| + |dobj|), after
 *                _converged (pdlp.py:523-530) holds for X on the original.
 * flags          0, or SOVLP_FLAG_NO_LDS_MATRIX: the one-workgroup engines
 *                (batch, single LP on one workgroup, serving slots) then keep
 *                no LP's matrix or step vectors in LDS.  By default an LP
 *                whose scaled matrix (both orientations, 16-bit indices),
 *                plans and step vectors fit the workgroup's staging LDS is
 *                held there for the PDHG loop, and an LP whose matrix does
 *                not fit keeps its step vectors there; the results are the
 *                same bit for bit either way. */
#define SOVLP_FLAG_NO_LDS_MATRIX 1
typedef struct {
  double eps_abs, eps_rel, feas_cap;
  int64_t max_iter;
  double time_limit;
  int32_t check_every, stepsize_every;
  double beta_sufficient, beta_necessary, beta_artificial, primal_weight_smoothing;
  int32_t refine_rounds, refine_stall;
  double refine_eps, refine_tol;
  int32_t refine_stall_open, flags;
} sovlp_params;

/* One LP in minimisation form (costs and offset already negated for a
 * maximisation; `maximise` only restores the textbook signs on output).
 * All pointers are device pointers.  Bounds use +-1e30 for infinity. */
typedef struct {
  int32_t m, n;
  int64_t nnz;
  const int32_t *rp, *ri;   /* CSR of A: rp has m+1 int32 offsets */
  const double *rx;
  const int32_t *cp, *ci;   /* CSC of A (= CSR of A^T): cp has n+1 entries */
  const double *cx;
  const double *c, *rl, *ru, *lo, *hi;
  const double *v0;         /* power-iteration start: default_rng(0).standard_normal(n) */
  const int32_t *tiles_a;   /* sovlp_plan_tiles of A with the engine's geometry */
  int32_t ntiles_a;
  const int32_t *tiles_t;   /* ... of A^T */
  int32_t ntiles_t;
  double obj_offset;
  int32_t maximise;
} sovlp_lp;

/* PDLPInfo, pdlp.py:119-132, plus engine bookkeeping. */
typedef struct {
  int32_t status;           /* Status codes, src/sovopt/core/problem.py:58-69; -1 running */
  int32_t stage;            /* 0 setup pending, 1 iterating, 2 report pending, 3 done */
  int64_t iterations, restarts, rejected_steps, history_len;
  double primal_res, dual_res, gap, primal_obj, dual_obj;
  double step_size, primal_weight, objective, scale_obj;
  int64_t device_ticks;     /* device clock ticks spent (sovlp_clock_rate) */
  int32_t refine_rows;      /* rows of the refinement record */
  int32_t refine_round;     /* last refinement round started (0: none) */
  double refine_scale;      /* Delta of that round */
  int64_t start_tick;       /* device clock when the solve started */
} sovlp_info;

/* Workspace bytes for one LP at `precision` 64 (FP64 iterates) or 32 (FP32
 * iterates and matrix, FP64 reductions and termination test), with room for
 * `history_cap` history rows. */
int sovlp_pdlp_workspace_bytes(const sovlp_lp *lp, int precision, int64_t history_cap,
                               int64_t *bytes);

/* Opaque engine handle (host memory; freed by sovlp_pdlp_destroy). */
typedef struct sovlp_pdlp sovlp_pdlp;

/* Prepares a solve in the caller's workspace (>= workspace_bytes) and records
 * one replay of it into a hipGraph.  `team` 0 = choose, 1 = one workgroup
 * (a replay is one launch running up to sovlp_phases_per_launch() phases of
 * the solve back to back), 2 = W workgroups per phase (W from the tile
 * count; a replay is 64 phase launches, each followed by a one-workgroup
 * reduction launch).  Synchronises. */
int sovlp_pdlp_create(const sovlp_lp *lp, const sovlp_params *params, int precision,
                      void *workspace, int64_t workspace_bytes, int64_t history_cap, int team,
                      sovlp_stream_t stream, sovlp_pdlp **out);

/* Replays the graph until the solve has advanced `iteration_budget`
 * iterations (counted over every refinement round) or terminated; with one
 * workgroup the device stops at the budget exactly, with W workgroups within
 * one replay of it.  `seconds_left` is the device time budget of the solve; it
 * is taken on the first call.  Synchronises. */
int sovlp_pdlp_advance(sovlp_pdlp *h, int64_t iteration_budget, double seconds_left,
                       sovlp_stream_t stream);

/* Copies the engine state to *info; synchronises the stream. */
int sovlp_pdlp_info(sovlp_pdlp *h, sovlp_info *info, sovlp_stream_t stream);

/* Device pointers of the report (x, y, reduced costs, FP64, original units)
 * and of the history rows (k, pres, dres, gap, pobj), valid once stage == 3. */
int sovlp_pdlp_outputs(sovlp_pdlp *h, double **x, double **y, double **d, double **history);

/* Device pointer of the refinement record: refine_rows rows of (iterations
 * so far, primal residual, dual residual, gap, objective of the minimisation
 * form) of X on the original problem, one per finished round. */
int sovlp_pdlp_refinement(sovlp_pdlp *h, double **rows);

/* Workgroups per phase chosen by sovlp_pdlp_create. */
int sovlp_pdlp_width(sovlp_pdlp *h, int *workgroups);

/* Device pointers of the scaling factors (FP64, m row and n column factors,
 * exact powers of two), valid once the solve has passed its scaling phases:
 * the row and column factors of scale_problem(prob, "pdlp"),
 * src/sovopt/numerics/scaling.py:374-478 (the objective factor is
 * sovlp_info.scale_obj). */
int sovlp_pdlp_scaling(sovlp_pdlp *h, double **row, double **col);

/* Phases a launch of the run kernel advances each LP by, at most (the
 * batched engine, the one-workgroup single-LP engine): a replay count that
 * bounds a solve of max_iter iterations is (61 max_iter + 400) / phases + 16,
 * an iteration taking at most 61 phases. */
int sovlp_phases_per_launch(int *phases);

int sovlp_pdlp_destroy(sovlp_pdlp *h);

/* ---- phase accounting (no reference counterpart) ----------------------
 * The engines count, per LP, how often each phase of the solve ran.  Slot
 * `s` < sovlp_phase_slots() is named by sovlp_phase_name(s) (INIT, ...,
 * B_FUSED, CHECK, ..., plus B_ADAPT_NEXT / B_FUSED_NEXT: the B phases that
 * also formed the next primal step).  With the byte model of each phase
 * this gives the algorithmic HBM bytes of a solve. */
int sovlp_phase_slots(int *slots);
const char *sovlp_phase_name(int slot);

/* counts[sovlp_phase_slots()] of the engine's LP; synchronises. */
int sovlp_pdlp_phase_counts(sovlp_pdlp *h, int64_t *counts, sovlp_stream_t stream);

/* ticks[sovlp_phase_slots()]: device clock ticks (sovlp_clock_rate) the LP
 * spent in each phase, from the end of the previous phase's transition to
 * the end of its own, as the one-workgroup run kernel measures them (zero
 * with W workgroups per phase); synchronises.  No reference counterpart. */
int sovlp_pdlp_phase_ticks(sovlp_pdlp *h, int64_t *ticks, sovlp_stream_t stream);

/* ---- batched PDLP: many independent LPs in one persistent launch ------
 * No reference counterpart: the reference solves one LP per call
 * (pdlp.py:238); its only batched PDHG shares one A across B&B nodes
 * (src/sovopt/mip/bnr.py:176-301).  Here each LP keeps its own matrix,
 * scaling, step size, restarts and termination.  Every launch advances every
 * unfinished LP by up to sovlp_phases_per_launch() phases, one workgroup per LP, and the launch is
 * replayed from a hipGraph, so the arithmetic of an LP does not depend on the
 * other LPs in the batch.
 * Tile plans must use sovlp_batch_geometry's tile size and row count. */
typedef struct sovlp_batch sovlp_batch;

/* Tile geometry of the batched engine for `precision` and `block` (64, 128,
 * 256 or 512 threads per LP): the row cap and nonzero cap of a tile. */
int sovlp_batch_geometry(int precision, int block, int *tile_rows, int *tile_nnz);

/* What the one-workgroup engines on `block` threads hold in LDS for the PDHG
 * loop of an LP of this shape (tiles_a / tiles_t: its tile counts for that
 * block's geometry; see sovlp_params.flags): *resident = 1 its scaled matrix
 * and step vectors, 2 only its step vectors dx and dy and its tile bounds (the
 * products then gather from LDS, the matrix streams from memory), 0 nothing.
 * No reference counterpart (the reference streams the matrix from memory in
 * every product, backend.py:118-152). */
int sovlp_lds_resident(int precision, int block, int m, int n, int64_t nnz, int tiles_a,
                       int tiles_t, int *resident);

/* Workspace for `count` LPs (host array of descriptors with device pointers). */
int sovlp_batch_workspace_bytes(int count, const sovlp_lp *lps, int precision,
                                int64_t history_cap, int64_t *bytes);

/* Lays out the batch in the workspace and records one replay into a
 * hipGraph: one launch, one workgroup of `block` threads per LP, each running
 * up to 65536 phases of its LP back to back.  `block` is 64, 128, 256 or 512.
 * Synchronises. */
int sovlp_batch_create(int count, const sovlp_lp *lps, const sovlp_params *params,
                       int precision, int block, void *workspace, int64_t workspace_bytes,
                       int64_t history_cap, sovlp_stream_t stream, sovlp_batch **out);

/* As sovlp_batch_create with a workgroup size per LP (blocks[count], each 64,
 * 128, 256 or 512; an autotuner's choice from the LP's own shape).  A replay
 * launches the LPs of each size in one launch of their own, one after the
 * other; each LP's tile plans use sovlp_batch_geometry of its own size. */
int sovlp_batch_create_mixed(int count, const sovlp_lp *lps, const sovlp_params *params,
                             int precision, const int32_t *blocks, void *workspace,
                             int64_t workspace_bytes, int64_t history_cap, sovlp_stream_t stream,
                             sovlp_batch **out);

/* Solves every LP of the batch from scratch: resets the states, then replays
 * the graph until every LP has terminated (or `max_graphs` > 0 replays).
 * `seconds_per_lp` is each LP's device time budget.  Synchronises. */
int sovlp_batch_run(sovlp_batch *h, double seconds_per_lp, int64_t max_graphs,
                    sovlp_stream_t stream, int64_t *graphs_run);

/* Copies the `count` engine states to infos[]; synchronises. */
int sovlp_batch_infos(sovlp_batch *h, sovlp_info *infos, sovlp_stream_t stream);

/* Device pointers of LP `index`'s report (FP64, original units). */
int sovlp_batch_outputs(sovlp_batch *h, int index, double **x, double **y, double **d);

/* Device pointer of LP `index`'s refinement record (as sovlp_pdlp_refinement). */
int sovlp_batch_refinement(sovlp_batch *h, int index, double **rows);

/* counts[count * sovlp_phase_slots()], LP-major; synchronises. */
int sovlp_batch_phase_counts(sovlp_batch *h, int64_t *counts, sovlp_stream_t stream);

/* ticks[count * sovlp_phase_slots()], LP-major: device clock ticks each LP
 * spent in each phase (as sovlp_pdlp_phase_ticks); synchronises. */
int sovlp_batch_phase_ticks(sovlp_batch *h, int64_t *ticks, sovlp_stream_t stream);

int sovlp_batch_destroy(sovlp_batch *h);

/* ---- serving: continuous batching over a slot table -------------------
 * No reference counterpart: the reference solves one LP per call
 * (src/sovopt/lp/pdlp.py:238) from a cold start (x0 = clip(0), y0 = 0,
 * pdlp.py:274-275).  A server holds `slots` LP slots that share every launch
 * of one recorded graph (one workgroup per slot, up to `phases_per_launch`
 * phases each, as the batched engine).  LPs are admitted into free slots
 * between launches, so requests that arrive while others run join the next
 * launch; each LP's arithmetic depends only on its own data and the
 * workgroup size, so its result is the same alone and in any mix.  The
 * caller owns all device memory: the server's table (workspace_bytes) and one
 * workspace per admitted LP (lp_bytes), which must stay allocated until the
 * slot is released.  A slot's tile plans use sovlp_batch_geometry's sizes for
 * the server's precision and block.  One host thread per server. */
typedef struct sovlp_server sovlp_server;

/* Bytes of the server's own device table for `slots` slots. */
int sovlp_server_workspace_bytes(int slots, int precision, int64_t *bytes);

/* The same for a team server (sovlp_server_create_team). */
int sovlp_server_team_workspace_bytes(int slots, int team, int precision, int64_t *bytes);

/* Workspace bytes of one admitted LP (history_cap history rows). */
int sovlp_server_lp_bytes(const sovlp_lp *lp, int precision, int64_t history_cap,
                          int64_t *bytes);

/* Creates an empty server (every slot free) and records its launch graph.
 * `block` 64, 128, 256 or 512; params->time_limit is each LP's device time budget,
 * counted from its first phase.  Synchronises. */
int sovlp_server_create(int slots, const sovlp_params *params, int precision, int block,
                        int phases_per_launch, void *workspace, int64_t workspace_bytes,
                        sovlp_stream_t stream, sovlp_server **out);

/* A server whose slots run each phase over `team` workgroups of 256 threads
 * (for LPs of tens of thousands of nonzeros and more, whose phases one
 * workgroup would run tile after tile): a replay is `pairs_per_launch` pairs
 * of launches, one phase of every occupied slot over slots x team
 * workgroups, then one workgroup per slot that reduces the team's partial
 * sums in a fixed order and takes the transition.  An LP's result depends on
 * its data and `team` only.  Admission plans the team's tile split on the
 * host (it synchronises the stream once).  Synchronises. */
int sovlp_server_create_team(int slots, int team, const sovlp_params *params, int precision,
                             int pairs_per_launch, void *workspace, int64_t workspace_bytes,
                             sovlp_stream_t stream, sovlp_server **out);

/* Admits an LP into free slot `slot` (asynchronous, stream-ordered before
 * the next step).  Warm start: x0 (n) and y0 (m) are device vectors in the
 * units and signs of a report (sovlp_server_outputs of an earlier solve of an
 * LP with the same matrix shape), omega0 > 0 a primal weight to start from;
 * NULL / 0 give the reference's cold start.  x0 and y0 must stay allocated
 * until the slot is released. */
int sovlp_server_admit(sovlp_server *h, int slot, const sovlp_lp *lp, const double *x0,
                       const double *y0, double omega0, void *workspace,
                       int64_t workspace_bytes, int64_t history_cap, sovlp_stream_t stream);

/* Replays the launch graph `launches` times (0: only poll), then copies the
 * slot states to the host.  finished[slots] (optional) receives 1 for every
 * occupied slot whose LP has terminated, *running the occupied slots still
 * running.  Synchronises. */
int sovlp_server_step(sovlp_server *h, int launches, sovlp_stream_t stream, int32_t *finished,
                      int32_t *running);

/* The slot's engine state as of the last step. */
int sovlp_server_info(sovlp_server *h, int slot, sovlp_info *info);

/* Device pointers of the slot's report (x, y, reduced costs; FP64, original
 * units), valid once the slot has finished and until it is released. */
int sovlp_server_outputs(sovlp_server *h, int slot, double **x, double **y, double **d);

/* Device pointer of the slot's refinement record (as sovlp_pdlp_refinement). */
int sovlp_server_refinement(sovlp_server *h, int slot, double **rows);
/* The slot's check history (device pointer; as sovlp_pdlp_outputs' history:
 * min(info.history_len, history_cap) rows of (iteration, unscaled primal
 * residual, dual residual, gap, primal objective)), as pdlp.py:433-441 logs it. */
int sovlp_server_history(sovlp_server *h, int slot, double **rows);

/* counts[sovlp_phase_slots()] of the slot's current LP; synchronises. */
int sovlp_server_phase_counts(sovlp_server *h, int slot, int64_t *counts,
                              sovlp_stream_t stream);

/* Frees a finished slot for the next admission (its LP's memory may then be
 * reused).  Fails while the slot's LP is still running. */
int sovlp_server_release(sovlp_server *h, int slot);

/* Device clock rate of the server's device in Hz (for tick counts). */
int sovlp_server_clock_hz(sovlp_server *h, double *hz);

int sovlp_server_destroy(sovlp_server *h);

/* ---------------------------------------------------------------------------
 * Crossover (host code): an optimal basis from a PDLP point.
 *
 * The variables are z = [x ; s] with A x - s = 0 (one logical per row, bounded
 * by the row bounds), the basis convention of the reference
 * (src/sovopt/lp/basis.py:75-79, 203-215): status codes BASIC 0, AT_LOWER 1,
 * AT_UPPER 2, FREE 3, FIXED 4, one per variable (n structurals, then m
 * logicals).  The basis is factorised densely: at most 8192 rows (error 4).
 * ------------------------------------------------------------------------- */

/* An LP on the host in CSC form, objective minimised; +-1e30 = no bound. */
typedef struct {
  int32_t m, n;
  int64_t nnz;
  const int64_t *cp;    /* n + 1 column pointers */
  const int32_t *ci;    /* row of each nonzero */
  const double *cx;     /* value of each nonzero */
  const double *c;      /* n costs */
  const double *col_lb, *col_ub;   /* n */
  const double *row_lb, *row_ub;   /* m */
} sovlp_host_lp;

/* Fields of the reference's SimplexParams (src/sovopt/lp/simplex.py:94-187)
 * that this simplex reads. */
typedef struct {
  int64_t max_iter;
  double time_limit;        /* seconds */
  double feas_tol;          /* primal feasibility, relative to 1 + |bound| */
  double opt_tol;           /* reduced-cost optimality, relative to 1 + |c_j| */
  double pivot_tol;         /* smallest |alpha| the ratio test takes */
  int32_t refactor_freq;    /* pivots between refactorisations */
  int32_t perturb_after;    /* degenerate pivots before the costs are perturbed (0: never) */
  double perturb_scale;
} sovlp_simplex_params;

typedef struct {
  int32_t status;           /* Status code: 0 OPTIMAL, 1 INFEASIBLE, 2 UNBOUNDED,
                               4 iteration limit, 5 time limit, 8 NUMERICAL */
  int32_t repairs;          /* columns the factorisations replaced by logicals */
  int64_t iterations;       /* pivots + bound flips */
  int64_t pivots;
  int64_t phase1_iterations;
  int64_t superbasics;      /* sovlp_crossover: nonbasics it started between their bounds */
  double objective;         /* c x */
  double primal_infeasibility;
  double dual_infeasibility;
  double seconds;
} sovlp_simplex_result;

/* Bounded primal revised simplex (phase 1 on the sum of infeasibilities, then
 * phase 2; Dantzig pricing, Harris ratio test, bound flips, cost perturbation
 * on degenerate stalls).  `warm_status` (n + m entries, exactly m BASIC, or
 * NULL: the all-logical basis with statuses from the costs) is the starting
 * basis.  Outputs: x (n), y (m; c - A^T y are the reduced costs, y the duals
 * of the rows), d (n reduced costs), the final status vector (n + m).
 * Replaces the simplex the reference's crossover finishes with
 * (src/sovopt/lp/simplex.py:1181-1300, solve_simplex(prob, params, warm_basis)). */
int sovlp_simplex_solve(const sovlp_host_lp *lp, const sovlp_simplex_params *params,
                        const uint8_t *warm_status, double *x, double *y, double *d,
                        uint8_t *status, sovlp_simplex_result *result);

/* A basis for the point x (n) and, if not NULL, duals y (m, tie-break only):
 * each variable of [x ; A x] scored by its distance to its nearest bound over
 * its range, then the structurals off their bounds pushed into the
 * all-logical basis best first, each pivoting on its largest entry in a row
 * that still holds its logical (candidates whose entries there are all at
 * most pivot_tol are skipped), so the basis is nonsingular by construction.
 * Writes the n + m statuses and the number of structurals placed.
 * Replaces basis_from_point (src/sovopt/lp/crossover.py:80-188). */
int sovlp_basis_from_point(const sovlp_host_lp *lp, const double *x, const double *y,
                           double pivot_tol, double tol_fixed, uint8_t *status,
                           int32_t *placed);

/* The crossover (src/sovopt/lp/crossover.py:192-222, crossover(prob, x, y)):
 * the basis of sovlp_basis_from_point for x0 (duals y0, may be NULL), then
 * the simplex from it with every nonbasic variable that x0 puts strictly
 * between its bounds held at that value (a superbasic) rather than moved to
 * a bound first; the simplex pushes those to bounds or into the basis
 * itself.  Starting at the point keeps the basic values at the point's, so
 * phase 1 has only the point's own infeasibility to remove.  Outputs as
 * sovlp_simplex_solve. */
int sovlp_crossover(const sovlp_host_lp *lp, const sovlp_simplex_params *params,
                    const double *x0, const double *y0, double *x, double *y, double *d,
                    uint8_t *status, sovlp_simplex_result *result);

/* Ranging of an optimal basis (the objective minimised, as given): the
 * basis the n + m statuses name is factorised, its point, duals y = B^-T c_B
 * and reduced costs recomputed, then
 *   value (n)        the basis's structural values;
 *   cost_lo/hi (n)   the interval of each c_j over which the basis stays
 *                    optimal: a nonbasic cost moves its own reduced cost, a
 *                    basic one every nonbasic reduced cost through its
 *                    tableau row (one BTRAN per basic structural);
 *   rhs, rhs_lo/hi (m) each row's active bound and the interval over which
 *                    the basis stays feasible (one FTRAN per binding row;
 *                    a non-binding row's bound may travel to its activity);
 *   activity, dual (m), binding (m: 1 when the row's logical is nonbasic).
 * Replaces compute_sensitivity (src/sovopt/lp/sensitivity.py:154-332). */
int sovlp_simplex_ranging(const sovlp_host_lp *lp, const uint8_t *status, double *value,
                          double *cost_lo, double *cost_hi, double *rhs, double *rhs_lo,
                          double *rhs_hi, double *activity, double *dual, uint8_t *binding);

/* Factorises the basis a status vector names (exactly m BASIC) and reports
 * how many of its columns had to be replaced by logicals (0: nonsingular).
 * The check of the reference's Basis.factorize / n_repairs
 * (tests/test_crossover.py:47-70). */
int sovlp_basis_repairs(const sovlp_host_lp *lp, const uint8_t *status, int32_t *repairs);

#ifdef __cplusplus
}
#endif
#endif /* SOVLP_H */
