/*
 * sevp.h -- C ABI of the MI355X SEVP (Stabilized Error Vector Propagation)
 * barotropic-pressure solver, a drop-in for the TIMCOM solver calls.
 *
 * Conventions shared by every entry point
 *   - Arrays are Fortran (column-major) arrays, passed as plain pointers.
 *     Coefficients AL/AB/AC/AR/AT (the reference's AX/AY/BB/CX/CY) and the
 *     right-hand side are float (REAL*4) with shape (I2, J2); the solution X
 *     and all influence matrices are double (REAL*8).  X carries one halo
 *     column and row on every side: shape (I2+2, J2+2).
 *   - IE holds the block row ends IE(1..NB0) with the reference's meaning
 *     (PREP/template/GLOBAL/src/prep_GLOBAL.f90:147-168): strictly increasing,
 *     IE(NB0) = J2+1.
 *   - Every function returns 0 on success or a negative SEVP_E* code;
 *     sevp_last_error() describes the most recent failure of the calling
 *     thread.  Numerical problems are not errors: zero pivots met by the
 *     Gauss-Jordan inversion are counted and reported, as the reference
 *     divides by them silently.
 *   - The "_dev" entry points take device pointers and enqueue on the plan's
 *     stream (or the stream given); the others take host pointers and return
 *     when the result is back in host memory.  The plan's own stream is a
 *     blocking stream, ordered with the null stream.
 *   - A plan is used by one host thread at a time.  Calls on one plan are
 *     ordered on its stream.
 */
#ifndef SEVP_H
#define SEVP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEVP_OK 0
#define SEVP_E_ARG (-1)      /* invalid argument or shape */
#define SEVP_E_HIP (-2)      /* HIP runtime error */
#define SEVP_E_NOMEM (-3)    /* device allocation failed */
#define SEVP_E_STATE (-4)    /* plan not ready (influence matrices missing) */
#define SEVP_E_NOTCONV (-5)  /* refinement did not reach the tolerance */
#define SEVP_E_IO (-6)       /* file missing, unreadable or not the expected record */

/* Solver form (the reference's FL_EVP_STP). */
#define SEVP_FORM_FULL_PERIODIC 0 /* REP1 + PRE1: full width, periodic east-west (FL_EVP_STP=0) */
#define SEVP_FORM_BIR 1           /* REPBIR + PRE on NBIR overlapping strips, multiplicative sweeps (FL_EVP_STP=1) */
#define SEVP_FORM_FULL 2          /* REP + PRE with M2=I2: full width, Dirichlet east-west halo columns */

/* Arithmetic of the influence-matrix corrections. */
#define SEVP_PREC_FP64 0 /* plain: reference operation order, bitwise equal to the reference */
#define SEVP_PREC_FP32 1 /* corrections as FP32 MFMA GEMMs; marching and residual FP64 */
#define SEVP_PREC_BF16 2 /* corrections as BF16-input, FP32-accumulate MFMA GEMMs (the fused pass
                            only: BF16 matrices, FP32 error rows rounded to BF16 in registers);
                            a trial: each refinement pass contracts the residual by ~0.2 at best */
#define SEVP_PREC_FP64_MFMA 3 /* corrections as FP64 MFMA split-K products: FP64 arithmetic in a
                                 blocked summation order (not bitwise the reference) */

typedef struct sevp_plan_s *sevp_plan;

typedef struct {
  int i2, j2;      /* interior grid size (I2 x J2) */
  int nb0;         /* number of block rows NB0 (at most 4096 for SEVP_FORM_BIR) */
  const int *ie;   /* IE(1..NB0), host pointer, copied */
  int form;        /* SEVP_FORM_* */
  int nbir;        /* number of BIR strips (SEVP_FORM_BIR); IBIR = 2*I2/NBIR */
  int nsweep;      /* BIR sweeps per solve (the reference runs 5) */
  int periodic;    /* BIR: refresh the periodic halo after the seam strip (FLAG_PERIODIC_WE) */
  int nrhs_max;    /* largest batch of right-hand sides solved in one call */
  int precision;   /* SEVP_PREC_* used by the correction products */
  int seam_fix;    /* BIR: 0 = the reference's seam halo gather (rows 1..J2,
                      MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1921-1923), which
                      leaves the seam strip a zero halo on row J2+1 so the
                      sweeps stall; 1 = gather rows 1..J2+1, sweeps converge */
} sevp_desc;

typedef struct {
  int iterations;        /* SEVP passes run (1 for a plain solve) */
  double max_rel_resid;  /* max over RHS of ||s - A x||_2 / ||s||_2 at exit, or -1 if not evaluated */
  double seconds;        /* wall time of the call */
} sevp_stats;

/* ---------------------------------------------------------------- plans */

/* Create a plan and upload the operator (host pointers, each I2*J2 floats).
 * Replaces the module allocation MAIN/template/GLOBAL/src/grid_var_GLOBAL.f90:
 * 537-592 (VAR_ALLOCATE_SEVP) and the operator read of
 * MAIN/template/GLOBAL/src/input_GLOBAL.f90:438-443. */
int sevp_plan_create(const sevp_desc *desc, const float *al, const float *ab, const float *ac,
                     const float *ar, const float *at, sevp_plan *out);
int sevp_plan_destroy(sevp_plan plan);

/* Build and invert the influence matrices on the device.  Replaces PREP's
 * INITFS strip loop (PREP/template/GLOBAL/src/prep_GLOBAL.f90:264-317) with
 * PRE (:457-513), PRE1 (:327-386) and MATINV (:516-568).  zero_pivots (may be
 * NULL) receives the number of zero pivots met. */
int sevp_build_influence(sevp_plan plan, int *zero_pivots);

/* Load / export influence matrices in the reference layout (the RINV, RINV1
 * records of the EVP file, PREP/template/GLOBAL/src/prep_GLOBAL.f90:320):
 *   BIR : RINV(IBIR*IBIR, NB0, NBIR), RINV1(IBIR*IBIR, NB0-1, NBIR)
 *   full: RINV(I2, I2, NB0),          RINV1(I2, I2, NB0-1)           */
int sevp_load_influence(sevp_plan plan, const double *rinv, const double *rinv1);
int sevp_get_influence(sevp_plan plan, double *rinv, double *rinv1);

/* Zero the seam-strip scratch XX (the reference's module array starts at
 * zero and keeps its interior between calls). */
int sevp_reset_seam(sevp_plan plan);

/* ---------------------------------------------------------------- solves */

/* One solve of nrhs independent right-hand sides with the plan's form, in the
 * reference's operation order (FP64 plans) -- replaces FS_SOLVER's solver
 * dispatch, MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1850-1966: the 5-sweep BIR
 * strip loop (FL_EVP_STP=1) or one REP1 pass (FL_EVP_STP=0).
 *   s : (I2, J2, nrhs) float;  x : (I2+2, J2+2, nrhs) double, initial guess in,
 *   solution out (halo columns and rows are the Dirichlet / periodic data). */
int sevp_solve(sevp_plan plan, const float *s, double *x, int nrhs);
int sevp_solve_dev(sevp_plan plan, const float *s_dev, double *x_dev, int nrhs, void *hip_stream);

/* Iterative refinement x <- x + SEVP(s - A x) with the residual in FP64, until
 * every RHS has ||s - A x||_2 / ||s||_2 <= tol or max_iter passes ran.  Full
 * forms only.  Returns SEVP_E_NOTCONV (with x holding the last iterate) if
 * the tolerance was not met.  rel_resid (host, nrhs doubles, may be NULL)
 * receives each RHS's final relative residual. */
int sevp_solve_refined_dev(sevp_plan plan, const float *s_dev, double *x_dev, int nrhs,
                           double tol, int max_iter, double *rel_resid, sevp_stats *stats);

/* Batched problems: many independent full-width plans (regional basins, each
 * with its own grid, operator, IE and right-hand sides) refined together.
 * There is no counterpart in the reference, which solves one domain per call
 * (REP, MAIN/template/GLOBAL/src/sevp.f90:335-438, or REP1 :11-107); each
 * problem's result is bitwise that of sevp_solve_refined_dev on its plan.
 * Every step of a pass is one launch over all problems and the pass is one
 * HIP graph, so the batch takes as many dependent launches as its tallest
 * problem, not as all of them together.
 * Requirements: full forms, influence matrices present, one precision for
 * all plans (SEVP_PREC_FP32 or SEVP_PREC_FP64_MFMA), I2 divisible by 4, and
 * one product slice length (all I2 <= 1536, or all I2 > 1536).  The batch
 * runs the split-K pass kernels and switches its plans to them
 * (sevp_set_pass).  The
 * batch uses the plans' scratch arrays: do not run a plan on its own while a
 * batched solve that holds it runs.  The plans must outlive the batch. */
typedef struct sevp_batch_s *sevp_batch;
int sevp_batch_create(const sevp_plan *plans, int nplans, sevp_batch *out);
int sevp_batch_destroy(sevp_batch batch);
/* s[i], x[i]: device arrays of problem i, as for sevp_solve_refined_dev with
 * nrhs[i] right-hand sides.  rel_resid (host, sum of nrhs[i] doubles in
 * problem order) and passes (host, nplans ints) may be NULL; stats gets the
 * most passes and the largest residual.  SEVP_E_NOTCONV if any problem
 * missed the tolerance. */
int sevp_batch_solve_refined_dev(sevp_batch batch, const float *const *s_dev,
                                 double *const *x_dev, const int *nrhs, double tol, int max_iter,
                                 double *rel_resid, int *passes, sevp_stats *stats);

/* Relative residual ||s - A x||_2 / ||s||_2 per RHS in FP64, with the
 * 5-point product of p_sqprod (MAIN/template/GLOBAL/src/sevp.f90:696-713). */
int sevp_residual_dev(sevp_plan plan, const float *s_dev, const double *x_dev, int nrhs,
                      double *rel_resid);

/* ------------------------------------- one-to-one reference entry points
 * Host pointers, the reference's argument lists and meaning; each call copies
 * its operands to the device and back.  Scratch arguments (DUM0, DUM1, DUM2,
 * H) are accepted for signature compatibility; the device keeps its own
 * scratch and does not write them. */

/* REPBIR, MAIN/template/GLOBAL/src/sevp.f90:218-328.  2 <= M2 <= 64, NBLK <= 4096. */
int sevp_repbir(const float *ax, const float *ay, const float *bb, const float *cx,
                const float *cy, const double *rinv, const double *rinv1, double *dum0,
                double *dum1, double *dum2, const float *f, double *h, double *x, const int *ie,
                int i0, int i2, int m0, int m2, int nblk);
/* REP1, MAIN/template/GLOBAL/src/sevp.f90:11-107. */
int sevp_rep1(const float *ax, const float *ay, const float *bb, const float *cx,
              const float *cy, const double *rinv, const double *rinv1, double *dum0,
              double *dum1, double *dum2, const float *f, double *h, double *x, const int *ie,
              int i0, int i2, int nblk);
/* REP, MAIN/template/GLOBAL/src/sevp.f90:335-438. */
int sevp_rep(const float *ax, const float *ay, const float *bb, const float *cx,
             const float *cy, const double *rinv, const double *rinv1, double *dum0,
             double *dum1, double *dum2, const float *f, double *h, double *x, const int *ie,
             int i0, int i2, int nblk);
/* REPBIR_PERIOD, MAIN/template/GLOBAL/src/sevp.f90:109-217.  Its periodic
 * copies index column I0 of both X and H, which is only meaningful when
 * M0 = I0 and M2 = I2 (its one call site, commented out at
 * MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1851); other shapes are rejected. */
int sevp_repbir_period(const float *ax, const float *ay, const float *bb, const float *cx,
                       const float *cy, const double *rinv, const double *rinv1, double *dum0,
                       double *dum1, double *dum2, const float *f, double *h, double *x,
                       const int *ie, int i0, int i2, int m0, int m2, int nblk);
/* PRE, PREP/template/GLOBAL/src/prep_GLOBAL.f90:457-513 (one strip). */
int sevp_pre(const float *ax, const float *ay, const float *bb, const float *cx,
             const float *cy, double *rinv, double *rinv1, double *h, const int *ie, int i2,
             int m0, int m2, int nblk);
/* PRE1, PREP/template/GLOBAL/src/prep_GLOBAL.f90:327-386. */
int sevp_pre1(const float *ax, const float *ay, const float *bb, const float *cx,
              const float *cy, double *rinv, double *rinv1, double *h, const int *ie, int i0,
              int i2, int nblk);
/* MATINV, PREP/template/GLOBAL/src/prep_GLOBAL.f90:516-568. */
int sevp_matinv(double *b, int n);

/* MATINV on a device matrix, in place, with no allocation.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The result is bitwise sevp_matinv's.  b_dev: n x n FP64
 * column-major, device.  zero_pivots_dev (device int, may be NULL) receives
 * the zero pivots met.  Enqueued on hip_stream (NULL: the null stream); the
 * call does not wait.  The inversion's scratch is a fixed 2 MiB device array
 * per GPU (orders up to 4096; larger ones are refused with SEVP_E_ARG),
 * shared by every call on that GPU: calls that overlap in time on one GPU
 * race. */
int sevp_matinv_dev_noalloc(double *b_dev, int n, int *zero_pivots_dev, void *hip_stream);

/* Enqueue this plan's work on `hip_stream` (NULL: the plan's own stream,
 * which is ordered with the null stream).
 * Callers that produce a plan's inputs on another stream set it here so the
 * solve is ordered after them. */
int sevp_set_stream(sevp_plan plan, void *hip_stream);

/* Replay each refinement pass as a captured HIP graph (default on); the
 * graph is captured on first use for a batch size and precision.  Off: the
 * pass's kernels are launched one by one. */
int sevp_set_graphs(sevp_plan plan, int on);

/* Split the right-hand sides of sevp_solve_refined_dev into `groups` (1..16)
 * independent groups whose SEVP passes run concurrently on separate streams
 * (default 1).  The result is bitwise the same for any grouping of FP64
 * passes; mixed-precision passes are deterministic for a given grouping. */
int sevp_set_rhs_groups(sevp_plan plan, int groups);

/* Kernels of a full-width pass with MFMA corrections (SEVP_PREC_FP32,
 * SEVP_PREC_FP64_MFMA).
 *   SEVP_PASS_SPLITK  four launches per block: march, split-K product into
 *                     slice partials, product, march; the consumers add the
 *                     partials.  Operands in lane tiles.  Any block height;
 *                     the batched solves use it.
 *   SEVP_PASS_FUSED   one launch per block and direction: each workgroup
 *                     sums whole dot products and marches the next block (or
 *                     forms the next closure residual) in its epilogue.
 *                     Column-major operands.  Block heights <= 4, I2
 *                     divisible by 4.  The default where it applies.
 * Both give the same FP64 marching and closure residuals; their products
 * add the same terms in different orders, so results differ at rounding
 * level.  Each is deterministic.  A products-only plan keeps the layout it
 * was built with.  No counterpart in the reference (REP1's corrections are
 * sequential loops, MAIN/template/GLOBAL/src/sevp.f90:11-107). */
#define SEVP_PASS_SPLITK 0
#define SEVP_PASS_FUSED 1
int sevp_set_pass(sevp_plan plan, int kind);
int sevp_get_pass(sevp_plan plan, int *kind);

/* Performance parameters of the MFMA corrections, chosen by sevp_autotune
 * or set directly:
 *   pass_kind   SEVP_PASS_FUSED or SEVP_PASS_SPLITK;
 *   waves       wavefronts per workgroup of the fused kernels (8 or 16): the
 *               share of K each wavefront sums;
 *   rhs_groups  RHS groups whose passes run concurrently (1..16);
 *   rows        16-row tiles (right-hand sides) per fused workgroup, 1 or 2
 *               (0: two for FP32 and BF16, one for FP64).
 * The results of each choice are deterministic; the pass kernels and the
 * waves change the products' summation order (rounding-level differences),
 * the row tiles and RHS groups do not.
 * sevp_autotune times refined solves of the caller's problem (s_dev, x_dev,
 * nrhs, tol, max_iter as for sevp_solve_refined_dev; x_dev is overwritten)
 * for every applicable choice, keeps the fastest and returns it with its
 * time per solve call.  Block height is a property of the plan (its IE):
 * musoc_solver_amd.autotune tunes it across plans.  No counterpart in the
 * reference. */
typedef struct {
  int pass_kind;
  int waves;
  int rhs_groups;
  int rows;
} sevp_tuning;
int sevp_get_tuning(sevp_plan plan, sevp_tuning *tuning);
int sevp_set_tuning(sevp_plan plan, const sevp_tuning *tuning);
int sevp_autotune(sevp_plan plan, const float *s_dev, double *x_dev, int nrhs, double tol,
                  int max_iter, sevp_tuning *best, double *best_ms);

/* Where the influence matrices live (set before sevp_build_influence or
 * sevp_load_influence; changing it drops the matrices built so far).
 *   SEVP_STORE_FP64      RINV / RINV1 in FP64 on the device (the default),
 *                        plus the copies the correction arithmetic needs.
 *   SEVP_STORE_PRODUCTS  only the FP32 product operands (RINV, RINV1 and the
 *                        coupling operators, 3 * NB0 * I2^2 * 4 bytes): the
 *                        FP64 matrices of one block at a time are built (or
 *                        uploaded) and rounded as soon as they are complete.
 *                        Full forms with SEVP_PREC_FP32 only; sevp_get_influence
 *                        and switching to another precision are refused.  For
 *                        grids whose FP64 set does not fit in HBM (3600 x 1800
 *                        with 2-row blocks: 186 GB of FP64 matrices, 140 GB of
 *                        FP32 operands).
 * No counterpart in the reference, which keeps RINV / RINV1 as REAL(8)
 * module arrays (MAIN/template/GLOBAL/src/grid_var_GLOBAL.f90:147-155). */
#define SEVP_STORE_FP64 0
#define SEVP_STORE_PRODUCTS 1
int sevp_set_storage(sevp_plan plan, int storage);

/* Switch the arithmetic of the correction products of an existing
 * full-width plan (SEVP_PREC_*); FP32 copies of the influence matrices are
 * made on first use.  Lets one influence-matrix build serve both the plain
 * and the mixed-precision solve. */
int sevp_set_precision(sevp_plan plan, int precision);

/* ---------------------------------------------------------- step 14
 * The barotropic correction of FS_GLOBAL around the solve
 * (MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1148-1284) on the device: per
 * ITMASK iteration, land masking of the top-layer velocities, W integrated
 * over the layers into the RHS S, the plan's solve (the FS_SOLVER call, BIR
 * or full width, warm-started from X), P0 += ODT*X, the U and V corrections
 * with their running sums SCR(:,:,3:4), and the largest top-layer land speed,
 * until it is <= TOLERANCE or MXMASK iterations have run.  The plan must have
 * been built for one right-hand side of this grid (i2, j2) or more.  All
 * array pointers are device pointers in the reference's shapes (Fortran
 * order), updated in place; values are bitwise those of the reference's
 * loops (REAL(4) arithmetic, REAL(8) where Fortran promotes). */
typedef struct {
  int i2, j2, k1;     /* I2, J2, K1 (K0 = K1 + 1) */
  int free_surface;   /* FLAG_FS: S = W(:,:,1) integrated upward, else S = -W(:,:,KB+1) */
  int periodic_we;    /* FLAG_PERIODIC_WE: correction of the seam face */
  int lopenw;         /* LOPENW: 1 also masks the periodic faces I = 1, I1 */
  int harness;        /* 1 reproduces the shipped source's U = 1, V = 0, W(:,:,K0) = 0 */
  int mxmask;         /* MXMASK, >= 1 */
  float tolerance;    /* TOLERANCE */
  float odt;          /* ODT */
} sevp_step14_desc;

typedef struct {      /* INTEGER(2) masks and REAL(4) metrics, read only */
  const short *kb, *iu0, *iv0;   /* (I0, J0) */
  const short *iu;               /* (I0, J0, K1) */
  const short *iv;               /* (I0, J1, K1) */
  const float *odx, *ody, *ocs;  /* (J0) */
  const float *csv, *odyv;       /* (J1) */
  const float *odz;              /* (K1) */
} sevp_step14_grid;

typedef struct {      /* updated in place */
  float *u;           /* (I0, J0, K1) */
  float *v;           /* (I0, J1, K1) */
  float *w;           /* (I0, J0, K0) */
  float *p0;          /* (I0, J0) */
  float *scr;         /* (I0, J0, 4); cleared at entry, as the reference does */
  float *s;           /* (I2, J2) the RHS */
  double *x;          /* (I0, J0) the solution, carried from call to call */
} sevp_step14_state;

/* itmask, tmp (optional): the iterations run and the final land speed. */
int sevp_step14_dev(sevp_plan plan, const sevp_step14_desc *desc, const sevp_step14_grid *grid,
                    const sevp_step14_state *state, int *itmask, float *tmp);

/* -------------------------------------------------------- measurement
 * Kernel-class timing with HIP events recorded on the plan's stream around
 * every launch of that class while enabled (used by the benchmark to measure
 * the dominant kernel's average duration over its timed region).  While it
 * is enabled, refinement passes launched outside a captured graph run their
 * RHS groups one after another on the plan's stream, so each duration is the
 * kernel's own. */
#define SEVP_KCLASS_CORR 0   /* influence-matrix correction products */
#define SEVP_KCLASS_MARCH 1  /* row marches (particular and homogeneous) */
#define SEVP_KCLASS_RESID 2  /* closure residuals and FP64 refinement residuals */
int sevp_profile_begin(sevp_plan plan, int kclass);
/* Stops recording; total_ms and launches receive the summed device time and
 * the number of launches timed since sevp_profile_begin. */
int sevp_profile_end(sevp_plan plan, double *total_ms, long *launches);
/* Correction products launched outside a captured graph also time
 * themselves from the GPU's real-time clock (first wavefront's start to last
 * wavefront's end, the span a kernel trace reports; their HIP events, which
 * sevp_profile_end sums, also span the dispatch latency).  Sum and count of
 * those spans since sevp_profile_begin. */
int sevp_profile_spans(sevp_plan plan, double *total_ms, long *launches);

/* -------------------------------------------------------- Krylov solvers
 * The reference's cross-check solvers (MAIN/template/GLOBAL/src/sevp.f90:
 * 443-694), used by FS_SOLVER to compare against the SEVP solution
 * (MAIN/template/GLOBAL/src/fs_GLOBAL.f90:1853,1968,2024-2026): BiCGSTAB
 * (p_bicgstab_le), Chebyshev iteration (p_csi) and two-sided Lanczos
 * eigenvalue bounds (lanczos).  Same vectors as the reference: b (I2, J2)
 * float, x (I2+2, J2+2) double with its halo rows as Dirichlet data.  The
 * unpreconditioned reference's CL..CT arguments have no counterpart. */
#define SEVP_KRYLOV_BICGSTAB 0
#define SEVP_KRYLOV_PCSI 1

typedef struct {
  int periodic;   /* FLAG_PERIODIC_WE: halo columns are periodic copies (solver_boundary_update) */
  int ref_order;  /* 1: inner products summed sequentially in p_inprod's order, so every iterate
                     is bitwise the reference's (one thread per sum: small grids); 0: fixed-order
                     parallel sums, deterministic */
  int max_iter;   /* 1000 in the reference */
  int restart;    /* BiCGSTAB restart length, 100 in the reference */
  double tol;     /* stop when ||r||_2 <= tol: the reference's REAL(4) 1e-8 (BiCGSTAB), 1e-7 (P-CSI) */
  double rtol;    /* BiCGSTAB inner loop: ||r|| / ||r_restart|| <= rtol, REAL(4) 1e-8 */
} sevp_krylov_opts;

typedef struct {
  int iterations;   /* the reference's ITER */
  int converged;    /* residual <= tol */
  double residual;  /* SQRT(TP) = ||b - A x||_2 over the interior at exit */
  double seconds;
} sevp_krylov_stats;

/* Fill opts with the reference's settings for solver (SEVP_KRYLOV_*). */
int sevp_krylov_defaults(sevp_krylov_opts *opts, int solver);

/* On a plan's operator; b_dev, x_dev device pointers; opts may be NULL. */
int sevp_bicgstab_dev(sevp_plan plan, const float *b_dev, double *x_dev,
                      const sevp_krylov_opts *opts, sevp_krylov_stats *stats);
int sevp_pcsi_dev(sevp_plan plan, double mineig, double maxeig, const float *b_dev, double *x_dev,
                  const sevp_krylov_opts *opts, sevp_krylov_stats *stats);
/* m Lanczos steps (the reference: 10); alpha (m) and sigma (m+1) may be NULL.
 * mineig is the smallest eigenvalue of FIGI's symmetric tridiagonal matrix
 * (the reference calls EISPACK RATQR for it), maxeig the reference's
 * row-sum bound. */
int sevp_lanczos_dev(sevp_plan plan, int m, const sevp_krylov_opts *opts, double *mineig,
                     double *maxeig, double *alpha, double *sigma);

/* One-to-one, host pointers (P_BICGSTAB_LE, P_CSI, LANCZOS of
 * MAIN/template/GLOBAL/src/sevp.f90:443,553,618); each call uploads the
 * operator. */
int sevp_p_bicgstab_le(const float *al, const float *ab, const float *ac, const float *ar,
                       const float *at, const float *b, double *x, int i2, int j2,
                       const sevp_krylov_opts *opts, sevp_krylov_stats *stats);
int sevp_p_csi(const float *al, const float *ab, const float *ac, const float *ar, const float *at,
               double mineig, double maxeig, const float *b, double *x, int i2, int j2,
               const sevp_krylov_opts *opts, sevp_krylov_stats *stats);
int sevp_lanczos(const float *al, const float *ab, const float *ac, const float *ar,
                 const float *at, int i2, int j2, int m, const sevp_krylov_opts *opts,
                 double *mineig, double *maxeig, double *alpha, double *sigma);

/* ------------------------------------------------------------- EVP file
 * The operator and influence matrices as one big-endian Fortran unformatted
 * sequential record (OPEN(..., CONVERT='BIG_ENDIAN', FORM='unformatted')).
 * PREP writes AL,AR,AB,AT,AC,RINV,RINV1,IE
 * (PREP/template/GLOBAL/src/prep_GLOBAL.f90:320); MAIN reads
 * AL,AR,AB,AT,AC,OAC,RINV,RINV1,IE (MAIN/template/GLOBAL/src/input_GLOBAL.f90:
 * 438-443), so the two layouts differ by OAC.  Shapes follow desc (i2, j2,
 * nb0, ie, form, nbir): coefficients (I2, J2) REAL(4); RINV, RINV1 REAL(8) as
 * in sevp_load_influence; IE(NB0) INTEGER(4).  Records longer than
 * 2147483639 bytes are split into subrecords with negated continuation
 * lengths (gfortran's convention); max_subrecord > 0 sets a smaller split
 * size (0 = the largest).  These calls run on the host only (no device is
 * touched); failures return SEVP_E_ARG (shapes) or SEVP_E_IO (the file) with
 * the message in sevp_last_error. */

/* Write the record; oac = NULL writes PREP's layout, else MAIN's.  desc->ie
 * supplies IE. */
int sevp_evp_write(const char *path, const sevp_desc *desc, const float *al, const float *ab,
                   const float *ac, const float *ar, const float *at, const double *rinv,
                   const double *rinv1, const float *oac, long max_subrecord);
/* Read the record.  with_oac: 0 = PREP's layout, 1 = MAIN's, -1 = whichever
 * the record length matches (*had_oac, if given, says which).  oac may be
 * NULL (the values are skipped).  A record whose length matches neither
 * layout is rejected with SEVP_E_IO. */
int sevp_evp_read(const char *path, const sevp_desc *desc, float *al, float *ab, float *ac,
                  float *ar, float *at, double *rinv, double *rinv1, int *ie, float *oac,
                  int with_oac, int *had_oac);

/* ----------------------------------------------------------- plan cache
 * Building the influence matrices is the expensive part of a plan (PREP's
 * whole job).  sevp_operator_hash hashes what they depend on: I2, J2, NB0,
 * IE, form, NBIR (BIR) and the bits of AL..AT (64-bit, host only).
 * sevp_plan_create_cached creates the plan and looks for
 * <cache_dir>/sevp_<hash>.evp, an EVP record in PREP's layout: when the
 * record's operator and IE equal the caller's bit for bit, its matrices are
 * loaded (*hit = 1); otherwise they are built and the file is written
 * (*hit = 0; -1 if it could not be written, which is not an error).  No
 * counterpart in the reference, where MAIN always reads the file PREP wrote. */
int sevp_operator_hash(const sevp_desc *desc, const float *al, const float *ab, const float *ac,
                       const float *ar, const float *at, unsigned long long *hash);
int sevp_plan_create_cached(const sevp_desc *desc, const float *al, const float *ab,
                            const float *ac, const float *ar, const float *at,
                            const char *cache_dir, sevp_plan *out, int *hit);

/* ------------------------------------------------------------- misc */
const char *sevp_last_error(void);
const char *sevp_version(void);
/* Device memory the plan holds, in bytes. */
size_t sevp_plan_device_bytes(sevp_plan plan);

#ifdef __cplusplus
}
#endif
#endif /* SEVP_H */
