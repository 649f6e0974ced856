/*
 * tailor.h -- C ABI of the MI355X BiGlobal assemble-and-solve path.
 *
 * One context owns one GPU's device memory and one HIP stream; every call is
 * stream-ordered on that stream.  A context is not reentrant: use one per
 * host thread.  Host buffers are always owned by the caller; device buffers
 * are owned by the context.  Every entry point returns a tailor_status
 * (0 = success); tailor_last_error() returns the message of the last
 * failure, which carries the reference's error text where the reference has
 * one.
 *
 * The entry points replace the reference's stage interfaces, which are C++
 * classes over PETSc objects (Stage(MPI_Comm)::run(const Recipe&,
 * ProblemData&, Diagnostics&) -> PetscErrorCode):
 *
 *   tailor_load_case     Prepare::initialize      Prepare.hpp:20-47,
 *                                                 Prepare.cpp:670-687
 *   tailor_assemble      Metrics::compute         Metrics.hpp:59-87
 *                        BaseFlowDerivatives::compute
 *                                                 BaseFlowDerivatives.hpp:86-111
 *                        GlobalOperator::assemble GlobalOperator.hpp:41-71
 *                        BoundaryConditions::apply
 *                                                 BoundaryConditions.hpp:46-65
 *   tailor_export_*      MatGetValues / VecGetArray reads of the PETSc
 *                        objects held in ProblemData (ProblemData.hpp:104-159)
 */
#ifndef TAILOR_H
#define TAILOR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tailor_ctx tailor_ctx;

/* Status codes mirror the PETSc error classes the reference raises. */
typedef enum {
  TAILOR_OK = 0,
  TAILOR_ERR_FILE_UNEXPECTED = 1, /* PETSC_ERR_FILE_UNEXPECTED: bad rule data */
  TAILOR_ERR_FP = 2,              /* PETSC_ERR_FP: non-finite values          */
  TAILOR_ERR_ARG_OUTOFRANGE = 3,  /* PETSC_ERR_ARG_OUTOFRANGE                 */
  TAILOR_ERR_SUP = 4,             /* PETSC_ERR_SUP: complex grid/base flow    */
  TAILOR_ERR_PLIB = 5,            /* PETSC_ERR_PLIB: internal inconsistency   */
  TAILOR_ERR_NOT_CONVERGED = 6,   /* PETSC_ERR_NOT_CONVERGED                  */
  TAILOR_ERR_USER_INPUT = 7,      /* PETSC_ERR_USER_INPUT                     */
  TAILOR_ERR_ARG_WRONGSTATE = 8,  /* PETSC_ERR_ARG_WRONGSTATE                 */
  TAILOR_ERR_MEM = 9,             /* device allocation failed                 */
  TAILOR_ERR_HIP = 10             /* HIP runtime error                        */
} tailor_status;

/* One-dimensional FD-q rule, row-major over (node, slot)
 * (FDQRuleData, ProblemData.hpp:26-96). */
typedef struct {
  int32_t node_count;
  int32_t q;                        /* stencil size is q + 1 */
  int32_t periodic;                 /* 0 bounded, 1 periodic half-open */
  double period;                    /* computational period (2 for [-1,1)) */
  const double *nodes;              /* node_count */
  const int64_t *stencil_indices;   /* node_count * (q+1) */
  const int64_t *stencil_offsets;   /* node_count * (q+1) */
  const double *weights;            /* 3 * node_count * (q+1): d0, d1, d2 */
} tailor_rule;

/* Everything the reference reads from config.yaml and the FD-q cache
 * (Recipe.hpp:17-61, Prepare.cpp:443-668).  Complex arrays are interleaved
 * (re, im) doubles in natural k_j_dof order: node = j*ny + i with i the xi
 * (wall-normal) index and j the eta (spanwise) index. */
typedef struct {
  int32_t ny, nz;
  tailor_rule xi, eta;
  const double *grid;       /* ny*nz*3 complex: x, y, z */
  const double *baseflow;   /* ny*nz*5 complex: rho, u, v, w, T */
  double spanwise_period;
  double reynolds, mach, prandtl, gamma;
  double gas_constant, reference_mu;   /* validated, not used by the physics */
  double reference_temperature, sutherland_constant;
  double alpha_re, alpha_im;
} tailor_case;

/* Diagnostics printed by the reference (Process.cpp:62-152). */
typedef struct {
  double min_jacobian, max_jacobian, max_grid_imag;
  double metric_norms[10];          /* xi_y xi_z eta_y eta_z xi_yy xi_zz xi_yz eta_yy eta_zz eta_yz */
  double max_baseflow_imag;
  double derivative_norms[25];      /* derivative-major [dy,dz,dyy,dzz,dyz] x [rho,u,v,w,T] */
  int64_t rows, block_size, mass_blocks, spatial_blocks;
  double mass_fro, spatial_fro;     /* ||M_Gamma||F, ||L||F */
  int32_t wall_nodes, farfield_nodes, incoming_modes, neutral_modes;
  int32_t min_incoming_per_node, max_incoming_per_node;
  double max_characteristic_imag, max_eigenvector_condition;
  double a_fro, b_fro;              /* ||A_bc||F, ||B_bc||F */
} tailor_assembly_diag;

/* Which operator an export or a re-assembly refers to. */
enum { TAILOR_OP_L = 0, TAILOR_OP_M = 1, TAILOR_OP_A = 2, TAILOR_OP_B = 3 };

/* Create a context on HIP device `device`. */
int tailor_create(tailor_ctx **ctx, int device);
void tailor_destroy(tailor_ctx *ctx);
const char *tailor_last_error(const tailor_ctx *ctx);

/* Validate rules (Prepare.cpp:328-437) and physics (Parser.cpp:110-145),
 * copy the case to the device and build the block-CSR structure:
 * 77 = (q_y+1)(q_z+1) block columns per block row, sorted ascending. */
int tailor_load_case(tailor_ctx *ctx, const tailor_case *c);

/* Change the streamwise wavenumber of the loaded case (alpha sweeps). */
int tailor_set_alpha(tailor_ctx *ctx, double alpha_re, double alpha_im);

/* Metrics -> base-flow derivatives -> A_bc = -L and B_bc = M with wall and
 * far-field rows replaced, all on the device.  diag may be NULL (then no
 * norms are computed and nothing but a status word is read back). */
int tailor_assemble(tailor_ctx *ctx, tailor_assembly_diag *diag);

/* Overwrite the device value buffers with L and M_Gamma (no boundary rows,
 * no sign change), the operators GlobalOperator::assemble leaves in
 * ProblemData.  Structure is the same as A_bc's. */
int tailor_assemble_unconstrained(tailor_ctx *ctx);

/* Block-CSR sizes: block rows (= nodes) and stored blocks of A_bc / L. */
int tailor_bsr_shape(const tailor_ctx *ctx, int64_t *block_rows, int64_t *nnzb);

/* Copy an operator to caller buffers.  For TAILOR_OP_L / TAILOR_OP_A:
 * rowptr (block_rows+1), colidx (nnzb), vals (nnzb*25 complex, 5x5
 * row-major blocks).  For TAILOR_OP_M / TAILOR_OP_B (block diagonal):
 * rowptr[i] = i, colidx[i] = i, vals (block_rows*25 complex).  Any pointer
 * may be NULL to skip it.  The value buffer holds whichever of L/A_bc
 * (resp. M/B_bc) was assembled last; asking for the other is an error. */
int tailor_export_bsr(tailor_ctx *ctx, int which, int32_t *rowptr, int32_t *colidx,
                      double *vals);

/* Copy the metric (N*10) and base-flow derivative (N*25) fields. */
int tailor_export_fields(tailor_ctx *ctx, double *metrics, double *derivatives);

/* Time `reps` launches of the value-writing assembly kernel on the
 * context's stream with HIP events; reports the mean per launch.  Also
 * reports the mean time of the full assembly (all kernels). */
int tailor_time_assembly(tailor_ctx *ctx, int reps, double *write_kernel_ms,
                         double *full_assembly_ms);

/* ------------------------------------------------------------------------
 * Shift-invert eigen-solve (replaces EigenSolver::solve, EigenSolver.cpp:31-89:
 * EPS GNHEP Krylov-Schur, ST SINVERT, KSP PREONLY, PC LU).
 * ---------------------------------------------------------------------- */

/* Sparse LU of K = A_bc - sigma B_bc (replaces the PCLU factorization inside
 * STSetUp).  Diagnostics of the symbolic and numeric phases. */
typedef struct {
  int32_t fronts, levels, max_front, perturbed_pivots;
  int64_t arena_bytes;     /* dense front storage in HBM */
  int64_t factor_entries;  /* complex entries of L and U */
  double flops;            /* real flops of the partial factorizations */
  double symbolic_ms, factor_ms;
} tailor_factor_diag;

int tailor_factor(tailor_ctx *ctx, double sigma_re, double sigma_im, tailor_factor_diag *diag);

/* The same factorization replayed from a HIP graph: at the first call (and
 * whenever the front precision, the GEMM tile choice or the panel kernel has
 * changed since) the norms of A_bc and B_bc, the pivot thresholds derived
 * from them and the shift, and every launch of the factorization are
 * captured from the context's streams; later calls write the shift to the
 * device and relaunch the graph (one host launch instead of some 3,000 at
 * 200x100).  Results are those of tailor_factor.  No reference counterpart
 * (performance only). */
int tailor_factor_graph(tailor_ctx *ctx, double sigma_re, double sigma_im,
                        tailor_factor_diag *diag);

/* Factor `count` distinct contexts of one device (each assembled; any
 * sizes) at the shifts sigmas[2k], sigmas[2k+1]: every context's
 * factorization graph (as tailor_factor_graph) is launched on that
 * context's stream from the calling thread, after the work already queued
 * on it, so the factorizations run concurrently; returns when all are done.
 * *batch_ms (may be NULL): the batch's time.  Errors are reported on the
 * context that failed (ctxs[0] for the batch itself).  No reference
 * counterpart (wavenumber sweeps; the reference runs one case per
 * process). */
int tailor_factor_batch(tailor_ctx *const *ctxs, int32_t count, const double *sigmas,
                        double *batch_ms);

/* Factor like tailor_factor, timing every frontal-update GEMM launch with
 * HIP events on the context stream: total factorization time, summed GEMM
 * time, the GEMMs' algorithmic flops (8 M N K per front and launch) and the
 * number of GEMM launches.  The factorization is bracketed by two empty
 * kernels, k_profile_window_begin / k_profile_window_end, so that a kernel
 * trace can select exactly its launches.  No reference counterpart
 * (measurement only). */
int tailor_profile_factor(tailor_ctx *ctx, double sigma_re, double sigma_im, double *factor_ms,
                          double *gemm_ms, double *gemm_flops, int32_t *gemm_launches);

/* The frontal GEMM time of one factorization at `sigma` by elimination
 * level and by update region, timed as in tailor_profile_factor.
 * level_ms / level_flops receive levels 0 .. max_levels - 1 (zero past the
 * plan's levels; *levels gets the plan's count), region_ms / region_flops
 * (8 entries each, may be NULL) the regions 0 INNER, 2 TRAIL,
 * 4 TRAILHEAD, 5 TRAILREST, 6 ZBLK, 7 PAIRROWS (entries 1 and 3 are always
 * zero: the block rows of the trailing columns and of Z are finished by
 * triangular sweeps, not GEMMs); flops count 8 per complex multiply-add.  No reference counterpart (profiling only). */
int tailor_profile_gemm(tailor_ctx *ctx, double sigma_re, double sigma_im, int32_t max_levels,
                        int32_t *levels, double *level_ms, double *level_flops,
                        double *region_ms, double *region_flops);

/* Autotuning of the frontal GEMM: factor at `sigma` once per configuration
 * c = 2 * shape + order, shape 0: 64x64 complex tile (256 threads, K
 * chunk 16), 1: 64x64 (K chunk 4, three waves per SIMD), 2: 64x64 (K chunk
 * 8, operands loaded straight into a three-stage LDS ring; FP32 fronts use
 * shape 1's tile), 3: 64x64 (K chunk 8, two waves per SIMD); order 0: tiles
 * in launch order, 1: XCD-aware grouped order.  The GEMM launches of each elimination level
 * are timed with HIP events and each level keeps its fastest configuration
 * for later factorizations of this context.  No reference counterpart
 * (performance only; results are unchanged up to rounding). */
typedef struct {
  int32_t levels;          /* elimination levels tuned (at most 64 reported) */
  int32_t order[64];       /* chosen configuration c per level */
  double gemm_ms[8];       /* total GEMM time per configuration c = 0..7 */
  double tuned_gemm_ms;    /* sum over levels of the chosen configurations' times */
} tailor_tune_report;
int tailor_autotune(tailor_ctx *ctx, double sigma_re, double sigma_im, tailor_tune_report *report);
/* Force configuration c (0..7, as above) on every level, replacing any
 * autotuned choice, for later factorizations of this context. */
int tailor_set_gemm_config(tailor_ctx *ctx, int32_t config);

/* Panel factorization kernel.  AUTO: the register-resident left-looking
 * panel where a level's fronts have at most 4096 fully summed rows, the
 * streaming right-looking panel above that; STREAMING: the streaming panel
 * on every level.  Both give the same pivots and factors up to rounding.
 * No reference counterpart (performance only). */
#define TAILOR_PANEL_AUTO 0
#define TAILOR_PANEL_STREAMING 1
int tailor_set_panel_kernel(tailor_ctx *ctx, int32_t kernel);

/* Two-level blocking of the factorization: on (the default), pairs of
 * 128-column blocks share one update of the rows below them, so the
 * trailing-update GEMMs run with K = 256 instead of 128 (the second block's
 * panels then run without look-ahead); off, every block updates on its own.
 * The factors agree to rounding (200x100: frontal GEMM 107 -> 102 ms per
 * factorization).  No reference counterpart (performance only). */
int tailor_set_block_pairs(tailor_ctx *ctx, int32_t on);

/* ------------------------------------------------------------------------
 * Batches: the factorizations of several contexts as one launch sequence.
 * A batch joins the fronts of `count` distinct contexts of one device (each
 * with a case loaded; any grid sizes) into one elimination forest whose
 * level l holds level l of every context (levels counted from the leaves).
 * Each launch of a level (front build, panels, pivot-block inverses, row
 * exchanges, frontal GEMMs, Z) covers the fronts of all contexts, so the
 * latency-bound stretches of one instance run beside the others' and a
 * batch takes as many launches as its deepest context alone.  The factors
 * land in each context's own front storage, as after tailor_factor, and its
 * solves and tailor_solve use them.  The batch refers to its contexts: they
 * must outlive it, keep the structure they had when it was made (a case of
 * another structure needs a new batch), share one front precision at each
 * factorization, and not be used by another thread during a batch call.
 * Errors are reported by tailor_batch_last_error.  No reference counterpart
 * (wavenumber sweeps; the reference runs one case per process). */
typedef struct tailor_batch tailor_batch;
int tailor_batch_create(tailor_batch **batch, tailor_ctx *const *ctxs, int32_t count);
void tailor_batch_destroy(tailor_batch *batch);
const char *tailor_batch_last_error(const tailor_batch *batch);
/* Factor every context k at sigmas[2k] + i sigmas[2k+1] (A_bc, B_bc
 * assembled), after the work already queued on the contexts' streams.
 * use_graph = 1: the launch sequence (per context the norms of A_bc and
 * B_bc, the pivot thresholds and the scatter of K into the fronts, then the
 * joined levels) is captured as a HIP graph at the first call, and again when
 * the front precision or the tile choice changed, and replayed; 0: the same
 * launches directly.  Returns when the factors are complete; *batch_ms (may
 * be NULL): the batch's time on its stream. */
int tailor_batch_factor(tailor_batch *batch, const double *sigmas, int32_t use_graph,
                        double *batch_ms);
/* As tailor_batch_factor, with every context's assembly (tailor_assemble
 * without diagnostics: A_bc and B_bc of the case and wavenumber it holds)
 * launched first in the same sequence, so one graph replay assembles and
 * factors the whole batch.  The contexts need a case loaded, not assembled.
 * A context whose assembly fails validation is reported with the
 * reference's message (its factors are not usable). */
int tailor_batch_assemble_factor(tailor_batch *batch, const double *sigmas, int32_t use_graph,
                                 double *batch_ms);
/* Autotuning of the batch's frontal GEMM tiles (as tailor_autotune, with
 * every context at sigma), kept for its later factorizations; until then the
 * batch uses the choice of its deepest context.  tailor_batch_set_gemm_config
 * forces configuration c (0..7) on every level. */
int tailor_batch_autotune(tailor_batch *batch, double sigma_re, double sigma_im,
                          tailor_tune_report *report);
int tailor_batch_set_gemm_config(tailor_batch *batch, int32_t config);

/* Precision of the factorization used by tailor_factor / tailor_solve.
 * FP64: complex FP64 fronts (the plain form).  MIXED: complex FP32 fronts
 * factored on FP32 MFMA, with FP64 right-hand sides, FP64 residuals and
 * FP64 iterative refinement of every solve (refine_max / refine_tol of the
 * calls); Krylov-Schur itself stays FP64.  No reference counterpart: PETSc
 * factors in the scalar type of the Mat (EigenSolver.cpp:61-67). */
enum { TAILOR_PRECISION_FP64 = 0, TAILOR_PRECISION_MIXED = 1 };
int tailor_set_precision(tailor_ctx *ctx, int32_t precision);

/* x = K^-1 b for natural-order interleaved complex vectors of length 5*N,
 * with up to refine_max steps of iterative refinement towards a relative
 * residual refine_tol; relres/steps report the outcome (may be NULL). */
int tailor_linear_solve(tailor_ctx *ctx, const double *b, double *x, int32_t refine_max,
                        double refine_tol, double *relres, int32_t *steps);

/* x1 = K^-1 b1 and x2 = K^-1 b2 (natural order, as tailor_linear_solve) in
 * one pass over the factors, without refinement: each factor entry is read
 * once for both right-hand sides.  No reference counterpart (one KSPSolve
 * per right-hand side, EigenSolver.cpp:61-67); for callers with two
 * independent right-hand sides at a time. */
int tailor_linear_solve_pair(tailor_ctx *ctx, const double *b1, const double *b2, double *x1,
                             double *x2);

/* x = K^-1 b like tailor_linear_solve, on DEVICE buffers b and x (natural
 * order, interleaved complex, 5*N each), with exactly refine_steps (0..64)
 * steps of FP64 iterative refinement, and without allocating device memory:
 * the call's work vectors live in a static 16 MiB device array of the
 * library, one per GPU, so it serves at most 262,144 unknowns (larger cases
 * get TAILOR_ERR_ARG_OUTOFRANGE).  Blocks until the result is in x.  Every
 * call on a GPU uses that GPU's one array: calls that overlap on one GPU
 * race.  No counterpart in the reference: for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). */
int tailor_linear_solve_noalloc(tailor_ctx *ctx, const double *b, double *x,
                                int32_t refine_steps);

/* Time `reps` unrefined solves x = K^-1 b (device-resident vectors) with
 * HIP events on the context stream: mean ms per solve and the bytes of
 * factor data one solve must read: E, L21 and U12 of every front and the
 * upper triangle of Z. */
int tailor_time_linear_solve(tailor_ctx *ctx, int32_t reps, double *ms_per_solve,
                             double *bytes_per_solve);

/* Time `reps` applications of the shift-invert residual operator
 * y = (A_bc - sigma B_bc) x (the SpMV of every refinement step and residual
 * check; sigma of the last factorization, else 0) with HIP events on the
 * context stream: mean ms per application and its algorithmic bytes
 * (A_bc values and block column indices, B_bc values, x read, y written).
 * No reference counterpart (MatMult inside EPSSolve, EigenSolver.cpp:72,
 * is not timed by the reference). */
int tailor_time_operator(tailor_ctx *ctx, int32_t reps, double *ms_per_apply,
                         double *bytes_per_apply);

/* Settings of EigenSolver in config.yaml (Parser.cpp:62-83). */
typedef struct {
  int32_t nev;             /* NumberOfEigenvalues */
  int32_t ncv;             /* subspace size (at most 128); 0 selects max(2 nev, nev + 15) */
  int32_t maxit;           /* MaximumIterations (restarts) */
  double tol;              /* Tolerance: a pair is accepted when ||A q - lambda B q|| / ||q||
                              and ||A q - lambda B q|| / (|lambda| ||q||) (SLEPc's relative
                              error, EigenSolver.cpp:77-82) are both <= tol */
  double omega_re, omega_im;  /* SearchCenterOmega; lambda_s = Im(omega) - i Re(omega) */
  int32_t refine_max;      /* iterative refinement steps per shift-invert solve */
  double refine_tol;       /* relative residual target of a refined solve */
  int32_t warm_start;      /* 1: start from the Ritz vectors of the previous solve of this
                              context (same structure), e.g. the previous wavenumber of a
                              sweep; 0: the fixed pseudo-random start vector */
  int32_t check_from;      /* 0: convergence is tested at the end of each pass of ncv
                              steps (the reference's behaviour); m > 0: also after every
                              second step from step m on, ending the pass there when the
                              Ritz estimates of the nev wanted pairs meet tol (no
                              counterpart in the reference: saves solves when a warm
                              start converges inside the first pass) */
} tailor_solve_opts;

typedef struct {
  int32_t requested, converged, iterations, solves, refinement_steps, perturbed_pivots;
  double lambda_target_re, lambda_target_im;
  double max_residual;        /* max ||A q - lambda B q|| / ||q|| over the requested pairs */
  double max_relative_error;  /* the same divided by |lambda| */
  double symbolic_ms, factor_ms, solve_ms;
  int64_t fronts, factor_bytes;
  double factor_flops;
} tailor_solve_diag;

/* Factor (if the shift changed) and run Krylov-Schur on the assembled
 * A_bc, B_bc.  TAILOR_ERR_NOT_CONVERGED if fewer than nev pairs meet tol. */
int tailor_solve(tailor_ctx *ctx, const tailor_solve_opts *opts, tailor_solve_diag *diag);

/* Number of Ritz pairs available from the last solve (sorted by distance
 * of lambda to the target). */
int tailor_eigenpair_count(const tailor_ctx *ctx);

/* Pair k: lambda (2 doubles), mode q (natural k_j_dof order, 5N complex,
 * unit 2-norm, largest entry real positive; may be NULL), residual
 * ||A q - lambda B q|| / ||q|| and relative error residual / |lambda|. */
int tailor_get_eigenpair(tailor_ctx *ctx, int32_t k, double *lambda, double *q, double *residual,
                         double *relative_error);

/* The rates device `device` sustains for the two resources the path's
 * kernels are bound by, each measured with a kernel that does nothing else:
 * dense FP64 MFMA (v_mfma_f64_16x16x4f64 back to back on independent
 * accumulators, two waves per SIMD, every CU) in TFLOP/s, and a streaming
 * HBM read of a 2 GiB buffer in GB/s (either pointer may be NULL).  The
 * attainable ceilings beside the published peaks; allocates and frees 2 GiB
 * when read_gbs is requested.  No reference counterpart (measurement only). */
int tailor_measure_peaks(int device, double *fp64_mfma_tflops, double *read_gbs);

/* ------------------------------------------------------------------------
 * Shift-proposal network (no reference counterpart).  A 3 -> 64 -> 64 -> 16
 * perceptron with GELU activations in FP32: normalised (alpha, Re, Ma) in,
 * normalised real/imaginary parts of the 8 leading omega out.  Forward,
 * backward (mean squared error) and Adam run as the library's own kernels on
 * the model's HIP stream.  X is n x 3 and targets/outputs n x 16, row-major.
 * ---------------------------------------------------------------------- */
typedef struct tailor_mlp tailor_mlp;
int tailor_mlp_param_count(void);   /* 5456: W1 b1 W2 b2 W3 b3, W row-major [out][in] */
int tailor_mlp_create(tailor_mlp **m, int device, uint64_t seed);  /* He-uniform init */
void tailor_mlp_destroy(tailor_mlp *m);
const char *tailor_mlp_last_error(const tailor_mlp *m);
int tailor_mlp_get_params(tailor_mlp *m, float *params);
int tailor_mlp_set_params(tailor_mlp *m, const float *params);  /* also resets Adam */
int tailor_mlp_predict(tailor_mlp *m, const float *X, int32_t n, float *Y);
/* Loss (mean over n x 16 entries) and its gradient w.r.t. every parameter. */
int tailor_mlp_gradients(tailor_mlp *m, const float *X, const float *T, int32_t n, float *grad,
                         float *loss);
/* `steps` full-batch Adam steps (beta 0.9 / 0.999, eps 1e-8); loss_history
 * (may be NULL) receives the loss before each step. */
int tailor_mlp_train(tailor_mlp *m, const float *X, const float *T, int32_t n, int32_t steps,
                     float lr, float *loss_history);

#ifdef __cplusplus
}
#endif

#endif /* TAILOR_H */
