// solver.hpp -- state of the factorization and eigen-solve of one context.
#pragma once

#include <complex>
#include <vector>

#include "../../include/tailor.h"
#include "dense.hpp"
#include "lu.hpp"
#include "symbolic.hpp"

namespace tailor {

struct SolverState {
  bool ready = false;     // symbolic analysis + device plan built
  unsigned long long plan_id = 0;  // unique over the process, set when the plan is built
  bool factored = false;
  bool fp32 = false;      // fronts held in complex FP32 (mixed precision)
  Symbolic sym;
  LuPlan plan;
  long long n = 0;
  cplx sigma = 0.0;
  int perturbed = 0;
  double factor_ms = 0.0, symbolic_ms = 0.0;
  int *d_order = nullptr, *d_pos = nullptr;
  int *d_xcol = nullptr;  // per stored block: its column node in elimination order (SpMV)
  double *V = nullptr, *work = nullptr, *ritz = nullptr, *partials = nullptr, *small = nullptr;
  size_t V_cap = 0, work_cap = 0, ritz_cap = 0, partials_cap = 0, small_cap = 0;
  double *hsteps = nullptr;  // Gram-Schmidt coefficients of one Krylov pass (device)
  size_t hsteps_cap = 0;
  // HIP graph of the factorization (norms -> pivot thresholds -> LU),
  // captured once per plan, front precision and tile choice, replayed with
  // any shift (written to the plan's parameter block before each replay)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  unsigned long long graph_id = 0;  // unique over the process
  bool graph_fp32 = false, graph_streaming = false, graph_pairs = true;
  std::vector<int> graph_shape, graph_order;
  // results
  int nritz = 0;
  int ritz_ld = 0;  // columns per plane of the Ritz vector buffer
  std::vector<cplx> lambda;
  std::vector<double> resid, relerr;
};

void solver_release(SolverState *S);
int solver_prepare(tailor_ctx *ctx);
int solver_factor(tailor_ctx *ctx, cplx sigma, FactorTiming *timing = nullptr);
// The same factorization replayed from the context's HIP graph (captured at
// the first call and whenever the precision or tile choice changed).
int solver_factor_graph(tailor_ctx *ctx, cplx sigma);
// The factorizations of `count` contexts of one device: each context's
// graph launched on its own stream from this thread, so they run
// concurrently; *ms: the batch's event time.
int solver_factor_batch(tailor_ctx *const *ctxs, int count, const cplx *sigmas, double *ms);
// ctx now holds the factors of K = A_bc - sigma B_bc made by a batch
// factorization (batch.cpp) in its front storage.
void solver_mark_factored(tailor_ctx *ctx, cplx sigma, bool fp32, int perturbed, double ms);
// Factor once per GEMM tile order with per-level timing and keep, for each
// elimination level, the order with the smallest GEMM time.
int solver_autotune(tailor_ctx *ctx, cplx sigma, tailor_tune_report *report);
int solver_linear_solve(tailor_ctx *ctx, const double *b_nat, double *x_nat, int refine_max,
                        double refine_tol, double *relres, int *steps);
int solver_krylov_schur(tailor_ctx *ctx, const tailor_solve_opts &o, tailor_solve_diag &diag);
int solver_time_solves(tailor_ctx *ctx, int reps, double *ms);
int solver_time_kmul(tailor_ctx *ctx, int reps, double *ms);
// x1 = K^-1 b1, x2 = K^-1 b2 (natural order, host vectors) in one pass over
// the factors, without refinement
int solver_linear_solve_pair(tailor_ctx *ctx, const double *b1_nat, const double *b2_nat,
                             double *x1_nat, double *x2_nat);
int solver_norm2(tailor_ctx *ctx, const double *xr, const double *xi, double *out);
int solver_dots(tailor_ctx *ctx, const double *Vr, const double *Vi, long long ldv, int ncols,
                const double *wr, const double *wi, cplx *h);

}  // namespace tailor
