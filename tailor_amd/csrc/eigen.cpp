// eigen.cpp -- shift-invert Krylov-Schur for A_bc q = lambda B_bc q on the
// GPU, replacing SLEPc's EPSKRYLOVSCHUR + STSINVERT + PCLU path
// (EigenSolver.cpp:31-89).
//
// Operator: Op = (A - sigma B)^-1 B with sigma = lambda_s = Im(omega_s) -
// i Re(omega_s) (EigenSolver.cpp:46-48); Ritz values theta of Op map to
// lambda = sigma + 1/theta, and the wanted pairs are those of largest
// |theta| (nearest sigma, EPS_TARGET_MAGNITUDE).  Krylov-Schur restarts
// (G. W. Stewart, SIAM J. Matrix Anal. Appl. 23 (2001) 601-614): Arnoldi with
// classical Gram-Schmidt and one reorthogonalisation, complex Schur form of
// the projected matrix, sorting by |theta|, restart keeping the converged
// pairs plus half of the rest (SLEPc's default restart ratio 0.5).  A pair
// is accepted when both its true residual ||A q - lambda B q|| / ||q|| and
// its relative error ||A q - lambda B q|| / (|lambda| ||q||) meet tol: the
// second is SLEPc's EPS_CONV_REL / EPS_ERROR_RELATIVE measure, which the
// reference stores as /spectrum/residual (EigenSolver.cpp:77-82).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>

#include "dense.hpp"
#include "internal.hpp"
#include "lu.hpp"
#include "solver.hpp"
#include "vec.hpp"

namespace tailor {

#define HIP_OK(expr)                                     \
  do {                                                   \
    hipError_t e_ = (expr);                              \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
  } while (0)

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

template <class T>
hipError_t grow(T **p, size_t &cap, size_t need) {
  if (*p && cap >= need) return hipSuccess;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  cap = 0;
  hipError_t e = hipMalloc(reinterpret_cast<void **>(p), need * sizeof(T));
  if (e == hipSuccess) cap = need;
  return e;
}

}  // namespace

namespace {

void release_graphs(SolverState *S) {
  if (S->graph_exec) (void)hipGraphExecDestroy(S->graph_exec);
  if (S->graph) (void)hipGraphDestroy(S->graph);
  S->graph_exec = nullptr;
  S->graph = nullptr;
  S->graph_id = 0;
}

std::atomic<unsigned long long> g_next_graph_id{1};
std::atomic<unsigned long long> g_next_plan_id{1};

}  // namespace

void solver_release(SolverState *S) {
  if (!S) return;
  release_graphs(S);
  lu_plan_destroy(S->plan);
  double *bufs[] = {S->V, S->work, S->ritz, S->partials, S->small, S->hsteps};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  int *ib[] = {S->d_order, S->d_pos, S->d_xcol};
  for (int *b : ib)
    if (b) (void)hipFree(b);
  delete S;
}

// Symbolic analysis and device plan for the loaded case (once per case).
int solver_prepare(tailor_ctx *ctx) {
  if (!ctx->solver) ctx->solver = new SolverState();
  SolverState &S = *ctx->solver;
  if (S.ready) return TAILOR_OK;
  const double t0 = now_ms();
  const auto &d = ctx->dims;
  const std::vector<int> &xoff = ctx->h_xoff, &eoff = ctx->h_eoff;
  try {
    S.sym = analyse(d.ny, d.nz, d.ssx, d.sse, xoff, eoff, 64);
  } catch (const std::exception &ex) {
    return fail(ctx, TAILOR_ERR_PLIB, std::string("symbolic analysis failed: ") + ex.what());
  }
  hipError_t e = lu_plan_create(S.sym, ctx->h_colidx, d.ssx * d.sse, S.plan, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "lu_plan_create");
  std::vector<int> order(S.sym.order), pos(S.sym.pos);
  HIP_OK(hipMalloc(reinterpret_cast<void **>(&S.d_order), order.size() * sizeof(int)));
  HIP_OK(hipMalloc(reinterpret_cast<void **>(&S.d_pos), pos.size() * sizeof(int)));
  HIP_OK(hipMemcpyAsync(S.d_order, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice,
                       ctx->stream));
  HIP_OK(hipMemcpyAsync(S.d_pos, pos.data(), pos.size() * sizeof(int), hipMemcpyHostToDevice,
                       ctx->stream));
  std::vector<int> xcol(ctx->h_colidx.size());
  for (size_t k = 0; k < xcol.size(); ++k) xcol[k] = pos[ctx->h_colidx[k]];
  HIP_OK(hipMalloc(reinterpret_cast<void **>(&S.d_xcol), std::max<size_t>(1, xcol.size()) * sizeof(int)));
  HIP_OK(hipMemcpyAsync(S.d_xcol, xcol.data(), xcol.size() * sizeof(int), hipMemcpyHostToDevice,
                       ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  S.n = 5LL * d.nodes;
  S.symbolic_ms = now_ms() - t0;
  S.plan_id = g_next_plan_id.fetch_add(1);
  S.ready = true;
  return TAILOR_OK;
}

int solver_factor(tailor_ctx *ctx, cplx sigma, FactorTiming *timing) {
  if (!ctx->assembled || ctx->vals_kind != TAILOR_OP_A || ctx->mvals_kind != TAILOR_OP_B)
    return fail(ctx, TAILOR_ERR_ARG_WRONGSTATE,
                "Boundary-constrained matrices are required before EPSSolve");
  if (int st = solver_prepare(ctx)) return st;
  SolverState &S = *ctx->solver;
  const auto &d = ctx->dims;
  LuInputs in;
  in.nnzb = d.nnzb;
  in.per = d.ssx * d.sse;
  in.avals = ctx->d_vals;
  in.colidx = ctx->d_colidx;
  in.bvals = ctx->d_mvals;
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipEventRecord(e0, ctx->stream));
  // sigma, then the pivot threshold and static perturbation scaled by the
  // RMS row norm of K, all on the device
  const bool fp32 = ctx->precision == TAILOR_PRECISION_MIXED;
  HIP_OK(lu_set_sigma(S.plan, sigma.real(), sigma.imag(), ctx->stream));
  HIP_OK(lu_factor_params(S.plan, in, fp32, ctx->d_partials, ctx->stream));
  HIP_OK(lu_factor(S.plan, in, fp32, ctx->stream, timing));
  HIP_OK(hipEventRecord(e1, ctx->stream));
  HIP_OK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  int counters[8];
  HIP_OK(hipMemcpyAsync(counters, S.plan.d_counters, sizeof counters, hipMemcpyDeviceToHost,
                       ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  S.perturbed = counters[0];
  S.factor_ms = ms;
  S.sigma = sigma;
  S.fp32 = fp32;
  S.factored = true;
  return TAILOR_OK;
}

namespace {

LuInputs factor_inputs(const tailor_ctx *ctx) {
  const auto &d = ctx->dims;
  LuInputs in;
  in.nnzb = d.nnzb;
  in.per = d.ssx * d.sse;
  in.avals = ctx->d_vals;
  in.colidx = ctx->d_colidx;
  in.bvals = ctx->d_mvals;
  return in;
}

// Capture the factorization of ctx as a HIP graph unless the one held was
// captured for the same front precision, tile shapes and orders and panel
// kernel.  The captured work: the norms of A_bc, B_bc and the pivot
// thresholds (from the shift in the parameter block), then every launch of
// lu_factor on the context stream and the plan's look-ahead stream.
int graph_ready(tailor_ctx *ctx) {
  SolverState &S = *ctx->solver;
  const bool fp32 = ctx->precision == TAILOR_PRECISION_MIXED;
  if (S.graph_exec && S.graph_fp32 == fp32 && S.graph_shape == S.plan.gemm_shape &&
      S.graph_order == S.plan.gemm_order && S.graph_streaming == S.plan.panel_streaming &&
      S.graph_pairs == S.plan.pair_blocks)
    return TAILOR_OK;
  release_graphs(&S);
  const LuInputs in = factor_inputs(ctx);
  // the look-ahead stream is made outside the capture
  HIP_OK(lu_plan_side(S.plan));
  HIP_OK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = lu_factor_params(S.plan, in, fp32, ctx->d_partials, ctx->stream);
  if (e == hipSuccess) e = lu_factor(S.plan, in, fp32, ctx->stream, nullptr);
  hipGraph_t g = nullptr;
  const hipError_t e_end = hipStreamEndCapture(ctx->stream, &g);
  if (e == hipSuccess) e = e_end;
  hipGraphExec_t x = nullptr;
  if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_fail(ctx, e, "capture of the factorization graph");
  }
  S.graph = g;
  S.graph_exec = x;
  S.graph_id = g_next_graph_id.fetch_add(1);
  S.graph_fp32 = fp32;
  S.graph_shape = S.plan.gemm_shape;
  S.graph_order = S.plan.gemm_order;
  S.graph_streaming = S.plan.panel_streaming;
  S.graph_pairs = S.plan.pair_blocks;
  return TAILOR_OK;
}

int factor_checks(tailor_ctx *ctx) {
  if (!ctx->assembled || ctx->vals_kind != TAILOR_OP_A || ctx->mvals_kind != TAILOR_OP_B)
    return fail(ctx, TAILOR_ERR_ARG_WRONGSTATE,
                "Boundary-constrained matrices are required before EPSSolve");
  return solver_prepare(ctx);
}

// After a factorization of ctx at sigma has completed on the device.
int factor_done(tailor_ctx *ctx, cplx sigma, double ms) {
  SolverState &S = *ctx->solver;
  int counters[8];
  HIP_OK(hipMemcpyAsync(counters, S.plan.d_counters, sizeof counters, hipMemcpyDeviceToHost,
                       ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  S.perturbed = counters[0];
  S.factor_ms = ms;
  S.sigma = sigma;
  S.fp32 = ctx->precision == TAILOR_PRECISION_MIXED;
  S.factored = true;
  return TAILOR_OK;
}

}  // namespace

void solver_mark_factored(tailor_ctx *ctx, cplx sigma, bool fp32, int perturbed, double ms) {
  SolverState &S = *ctx->solver;
  S.perturbed = perturbed;
  S.factor_ms = ms;
  S.sigma = sigma;
  S.fp32 = fp32;
  S.factored = true;
}

int solver_factor_graph(tailor_ctx *ctx, cplx sigma) {
  if (int st = factor_checks(ctx)) return st;
  if (int st = graph_ready(ctx)) return st;
  SolverState &S = *ctx->solver;
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipEventRecord(e0, ctx->stream));
  HIP_OK(lu_set_sigma(S.plan, sigma.real(), sigma.imag(), ctx->stream));
  HIP_OK(hipGraphLaunch(S.graph_exec, ctx->stream));
  HIP_OK(hipEventRecord(e1, ctx->stream));
  HIP_OK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return factor_done(ctx, sigma, ms);
}

int solver_factor_batch(tailor_ctx *const *ctxs, int count, const cplx *sigmas, double *ms_out) {
  for (int k = 0; k < count; ++k) {
    if (int st = factor_checks(ctxs[k])) return st;
    if (int st = graph_ready(ctxs[k])) return st;
  }
  // every context's graph on its own stream, all launched from this thread:
  // the streams run concurrently (one launch of a graph holding the
  // contexts' graphs as parallel child graphs measured no faster than the
  // factorizations one after another); the batch's time runs from the first
  // context's start event to the last completion
  tailor_ctx *ctx = ctxs[0];  // HIP_OK reports on the batch's first context
  std::vector<hipEvent_t> done(count, nullptr);
  hipEvent_t e0;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventRecord(e0, ctx->stream));
  for (int k = 0; k < count; ++k) {
    tailor_ctx *c = ctxs[k];
    SolverState &S = *c->solver;
    if (k > 0) HIP_OK(hipStreamWaitEvent(c->stream, e0, 0));
    HIP_OK(lu_set_sigma(S.plan, sigmas[k].real(), sigmas[k].imag(), c->stream));
    HIP_OK(hipGraphLaunch(S.graph_exec, c->stream));
    HIP_OK(hipEventCreate(&done[k]));
    HIP_OK(hipEventRecord(done[k], c->stream));
  }
  float ms = 0.f;
  for (int k = 0; k < count; ++k) {
    HIP_OK(hipEventSynchronize(done[k]));
    float t = 0.f;
    HIP_OK(hipEventElapsedTime(&t, e0, done[k]));
    ms = std::max(ms, t);
  }
  (void)hipEventDestroy(e0);
  for (hipEvent_t ev : done) (void)hipEventDestroy(ev);
  if (ms_out) *ms_out = ms;
  for (int k = 0; k < count; ++k)
    if (int st = factor_done(ctxs[k], sigmas[k], ms)) return st;
  return TAILOR_OK;
}

int solver_autotune(tailor_ctx *ctx, cplx sigma, tailor_tune_report *report) {
  if (int st = solver_prepare(ctx)) return st;
  SolverState &S = *ctx->solver;
  const size_t nl = S.plan.levels.size();
  // every (tile shape, tile order) configuration, applied to all levels
  const int nconf = kGemmShapes * kOrderCount;
  std::vector<std::vector<double>> t(nconf);
  std::vector<double> total(nconf, 0.0);
  for (int c = 0; c < nconf; ++c) {
    S.plan.gemm_shape.assign(nl, c / kOrderCount);
    S.plan.gemm_order.assign(nl, c % kOrderCount);
    FactorTiming tm;
    if (int st = solver_factor(ctx, sigma, &tm)) return st;
    t[c] = tm.level_gemm_ms;
    total[c] = tm.gemm_ms;
  }
  double tuned = 0.0;
  for (size_t l = 0; l < nl; ++l) {
    int best = 0;
    for (int c = 1; c < nconf; ++c)
      if (t[c][l] < t[best][l]) best = c;
    S.plan.gemm_shape[l] = best / kOrderCount;
    S.plan.gemm_order[l] = best % kOrderCount;
    tuned += t[best][l];
  }
  if (report) {
    report->levels = static_cast<int32_t>(nl);
    for (size_t l = 0; l < nl && l < 64; ++l)
      report->order[l] = S.plan.gemm_shape[l] * kOrderCount + S.plan.gemm_order[l];
    for (int c = 0; c < nconf && c < 8; ++c) report->gemm_ms[c] = total[c];
    report->tuned_gemm_ms = tuned;
  }
  return TAILOR_OK;
}

namespace {

// y = K^-1 b with optional iterative refinement (b, y: elimination order,
// split planes of length n; scratch: 3 vectors).
int refined_solve(tailor_ctx *ctx, const double *br, const double *bi, double *yr, double *yi,
                  int refine_max, double refine_tol, int *steps, double *relres) {
  SolverState &S = *ctx->solver;
  const long long n = S.n;
  double *rr = S.work + 6 * n, *ri = rr + n;  // residual
  double *dr = S.work + 8 * n, *di = dr + n;  // correction
  HIP_OK(lu_solve(S.plan, br, bi, yr, yi, S.fp32, ctx->stream));
  if (steps) *steps = 0;
  if (refine_max <= 0 && !relres) return TAILOR_OK;
  const auto &d = ctx->dims;
  double bnorm = -1.0;
  for (int it = 0; it <= refine_max; ++it) {
    // r = b - K y
    HIP_OK(vec_kmul(ctx->stream, d.nodes, d.ssx * d.sse, ctx->d_vals, S.d_xcol, ctx->d_mvals,
                    S.d_pos, S.sigma.real(), S.sigma.imag(), yr, yi, rr, ri));
    HIP_OK(vec_axpby(ctx->stream, n, 1.0, 0.0, br, bi, -1.0, 0.0, rr, ri));
    double nr = 0.0, nb = 0.0;
    if (int st = solver_norm2(ctx, rr, ri, &nr)) return st;
    if (bnorm < 0.0) {
      if (int st = solver_norm2(ctx, br, bi, &nb)) return st;
      bnorm = nb;
    }
    const double rel = bnorm > 0.0 ? nr / bnorm : nr;
    if (relres) *relres = rel;
    if (it == refine_max || rel <= refine_tol) break;
    HIP_OK(lu_solve(S.plan, rr, ri, dr, di, S.fp32, ctx->stream));
    HIP_OK(vec_axpby(ctx->stream, n, 1.0, 0.0, dr, di, 1.0, 0.0, yr, yi));
    if (steps) ++*steps;
  }
  return TAILOR_OK;
}

}  // namespace

int solver_norm2(tailor_ctx *ctx, const double *xr, const double *xi, double *out) {
  SolverState &S = *ctx->solver;
  std::vector<cplx> h(1);
  if (int st = solver_dots(ctx, xr, xi, S.n, 1, xr, xi, h.data())) return st;
  *out = std::sqrt(std::max(0.0, h[0].real()));
  return TAILOR_OK;
}

int solver_dots(tailor_ctx *ctx, const double *Vr, const double *Vi, long long ldv, int ncols,
                const double *wr, const double *wi, cplx *h) {
  SolverState &S = *ctx->solver;
  HIP_OK(vec_dots(ctx->stream, S.n, ncols, Vr, Vi, ldv, wr, wi, S.partials, kDotBlocks));
  std::vector<double> part(2 * static_cast<size_t>(kDotBlocks) * ncols);
  HIP_OK(hipMemcpyAsync(part.data(), S.partials, part.size() * sizeof(double), hipMemcpyDeviceToHost,
                        ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  for (int j = 0; j < ncols; ++j) {
    double sr = 0.0, si = 0.0;
    for (int b = 0; b < kDotBlocks; ++b) {
      sr += part[2 * (static_cast<size_t>(b) * ncols + j)];
      si += part[2 * (static_cast<size_t>(b) * ncols + j) + 1];
    }
    h[j] = cplx(sr, si);
  }
  return TAILOR_OK;
}

int solver_linear_solve(tailor_ctx *ctx, const double *b_nat, double *x_nat, int refine_max,
                        double refine_tol, double *relres, int *steps) {
  SolverState &S = *ctx->solver;
  const long long n = S.n;
  HIP_OK(grow(&S.work, S.work_cap, static_cast<size_t>(12 * n)));
  HIP_OK(grow(&S.partials, S.partials_cap, static_cast<size_t>(2 * kDotBlocks * 64)));
  double *nat = S.work + 10 * n;  // interleaved natural vector (2n)
  double *br = S.work, *bi = br + n, *xr = S.work + 2 * n, *xi = xr + n;
  HIP_OK(hipMemcpyAsync(nat, b_nat, 2 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIP_OK(vec_nat_to_elim(ctx->stream, n, S.plan.d_perm, nat, br, bi));
  if (int st = refined_solve(ctx, br, bi, xr, xi, refine_max, refine_tol, steps, relres)) return st;
  HIP_OK(vec_elim_to_nat(ctx->stream, n, S.plan.d_perm, xr, xi, nat));
  HIP_OK(hipMemcpyAsync(x_nat, nat, 2 * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return TAILOR_OK;
}

int solver_linear_solve_pair(tailor_ctx *ctx, const double *b1_nat, const double *b2_nat,
                             double *x1_nat, double *x2_nat) {
  SolverState &S = *ctx->solver;
  const long long n = S.n;
  HIP_OK(grow(&S.work, S.work_cap, static_cast<size_t>(12 * n)));
  double *nat = S.work + 10 * n;  // interleaved natural vector (2n)
  double *b1r = S.work, *b1i = b1r + n, *b2r = S.work + 2 * n, *b2i = b2r + n;
  double *x1r = S.work + 4 * n, *x1i = x1r + n, *x2r = S.work + 6 * n, *x2i = x2r + n;
  HIP_OK(hipMemcpyAsync(nat, b1_nat, 2 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIP_OK(vec_nat_to_elim(ctx->stream, n, S.plan.d_perm, nat, b1r, b1i));
  HIP_OK(hipMemcpyAsync(nat, b2_nat, 2 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  HIP_OK(vec_nat_to_elim(ctx->stream, n, S.plan.d_perm, nat, b2r, b2i));
  HIP_OK(lu_solve2(S.plan, b1r, b1i, b2r, b2i, x1r, x1i, x2r, x2i, S.fp32, ctx->stream));
  HIP_OK(vec_elim_to_nat(ctx->stream, n, S.plan.d_perm, x1r, x1i, nat));
  HIP_OK(hipMemcpyAsync(x1_nat, nat, 2 * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  HIP_OK(vec_elim_to_nat(ctx->stream, n, S.plan.d_perm, x2r, x2i, nat));
  HIP_OK(hipMemcpyAsync(x2_nat, nat, 2 * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(hipStreamSynchronize(ctx->stream));
  return TAILOR_OK;
}

int solver_time_solves(tailor_ctx *ctx, int reps, double *ms) {
  SolverState &S = *ctx->solver;
  const long long n = S.n;
  HIP_OK(grow(&S.work, S.work_cap, static_cast<size_t>(12 * n)));
  double *br = S.work, *bi = br + n, *xr = S.work + 2 * n, *xi = xr + n;
  HIP_OK(lu_fill(ctx->stream, br, 2 * n, 1.0));
  HIP_OK(lu_solve(S.plan, br, bi, xr, xi, S.fp32, ctx->stream));  // warm-up
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipEventRecord(e0, ctx->stream));
  for (int r = 0; r < reps; ++r) HIP_OK(lu_solve(S.plan, br, bi, xr, xi, S.fp32, ctx->stream));
  HIP_OK(hipEventRecord(e1, ctx->stream));
  HIP_OK(hipEventSynchronize(e1));
  float t = 0.f;
  HIP_OK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = static_cast<double>(t) / reps;
  return TAILOR_OK;
}

// y = (A - sigma B) x on the assembled operator, `reps` launches timed with
// HIP events on the context stream (x = 1, elimination order)
int solver_time_kmul(tailor_ctx *ctx, int reps, double *ms) {
  if (int st = solver_prepare(ctx)) return st;
  SolverState &S = *ctx->solver;
  const long long n = S.n;
  const auto &d = ctx->dims;
  HIP_OK(grow(&S.work, S.work_cap, static_cast<size_t>(12 * n)));
  double *xr = S.work, *xi = xr + n, *yr = S.work + 2 * n, *yi = yr + n;
  HIP_OK(lu_fill(ctx->stream, xr, 2 * n, 1.0));
  auto launch = [&]() {
    return vec_kmul(ctx->stream, d.nodes, d.ssx * d.sse, ctx->d_vals, S.d_xcol, ctx->d_mvals,
                    S.d_pos, S.sigma.real(), S.sigma.imag(), xr, xi, yr, yi);
  };
  HIP_OK(launch());  // warm-up
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  HIP_OK(hipEventRecord(e0, ctx->stream));
  for (int r = 0; r < reps; ++r) HIP_OK(launch());
  HIP_OK(hipEventRecord(e1, ctx->stream));
  HIP_OK(hipEventSynchronize(e1));
  float t = 0.f;
  HIP_OK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = static_cast<double>(t) / reps;
  return TAILOR_OK;
}

// ----------------------------------------------------------- Krylov-Schur

int solver_krylov_schur(tailor_ctx *ctx, const tailor_solve_opts &o, tailor_solve_diag &diag) {
  SolverState &S = *ctx->solver;
  const long long n = S.n;
  const int nev = o.nev;
  int ncv = o.ncv > 0 ? o.ncv : std::max(2 * nev, nev + 15);
  ncv = std::min<long long>(ncv, n - 1);
  if (ncv > 128) ncv = 128;
  if (nev <= 0 || nev >= ncv) return fail(ctx, TAILOR_ERR_USER_INPUT, "invalid nev / ncv");
  // the previous solve's Ritz vectors survive until the start vector is built
  const int prev_ritz = (o.warm_start && S.ritz && S.nritz > 0) ? std::min(S.nritz, nev) : 0;
  const int prev_ld = S.ritz_ld;
  HIP_OK(grow(&S.V, S.V_cap, static_cast<size_t>(2 * n * (ncv + 1))));
  HIP_OK(grow(&S.work, S.work_cap, static_cast<size_t>(12 * n)));
  // dot partials for up to ncv + 1 columns, and the small host-computed
  // matrices (QY: ncv x ncheck, Q_l: ncv x l, both below (ncv + 1)^2)
  HIP_OK(grow(&S.partials, S.partials_cap, static_cast<size_t>(2 * kDotBlocks * (ncv + 1))));
  HIP_OK(grow(&S.small, S.small_cap, static_cast<size_t>(2 * (ncv + 1) * (ncv + 1))));
  // per Arnoldi step j: h (j + 1 complex), the reorthogonalisation's h2,
  // and |w|^2, at hsteps + j * hstride (read back once per pass)
  const size_t hstride = 4 * static_cast<size_t>(ncv + 1) + 2;
  HIP_OK(grow(&S.hsteps, S.hsteps_cap, hstride * ncv));
  std::vector<double> hhost(hstride * ncv);
  double *Vr = S.V, *Vi = S.V + n * (ncv + 1);
  auto col_r = [&](int j) { return Vr + static_cast<long long>(j) * n; };
  auto col_i = [&](int j) { return Vi + static_cast<long long>(j) * n; };
  double *wr = S.work, *wi = wr + n;     // Arnoldi vector
  double *tr = S.work + 2 * n, *ti = tr + n;  // B v
  double *yr = S.work + 4 * n, *yi = yr + n;  // residual checks
  const auto &d = ctx->dims;

  // deterministic start vector (SplitMix64 stream, elimination order)
  {
    std::vector<double> v0(2 * n);
    unsigned long long st = 0x5eed1234ULL;
    for (long long k = 0; k < 2 * n; ++k) {
      st += 0x9E3779B97F4A7C15ULL;
      unsigned long long z = st;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      z ^= z >> 31;
      v0[k] = static_cast<double>(z >> 11) * 0x1.0p-53 - 0.5;
    }
    HIP_OK(hipMemcpyAsync(col_r(0), v0.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIP_OK(hipMemcpyAsync(col_i(0), v0.data() + n, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    double nv;
    if (int s2 = solver_norm2(ctx, col_r(0), col_i(0), &nv)) return s2;
    HIP_OK(vec_axpby(ctx->stream, n, 0.0, 0.0, col_r(0), col_i(0), 1.0 / nv, 0.0, col_r(0), col_i(0)));
    // warm start: v0 = sum of the previous unit Ritz vectors + 1e-2 * random
    if (prev_ritz > 0) {
      HIP_OK(vec_axpby(ctx->stream, n, 0.0, 0.0, col_r(0), col_i(0), 1e-2, 0.0, col_r(0), col_i(0)));
      for (int i = 0; i < prev_ritz; ++i) {
        const double *Rr = S.ritz + i * n, *Ri = S.ritz + n * prev_ld + i * n;
        double nr;
        if (int s2 = solver_norm2(ctx, Rr, Ri, &nr)) return s2;
        if (!(nr > 0.0)) continue;
        HIP_OK(vec_axpby(ctx->stream, n, 1.0 / nr, 0.0, Rr, Ri, 1.0, 0.0, col_r(0), col_i(0)));
      }
      if (int s2 = solver_norm2(ctx, col_r(0), col_i(0), &nv)) return s2;
      HIP_OK(vec_axpby(ctx->stream, n, 0.0, 0.0, col_r(0), col_i(0), 1.0 / nv, 0.0, col_r(0), col_i(0)));
    }
  }
  S.nritz = 0;
  HIP_OK(grow(&S.ritz, S.ritz_cap, static_cast<size_t>(2 * n * ncv)));

  CMat Sm(ncv + 1, ncv);  // projected matrix (rows 0..ncv)
  int nout_pass = ncv;     // columns of the last pass
  int k = 0;
  int solves = 0, refine_total = 0, restarts = 0;
  std::vector<cplx> theta;
  std::vector<cplx> b;
  CMat T, Q;
  int nconv = 0;
  std::vector<double> resid(ncv, 1.0), relerr(ncv, 1.0);
  const double t0 = now_ms();
  const double tol = o.tol;
  double best_res = 1e300;
  int stagnant = 0;
  int reason = -1;  // SLEPc's EPS_DIVERGED_ITS unless the residuals stagnate
  // read back the coefficients of Arnoldi steps [from, to) into Sm
  auto fill_steps = [&](int from, int to) -> int {
    if (to <= from) return TAILOR_OK;
    HIP_OK(hipMemcpyAsync(hhost.data() + hstride * from, S.hsteps + hstride * from,
                          hstride * (to - from) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(hipStreamSynchronize(ctx->stream));
    for (int j = from; j < to; ++j) {
      const double *hj = hhost.data() + hstride * j;
      const double *h2j = hj + 2 * (ncv + 1), *w2j = hj + 4 * (ncv + 1);
      for (int i = 0; i <= j; ++i)
        Sm(i, j) = cplx(hj[2 * i], hj[2 * i + 1]) + cplx(h2j[2 * i], h2j[2 * i + 1]);
      const double beta = std::sqrt(std::max(0.0, w2j[0]));
      Sm(j + 1, j) = beta;
      if (beta == 0.0) return fail(ctx, TAILOR_ERR_NOT_CONVERGED, "Krylov subspace broke down");
    }
    return TAILOR_OK;
  };
  // Schur form of the leading m x m projected matrix sorted by |theta|
  // descending, its coupling row b = S(m, :) Q and the Ritz values
  auto project = [&](int m) -> int {
    T = CMat(m, m);
    for (int jj = 0; jj < m; ++jj)
      for (int ii = 0; ii < m; ++ii) T(ii, jj) = Sm(ii, jj);
    if (!schur_decompose(T, Q)) return fail(ctx, TAILOR_ERR_NOT_CONVERGED, "projected Schur form failed");
    for (int pos = 0; pos < m; ++pos) {
      int best = pos;
      for (int q = pos + 1; q < m; ++q)
        if (std::abs(T(q, q)) > std::abs(T(best, best))) best = q;
      if (best != pos) {
        std::vector<bool> sel(m, false);
        // move `best` to `pos`: select it together with the already placed ones
        for (int q = 0; q < pos; ++q) sel[q] = true;
        sel[best] = true;
        schur_reorder(T, Q, sel);
      }
    }
    b.assign(m, 0.0);
    for (int jj = 0; jj < m; ++jj) {
      cplx s = 0.0;
      for (int q = 0; q < m; ++q) s += Sm(m, q) * Q(q, jj);
      b[jj] = s;
    }
    theta.assign(m, 0.0);
    for (int i = 0; i < m; ++i) theta[i] = T(i, i);
    return TAILOR_OK;
  };
  // Ritz estimate |b y_i| / |theta_i| of pair i of the current projection
  auto estimate = [&](int i, const std::vector<cplx> &y) {
    cplx s = 0.0;
    for (size_t q = 0; q < y.size(); ++q) s += b[q] * y[q];
    return std::abs(s) / std::abs(theta[i]);
  };
  for (restarts = 0; restarts < o.maxit; ++restarts) {
    int m = ncv;       // columns of this pass
    int filled = k;    // columns of Sm holding this pass's coefficients
    bool projected = false;
    for (int j = k; j < ncv; ++j) {
      // w = Op v_j
      HIP_OK(vec_bmul(ctx->stream, d.nodes, S.d_order, ctx->d_mvals, col_r(j), col_i(j), tr, ti));
      int steps = 0;
      if (int st = refined_solve(ctx, tr, ti, wr, wi, o.refine_max, o.refine_tol, &steps, nullptr))
        return st;
      ++solves;
      refine_total += steps;
      // CGS with one reorthogonalisation, entirely on the device: the
      // coefficients stay in hsteps and the host reads them after the pass
      double *hj = S.hsteps + hstride * j;
      double *h2j = hj + 2 * (ncv + 1), *w2j = hj + 4 * (ncv + 1);
      HIP_OK(vec_dots_device(ctx->stream, n, j + 1, Vr, Vi, n, wr, wi, S.partials, kDotBlocks, hj));
      HIP_OK(vec_update(ctx->stream, n, j + 1, Vr, Vi, n, hj, wr, wi));
      HIP_OK(vec_dots_device(ctx->stream, n, j + 1, Vr, Vi, n, wr, wi, S.partials, kDotBlocks, h2j));
      HIP_OK(vec_update(ctx->stream, n, j + 1, Vr, Vi, n, h2j, wr, wi));
      HIP_OK(vec_dots_device(ctx->stream, n, 1, wr, wi, n, wr, wi, S.partials, kDotBlocks, w2j));
      HIP_OK(vec_scale_inv_norm(ctx->stream, n, w2j, wr, wi, col_r(j + 1), col_i(j + 1)));
      // early test (check_from > 0): every second step from check_from on,
      // end the pass here when the wanted pairs' Ritz estimates meet tol
      const int cols = j + 1;
      if (o.check_from > 0 && cols >= o.check_from && cols > nev + 1 && cols < ncv &&
          (cols - o.check_from) % 2 == 0) {
        if (int st = fill_steps(filled, cols)) return st;
        filled = cols;
        if (int st = project(cols)) return st;
        bool ok = true;
        for (int i = 0; i < nev && ok; ++i) ok = estimate(i, triangular_eigenvector(T, i)) <= tol;
        if (ok) {
          m = cols;
          projected = true;
          break;
        }
      }
    }
    if (int st = fill_steps(filled, m)) return st;
    if (!projected)
      if (int st = project(m)) return st;
    // Ritz estimates and true residuals of the leading pairs
    nconv = 0;
    const int ncheck = std::min(m - 1, nev + 2);
    std::vector<std::vector<cplx>> Y(ncheck);
    bool all_est_ok = true;
    for (int i = 0; i < ncheck; ++i) {
      Y[i] = triangular_eigenvector(T, i);
      if (i < nev && estimate(i, Y[i]) > tol) all_est_ok = false;
    }
    if (all_est_ok || restarts == o.maxit - 1) {
      // Ritz vectors x_i = V Q y_i and their true residuals
      CMat QY(m, ncheck);
      for (int i = 0; i < ncheck; ++i)
        for (int q = 0; q < m; ++q) {
          cplx s = 0.0;
          for (int r = 0; r < m; ++r) s += Q(q, r) * Y[i][r];
          QY(q, i) = s;
        }
      HIP_OK(hipMemcpyAsync(S.small, QY.a.data(), QY.a.size() * sizeof(cplx), hipMemcpyHostToDevice, ctx->stream));
      double *Rr = S.ritz, *Ri = S.ritz + n * ncv;
      HIP_OK(vec_combine(ctx->stream, n, m, ncheck, Vr, Vi, n, S.small, Rr, Ri, n));
      for (int i = 0; i < ncheck; ++i) {
        const cplx lam = S.sigma + 1.0 / theta[i];
        HIP_OK(vec_kmul(ctx->stream, d.nodes, d.ssx * d.sse, ctx->d_vals, S.d_xcol, ctx->d_mvals,
                        S.d_pos, lam.real(), lam.imag(), Rr + i * n, Ri + i * n, yr, yi));
        double nr, nx;
        if (int st = solver_norm2(ctx, yr, yi, &nr)) return st;
        if (int st = solver_norm2(ctx, Rr + i * n, Ri + i * n, &nx)) return st;
        resid[i] = nr / nx;
        relerr[i] = resid[i] / std::max(std::abs(lam), 1e-300);
      }
      nconv = 0;
      while (nconv < ncheck && resid[nconv] <= tol && relerr[nconv] <= tol) ++nconv;
      if (nconv >= nev || restarts == o.maxit - 1) {
        nout_pass = m;
        break;
      }
      // the Ritz estimates are below tol but the true residuals are not:
      // stop when they no longer improve (accuracy floor of the solves)
      double worst = 0.0;
      for (int i = 0; i < nev; ++i) worst = std::max(worst, std::max(resid[i], relerr[i]));
      if (worst < 0.5 * best_res) {
        best_res = worst;
        stagnant = 0;
      } else if (++stagnant >= 3) {
        nout_pass = m;
        reason = -2;  // EPS_DIVERGED_BREAKDOWN: the accuracy floor of the solves
        break;
      }
    }
    // a pass that ended early on its Ritz estimates but not on the true
    // residuals is extended from where it stopped rather than restarted
    if (m < ncv) {
      k = m;
      continue;
    }
    // restart: keep the leading l Schur vectors
    int conv_est = 0;
    for (int i = 0; i < ncheck; ++i) {
      if (estimate(i, Y[i]) <= tol) ++conv_est;
      else break;
    }
    int l = conv_est + std::max(1, (m - conv_est) / 2);
    l = std::min(l, m - 1);
    {
      CMat Ql(m, l);
      for (int jj = 0; jj < l; ++jj)
        for (int q = 0; q < m; ++q) Ql(q, jj) = Q(q, jj);
      HIP_OK(hipMemcpyAsync(S.small, Ql.a.data(), Ql.a.size() * sizeof(cplx), hipMemcpyHostToDevice, ctx->stream));
      double *Rr = S.ritz, *Ri = S.ritz + n * ncv;
      HIP_OK(vec_combine(ctx->stream, n, m, l, Vr, Vi, n, S.small, Rr, Ri, n));
      HIP_OK(hipMemcpyAsync(Vr, Rr, l * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      HIP_OK(hipMemcpyAsync(Vi, Ri, l * n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      HIP_OK(hipMemcpyAsync(col_r(l), col_r(m), n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      HIP_OK(hipMemcpyAsync(col_i(l), col_i(m), n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    }
    Sm = CMat(ncv + 1, ncv);
    for (int jj = 0; jj < l; ++jj) {
      for (int ii = 0; ii <= jj; ++ii) Sm(ii, jj) = T(ii, jj);
      Sm(l, jj) = b[jj];
    }
    k = l;
    nout_pass = m;
  }
  HIP_OK(hipStreamSynchronize(ctx->stream));
  // results: leading pairs in order of |theta|
  S.lambda.clear();
  S.resid.clear();
  S.relerr.clear();
  const int nout = std::min(nout_pass - 1, std::max(nconv, 0));
  for (int i = 0; i < nout; ++i) {
    S.lambda.push_back(S.sigma + 1.0 / theta[i]);
    S.resid.push_back(resid[i]);
    S.relerr.push_back(relerr[i]);
  }
  S.nritz = nout;
  S.ritz_ld = ncv;
  diag.requested = nev;
  diag.converged = nconv;
  diag.iterations = restarts + 1;
  diag.solves = solves;
  diag.refinement_steps = refine_total;
  diag.perturbed_pivots = S.perturbed;
  diag.lambda_target_re = S.sigma.real();
  diag.lambda_target_im = S.sigma.imag();
  diag.max_residual = 0.0;
  diag.max_relative_error = 0.0;
  for (int i = 0; i < std::min(nconv, nev); ++i) {
    diag.max_residual = std::max(diag.max_residual, resid[i]);
    diag.max_relative_error = std::max(diag.max_relative_error, relerr[i]);
  }
  diag.solve_ms = now_ms() - t0;
  diag.factor_ms = S.factor_ms;
  diag.symbolic_ms = S.symbolic_ms;
  diag.fronts = S.sym.nsn;
  diag.factor_bytes = static_cast<int64_t>(S.plan.arena_doubles) * (S.fp32 ? 4 : 8);
  diag.factor_flops = static_cast<double>(S.sym.flops);
  // the reference's message (EigenSolver.cpp:84-88)
  if (nconv < nev)
    return fail(ctx, TAILOR_ERR_NOT_CONVERGED,
                "SLEPc failed to converge (reason " + std::to_string(reason) + ", " +
                    std::to_string(nconv) + " eigenpairs after " + std::to_string(restarts + 1) +
                    " iterations)");
  return TAILOR_OK;
}

}  // namespace tailor
