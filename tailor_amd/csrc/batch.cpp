// batch.cpp -- factorization of several contexts as one elimination forest.
//
// A batch joins the fronts of `count` contexts of one device (any grid
// sizes) into one plan whose level l holds level l of every context (levels
// count from the leaves, so a parent still follows its children).  Every
// launch of the factorization (front build, panels, pivot-block inverses,
// row exchanges, frontal GEMMs, Z) then covers the fronts of all contexts at
// that level: the latency-bound stretches of one instance (top-level panel
// chains on one workgroup per front, small levels) run side by side with
// the others' instead of one after another, and the launch count of a
// batch is that of its deepest context.  The launch sequence is captured
// once as a HIP graph and replayed with new shifts.  The factors land in
// each context's own front storage, so its solves and Krylov-Schur run on
// it as after tailor_factor.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "internal.hpp"
#include "lu.hpp"
#include "solver.hpp"

struct tailor_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  std::vector<tailor_ctx *> ctxs;
  std::vector<unsigned long long> plan_ids;  // each context's plan when the batch was made
  tailor::LuPlan plan;                 // the joined plan for front precision plan_fp32
  bool plan_built = false, plan_fp32 = false;
  std::vector<int> fbase;
  double *partials = nullptr;          // 1024 doubles of norm partials per context
  // GEMM tile configuration chosen by tailor_batch_autotune or
  // tailor_batch_set_gemm_config (kept when the plan is rebuilt)
  std::vector<int> cfg_shape, cfg_order;
  // captured launch sequences: [0] the joined factorization, [1] every
  // context's assembly followed by it
  struct Graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool fp32 = false, streaming = false, pairs = false;
    std::vector<int> shape, order;
  } graphs[2];
  hipEvent_t ev_start = nullptr, ev_done = nullptr;
  std::vector<hipEvent_t> ev_ready;    // one per context: its queued work before the batch
};

namespace {

int bfail(tailor_batch *b, int code, const std::string &msg) {
  if (b) b->error = msg;
  return code;
}

int bhip(tailor_batch *b, hipError_t e, const char *what) {
  return bfail(b, e == hipErrorOutOfMemory ? TAILOR_ERR_MEM : TAILOR_ERR_HIP,
               std::string(what) + ": " + hipGetErrorString(e));
}

#define BATCH_OK(expr)                                   \
  do {                                                   \
    hipError_t e_ = (expr);                              \
    if (e_ != hipSuccess) return bhip(bt, e_, #expr);    \
  } while (0)

void release_graph(tailor_batch *bt) {
  for (auto &g : bt->graphs) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
  }
}

std::vector<tailor::BatchPart> parts_of(const tailor_batch *bt) {
  std::vector<tailor::BatchPart> parts;
  for (tailor_ctx *c : bt->ctxs) parts.push_back({&c->solver->plan, &c->solver->sym});
  return parts;
}

std::vector<tailor::LuInputs> inputs_of(const tailor_batch *bt) {
  std::vector<tailor::LuInputs> in;
  for (tailor_ctx *c : bt->ctxs) {
    tailor::LuInputs x;
    x.nnzb = c->dims.nnzb;
    x.per = c->dims.ssx * c->dims.sse;
    x.avals = c->d_vals;
    x.colidx = c->d_colidx;
    x.bvals = c->d_mvals;
    in.push_back(x);
  }
  return in;
}

std::vector<double *> partials_of(const tailor_batch *bt) {
  std::vector<double *> p;
  for (size_t k = 0; k < bt->ctxs.size(); ++k) p.push_back(bt->partials + 1024 * k);
  return p;
}

// Every context with a case loaded and, unless the batch assembles them
// itself, assembled (A_bc, B_bc); all at the structure the batch was made
// with and at one front precision; the joined plan built for it.
int batch_ready(tailor_batch *bt, bool assembles, bool *fp32_out) {
  const int prec = bt->ctxs[0]->precision;
  for (size_t k = 0; k < bt->ctxs.size(); ++k) {
    tailor_ctx *c = bt->ctxs[k];
    if (!c->loaded)
      return bfail(bt, TAILOR_ERR_ARG_WRONGSTATE,
                   "Prepared field data are required before matrix assembly (context " +
                       std::to_string(k) + " of the batch)");
    if (!assembles &&
        (!c->assembled || c->vals_kind != TAILOR_OP_A || c->mvals_kind != TAILOR_OP_B))
      return bfail(bt, TAILOR_ERR_ARG_WRONGSTATE,
                   "Boundary-constrained matrices are required before EPSSolve (context " +
                       std::to_string(k) + " of the batch)");
    if (!c->solver || !c->solver->ready || c->solver->plan_id != bt->plan_ids[k])
      return bfail(bt, TAILOR_ERR_ARG_WRONGSTATE,
                   "context " + std::to_string(k) +
                       " of the batch was loaded with another structure; create the batch again");
    if (c->precision != prec)
      return bfail(bt, TAILOR_ERR_ARG_OUTOFRANGE,
                   "the contexts of a batch must share one front precision");
  }
  const bool fp32 = prec == TAILOR_PRECISION_MIXED;
  *fp32_out = fp32;
  if (bt->plan_built && bt->plan_fp32 == fp32) return TAILOR_OK;
  release_graph(bt);
  bt->plan_built = false;
  BATCH_OK(tailor::lu_plan_batch(parts_of(bt), fp32, bt->plan, bt->fbase, bt->stream));
  if (!bt->cfg_shape.empty()) {
    bt->plan.gemm_shape = bt->cfg_shape;
    bt->plan.gemm_order = bt->cfg_order;
  }
  bt->plan_built = true;
  bt->plan_fp32 = fp32;
  return TAILOR_OK;
}

// The batch stream waits for the work already queued on every context's
// stream (their assembly).
int wait_for_contexts(tailor_batch *bt) {
  for (size_t k = 0; k < bt->ctxs.size(); ++k) {
    BATCH_OK(hipEventRecord(bt->ev_ready[k], bt->ctxs[k]->stream));
    BATCH_OK(hipStreamWaitEvent(bt->stream, bt->ev_ready[k], 0));
  }
  return TAILOR_OK;
}

int set_shifts(tailor_batch *bt, const tailor::cplx *sig) {
  for (size_t k = 0; k < bt->ctxs.size(); ++k)
    BATCH_OK(tailor::lu_set_sigma(bt->plan, sig[k].real(), sig[k].imag(), bt->stream,
                                  static_cast<int>(k)));
  return TAILOR_OK;
}

// After the batch's factorization has completed: every context holds its
// factors at its shift.
int mark_factored(tailor_batch *bt, const tailor::cplx *sig, bool fp32, double ms) {
  std::vector<int> counters(std::max<size_t>(8, bt->ctxs.size()), 0);
  BATCH_OK(hipMemcpyAsync(counters.data(), bt->plan.d_counters, bt->ctxs.size() * sizeof(int),
                          hipMemcpyDeviceToHost, bt->stream));
  BATCH_OK(hipStreamSynchronize(bt->stream));
  for (size_t k = 0; k < bt->ctxs.size(); ++k)
    tailor::solver_mark_factored(bt->ctxs[k], sig[k], fp32, counters[k], ms);
  return TAILOR_OK;
}

// The batch's work as a launch sequence on its stream: with `assembles`,
// every context's assembly (its wavenumber and period already in its
// parameter block), then the joined factorization (the shifts already in the
// batch's parameter blocks).
hipError_t enqueue(tailor_batch *bt, bool assembles, bool fp32) {
  if (assembles)
    for (tailor_ctx *c : bt->ctxs) {
      hipError_t e = tailor::enqueue_assembly(c, true, bt->stream);
      if (e != hipSuccess) return e;
    }
  return tailor::lu_factor_batch(bt->plan, parts_of(bt), bt->fbase, inputs_of(bt),
                                 partials_of(bt), fp32, bt->stream, nullptr);
}

int graph_ready(tailor_batch *bt, bool assembles, bool fp32) {
  const tailor::LuPlan &U = bt->plan;
  auto &G = bt->graphs[assembles ? 1 : 0];
  if (G.exec && G.fp32 == fp32 && G.shape == U.gemm_shape && G.order == U.gemm_order &&
      G.streaming == U.panel_streaming && G.pairs == U.pair_blocks)
    return TAILOR_OK;
  if (G.exec) (void)hipGraphExecDestroy(G.exec);
  if (G.graph) (void)hipGraphDestroy(G.graph);
  G.exec = nullptr;
  G.graph = nullptr;
  BATCH_OK(hipStreamBeginCapture(bt->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = enqueue(bt, assembles, fp32);
  hipGraph_t g = nullptr;
  const hipError_t e_end = hipStreamEndCapture(bt->stream, &g);
  if (e == hipSuccess) e = e_end;
  hipGraphExec_t x = nullptr;
  if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return bhip(bt, e, "capture of the batch graph");
  }
  G.graph = g;
  G.exec = x;
  G.fp32 = fp32;
  G.shape = U.gemm_shape;
  G.order = U.gemm_order;
  G.streaming = U.panel_streaming;
  G.pairs = U.pair_blocks;
  return TAILOR_OK;
}

// One factorization of the batch at `sig`, after (with `assembles`) the
// assembly of every context; blocks until done.
int run(tailor_batch *bt, const tailor::cplx *sig, bool assembles, bool use_graph,
        double *batch_ms) {
  BATCH_OK(hipSetDevice(bt->device));
  bool fp32 = false;
  if (int st = batch_ready(bt, assembles, &fp32)) return st;
  if (use_graph)
    if (int st = graph_ready(bt, assembles, fp32)) return st;
  if (int st = wait_for_contexts(bt)) return st;
  BATCH_OK(hipEventRecord(bt->ev_start, bt->stream));
  if (assembles)
    for (tailor_ctx *c : bt->ctxs) BATCH_OK(tailor::launch_case_scalars(c, bt->stream));
  if (int st = set_shifts(bt, sig)) return st;
  if (use_graph)
    BATCH_OK(hipGraphLaunch(bt->graphs[assembles ? 1 : 0].exec, bt->stream));
  else
    BATCH_OK(enqueue(bt, assembles, fp32));
  BATCH_OK(hipEventRecord(bt->ev_done, bt->stream));
  BATCH_OK(hipEventSynchronize(bt->ev_done));
  float ms = 0.f;
  BATCH_OK(hipEventElapsedTime(&ms, bt->ev_start, bt->ev_done));
  if (batch_ms) *batch_ms = ms;
  if (assembles)
    for (size_t k = 0; k < bt->ctxs.size(); ++k)
      if (int st = tailor::finish_assembly(bt->ctxs[k], true, nullptr))
        return bfail(bt, st, "context " + std::to_string(k) + " of the batch: " +
                                 bt->ctxs[k]->error);
  return mark_factored(bt, sig, fp32, ms);
}

int check_shifts(tailor_batch *bt, const double *sigmas, std::vector<tailor::cplx> &sig) {
  if (!sigmas) return bfail(bt, TAILOR_ERR_USER_INPUT, "shifts are required");
  sig.resize(bt->ctxs.size());
  for (size_t k = 0; k < bt->ctxs.size(); ++k) {
    if (!std::isfinite(sigmas[2 * k]) || !std::isfinite(sigmas[2 * k + 1]))
      return bfail(bt, TAILOR_ERR_USER_INPUT, "shift must be finite");
    sig[k] = tailor::cplx(sigmas[2 * k], sigmas[2 * k + 1]);
  }
  return TAILOR_OK;
}

}  // namespace

extern "C" {

int tailor_batch_create(tailor_batch **out, tailor_ctx *const *ctxs, int32_t count) {
  if (!out) return TAILOR_ERR_USER_INPUT;
  *out = nullptr;
  if (!ctxs || count <= 0 || count > 64) return TAILOR_ERR_USER_INPUT;
  for (int k = 0; k < count; ++k) {
    if (!ctxs[k]) return TAILOR_ERR_USER_INPUT;
    for (int q = 0; q < k; ++q)
      if (ctxs[q] == ctxs[k]) return TAILOR_ERR_USER_INPUT;
    if (ctxs[k]->device != ctxs[0]->device) return TAILOR_ERR_ARG_OUTOFRANGE;
    if (!ctxs[k]->loaded) return TAILOR_ERR_ARG_WRONGSTATE;
  }
  auto *bt = new tailor_batch();
  bt->device = ctxs[0]->device;
  bt->ctxs.assign(ctxs, ctxs + count);
  int st = TAILOR_OK;
  hipError_t e = hipSetDevice(bt->device);
  for (int k = 0; k < count && e == hipSuccess && st == TAILOR_OK; ++k) {
    st = tailor::solver_prepare(ctxs[k]);
    if (st == TAILOR_OK) bt->plan_ids.push_back(ctxs[k]->solver->plan_id);
  }
  if (e == hipSuccess && st == TAILOR_OK)
    e = hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking);
  if (e == hipSuccess && st == TAILOR_OK)
    e = hipMalloc(reinterpret_cast<void **>(&bt->partials), 1024 * sizeof(double) * count);
  if (e == hipSuccess && st == TAILOR_OK) e = hipEventCreate(&bt->ev_start);
  if (e == hipSuccess && st == TAILOR_OK) e = hipEventCreate(&bt->ev_done);
  for (int k = 0; k < count && e == hipSuccess && st == TAILOR_OK; ++k) {
    hipEvent_t ev = nullptr;
    e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) bt->ev_ready.push_back(ev);
  }
  if (e != hipSuccess || st != TAILOR_OK) {
    const int code = st != TAILOR_OK ? st
                                     : (e == hipErrorOutOfMemory ? TAILOR_ERR_MEM : TAILOR_ERR_HIP);
    tailor_batch_destroy(bt);
    return code;
  }
  *out = bt;
  return TAILOR_OK;
}

void tailor_batch_destroy(tailor_batch *bt) {
  if (!bt) return;
  (void)hipSetDevice(bt->device);
  if (bt->stream) (void)hipStreamSynchronize(bt->stream);
  release_graph(bt);
  tailor::lu_plan_destroy(bt->plan);
  if (bt->partials) (void)hipFree(bt->partials);
  if (bt->ev_start) (void)hipEventDestroy(bt->ev_start);
  if (bt->ev_done) (void)hipEventDestroy(bt->ev_done);
  for (hipEvent_t ev : bt->ev_ready) (void)hipEventDestroy(ev);
  if (bt->stream) (void)hipStreamDestroy(bt->stream);
  delete bt;
}

const char *tailor_batch_last_error(const tailor_batch *bt) {
  return bt ? bt->error.c_str() : "null batch";
}

int tailor_batch_factor(tailor_batch *bt, const double *sigmas, int32_t use_graph,
                        double *batch_ms) {
  if (!bt) return TAILOR_ERR_USER_INPUT;
  std::vector<tailor::cplx> sig;
  if (int st = check_shifts(bt, sigmas, sig)) return st;
  return run(bt, sig.data(), false, use_graph != 0, batch_ms);
}

int tailor_batch_assemble_factor(tailor_batch *bt, const double *sigmas, int32_t use_graph,
                                 double *batch_ms) {
  if (!bt) return TAILOR_ERR_USER_INPUT;
  std::vector<tailor::cplx> sig;
  if (int st = check_shifts(bt, sigmas, sig)) return st;
  return run(bt, sig.data(), true, use_graph != 0, batch_ms);
}

int tailor_batch_autotune(tailor_batch *bt, double sigma_re, double sigma_im,
                          tailor_tune_report *report) {
  if (!bt) return TAILOR_ERR_USER_INPUT;
  if (!std::isfinite(sigma_re) || !std::isfinite(sigma_im))
    return bfail(bt, TAILOR_ERR_USER_INPUT, "shift must be finite");
  BATCH_OK(hipSetDevice(bt->device));
  bool fp32 = false;
  if (int st = batch_ready(bt, false, &fp32)) return st;
  if (int st = wait_for_contexts(bt)) return st;
  std::vector<tailor::cplx> sig(bt->ctxs.size(), tailor::cplx(sigma_re, sigma_im));
  tailor::LuPlan &U = bt->plan;
  const size_t nl = U.levels.size();
  const int nconf = tailor::kGemmShapes * tailor::kOrderCount;
  std::vector<std::vector<double>> t(nconf);
  std::vector<double> total(nconf, 0.0);
  for (int c = 0; c < nconf; ++c) {
    U.gemm_shape.assign(nl, c / tailor::kOrderCount);
    U.gemm_order.assign(nl, c % tailor::kOrderCount);
    if (int st = set_shifts(bt, sig.data())) return st;
    tailor::FactorTiming tm;
    BATCH_OK(tailor::lu_factor_batch(U, parts_of(bt), bt->fbase, inputs_of(bt), partials_of(bt),
                                     fp32, bt->stream, &tm));
    t[c] = tm.level_gemm_ms;
    total[c] = tm.gemm_ms;
  }
  BATCH_OK(hipStreamSynchronize(bt->stream));
  double tuned = 0.0;
  for (size_t l = 0; l < nl; ++l) {
    int best = 0;
    for (int c = 1; c < nconf; ++c)
      if (t[c][l] < t[best][l]) best = c;
    U.gemm_shape[l] = best / tailor::kOrderCount;
    U.gemm_order[l] = best % tailor::kOrderCount;
    tuned += t[best][l];
  }
  bt->cfg_shape = U.gemm_shape;
  bt->cfg_order = U.gemm_order;
  if (report) {
    report->levels = static_cast<int32_t>(nl);
    for (size_t l = 0; l < nl && l < 64; ++l)
      report->order[l] = U.gemm_shape[l] * tailor::kOrderCount + U.gemm_order[l];
    for (int c = 0; c < nconf && c < 8; ++c) report->gemm_ms[c] = total[c];
    report->tuned_gemm_ms = tuned;
  }
  // the contexts hold factors of the last configuration at sigma
  return mark_factored(bt, sig.data(), fp32, 0.0);
}

int tailor_batch_set_gemm_config(tailor_batch *bt, int32_t config) {
  if (!bt) return TAILOR_ERR_USER_INPUT;
  if (config < 0 || config >= tailor::kGemmShapes * tailor::kOrderCount)
    return bfail(bt, TAILOR_ERR_ARG_OUTOFRANGE, "GEMM configuration must be 0..7");
  size_t nl = 0;
  for (tailor_ctx *c : bt->ctxs) nl = std::max(nl, c->solver->plan.levels.size());
  bt->cfg_shape.assign(nl, config / tailor::kOrderCount);
  bt->cfg_order.assign(nl, config % tailor::kOrderCount);
  if (bt->plan_built) {
    bt->plan.gemm_shape = bt->cfg_shape;
    bt->plan.gemm_order = bt->cfg_order;
  }
  return TAILOR_OK;
}

}  // extern "C"
