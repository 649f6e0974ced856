// noalloc.hip -- the linear solve x = K^-1 b without any device allocation.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).  The four
// elimination-order work vectors of a refined solve (b, y, residual,
// correction; split real/imaginary planes) live in one file-scope device
// array, `scratch`, of kScratchUnknowns complex unknowns.  Each GPU holds its
// own copy of the array at its own address from the first use of the
// library on it until the process ends; tailor_create looks the address up
// once per device (noalloc_setup) and host code uses only that address.
//
// Every call on a GPU shares that GPU's array and leaves its contents
// behind: calls that overlap on one GPU (two host threads, or two contexts
// whose streams run concurrently) race on it.  Nothing in the array needs to
// start at zero: every element a call reads was written earlier in the same
// call.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "internal.hpp"
#include "lu.hpp"
#include "solver.hpp"
#include "vec.hpp"

namespace {

constexpr long long kScratchUnknowns = 262144;  // largest 5 * nodes a call may have
constexpr long long kScratchDoubles = 8 * kScratchUnknowns;  // 4 vectors x 2 planes = 16 MiB
constexpr int kMaxDevices = 64;
constexpr int kMaxRefineSteps = 64;

std::mutex g_mutex;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch` per device

}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace tailor {

// Called by tailor_create with `device` current: look up this device's copy
// of `scratch` once.
hipError_t noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[device]) return hipSuccess;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = static_cast<double *>(p);
  return hipSuccess;
}

}  // namespace tailor

extern "C" int tailor_linear_solve_noalloc(tailor_ctx *ctx, const double *b, double *x,
                                           int32_t refine_steps) {
  using namespace tailor;
  if (!ctx) return TAILOR_ERR_USER_INPUT;
  if (!b || !x) return fail(ctx, TAILOR_ERR_USER_INPUT, "b and x must be device pointers");
  if (refine_steps < 0 || refine_steps > kMaxRefineSteps)
    return fail(ctx, TAILOR_ERR_ARG_OUTOFRANGE,
                "refine_steps must be 0.." + std::to_string(kMaxRefineSteps));
  if (!ctx->loaded) return fail(ctx, TAILOR_ERR_ARG_WRONGSTATE, "no case is loaded");
  const long long n = 5LL * ctx->dims.nodes;
  if (n > kScratchUnknowns)
    return fail(ctx, TAILOR_ERR_ARG_OUTOFRANGE,
                "tailor_linear_solve_noalloc: " + std::to_string(n) +
                    " unknowns exceed its scratch limit of " + std::to_string(kScratchUnknowns) +
                    " (use tailor_linear_solve)");
  if (!ctx->solver || !ctx->solver->factored)
    return fail(ctx, TAILOR_ERR_ARG_WRONGSTATE,
                "tailor_factor must precede tailor_linear_solve_noalloc");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
  double *base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    base = ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
  }
  if (!base) return fail(ctx, TAILOR_ERR_ARG_WRONGSTATE, "no-allocation scratch is not set up");
  SolverState &S = *ctx->solver;
  const auto &d = ctx->dims;
  hipStream_t st = ctx->stream;
  double *br = base, *bi = br + n;          // b, elimination order
  double *yr = base + 2 * n, *yi = yr + n;  // solution
  double *rr = base + 4 * n, *ri = rr + n;  // residual b - K y
  double *dr = base + 6 * n, *di = dr + n;  // correction K^-1 r
  // the same kernels, in the same order, as tailor_linear_solve with
  // refine_max = refine_steps and refine_tol = 0 (which never stops early)
  if ((e = vec_nat_to_elim(st, n, S.plan.d_perm, b, br, bi)) != hipSuccess)
    return hip_fail(ctx, e, "vec_nat_to_elim");
  if ((e = lu_solve(S.plan, br, bi, yr, yi, S.fp32, st)) != hipSuccess)
    return hip_fail(ctx, e, "lu_solve");
  for (int it = 0; it < refine_steps; ++it) {
    if ((e = vec_kmul(st, d.nodes, d.ssx * d.sse, ctx->d_vals, S.d_xcol, ctx->d_mvals,
                      S.d_pos, S.sigma.real(), S.sigma.imag(), yr, yi, rr, ri)) != hipSuccess)
      return hip_fail(ctx, e, "vec_kmul");
    if ((e = vec_axpby(st, n, 1.0, 0.0, br, bi, -1.0, 0.0, rr, ri)) != hipSuccess)
      return hip_fail(ctx, e, "vec_axpby");
    if ((e = lu_solve(S.plan, rr, ri, dr, di, S.fp32, st)) != hipSuccess)
      return hip_fail(ctx, e, "lu_solve");
    if ((e = vec_axpby(st, n, 1.0, 0.0, dr, di, 1.0, 0.0, yr, yi)) != hipSuccess)
      return hip_fail(ctx, e, "vec_axpby");
  }
  if ((e = vec_elim_to_nat(st, n, S.plan.d_perm, yr, yi, x)) != hipSuccess)
    return hip_fail(ctx, e, "vec_elim_to_nat");
  if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(ctx, e, "hipStreamSynchronize");
  return TAILOR_OK;
}
