// lu.hpp -- device-side plan of the multifrontal LU and its launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "symbolic.hpp"

namespace tailor {

// Per-front descriptor read by the kernels.
struct FrontDev {
  long long off;   // arena offset (doubles)
  long long woff;  // workspace offset (per-front vector of m entries)
  int m, p, c;     // order (p + c rounded up to kFrontAlign), pivots, update rows
  int pe;          // leading dimension of the E and Z planes (p rounded up to kFrontAlign)
  int piv_off;     // offset into the pivot array
  int dof0;        // first elimination-order dof of the front's pivots
  int dinv_off;    // first panel slot of the front in its level's inverse workspace
  int inst;        // instance of a batch (its parameter block and pivot counter); 0 alone
};

struct LevelPlan {
  int list_off = 0, count = 0;
  int pmax = 0, cmax = 0, mmax = 0;
  long long ablk_off = 0, ablk_count = 0;  // the level's BSR blocks in d_ablk (context plans)
};

// Everything the numeric phase needs, resident on the device.
struct LuPlan {
  int nsn = 0;
  int nb = 32;
  long long n = 0;              // unknowns
  long long arena_doubles = 0;
  long long w_len = 0;          // per-plane length of the front workspace
  std::vector<LevelPlan> levels;
  std::vector<int> gemm_order;     // per level (kOrder*); empty = linear everywhere
  std::vector<int> gemm_shape;     // per level tile shape (the kShape* codes of lu.hip)
  bool panel_streaming = false;    // force the streaming panel kernel on every level
  bool pair_blocks = true;         // two-level blocking: pairs of blocks share one update (K = 256)
  std::vector<FrontDev> h_front;   // host copies (flop accounting)
  std::vector<int> h_level_list;
  FrontDev *d_front = nullptr;
  int *d_level_list = nullptr;
  // front build: per front kBuildInts ints (children, -1 where absent, and the
  // offsets of their maps in d_inv); d_inv maps a front's nodes to the
  // child's update nodes (-1 where the child has none)
  int *d_build = nullptr;
  int *d_inv = nullptr;
  std::vector<int> h_build, h_inv;  // host copies (joined into batch plans)
  int *d_ablk = nullptr;            // BSR blocks ordered by the level of their front
  int *d_parent = nullptr;
  int *d_rel = nullptr;
  long long *d_rel_off = nullptr;
  int *d_child_ptr = nullptr;
  int *d_child_list = nullptr;
  int *d_upd_pos = nullptr;
  long long *d_upd_off = nullptr;
  int *d_amap = nullptr;        // 3 ints per BSR block
  int *d_ipiv = nullptr;
  int *d_counters = nullptr;    // [0] perturbed pivots
  int *d_perm = nullptr;        // elimination dof -> natural dof (n entries)
  double *d_fparams = nullptr;  // kFpCount doubles per instance: sigma and the pivot thresholds
  double *arena = nullptr;
  bool owns_arena = true;       // false for a batch plan, whose fronts live in its parts' arenas
  int instances = 1;            // parameter blocks / pivot counters (one per batch part)
  // look-ahead TRAILREST (default priority): a batch plan's is a plain
  // stream made with the plan; a context plan's is CU-masked and made at its
  // first factorization (lu_plan_side), so that a context whose factorizations
  // only run in batches holds no queue of its own for it
  mutable hipStream_t side = nullptr;
  hipEvent_t ev_head = nullptr, ev_panel = nullptr;
  double *dinv = nullptr;       // inverted pivot blocks: dinv_slots panel slots
  long long dinv_slots = 0;     // most panels of any level
  int *d_bperm = nullptr;       // block row permutations: one list per front of a level
  int count_max = 0;            // most fronts of any level
  double *W = nullptr;          // 2 * w_len
  double *Y = nullptr;          // 2 * n
  double *W2 = nullptr;         // the second right-hand side's W and Y (lu_solve2)
  double *Y2 = nullptr;
};

struct LuInputs {
  long long nnzb = 0;
  int per = 0;                  // blocks per block row
  const double *avals = nullptr;  // A_bc, interleaved complex, 25 per block
  const int *colidx = nullptr;
  const double *bvals = nullptr;  // B_bc block diagonal, interleaved complex
};

// The factorization's parameter block (device): the shift and the pivot
// threshold / static perturbation derived from it and from |A|, |B|.
enum { kFpSigRe = 0, kFpSigIm = 1, kFpTiny = 2, kFpPerturb = 3, kFpCount = 4 };

constexpr int kLuNb = 32;        // panel width
// Front orders are rounded up to, and front offsets aligned to, multiples of
// kFrontAlign elements, so that every column of F starts on a 128-byte line
// (FP32 fronts; 256 bytes with FP64).  The extra rows and columns belong to
// the update part (rows and columns [p + c, m)): the front build writes them
// as zeros, no pivot, child update, scatter or solve reaches them, and the
// updates keep them zero.  The p x p planes E and Z are stored with leading dimension pe
// (p rounded up likewise); their rows [p, pe) stay zero and unused.
constexpr int kFrontAlign = 32;
constexpr int kBuildInts = 4;  // child 0, child 1, their map offsets
constexpr int kLuNbOuter = 128;  // outer block width (K of the trailing update)
// Compute units kept free of the look-ahead stream's TRAILREST updates (the
// highest-numbered ones), so that the panel chain on the main stream is
// dispatched at once: 16 of 256 measured 134.0 -> 130.0 ms per 200x100
// factorization and +2.6 % on the eight-context sweep; 32 gave the same.
constexpr int kSideReservedCus = 16;
constexpr int kLuInvSlot = 4 * kLuNb * kLuNb;      // inverted pivot blocks of one panel
constexpr int kLuPermStride = 1 + 4 * kLuNbOuter;  // one front's block row permutation

// Build the device plan from the symbolic analysis and the BSR structure
// (host arrays).  Allocates device memory; returns hipSuccess or an error.
hipError_t lu_plan_create(const Symbolic &S, const std::vector<int> &colidx, int per,
                          LuPlan &P, hipStream_t s);
void lu_plan_destroy(LuPlan &P);

// A batch plan: the fronts of several plans (any structures) as one
// elimination forest, level l holding level l of every part (levels count
// from the leaves), so that each launch of a level covers the fronts of all
// parts.  Fronts stay in their parts' arenas: a batch front's offset is taken
// from the lowest part arena in units of the front type (FP32 fronts when
// fp32), which the flat device address space makes exact.  Front f of part b
// is batch front fbase[b] + f; its `inst` is b.  The batch owns its pivot,
// inverse and permutation workspaces, parameter blocks and look-ahead stream.
struct BatchPart {
  const LuPlan *plan;
  const Symbolic *sym;
};
hipError_t lu_plan_batch(const std::vector<BatchPart> &parts, bool fp32, LuPlan &U,
                         std::vector<int> &fbase, hipStream_t s);

hipError_t lu_fill(hipStream_t s, double *p, long long n, double v);
// Make P's look-ahead stream if it has none yet (before a graph capture of
// its factorization; lu_factor calls it too).
hipError_t lu_plan_side(const LuPlan &P);
// Per-kernel timing of one factorization (HIP events around every frontal
// GEMM launch) and the algorithmic flops of those launches (8 M N K per
// front and launch: a complex multiply-add is 8 real flops).
struct FactorTiming {
  double gemm_ms = 0.0;
  double gemm_flops = 0.0;
  int gemm_launches = 0;
  std::vector<double> level_gemm_ms;     // GEMM time per elimination level
  std::vector<double> level_gemm_flops;  // and its flops (8 per complex multiply-add)
  double region_gemm_ms[8] = {};         // by update region (the kGemm* codes of lu.hip)
  double region_gemm_flops[8] = {};
};

// Tile orders and shapes of the frontal GEMM (chosen per level by the
// autotuner).
enum { kOrderLinear = 0, kOrderXcdGrouped = 1, kOrderCount = 2 };
constexpr int kGemmShapes = 4;

// The shift sigma of the next factorization into the parameter block
// (instance `inst` of a batch plan).
hipError_t lu_set_sigma(const LuPlan &P, double sig_re, double sig_im, hipStream_t s,
                        int inst = 0);
// The pivot threshold and perturbation from sigma (already in the block) and
// the Frobenius norms of A_bc and B_bc, on the device (`partials`: 1024
// doubles of scratch).
hipError_t lu_factor_params(const LuPlan &P, const LuInputs &in, bool fp32, double *partials,
                            hipStream_t s);
// LU of K = A - sigma B with the parameter block's sigma and thresholds:
// fronts in complex FP64, or complex FP32 (fp32) in the same arena.  With
// `timing`, every GEMM launch is bracketed by events (slightly slower) and
// the factorization runs on one stream; without, it is a pure launch
// sequence on s and the plan's side stream (joined back to s), so it can be
// captured in a HIP graph.
hipError_t lu_factor(const LuPlan &P, const LuInputs &in, bool fp32, hipStream_t s,
                     FactorTiming *timing = nullptr);
// One factorization of every part of a batch plan U (lu_plan_batch with the
// same fp32) at the shifts already in U's parameter blocks: per part the
// norms of A_bc and B_bc and the pivot thresholds (in[b], partials[b]: 1024
// doubles of scratch per part), then one launch sequence over the levels of
// U, each level's fronts built from their children with every part's K
// added.  As lu_factor, a pure launch sequence on s and U's side stream
// without `timing`.
hipError_t lu_factor_batch(const LuPlan &U, const std::vector<BatchPart> &parts,
                           const std::vector<int> &fbase, const std::vector<LuInputs> &in,
                           const std::vector<double *> &partials, bool fp32, hipStream_t s,
                           FactorTiming *timing = nullptr);
// x = K^-1 b, vectors in elimination dof order, split planes (re, im)
hipError_t lu_solve(const LuPlan &P, const double *br, const double *bi, double *xr, double *xi,
                    bool fp32, hipStream_t s);
// x1 = K^-1 b1 and x2 = K^-1 b2 in one pass over the factors (each factor
// entry is read once for both); the vectors as for lu_solve, all distinct
hipError_t lu_solve2(const LuPlan &P, const double *b1r, const double *b1i, const double *b2r,
                     const double *b2i, double *x1r, double *x1i, double *x2r, double *x2i,
                     bool fp32, hipStream_t s);

}  // namespace tailor
