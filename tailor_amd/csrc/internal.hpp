// Internal declarations shared by the HIP translation units of libtailor.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tailor.h"

namespace tailor {

struct SolverState;

struct dcx {
  double re, im;
};

// Status bits raised by the assembly kernels (checked on the host in the
// reference's PetscCheck order).
enum : unsigned {
  kGridNonFinite = 1u << 0,       // Metrics.cpp:208-209
  kJacobian = 1u << 1,            // Metrics.cpp:210-214
  kSecondNonFinite = 1u << 2,     // Metrics.cpp:346-347
  kBaseNonFinite = 1u << 3,       // BaseFlowDerivatives.cpp:239-240
  kBasePositive = 1u << 4,        // BaseFlowDerivatives.cpp:241-243
  kDerivNonFinite = 1u << 5,      // BaseFlowDerivatives.cpp:244-245
  kCoefficient = 1u << 6,         // GlobalOperator.cpp:201-205
  kCharacteristic = 1u << 7,      // BoundaryConditions.cpp:448-454
};

// Device-side scalar diagnostics (reduced with atomics).
struct DeviceDiag {
  unsigned flags;
  int char_status;        // max CharacteristicStatus over far-field nodes
  int incoming, neutral;  // totals
  int min_incoming, max_incoming;
  int pad[2];
  double min_jac, max_jac;
  double max_grid_imag, max_base_imag;
  double max_char_imag, max_cond;
};

struct CaseDims {
  int ny, nz, ssx, sse;
  int hx, hy;  // largest |stencil offset| in xi and eta
  long long nodes, nnzb;
};

}  // namespace tailor

struct tailor_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  bool loaded = false;
  bool assembled = false;
  int vals_kind = -1;   // TAILOR_OP_L or TAILOR_OP_A
  int mvals_kind = -1;  // TAILOR_OP_M or TAILOR_OP_B

  tailor::CaseDims dims{};
  double period = 0.0;
  double recipe[6] = {0, 0, 0, 0, 0, 0};  // Re, Ma, Pr, gamma, Tref, S
  double alpha_re = 0.0, alpha_im = 0.0;

  // device copies of the case
  int *d_xoff = nullptr;     // ny*ssx stencil offsets (bounded xi)
  double *d_xw = nullptr;    // 3*ny*ssx weights d0,d1,d2
  int *d_eoff = nullptr;     // nz*sse offsets (periodic eta)
  double *d_ew = nullptr;    // 3*nz*sse
  double *d_grid = nullptr;  // nodes*6
  double *d_base = nullptr;  // nodes*10
  // fields
  double *d_metrics = nullptr;  // nodes*10
  double *d_deriv = nullptr;    // nodes*25
  // operators (BSR, 5x5 row-major complex blocks)
  int *d_rowptr = nullptr;
  int *d_colidx = nullptr;
  double *d_vals = nullptr;   // nnzb*50
  double *d_mvals = nullptr;  // nodes*50
  double *d_coef = nullptr;   // nodes*250: per-node coefficient records (assembly)
  int *d_coef_status = nullptr;  // nodes
  tailor::DeviceDiag *d_diag = nullptr;
  double *d_bcdata = nullptr;    // nz * 106: P, R^-1, incoming mask, status
  double *d_recipe = nullptr;    // kRecipeDoubles: Re, Ma, Pr, gamma, Tref, S, alpha, period
  double *d_partials = nullptr;  // reduction scratch, kPartials doubles

  // host copy of the structure and of the rule offsets
  std::vector<int> h_rowptr, h_colidx;
  std::vector<int> h_xoff, h_eoff;
  // factorization and eigen-solve state (solver.hpp)
  tailor::SolverState *solver = nullptr;
  int precision = 0;  // TAILOR_PRECISION_FP64 or TAILOR_PRECISION_MIXED
};

namespace tailor {

// The case's device parameter block (d_recipe): the physics of the recipe,
// then the streamwise wavenumber and the spanwise period, which the assembly
// kernels read from it (not from launch arguments), so a captured assembly
// replays with the values written before the replay.
enum {
  kRecipeRe = 0, kRecipeMa = 1, kRecipePr = 2, kRecipeGamma = 3, kRecipeTref = 4, kRecipeS = 5,
  kRecipeAlphaRe = 6, kRecipeAlphaIm = 7, kRecipePeriod = 8, kRecipeDoubles = 16
};

// Widest stencils (points per 1-D rule) the operator writer takes: its
// run-time-sized instance holds this many per node in LDS; the derivative
// tile's LDS bound is tighter for most pairs.
constexpr int kOperatorMaxSsx = 27, kOperatorMaxSse = 19;

constexpr int kPartialBlocks = 512;
constexpr int kPartialComps = 32;

// error helpers
int fail(tailor_ctx *ctx, int code, const std::string &msg);
int hip_fail(tailor_ctx *ctx, hipError_t e, const char *what);

// assembly launchers (assembly.hip), on stream s
hipError_t launch_case_scalars(tailor_ctx *ctx, hipStream_t s);  // alpha, period -> d_recipe
hipError_t launch_reset_diag(tailor_ctx *ctx, hipStream_t s);
hipError_t launch_metrics(tailor_ctx *ctx, hipStream_t s);
hipError_t launch_derivatives(tailor_ctx *ctx, hipStream_t s);
hipError_t launch_characteristics(tailor_ctx *ctx, hipStream_t s);
hipError_t launch_coefficients(tailor_ctx *ctx, hipStream_t s);  // per-node LNS coefficient records
hipError_t launch_operator(tailor_ctx *ctx, bool constrained, hipStream_t s);

// The assembly of A_bc, B_bc (constrained) or L, M as a launch sequence on
// s, after launch_case_scalars; nothing is read back (capi.cpp).
hipError_t enqueue_assembly(tailor_ctx *ctx, bool constrained, hipStream_t s);
// After an enqueued assembly has completed: read its status word back,
// record which operators the value buffers hold and turn the status bits into
// the reference's messages (capi.cpp).
int finish_assembly(tailor_ctx *ctx, bool constrained, DeviceDiag *host_diag);
hipError_t launch_sumsq(tailor_ctx *ctx, const double *data, long long rows, int ncomp,
                        long long stride, double *host_out);

// device address of the no-allocation solve's scratch array on `device`,
// looked up once per device (noalloc.hip); `device` must be current
hipError_t noalloc_setup(int device);

}  // namespace tailor
