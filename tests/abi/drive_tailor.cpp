// drive_tailor.cpp -- the C ABI of libtailor driven from C++, as a
// reference-side caller would bind it (the runOnGpu sequence of the
// integration notes): create, load_case, assemble, solve, get_eigenpair,
// destroy.  Test infrastructure: it reads a case written by
// tests/test_gpu_abi_driver.py as raw little-endian arrays and prints one
// line per eigenpair.
//
//   drive_tailor <case.bin>
//
// File layout: int32 ny, nz, ssx, sse, nev, maxit; float64 xi period, eta
// period, spanwise period, Re, Ma, Pr, gamma, R, mu_ref, Tref, S, alpha (re, im),
// omega_s (re, im), tol; float64 xi nodes (ny), eta nodes (nz); int64 xi
// stencil indices and offsets (ny*ssx each), eta indices and offsets
// (nz*sse each); float64 xi weights (3*ny*ssx), eta weights (3*nz*sse),
// grid (ny*nz*3 complex), base flow (ny*nz*5 complex).
// Output: "pair k lambda_re lambda_im residual relerr norm q_re q_im" with
// q the largest-magnitude entry of the returned mode; exit status = the
// first failing tailor_status.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "tailor.h"

namespace {

struct Reader {
  std::ifstream in;
  explicit Reader(const char *path) : in(path, std::ios::binary) {
    if (!in) throw std::runtime_error(std::string("cannot open ") + path);
  }
  template <class T>
  T one() {
    T v{};
    in.read(reinterpret_cast<char *>(&v), sizeof v);
    if (!in) throw std::runtime_error("case file too short");
    return v;
  }
  template <class T>
  std::vector<T> many(size_t n) {
    std::vector<T> v(n);
    in.read(reinterpret_cast<char *>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
    if (!in) throw std::runtime_error("case file too short");
    return v;
  }
};

struct Failure {
  int status;
  std::string message;
};

void check(tailor_ctx *ctx, int st) {
  if (st != TAILOR_OK) throw Failure{st, tailor_last_error(ctx)};
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: drive_tailor <case.bin>\n");
    return 2;
  }
  tailor_ctx *ctx = nullptr;
  try {
    Reader r(argv[1]);
    const int32_t ny = r.one<int32_t>(), nz = r.one<int32_t>();
    const int32_t ssx = r.one<int32_t>(), sse = r.one<int32_t>();
    const int32_t nev = r.one<int32_t>(), maxit = r.one<int32_t>();
    const double xi_period = r.one<double>(), eta_period = r.one<double>();
    const double span = r.one<double>();
    const auto phys = r.many<double>(8);  // Re, Ma, Pr, gamma, R, mu_ref, Tref, S
    const auto alpha = r.many<double>(2), omega = r.many<double>(2);
    const double tol = r.one<double>();
    const auto xnodes = r.many<double>(ny), enodes = r.many<double>(nz);
    const auto xidx = r.many<int64_t>(static_cast<size_t>(ny) * ssx);
    const auto xoff = r.many<int64_t>(static_cast<size_t>(ny) * ssx);
    const auto eidx = r.many<int64_t>(static_cast<size_t>(nz) * sse);
    const auto eoff = r.many<int64_t>(static_cast<size_t>(nz) * sse);
    const auto xw = r.many<double>(3 * static_cast<size_t>(ny) * ssx);
    const auto ew = r.many<double>(3 * static_cast<size_t>(nz) * sse);
    const size_t nodes = static_cast<size_t>(ny) * nz;
    const auto grid = r.many<double>(nodes * 6), base = r.many<double>(nodes * 10);

    if (tailor_create(&ctx, /*device=*/0) != TAILOR_OK) {
      std::fprintf(stderr, "tailor_create failed\n");
      return TAILOR_ERR_HIP;
    }
    tailor_case c{};
    c.ny = ny;
    c.nz = nz;
    c.xi = {ny, ssx - 1, 0, xi_period, xnodes.data(), xidx.data(), xoff.data(), xw.data()};
    c.eta = {nz, sse - 1, 1, eta_period, enodes.data(), eidx.data(), eoff.data(), ew.data()};
    c.grid = grid.data();
    c.baseflow = base.data();
    c.spanwise_period = span;
    c.reynolds = phys[0];
    c.mach = phys[1];
    c.prandtl = phys[2];
    c.gamma = phys[3];
    c.gas_constant = phys[4];
    c.reference_mu = phys[5];
    c.reference_temperature = phys[6];
    c.sutherland_constant = phys[7];
    c.alpha_re = alpha[0];
    c.alpha_im = alpha[1];
    check(ctx, tailor_load_case(ctx, &c));
    tailor_assembly_diag ad{};
    check(ctx, tailor_assemble(ctx, &ad));
    std::printf("assembled rows %lld blocks %lld a_fro %.17g b_fro %.17g\n",
                static_cast<long long>(ad.rows), static_cast<long long>(ad.spatial_blocks),
                ad.a_fro, ad.b_fro);
    tailor_solve_opts o{};
    o.nev = nev;
    o.ncv = 0;
    o.maxit = maxit;
    o.tol = tol;
    o.omega_re = omega[0];
    o.omega_im = omega[1];
    tailor_solve_diag sd{};
    check(ctx, tailor_solve(ctx, &o, &sd));
    std::printf("solved converged %d iterations %d solves %d\n", sd.converged, sd.iterations,
                sd.solves);
    const int count = tailor_eigenpair_count(ctx);
    std::vector<double> q(2 * 5 * nodes);
    for (int k = 0; k < count; ++k) {
      double lam[2], res = 0.0, rel = 0.0;
      check(ctx, tailor_get_eigenpair(ctx, k, lam, q.data(), &res, &rel));
      double norm2 = 0.0, best = -1.0, pr = 0.0, pi = 0.0;
      for (size_t i = 0; i < 5 * nodes; ++i) {
        const double a = q[2 * i] * q[2 * i] + q[2 * i + 1] * q[2 * i + 1];
        norm2 += a;
        if (a > best) {
          best = a;
          pr = q[2 * i];
          pi = q[2 * i + 1];
        }
      }
      std::printf("pair %d %.17g %.17g %.6e %.6e %.17g %.17g %.17g\n", k, lam[0], lam[1], res, rel,
                  std::sqrt(norm2), pr, pi);
    }
    tailor_destroy(ctx);
    return 0;
  } catch (const Failure &f) {
    std::fprintf(stderr, "tailor error %d: %s\n", f.status, f.message.c_str());
    if (ctx) tailor_destroy(ctx);
    return f.status;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    if (ctx) tailor_destroy(ctx);
    return 2;
  }
}
