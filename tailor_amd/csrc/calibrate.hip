// calibrate.hip -- the rates this GPU sustains for the two resources the
// path's kernels are bound by, measured with kernels that do nothing else:
//
//   FP64 MFMA: v_mfma_f64_16x16x4f64 back to back on eight independent
//              accumulators per wave, two resident waves per SIMD, every CU;
//   HBM read:  a grid-strided 32-byte-per-lane streaming read of a buffer
//              far larger than the caches.
//
// They give the attainable ceilings beside the published peaks (78.6 TFLOP/s
// dense FP64, 8 TB/s).  The accumulators must stay in the registers the MFMA
// writes (the build compiles this file with the VGPR form of the MFMA
// builtins): copied to and from AGPRs around every round of MFMAs, the same
// loop measures ~47 TFLOP/s instead.  On the MI355X measured here the MFMA
// issues one 16x16x4 FP64 instruction per 64 shader cycles per SIMD with one
// or more waves, ~77 TFLOP/s at 2.4 GHz, and a streaming read reaches ~6 TB/s.
#include <hip/hip_runtime.h>

#include "../../include/tailor.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kCalAcc = 8;  // independent accumulators per wave

__global__ __launch_bounds__(256) void k_cal_mfma(double *out, int iters, double a0, double b0) {
  d4 acc[kCalAcc];
#pragma unroll
  for (int k = 0; k < kCalAcc; ++k) acc[k] = d4{0, 0, 0, 0};
  const double a = a0 + threadIdx.x * 1e-9, b = b0 - threadIdx.x * 1e-9;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < kCalAcc; ++k)
      acc[k] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[k], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kCalAcc; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
  if (s == 1234.5) out[0] = s;  // never true; keeps the chain live
}

__global__ __launch_bounds__(256) void k_cal_read(const double4 *__restrict__ p, long long n4,
                                                  double *out) {
  double acc = 0.0;
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n4;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const double4 v = p[i];
    acc += v.x + v.y + v.z + v.w;
  }
  if (acc == 1234.5) out[0] = acc;  // never true: the buffer holds zeros
}

__global__ void k_cal_zero(double *p, long long n) {
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<long long>(gridDim.x) * blockDim.x)
    p[i] = 0.0;
}

}  // namespace

extern "C" int tailor_measure_peaks(int device, double *fp64_mfma_tflops, double *read_gbs) {
  if (!fp64_mfma_tflops && !read_gbs) return TAILOR_ERR_USER_INPUT;
  if (hipSetDevice(device) != hipSuccess) return TAILOR_ERR_HIP;
  hipDeviceProp_t prop{};
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return TAILOR_ERR_HIP;
  const int cus = prop.multiProcessorCount;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double *out = nullptr, *buf = nullptr;
  const long long bytes = 2LL << 30;  // 2 GiB: far beyond the caches
  int st = TAILOR_OK;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&out), 64) != hipSuccess)
    st = TAILOR_ERR_HIP;
  auto timed = [&](auto launch, int reps) -> double {
    launch();  // warm-up
    float ms = 0.f;
    if (hipEventRecord(e0, s) != hipSuccess) return -1.0;
    for (int r = 0; r < reps; ++r) launch();
    if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      return -1.0;
    return static_cast<double>(ms) / reps;
  };
  if (st == TAILOR_OK && fp64_mfma_tflops) {
    const int iters = 8000, grid = 2 * cus;  // two workgroups of four waves per CU
    const double ms = timed([&] {
      hipLaunchKernelGGL(k_cal_mfma, dim3(grid), dim3(256), 0, s, out, iters, 1.0, 1.0);
    }, 3);
    if (ms <= 0.0) st = TAILOR_ERR_HIP;
    else *fp64_mfma_tflops = 2048.0 * kCalAcc * iters * grid * 4.0 / (ms * 1e-3) / 1e12;
  }
  if (st == TAILOR_OK && read_gbs) {
    if (hipMalloc(reinterpret_cast<void **>(&buf), bytes) != hipSuccess) {
      st = TAILOR_ERR_MEM;
    } else {
      hipLaunchKernelGGL(k_cal_zero, dim3(4096), dim3(256), 0, s, buf, bytes / 8);
      const double ms = timed([&] {
        hipLaunchKernelGGL(k_cal_read, dim3(16 * cus), dim3(256), 0, s,
                           reinterpret_cast<const double4 *>(buf), bytes / 32, out);
      }, 10);
      if (ms <= 0.0) st = TAILOR_ERR_HIP;
      else *read_gbs = static_cast<double>(bytes) / (ms * 1e-3) / 1e9;
    }
  }
  if (buf) (void)hipFree(buf);
  if (out) (void)hipFree(out);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  return st;
}
