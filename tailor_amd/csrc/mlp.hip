// mlp.hip -- the shift-proposal network: a 3 -> 64 -> 64 -> 16 perceptron
// with GELU activations in FP32, trained and evaluated by its own kernels.
//
// Inputs are normalised (alpha, Re, Ma); outputs are the normalised real and
// imaginary parts of the 8 leading omega.  Parameters are one float buffer:
//   W1[64][3] b1[64] W2[64][64] b2[64] W3[16][64] b3[16]   (5456 floats)
// Training minimises the mean squared error over the batch with Adam
// (Kingma & Ba 2015) on full batches.  Nothing here has a counterpart in the
// reference; the network only proposes shift centres for the eigen-solve,
// whose true residual decides whether a proposal is accepted.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/tailor.h"

namespace {

constexpr int kIn = 3, kH = 64, kOut = 16;
constexpr int oW1 = 0, ob1 = oW1 + kH * kIn, oW2 = ob1 + kH, ob2 = oW2 + kH * kH,
              oW3 = ob2 + kH, ob3 = oW3 + kOut * kH, kParams = ob3 + kOut;

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
  // d/dz [z Phi(z)] = Phi(z) + z phi(z)
  const float Phi = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
  const float phi = 0.39894228040143268f * expf(-0.5f * z * z);
  return Phi + z * phi;
}

// Z[s][o] = b[o] + sum_i W[o][i] X[s][i];  A = act ? gelu(Z) : Z.  Thread per (s, o).
__global__ void k_dense_fwd(const float *__restrict__ W, const float *__restrict__ b,
                            const float *__restrict__ X, int n, int nin, int nout, int act,
                            float *__restrict__ Z, float *__restrict__ A) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nout) return;
  const int s = t / nout, o = t % nout;
  float z = b[o];
  for (int i = 0; i < nin; ++i) z += W[o * nin + i] * X[s * nin + i];
  if (Z) Z[t] = z;
  A[t] = act ? gelu(z) : z;
}

// dY = 2 (Y - T) / (n * kOut), and the squared-error partial sums.
__global__ void k_mse(const float *__restrict__ Y, const float *__restrict__ T, int n,
                      float *__restrict__ dY, float *__restrict__ loss_part) {
  __shared__ float red[256];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.0f;
  if (t < n * kOut) {
    const float e = Y[t] - T[t];
    dY[t] = 2.0f * e / static_cast<float>(n * kOut);
    e2 = e * e;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// Gradients of one dense layer from dZ (n x nout) and its input X (n x nin):
// dW[o][i] = sum_s dZ[s][o] X[s][i], db[o] = sum_s dZ[s][o].  Thread per
// (o, i) plus one per bias; the batch sum runs in a fixed order.
__global__ void k_dense_grad(const float *__restrict__ dZ, const float *__restrict__ X, int n,
                             int nin, int nout, float *__restrict__ dW, float *__restrict__ db) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nout * nin) {
    const int o = t / nin, i = t % nin;
    float g = 0.0f;
    for (int s = 0; s < n; ++s) g += dZ[s * nout + o] * X[s * nin + i];
    dW[t] = g;
  } else if (t < nout * nin + nout) {
    const int o = t - nout * nin;
    float g = 0.0f;
    for (int s = 0; s < n; ++s) g += dZ[s * nout + o];
    db[o] = g;
  }
}

// dZprev[s][i] = gelu'(Zprev[s][i]) * sum_o dZ[s][o] W[o][i].  Thread per (s, i).
__global__ void k_dense_back(const float *__restrict__ dZ, const float *__restrict__ W,
                             const float *__restrict__ Zprev, int n, int nin, int nout,
                             float *__restrict__ dZprev) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nin) return;
  const int s = t / nin, i = t % nin;
  float g = 0.0f;
  for (int o = 0; o < nout; ++o) g += dZ[s * nout + o] * W[o * nin + i];
  dZprev[t] = g * gelu_grad(Zprev[t]);
}

// Adam with bias correction; step is 1-based.
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, int count, float lr, float b1, float b2, float eps,
                       int step) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const float gt = g[t];
  const float mt = b1 * m[t] + (1.0f - b1) * gt;
  const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
  m[t] = mt;
  v[t] = vt;
  const float mh = mt / (1.0f - powf(b1, static_cast<float>(step)));
  const float vh = vt / (1.0f - powf(b2, static_cast<float>(step)));
  p[t] -= lr * mh / (sqrtf(vh) + eps);
}

inline unsigned blocks(long long n, int t) { return static_cast<unsigned>((n + t - 1) / t); }

}  // namespace

struct tailor_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string error;
  float *params = nullptr, *grads = nullptr, *m = nullptr, *v = nullptr;
  int step = 0;
  // batch buffers (grown on demand)
  int cap = 0;
  float *X = nullptr, *T = nullptr, *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr,
        *Y = nullptr, *dY = nullptr, *dZ2 = nullptr, *dZ1 = nullptr, *loss = nullptr;
};

namespace {

int mfail(tailor_mlp *m, int st, const std::string &msg) {
  m->error = msg;
  return st;
}

#define MLP_HIP(m, expr)                                                              \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) return mfail(m, TAILOR_ERR_HIP, std::string(#expr ": ") + \
                                                         hipGetErrorString(e_));      \
  } while (0)

void release(float *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

int ensure_batch(tailor_mlp *m, int n) {
  if (n <= m->cap) return TAILOR_OK;
  float **bufs[] = {&m->X, &m->T, &m->Z1, &m->A1, &m->Z2, &m->A2, &m->Y, &m->dY, &m->dZ2, &m->dZ1, &m->loss};
  for (float **b : bufs) release(*b);
  const size_t sz[] = {size_t(n) * kIn, size_t(n) * kOut, size_t(n) * kH, size_t(n) * kH,
                       size_t(n) * kH, size_t(n) * kH, size_t(n) * kOut, size_t(n) * kOut,
                       size_t(n) * kH, size_t(n) * kH, size_t(blocks(size_t(n) * kOut, 256))};
  for (int k = 0; k < 11; ++k)
    MLP_HIP(m, hipMalloc(reinterpret_cast<void **>(bufs[k]), sz[k] * sizeof(float)));
  m->cap = n;
  return TAILOR_OK;
}

hipError_t forward(tailor_mlp *m, int n) {
  const float *P = m->params;
  hipLaunchKernelGGL(k_dense_fwd, dim3(blocks(n * kH, 256)), dim3(256), 0, m->stream, P + oW1,
                     P + ob1, m->X, n, kIn, kH, 1, m->Z1, m->A1);
  hipLaunchKernelGGL(k_dense_fwd, dim3(blocks(n * kH, 256)), dim3(256), 0, m->stream, P + oW2,
                     P + ob2, m->A1, n, kH, kH, 1, m->Z2, m->A2);
  hipLaunchKernelGGL(k_dense_fwd, dim3(blocks(n * kOut, 256)), dim3(256), 0, m->stream, P + oW3,
                     P + ob3, m->A2, n, kH, kOut, 0, nullptr, m->Y);
  return hipGetLastError();
}

// Loss and gradients into m->grads for the batch in m->X / m->T.
hipError_t backward(tailor_mlp *m, int n) {
  const float *P = m->params;
  float *G = m->grads;
  hipLaunchKernelGGL(k_mse, dim3(blocks(n * kOut, 256)), dim3(256), 0, m->stream, m->Y, m->T, n,
                     m->dY, m->loss);
  hipLaunchKernelGGL(k_dense_grad, dim3(blocks(kOut * kH + kOut, 256)), dim3(256), 0, m->stream,
                     m->dY, m->A2, n, kH, kOut, G + oW3, G + ob3);
  hipLaunchKernelGGL(k_dense_back, dim3(blocks(n * kH, 256)), dim3(256), 0, m->stream, m->dY,
                     P + oW3, m->Z2, n, kH, kOut, m->dZ2);
  hipLaunchKernelGGL(k_dense_grad, dim3(blocks(kH * kH + kH, 256)), dim3(256), 0, m->stream,
                     m->dZ2, m->A1, n, kH, kH, G + oW2, G + ob2);
  hipLaunchKernelGGL(k_dense_back, dim3(blocks(n * kH, 256)), dim3(256), 0, m->stream, m->dZ2,
                     P + oW2, m->Z1, n, kH, kH, m->dZ1);
  hipLaunchKernelGGL(k_dense_grad, dim3(blocks(kH * kIn + kH, 256)), dim3(256), 0, m->stream,
                     m->dZ1, m->X, n, kIn, kH, G + oW1, G + ob1);
  return hipGetLastError();
}

int read_loss(tailor_mlp *m, int n, float *out) {
  const int nb = static_cast<int>(blocks(n * kOut, 256));
  std::vector<float> part(nb);
  MLP_HIP(m, hipMemcpyAsync(part.data(), m->loss, nb * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  double s = 0.0;
  for (float v : part) s += v;
  *out = static_cast<float>(s / (static_cast<double>(n) * kOut));
  return TAILOR_OK;
}

int upload_batch(tailor_mlp *m, const float *X, const float *T, int n) {
  if (int st = ensure_batch(m, n)) return st;
  MLP_HIP(m, hipMemcpyAsync(m->X, X, size_t(n) * kIn * sizeof(float), hipMemcpyHostToDevice, m->stream));
  if (T)
    MLP_HIP(m, hipMemcpyAsync(m->T, T, size_t(n) * kOut * sizeof(float), hipMemcpyHostToDevice, m->stream));
  return TAILOR_OK;
}

}  // namespace

extern "C" {

int tailor_mlp_param_count(void) { return kParams; }

int tailor_mlp_create(tailor_mlp **out, int device, uint64_t seed) {
  if (!out) return TAILOR_ERR_USER_INPUT;
  *out = nullptr;
  tailor_mlp *m = new tailor_mlp();
  m->device = device;
  auto bad = [&](hipError_t e) {
    m->error = hipGetErrorString(e);
    tailor_mlp_destroy(m);
    return TAILOR_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipSetDevice(device)) != hipSuccess) return bad(e);
  if ((e = hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking)) != hipSuccess) return bad(e);
  float **bufs[] = {&m->params, &m->grads, &m->m, &m->v};
  for (float **b : bufs)
    if ((e = hipMalloc(reinterpret_cast<void **>(b), kParams * sizeof(float))) != hipSuccess) return bad(e);
  // He-uniform weights from a SplitMix64 stream, zero biases
  std::vector<float> h(kParams, 0.0f);
  unsigned long long st = seed;
  auto uni = [&]() {
    st += 0x9E3779B97F4A7C15ULL;
    unsigned long long z = st;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return static_cast<double>(z >> 11) * 0x1.0p-53;
  };
  auto fill = [&](int off, int fan_out, int fan_in) {
    const double a = std::sqrt(6.0 / fan_in);
    for (int k = 0; k < fan_out * fan_in; ++k) h[off + k] = static_cast<float>((2.0 * uni() - 1.0) * a);
  };
  fill(oW1, kH, kIn);
  fill(oW2, kH, kH);
  fill(oW3, kOut, kH);
  std::vector<float> z(kParams, 0.0f);
  if ((e = hipMemcpyAsync(m->params, h.data(), kParams * sizeof(float), hipMemcpyHostToDevice,
                          m->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(m->m, z.data(), kParams * sizeof(float), hipMemcpyHostToDevice,
                          m->stream)) != hipSuccess ||
      (e = hipMemcpyAsync(m->v, z.data(), kParams * sizeof(float), hipMemcpyHostToDevice,
                          m->stream)) != hipSuccess ||
      (e = hipStreamSynchronize(m->stream)) != hipSuccess)
    return bad(e);
  *out = m;
  return TAILOR_OK;
}

void tailor_mlp_destroy(tailor_mlp *m) {
  if (!m) return;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  float **bufs[] = {&m->params, &m->grads, &m->m, &m->v, &m->X, &m->T, &m->Z1, &m->A1,
                    &m->Z2, &m->A2, &m->Y, &m->dY, &m->dZ2, &m->dZ1, &m->loss};
  for (float **b : bufs) release(*b);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

const char *tailor_mlp_last_error(const tailor_mlp *m) { return m ? m->error.c_str() : "null model"; }

int tailor_mlp_get_params(tailor_mlp *m, float *out) {
  if (!m || !out) return TAILOR_ERR_USER_INPUT;
  MLP_HIP(m, hipSetDevice(m->device));
  MLP_HIP(m, hipMemcpyAsync(out, m->params, kParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  return TAILOR_OK;
}

int tailor_mlp_set_params(tailor_mlp *m, const float *in) {
  if (!m || !in) return TAILOR_ERR_USER_INPUT;
  MLP_HIP(m, hipSetDevice(m->device));
  MLP_HIP(m, hipMemcpyAsync(m->params, in, kParams * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  m->step = 0;  // a new parameter set restarts the optimiser
  std::vector<float> z(kParams, 0.0f);
  MLP_HIP(m, hipMemcpyAsync(m->m, z.data(), kParams * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLP_HIP(m, hipMemcpyAsync(m->v, z.data(), kParams * sizeof(float), hipMemcpyHostToDevice, m->stream));
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  return TAILOR_OK;
}

int tailor_mlp_predict(tailor_mlp *m, const float *X, int32_t n, float *Y) {
  if (!m || !X || !Y || n <= 0) return TAILOR_ERR_USER_INPUT;
  MLP_HIP(m, hipSetDevice(m->device));
  if (int st = upload_batch(m, X, nullptr, n)) return st;
  MLP_HIP(m, forward(m, n));
  MLP_HIP(m, hipMemcpyAsync(Y, m->Y, size_t(n) * kOut * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  return TAILOR_OK;
}

int tailor_mlp_gradients(tailor_mlp *m, const float *X, const float *T, int32_t n, float *grad,
                         float *loss) {
  if (!m || !X || !T || n <= 0) return TAILOR_ERR_USER_INPUT;
  MLP_HIP(m, hipSetDevice(m->device));
  if (int st = upload_batch(m, X, T, n)) return st;
  MLP_HIP(m, forward(m, n));
  MLP_HIP(m, backward(m, n));
  if (grad)
    MLP_HIP(m, hipMemcpyAsync(grad, m->grads, kParams * sizeof(float), hipMemcpyDeviceToHost, m->stream));
  float l = 0.0f;
  if (int st = read_loss(m, n, &l)) return st;
  if (loss) *loss = l;
  return TAILOR_OK;
}

int tailor_mlp_train(tailor_mlp *m, const float *X, const float *T, int32_t n, int32_t steps,
                     float lr, float *loss_history) {
  if (!m || !X || !T || n <= 0 || steps <= 0 || !(lr > 0.0f)) return TAILOR_ERR_USER_INPUT;
  MLP_HIP(m, hipSetDevice(m->device));
  if (int st = upload_batch(m, X, T, n)) return st;
  for (int k = 0; k < steps; ++k) {
    MLP_HIP(m, forward(m, n));
    MLP_HIP(m, backward(m, n));
    ++m->step;
    hipLaunchKernelGGL(k_adam, dim3(blocks(kParams, 256)), dim3(256), 0, m->stream, m->params,
                       m->grads, m->m, m->v, kParams, lr, 0.9f, 0.999f, 1e-8f, m->step);
    MLP_HIP(m, hipGetLastError());
    if (loss_history) {
      float l = 0.0f;
      if (int st = read_loss(m, n, &l)) return st;
      loss_history[k] = l;
    }
  }
  MLP_HIP(m, hipStreamSynchronize(m->stream));
  return TAILOR_OK;
}

}  // extern "C"
