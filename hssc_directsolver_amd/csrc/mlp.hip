// Rank-prediction network: MLP 8 -> 64 -> 64 -> 1 with GELU, FP32, trained and evaluated on
// the GPU by the kernels below (matrix products on the MFMA GEMM of kernels.hip).
//
// It replaces the closed-form rank guesses of HSS2D/predict_rank.m:5-23 and the level
// extrapolation of HSS2D/interpolation_operator.m:101-117: given geometric features of a box
// it predicts log2 of the box's ID rank, which sizes the randomized sketch of the box's
// interpolative decomposition (the adaptive ID stays as the fallback and judge).
//
// Layout: a batch of B samples is column-major B x F (F = 8 features); weights
// W1 (8 x 64), W2 (64 x 64), w3 (64 x 1) act from the right (Z = X W + 1 b^T).
// Features and the target are standardised with (mu, sd) kept in the model.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "stats.h"

namespace {

constexpr int kIn = 8, kHid = 64;
constexpr int kParams = kIn * kHid + kHid + kHid * kHid + kHid + kHid + 1;  // 4801

struct Offs {
  // offsets of the parameter blocks in the flat parameter vector
  static constexpr int W1 = 0, b1 = W1 + kIn * kHid, W2 = b1 + kHid, b2 = W2 + kHid * kHid,
                       w3 = b2 + kHid, b3 = w3 + kHid;
};

#define MCHK(expr)                                                                   \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) return hssc::fail(HSSC_ERR_HIP, std::string(#expr) + ": " + \
                                                              hipGetErrorString(e_)); \
  } while (0)

// GELU with the tanh approximation: g(z) = z/2 (1 + tanh(s (z + 0.044715 z^3))),
// s = sqrt(2/pi).
__device__ __forceinline__ float gelu(float z) {
  const float s = 0.7978845608f, u = s * (z + 0.044715f * z * z * z);
  return 0.5f * z * (1.0f + tanhf(u));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float s = 0.7978845608f, u = s * (z + 0.044715f * z * z * z);
  const float t = tanhf(u);
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * s * (1.0f + 3.0f * 0.044715f * z * z);
}

// X (B x F) standardised by (mu, sd) into Xs.
__global__ void standardise_kernel(int B, const float* __restrict__ X, const float* __restrict__ mu,
                                   const float* __restrict__ sd, float* __restrict__ Xs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * kIn) return;
  const int f = i / B;
  Xs[i] = (X[i] - mu[f]) / sd[f];
}

// Z (B x H) += 1 b^T, A = gelu(Z); Z keeps the pre-activation for the backward pass.
__global__ void bias_gelu_kernel(int B, int H, float* __restrict__ Z, const float* __restrict__ b,
                                 float* __restrict__ A) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * H) return;
  const float z = Z[i] + b[i / B];
  Z[i] = z;
  A[i] = gelu(z);
}

// y = y + b3 (output bias)
__global__ void add_scalar_kernel(int B, float* __restrict__ y, const float* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < B) y[i] += b[0];
}

// d = 2 (y - t) / B ; loss partials sum (y - t)^2 per block
__global__ void mse_grad_kernel(int B, const float* __restrict__ y, const float* __restrict__ t,
                                float* __restrict__ d, float* __restrict__ part) {
  __shared__ float red[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float e = 0.0f;
  if (i < B) {
    const float r = y[i] - t[i];
    d[i] = 2.0f * r / float(B);
    e = r * r;
  }
  for (int o = 32; o > 0; o >>= 1) e += __shfl_xor(e, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// G (B x H) *= gelu'(Z)
__global__ void gelu_back_kernel(int B, int H, const float* __restrict__ Z, float* __restrict__ G) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < B * H) G[i] *= gelu_grad(Z[i]);
}

// g[h] = sum_i G(i, h) for a B x H matrix (one workgroup per column, fixed summation order)
__global__ void colsum_kernel(int B, const float* __restrict__ G, float* __restrict__ g) {
  __shared__ float red[4];
  const int h = blockIdx.x;
  float s = 0.0f;
  for (int i = threadIdx.x; i < B; i += blockDim.x) s += G[i + int64_t(h) * B];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) g[h] = red[0] + red[1] + red[2] + red[3];
}

// Adam (Kingma & Ba): m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
// p -= lr (m / (1-b1^t)) / (sqrt(v / (1-b2^t)) + eps)
__global__ void adam_kernel(int n, float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, float lr, float c1,
                            float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi, v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

inline dim3 g1(int n) { return dim3((n + 255) / 256); }

}  // namespace

struct hssc_mlp {
  float* P = nullptr;   // parameters (kParams)
  float* G = nullptr;   // gradients
  float* Mm = nullptr;  // Adam first moments
  float* Vm = nullptr;  // Adam second moments
  float* mu = nullptr;  // feature mean (8), then target mean, sd (8), target sd
  float hmu[kIn + 1] = {0}, hsd[kIn + 1] = {1, 1, 1, 1, 1, 1, 1, 1, 1};
  int steps = 0;
  // batch workspace
  int cap = 0;
  float *Xs = nullptr, *Z1 = nullptr, *A1 = nullptr, *Z2 = nullptr, *A2 = nullptr, *y = nullptr,
        *d = nullptr, *t = nullptr, *Xd = nullptr, *D2 = nullptr, *D1 = nullptr, *part = nullptr;
  float* kpart = nullptr;  // partial products of the split-K weight gradients
  int64_t kpart_cap = 0;
  void free_ws() {
    for (float* q : {Xs, Z1, A1, Z2, A2, y, d, t, Xd, D2, D1, part, kpart})
      if (q) (void)hipFree(q);
    Xs = Z1 = A1 = Z2 = A2 = y = d = t = Xd = D2 = D1 = part = kpart = nullptr;
    cap = 0, kpart_cap = 0;
  }
  ~hssc_mlp() {
    free_ws();
    for (float* q : {P, G, Mm, Vm, mu})
      if (q) (void)hipFree(q);
  }
};

namespace {

// Weight gradients are sums over the batch (K = B): the batch is split into chunks of this
// many samples, one workgroup each, and the partial products are summed in chunk order.
constexpr int kGradChunk = 128;

// C (m x n) = A^T B over the batch, split over K.
void grad_gemm(hssc_mlp* m, int64_t rows, int64_t cols, int B, const float* A, const float* Bm,
               float* C) {
  hssc::k::gemm_tn_splitk<float>(0, rows, cols, B, A, B, 0, Bm, B, 0, C, rows, 0, 1, m->kpart,
                                  kGradChunk, m->kpart_cap);
}

int ensure(hssc_mlp* m, int B) {
  if (B <= m->cap) return HSSC_OK;
  m->free_ws();
  auto al = [&](float** p, size_t n) { return hipMalloc(reinterpret_cast<void**>(p), n * sizeof(float)); };
  const size_t b = size_t(B);
  if (al(&m->Xs, b * kIn) || al(&m->Z1, b * kHid) || al(&m->A1, b * kHid) ||
      al(&m->Z2, b * kHid) || al(&m->A2, b * kHid) || al(&m->y, b) || al(&m->d, b) ||
      al(&m->t, b) || al(&m->Xd, b * kIn) || al(&m->D2, b * kHid) || al(&m->D1, b * kHid) ||
      al(&m->part, (b + 255) / 256 + 1) ||
      al(&m->kpart, size_t((B + kGradChunk - 1) / kGradChunk) * kHid * kHid))
    return hssc::fail(HSSC_ERR_OUT_OF_MEMORY, "mlp workspace");
  m->kpart_cap = int64_t((B + kGradChunk - 1) / kGradChunk) * kHid * kHid;
  m->cap = B;
  return HSSC_OK;
}

// Forward pass on the standardised batch already in m->Xs; output (standardised) in m->y.
void forward(hssc_mlp* m, int B) {
  using hssc::k::gemm;
  const float* P = m->P;
  gemm<float>(0, false, false, B, kHid, kIn, 1.0f, m->Xs, B, 0, P + Offs::W1, kIn, 0, 0.0f, m->Z1,
              B, 0, 1);
  hipLaunchKernelGGL(bias_gelu_kernel, g1(B * kHid), dim3(256), 0, 0, B, kHid, m->Z1,
                     P + Offs::b1, m->A1);
  gemm<float>(0, false, false, B, kHid, kHid, 1.0f, m->A1, B, 0, P + Offs::W2, kHid, 0, 0.0f,
              m->Z2, B, 0, 1);
  hipLaunchKernelGGL(bias_gelu_kernel, g1(B * kHid), dim3(256), 0, 0, B, kHid, m->Z2,
                     P + Offs::b2, m->A2);
  gemm<float>(0, false, false, B, 1, kHid, 1.0f, m->A2, B, 0, P + Offs::w3, kHid, 0, 0.0f, m->y, B,
              0, 1);
  hipLaunchKernelGGL(add_scalar_kernel, g1(B), dim3(256), 0, 0, B, m->y, P + Offs::b3);
}

// Backward pass of the mean squared error against the standardised targets in m->t; the
// gradient lands in m->G, the summed squared error in *sse.
int backward(hssc_mlp* m, int B, double* sse) {
  using hssc::k::gemm;
  const float* P = m->P;
  float* G = m->G;
  const int nb = (B + 255) / 256;
  hipLaunchKernelGGL(mse_grad_kernel, dim3(nb), dim3(256), 0, 0, B, m->y, m->t, m->d, m->part);
  // output layer: dw3 = A2^T d, db3 = sum d, dA2 = d w3^T
  grad_gemm(m, kHid, 1, B, m->A2, m->d, G + Offs::w3);
  hipLaunchKernelGGL(colsum_kernel, dim3(1), dim3(256), 0, 0, B, m->d, G + Offs::b3);
  gemm<float>(0, false, true, B, kHid, 1, 1.0f, m->d, B, 0, P + Offs::w3, kHid, 0, 0.0f, m->D2, B,
              0, 1);
  hipLaunchKernelGGL(gelu_back_kernel, g1(B * kHid), dim3(256), 0, 0, B, kHid, m->Z2, m->D2);
  // hidden layer 2: dW2 = A1^T D2, db2 = colsum D2, dA1 = D2 W2^T
  grad_gemm(m, kHid, kHid, B, m->A1, m->D2, G + Offs::W2);
  hipLaunchKernelGGL(colsum_kernel, dim3(kHid), dim3(256), 0, 0, B, m->D2, G + Offs::b2);
  gemm<float>(0, false, true, B, kHid, kHid, 1.0f, m->D2, B, 0, P + Offs::W2, kHid, 0, 0.0f, m->D1,
              B, 0, 1);
  hipLaunchKernelGGL(gelu_back_kernel, g1(B * kHid), dim3(256), 0, 0, B, kHid, m->Z1, m->D1);
  // hidden layer 1: dW1 = Xs^T D1, db1 = colsum D1
  grad_gemm(m, kIn, kHid, B, m->Xs, m->D1, G + Offs::W1);
  hipLaunchKernelGGL(colsum_kernel, dim3(kHid), dim3(256), 0, 0, B, m->D1, G + Offs::b1);
  if (sse) {
    std::vector<float> part(nb);
    MCHK(hipMemcpy(part.data(), m->part, nb * sizeof(float), hipMemcpyDeviceToHost));
    double s = 0;
    for (float p : part) s += p;
    *sse = s;
  }
  return HSSC_OK;
}

int load_batch(hssc_mlp* m, const float* X, const float* t, int B) {
  if (int s = ensure(m, B)) return s;
  MCHK(hipMemcpy(m->Xd, X, size_t(B) * kIn * sizeof(float), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(standardise_kernel, g1(B * kIn), dim3(256), 0, 0, B, m->Xd, m->mu,
                     m->mu + kIn + 1, m->Xs);
  if (t) {
    std::vector<float> ts(B);
    for (int i = 0; i < B; ++i) ts[i] = (t[i] - m->hmu[kIn]) / m->hsd[kIn];
    MCHK(hipMemcpy(m->t, ts.data(), size_t(B) * sizeof(float), hipMemcpyHostToDevice));
  }
  return HSSC_OK;
}

int upload_scaling(hssc_mlp* m) {
  float buf[2 * (kIn + 1)];
  std::memcpy(buf, m->hmu, sizeof(m->hmu));
  std::memcpy(buf + kIn + 1, m->hsd, sizeof(m->hsd));
  MCHK(hipMemcpy(m->mu, buf, sizeof(buf), hipMemcpyHostToDevice));
  return HSSC_OK;
}

}  // namespace

extern "C" {

int hssc_mlp_create(uint64_t seed, hssc_mlp** out) {
  if (!out) return hssc::fail(HSSC_ERR_INVALID_ARG, "null output");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
    return hssc::fail(HSSC_ERR_HIP, "no HIP device is available");
  auto* m = new hssc_mlp();
  auto al = [&](float** p, size_t n) { return hipMalloc(reinterpret_cast<void**>(p), n * sizeof(float)); };
  if (al(&m->P, kParams) || al(&m->G, kParams) || al(&m->Mm, kParams) || al(&m->Vm, kParams) ||
      al(&m->mu, 2 * (kIn + 1))) {
    delete m;
    return hssc::fail(HSSC_ERR_OUT_OF_MEMORY, "mlp parameters");
  }
  // Glorot-uniform weights from a splitmix64 stream, zero biases
  std::vector<float> p(kParams, 0.0f);
  uint64_t s = seed;
  auto unif = [&]() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return double(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int fan_in, int fan_out) {
    const double a = std::sqrt(6.0 / double(fan_in + fan_out));
    for (int i = 0; i < fan_in * fan_out; ++i) p[off + i] = float((2 * unif() - 1) * a);
  };
  fill(Offs::W1, kIn, kHid);
  fill(Offs::W2, kHid, kHid);
  fill(Offs::w3, kHid, 1);
  if (hipMemcpy(m->P, p.data(), kParams * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->Mm, 0, kParams * sizeof(float)) != hipSuccess ||
      hipMemset(m->Vm, 0, kParams * sizeof(float)) != hipSuccess || upload_scaling(m) != HSSC_OK) {
    delete m;
    return hssc::fail(HSSC_ERR_HIP, "mlp initialisation");
  }
  *out = m;
  return HSSC_OK;
}

void hssc_mlp_destroy(hssc_mlp* m) { delete m; }

int hssc_mlp_num_params(void) { return kParams; }

int hssc_mlp_get_params(const hssc_mlp* m, float* out) {
  if (!m || !out) return hssc::fail(HSSC_ERR_INVALID_ARG, "null argument");
  MCHK(hipMemcpy(out, m->P, kParams * sizeof(float), hipMemcpyDeviceToHost));
  return HSSC_OK;
}

int hssc_mlp_set_params(hssc_mlp* m, const float* in) {
  if (!m || !in) return hssc::fail(HSSC_ERR_INVALID_ARG, "null argument");
  MCHK(hipMemcpy(m->P, in, kParams * sizeof(float), hipMemcpyHostToDevice));
  return HSSC_OK;
}

int hssc_mlp_set_scaling(hssc_mlp* m, const float* mu, const float* sd) {
  if (!m || !mu || !sd) return hssc::fail(HSSC_ERR_INVALID_ARG, "null argument");
  for (int i = 0; i <= kIn; ++i) {
    if (!(sd[i] > 0)) return hssc::fail(HSSC_ERR_INVALID_ARG, "scales must be positive");
    m->hmu[i] = mu[i], m->hsd[i] = sd[i];
  }
  return upload_scaling(m);
}

int hssc_mlp_predict(hssc_mlp* m, const float* X, int32_t B, float* y) {
  if (!m || !X || !y || B < 1) return hssc::fail(HSSC_ERR_INVALID_ARG, "invalid mlp_predict arguments");
  if (int s = load_batch(m, X, nullptr, B)) return s;
  HSSC_TIMED("mlp_forward_f32", 0, 2.0 * B * (kIn * kHid + kHid * kHid + kHid), 0.0, forward(m, B));
  MCHK(hipMemcpy(y, m->y, size_t(B) * sizeof(float), hipMemcpyDeviceToHost));
  for (int i = 0; i < B; ++i) y[i] = y[i] * m->hsd[kIn] + m->hmu[kIn];
  return HSSC_OK;
}

int hssc_mlp_grads(hssc_mlp* m, const float* X, const float* t, int32_t B, double* loss,
                   float* grads) {
  if (!m || !X || !t || B < 1) return hssc::fail(HSSC_ERR_INVALID_ARG, "invalid mlp_grads arguments");
  if (int s = load_batch(m, X, t, B)) return s;
  forward(m, B);
  double sse = 0;
  if (int s = backward(m, B, &sse)) return s;
  if (loss) *loss = sse / double(B);
  if (grads) MCHK(hipMemcpy(grads, m->G, kParams * sizeof(float), hipMemcpyDeviceToHost));
  return HSSC_OK;
}

int hssc_mlp_train(hssc_mlp* m, const float* X, const float* t, int32_t B, int32_t epochs,
                   float lr, double* loss) {
  if (!m || !X || !t || B < 1 || epochs < 1 || !(lr > 0))
    return hssc::fail(HSSC_ERR_INVALID_ARG, "invalid mlp_train arguments");
  if (int s = load_batch(m, X, t, B)) return s;
  double sse = 0;
  for (int e = 0; e < epochs; ++e) {
    forward(m, B);
    if (int s = backward(m, B, e + 1 == epochs ? &sse : nullptr)) return s;
    ++m->steps;
    const float c1 = 1.0f - std::pow(0.9f, float(m->steps));
    const float c2 = 1.0f - std::pow(0.999f, float(m->steps));
    hipLaunchKernelGGL(adam_kernel, g1(kParams), dim3(256), 0, 0, kParams, m->P, m->G, m->Mm,
                       m->Vm, lr, c1, c2);
  }
  MCHK(hipDeviceSynchronize());
  MCHK(hipGetLastError());
  if (loss) *loss = sse / double(B);
  return HSSC_OK;
}

}  // extern "C"
