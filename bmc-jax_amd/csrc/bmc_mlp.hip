// The learned Levenberg-Marquardt proposal: a small MLP trained and run with hand-written BF16 MFMA
// GEMMs (FP32 accumulate) on CDNA4. No counterpart in the reference; it supplies the LM start x0.
//
// Layout: activations row-major [batch, features]; weights W_l [out, in] (FP32 master copy plus a BF16
// copy for the matrix cores), biases [out]. Hidden layers use GELU (0.5 x (1 + erf(x / sqrt 2))), the
// last layer is linear; the loss is the mean over batch and outputs of the squared error.
//
// GEMM: one kernel for C[M, N] = A[M, K] B[K, N] with arbitrary strides on A and B (so the forward
// X W^T, the input gradient dY W and the weight gradient dY^T X are the same kernel). A 256-thread block
// computes a 64 x 64 tile; each wave a 32 x 32 quarter as 2 x 2 tiles of v_mfma_f32_16x16x16_bf16. K is
// walked 32 at a time through LDS. Lane layout of the 16x16x16 BF16 MFMA (64 lanes, l = lane):
//   A fragment: row l % 16, k = 4 (l / 16) .. +3       B fragment: column l % 16, k = 4 (l / 16) .. +3
//   D fragment: column l % 16, rows 4 (l / 16) .. +3
#include <math.h>

#include <algorithm>
#include <string.h>

#include <string>
#include <vector>

#include "bmc_internal.h"

namespace bmc {
namespace mlp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16) | ((u & 0xffffu) ? 0x40 : 0);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

constexpr int BM = 64, BN = 64, BK = 32;

struct GemmArgs {
  int M, N, K;
  const uint16_t* A;  // bf16
  long sam, sak;      // A[m, k] = A[m * sam + k * sak]
  const uint16_t* B;  // bf16
  long sbk, sbn;      // B[k, n] = B[k * sbk + n * sbn]
  float* C;           // [M, ldc]
  long ldc;
  const float* bias;  // [N] or null
  int act;            // 0 none, 1 GELU
  float* Cpre;        // pre-activation copy [M, ldc] or null
  uint16_t* Cbf;      // bf16 copy of C [M, ldc] or null
  float alpha;
  int accumulate;     // C += result
};

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
  // LDS tiles, padded by 4 bf16 per row against bank conflicts
  __shared__ uint16_t sA[BM][BK + 4];
  __shared__ uint16_t sB[BN][BK + 4];  // stored transposed: sB[n][k]
  const int tid = threadIdx.x;
  const int wave = tid / 64, lane = tid % 64;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < g.K; k0 += BK) {
    // stage A[m0:m0+64, k0:k0+32] and B[k0:k0+32, n0:n0+64]: 2048 elements each, 8 per thread
#pragma unroll
    for (int r = 0; r < 8; r++) {
      const int e = tid + r * 256;
      const int mi = e / BK, ki = e % BK;
      const int gm = m0 + mi, gk = k0 + ki;
      sA[mi][ki] = (gm < g.M && gk < g.K) ? g.A[gm * g.sam + gk * g.sak] : (uint16_t)0;
      const int ni = e / BK, kj = e % BK;
      const int gn = n0 + ni, gk2 = k0 + kj;
      sB[ni][kj] = (gn < g.N && gk2 < g.K) ? g.B[gk2 * g.sbk + gn * g.sbn] : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const int kb = kk + 4 * (lane / 16);
      bf16x4 a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; i++) {
        const int row = wm + 16 * i + (lane % 16);
        a[i] = bf16x4{(short)sA[row][kb], (short)sA[row][kb + 1], (short)sA[row][kb + 2], (short)sA[row][kb + 3]};
      }
#pragma unroll
      for (int j = 0; j < 2; j++) {
        const int col = wn + 16 * j + (lane % 16);
        b[j] = bf16x4{(short)sB[col][kb], (short)sB[col][kb + 1], (short)sB[col][kb + 2], (short)sB[col][kb + 3]};
      }
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int gm = m0 + wm + 16 * i + 4 * (lane / 16) + r;
        const int gn = n0 + wn + 16 * j + (lane % 16);
        if (gm < g.M && gn < g.N) {
          float v = g.alpha * acc[i][j][r];
          const long o = gm * g.ldc + gn;
          if (g.accumulate) v += g.C[o];
          if (g.bias) v += g.bias[gn];
          if (g.Cpre) g.Cpre[o] = v;
          if (g.act == 1) v = gelu(v);
          g.C[o] = v;
          if (g.Cbf) g.Cbf[o] = f2bf(v);
        }
      }
}

__global__ void f32_to_bf16_kernel(const float* x, uint16_t* y, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// dL/dY of the mean squared error, L = mean((Y - T)^2) over batch x outputs; also accumulates L.
__global__ void mse_grad_kernel(const float* Y, const float* T, int B, int D, float* dY, uint16_t* dYbf,
                                float* loss) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < (long)B * D) {
    const float d = Y[i] - T[i];
    const float gr = 2.0f * d / (float)((long)B * D);
    dY[i] = gr;
    dYbf[i] = f2bf(gr);
    l = d * d / (float)((long)B * D);
  }
  for (int m = 32; m > 0; m >>= 1) l += __shfl_xor(l, m, 64);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, l);
}

// dZ = dA * gelu'(pre) (elementwise), with a bf16 copy for the next GEMM.
__global__ void gelu_back_kernel(const float* dA, const float* pre, long n, float* dZ, uint16_t* dZbf) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float v = dA[i] * gelu_grad(pre[i]);
    dZ[i] = v;
    dZbf[i] = f2bf(v);
  }
}

// db[n] = sum_b dZ[b, n] in two passes with a fixed summation order (bitwise reproducible training):
// partial sums over chunks of kBiasRows rows, 64 columns x 4 row lanes per block so that every row is
// read in 256-byte runs, then one thread per column adds the chunks.
constexpr int kBiasRows = 256;

__global__ void __launch_bounds__(256) bias_grad_partial_kernel(const float* dZ, int B, int N, float* part) {
  const int c = blockIdx.x * 64 + (threadIdx.x & 63);
  const int r1 = min(B, (int)(blockIdx.y + 1) * kBiasRows);
  float s = 0.f;
  if (c < N)
    for (int r = blockIdx.y * kBiasRows + (threadIdx.x >> 6); r < r1; r += 4) s += dZ[(long)r * N + c];
  __shared__ float sh[4][64];
  sh[threadIdx.x >> 6][threadIdx.x & 63] = s;
  __syncthreads();
  if (threadIdx.x < 64 && c < N)
    part[(long)blockIdx.y * N + c] = (sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
}

__global__ void bias_grad_sum_kernel(const float* part, int chunks, int N, float* db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int k = 0; k < chunks; k++) s += part[(long)k * N + n];
  db[n] = s;
}

// Adam with bias correction (the update of design/optimize.py:95-104, applied to the MLP), and the
// bf16 copy of the new weights.
__global__ void adam_kernel(float* w, const float* g, float* m, float* v, long n, float lr, float b1, float b2,
                            float eps, float c1, float c2, uint16_t* wbf, long nbf) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wn = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wn;
  if (wbf && i < nbf) wbf[i] = f2bf(wn);
}

int gemm(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return BMC_OK;
  dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM);
  hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, s, g);
  BMC_HIP_CHECK(hipGetLastError());
  return BMC_OK;
}

}  // namespace mlp
}  // namespace bmc

using namespace bmc;
using namespace bmc::mlp;

struct bmc_mlp {
  int L = 0;                 // layers
  std::vector<int> dims;     // L + 1
  int max_batch = 0;
  std::vector<long> woff, boff;  // offsets of W_l and b_l in the flat parameter vector
  long n_params = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // fp32 [n_params]
  uint16_t* wbf = nullptr;                                        // bf16 copy [n_params]
  // activations: per layer output pre-activation and post (fp32 + bf16), input bf16
  std::vector<float*> pre, act, dact;
  std::vector<uint16_t*> act_bf, dact_bf;
  uint16_t* x_bf = nullptr;
  float* loss = nullptr;
  float* bias_part = nullptr;  // [ceil(max_batch / kBiasRows), max width] partial bias gradients
  int step = 0;
};

namespace {

void mlp_free(bmc_mlp* h) {
  if (!h) return;
  auto F = [](void* p) {
    if (p) (void)hipFree(p);
  };
  F(h->w);
  F(h->g);
  F(h->m);
  F(h->v);
  F(h->wbf);
  for (auto p : h->pre) F(p);
  for (auto p : h->act) F(p);
  for (auto p : h->dact) F(p);
  for (auto p : h->act_bf) F(p);
  for (auto p : h->dact_bf) F(p);
  F(h->x_bf);
  F(h->loss);
  F(h->bias_part);
  delete h;
}

int forward(bmc_mlp* h, const float* X, int B, hipStream_t s) {
  const long nin = (long)B * h->dims[0];
  hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((nin + 255) / 256), dim3(256), 0, s, X, h->x_bf, nin);
  BMC_HIP_CHECK(hipGetLastError());
  for (int l = 0; l < h->L; l++) {
    const int K = h->dims[l], N = h->dims[l + 1];
    GemmArgs g{};
    g.M = B;
    g.N = N;
    g.K = K;
    g.A = l == 0 ? h->x_bf : h->act_bf[l - 1];
    g.sam = K;
    g.sak = 1;
    g.B = h->wbf + h->woff[l];  // W [N, K]: B[k, n] = W[n, k]
    g.sbk = 1;
    g.sbn = K;
    g.C = h->act[l];
    g.ldc = N;
    g.bias = h->w + h->boff[l];
    g.act = (l + 1 < h->L) ? 1 : 0;
    g.Cpre = h->pre[l];
    g.Cbf = h->act_bf[l];
    g.alpha = 1.f;
    int rc = gemm(g, s);
    if (rc) return rc;
  }
  return BMC_OK;
}

// Gradients of the MSE loss w.r.t. every parameter into h->g (after forward at the same batch).
int backward(bmc_mlp* h, const float* T, int B, hipStream_t s) {
  const int L = h->L;
  const int D = h->dims[L];
  const long nout = (long)B * D;
  hipLaunchKernelGGL(mse_grad_kernel, dim3((nout + 255) / 256), dim3(256), 0, s, h->act[L - 1], T, B, D,
                     h->dact[L - 1], h->dact_bf[L - 1], h->loss);
  BMC_HIP_CHECK(hipGetLastError());
  for (int l = L - 1; l >= 0; l--) {
    const int K = h->dims[l], N = h->dims[l + 1];
    // dZ_l: for hidden layers dact * gelu'(pre); the last layer is linear
    float* dZ = h->dact[l];
    uint16_t* dZbf = h->dact_bf[l];
    if (l + 1 < L) {
      const long n = (long)B * N;
      hipLaunchKernelGGL(gelu_back_kernel, dim3((n + 255) / 256), dim3(256), 0, s, h->dact[l], h->pre[l], n, dZ,
                         dZbf);
      BMC_HIP_CHECK(hipGetLastError());
    }
    // dW_l [N, K] = dZ^T [N, B] X_l [B, K]
    GemmArgs gw{};
    gw.M = N;
    gw.N = K;
    gw.K = B;
    gw.A = dZbf;  // A[n, b] = dZ[b, n]
    gw.sam = 1;
    gw.sak = N;
    gw.B = l == 0 ? h->x_bf : h->act_bf[l - 1];  // B[b, k] = X[b, k]
    gw.sbk = K;
    gw.sbn = 1;
    gw.C = h->g + h->woff[l];
    gw.ldc = K;
    gw.alpha = 1.f;
    int rc = gemm(gw, s);
    if (rc) return rc;
    const int chunks = (B + kBiasRows - 1) / kBiasRows;
    hipLaunchKernelGGL(bias_grad_partial_kernel, dim3((N + 63) / 64, chunks), dim3(256), 0, s, dZ, B, N,
                       h->bias_part);
    hipLaunchKernelGGL(bias_grad_sum_kernel, dim3((N + 255) / 256), dim3(256), 0, s, h->bias_part, chunks, N,
                       h->g + h->boff[l]);
    BMC_HIP_CHECK(hipGetLastError());
    if (l > 0) {
      // dX_l [B, K] = dZ [B, N] W [N, K], into dact[l - 1]
      GemmArgs gx{};
      gx.M = B;
      gx.N = K;
      gx.K = N;
      gx.A = dZbf;
      gx.sam = N;
      gx.sak = 1;
      gx.B = h->wbf + h->woff[l];  // B[n, k] = W[n, k]
      gx.sbk = K;
      gx.sbn = 1;
      gx.C = h->dact[l - 1];
      gx.ldc = K;
      gx.alpha = 1.f;
      if ((rc = gemm(gx, s))) return rc;
    }
  }
  return BMC_OK;
}

}  // namespace

extern "C" {

int bmc_mlp_create(int n_layers, const int32_t* dims, int max_batch, uint64_t seed, bmc_mlp** out) {
  if (!out || n_layers < 1 || !dims || max_batch < 1) {
    set_error("bmc_mlp_create: bad arguments");
    return BMC_EINVAL;
  }
  bmc_mlp* h = new bmc_mlp();
  h->L = n_layers;
  h->dims.assign(dims, dims + n_layers + 1);
  h->max_batch = max_batch;
  for (int l = 0; l < n_layers; l++) {
    if (dims[l] < 1 || dims[l + 1] < 1) {
      delete h;
      set_error("bmc_mlp_create: layer widths must be positive");
      return BMC_EINVAL;
    }
    h->woff.push_back(h->n_params);
    h->n_params += (long)dims[l] * dims[l + 1];
    h->boff.push_back(h->n_params);
    h->n_params += dims[l + 1];
  }
  auto A = [&](void** p, size_t bytes) { return hipMalloc(p, bytes) == hipSuccess; };
  bool ok = A((void**)&h->w, 4 * h->n_params) && A((void**)&h->g, 4 * h->n_params) &&
            A((void**)&h->m, 4 * h->n_params) && A((void**)&h->v, 4 * h->n_params) &&
            A((void**)&h->wbf, 2 * h->n_params) && A((void**)&h->x_bf, 2L * max_batch * dims[0]) &&
            A((void**)&h->loss, 4) &&
            A((void**)&h->bias_part, 4L * ((max_batch + kBiasRows - 1) / kBiasRows) *
                                         *std::max_element(dims + 1, dims + n_layers + 1));
  for (int l = 0; ok && l < n_layers; l++) {
    const size_t n = (size_t)max_batch * dims[l + 1];
    float *a, *b, *c;
    uint16_t *d, *e;
    ok = A((void**)&a, 4 * n) && A((void**)&b, 4 * n) && A((void**)&c, 4 * n) && A((void**)&d, 2 * n) &&
         A((void**)&e, 2 * n);
    if (ok) {
      h->pre.push_back(a);
      h->act.push_back(b);
      h->dact.push_back(c);
      h->act_bf.push_back(d);
      h->dact_bf.push_back(e);
    }
  }
  if (!ok) {
    mlp_free(h);
    set_error("bmc_mlp_create: device allocation failed");
    return BMC_ENOMEM;
  }
  // He-uniform weights from a seeded host generator, zero biases
  std::vector<float> w(h->n_params, 0.f);
  uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
  auto rnd = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return ((st >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
  };
  for (int l = 0; l < n_layers; l++) {
    const double lim = sqrt(6.0 / dims[l]);
    for (long i = 0; i < (long)dims[l] * dims[l + 1]; i++) w[h->woff[l] + i] = (float)(lim * rnd());
  }
  int rc = bmc_mlp_set_params(h, w.data());
  if (rc) {
    mlp_free(h);
    return rc;
  }
  (void)hipMemset(h->m, 0, 4 * h->n_params);
  (void)hipMemset(h->v, 0, 4 * h->n_params);
  *out = h;
  return BMC_OK;
}

int bmc_mlp_destroy(bmc_mlp* h) {
  mlp_free(h);
  return BMC_OK;
}

int64_t bmc_mlp_n_params(const bmc_mlp* h) { return h ? h->n_params : -1; }

int bmc_mlp_set_params(bmc_mlp* h, const float* host_params) {
  if (!h || !host_params) {
    set_error("bmc_mlp_set_params: null argument");
    return BMC_EINVAL;
  }
  BMC_HIP_CHECK(hipMemcpy(h->w, host_params, 4 * h->n_params, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((h->n_params + 255) / 256), dim3(256), 0, 0, h->w, h->wbf, h->n_params);
  BMC_HIP_CHECK(hipGetLastError());
  BMC_HIP_CHECK(hipDeviceSynchronize());
  h->step = 0;
  return BMC_OK;
}

int bmc_mlp_get_params(const bmc_mlp* h, float* host_params) {
  if (!h || !host_params) {
    set_error("bmc_mlp_get_params: null argument");
    return BMC_EINVAL;
  }
  BMC_HIP_CHECK(hipMemcpy(host_params, h->w, 4 * h->n_params, hipMemcpyDeviceToHost));
  return BMC_OK;
}

int bmc_mlp_forward(bmc_mlp* h, const float* x_dev, int batch, float* y_dev, void* stream) {
  if (!h || !x_dev || !y_dev || batch < 0 || batch > h->max_batch) {
    set_error("bmc_mlp_forward: bad arguments (batch above max_batch?)");
    return BMC_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  int rc = forward(h, x_dev, batch, s);
  if (rc) return rc;
  const int D = h->dims[h->L];
  BMC_HIP_CHECK(hipMemcpyAsync(y_dev, h->act[h->L - 1], 4L * batch * D, hipMemcpyDeviceToDevice, s));
  return BMC_OK;
}

int bmc_mlp_gradients(bmc_mlp* h, const float* x_dev, const float* t_dev, int batch, float* grad_host,
                      float* loss_host, void* stream) {
  if (!h || !x_dev || !t_dev || batch < 1 || batch > h->max_batch) {
    set_error("bmc_mlp_gradients: bad arguments");
    return BMC_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  BMC_HIP_CHECK(hipMemsetAsync(h->loss, 0, 4, s));
  int rc = forward(h, x_dev, batch, s);
  if (!rc) rc = backward(h, t_dev, batch, s);
  if (rc) return rc;
  if (grad_host) BMC_HIP_CHECK(hipMemcpyAsync(grad_host, h->g, 4 * h->n_params, hipMemcpyDeviceToHost, s));
  if (loss_host) BMC_HIP_CHECK(hipMemcpyAsync(loss_host, h->loss, 4, hipMemcpyDeviceToHost, s));
  BMC_HIP_CHECK(hipStreamSynchronize(s));
  return BMC_OK;
}

int bmc_mlp_train_step(bmc_mlp* h, const float* x_dev, const float* t_dev, int batch, float lr, float* loss_dev,
                       void* stream) {
  if (!h || !x_dev || !t_dev || batch < 1 || batch > h->max_batch) {
    set_error("bmc_mlp_train_step: bad arguments");
    return BMC_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  float* lacc = loss_dev ? loss_dev : h->loss;
  int rc = forward(h, x_dev, batch, s);
  if (!rc) {
    // the loss is accumulated into *lacc (the caller clears it)
    float* keep = h->loss;
    h->loss = lacc;
    rc = backward(h, t_dev, batch, s);
    h->loss = keep;
  }
  if (rc) return rc;
  h->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - powf(b1, (float)h->step), c2 = 1.f - powf(b2, (float)h->step);
  hipLaunchKernelGGL(adam_kernel, dim3((h->n_params + 255) / 256), dim3(256), 0, s, h->w, h->g, h->m, h->v,
                     h->n_params, lr, b1, b2, 1e-8f, c1, c2, h->wbf, h->n_params);
  BMC_HIP_CHECK(hipGetLastError());
  return BMC_OK;
}

}  // extern "C"
