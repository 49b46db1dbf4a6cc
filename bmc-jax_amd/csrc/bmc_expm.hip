// Dense matrix exponential of a batch of small matrices, for callers that hold the matrix themselves
// (src/bmc_jax/simulation/expm.py:26-83 of the reference: expm_library and expm_plain, and the time step
// propagate built on them, :86-119). Inside a simulation the kernels of bmc_sim.hip take each step with
// their own truncated Taylor series; this file serves the host-facing functions.
//
// Algorithm: Pade approximation with scaling and squaring (Higham, "The scaling and squaring method for
// the matrix exponential revisited", SIAM J. Matrix Anal. Appl. 26 (2005) 1179-1193), with the order
// and scaling rule of the reference's exponentials:
//   l1 = ||A||_1, s = max(0, floor(log2(l1 / theta_max))), the Pade order from l1 (unscaled) against the
//   published thresholds, R = (V - U)^-1 (V + U) of A / 2^s, then s squarings; s > 16 gives NaN.
//   float64: orders 3, 5, 7, 9, 13 (theta_13 = 5.371920351148152);
//   float32: orders 3, 5, 7 (theta_7 = 3.925724783138660), all arithmetic in float32 (expm_plain).
// The solve is Gauss-Jordan elimination with partial pivoting, as expm_plain writes it (expm.py:36-47).
//
// One workgroup of 256 threads per matrix, every intermediate in LDS; n <= BMC_EXPM_MAX_N.
#include "bmc_internal.h"

namespace bmc {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxSquarings = 16;

// Pade numerator / denominator coefficients b_0..b_m (Higham 2005, Table 2.3 / Algorithm 2.3).
__constant__ double kB3[4] = {120., 60., 12., 1.};
__constant__ double kB5[6] = {30240., 15120., 3360., 420., 30., 1.};
__constant__ double kB7[8] = {17297280., 8648640., 1995840., 277200., 25200., 1512., 56., 1.};
__constant__ double kB9[10] = {17643225600., 8821612800., 2075673600., 302702400., 30270240., 2162160., 110880.,
                               3960., 90., 1.};
__constant__ double kB13[14] = {64764752532480000., 32382376266240000., 7771770303897600., 1187353796428800.,
                                129060195264000., 10559470521600., 670442572800., 33522128640., 1323241920.,
                                40840800., 960960., 16380., 182., 1.};

template <typename R>
struct Thresholds;
template <>
struct Thresholds<double> {
  static constexpr int n = 4;  // orders 3, 5, 7, 9 below their thresholds; 13 above
  __device__ static double at(int i) {
    const double t[4] = {1.495585217958292e-2, 2.539398330063230e-1, 9.504178996162932e-1, 2.097847961257068e0};
    return t[i];
  }
  static constexpr double max_norm = 5.371920351148152;
};
template <>
struct Thresholds<float> {
  static constexpr int n = 2;  // orders 3, 5 below their thresholds; 7 above
  __device__ static double at(int i) {
    const double t[2] = {4.258730016922831e-1, 1.880152677804762e0};
    return t[i];
  }
  static constexpr double max_norm = 3.925724783138660;
};

// C = A B for n x n row-major matrices in LDS. Every thread of the block calls it; the caller syncs.
template <typename R>
__device__ void matmul(R* C, const R* A, const R* B, int n) {
  for (int idx = threadIdx.x; idx < n * n; idx += kThreads) {
    const int i = idx / n, j = idx - i * n;
    R s = R(0);
    for (int k = 0; k < n; ++k) s += A[i * n + k] * B[k * n + j];
    C[idx] = s;
  }
}

// D = sum_t c_t X_t (+ c_I on the diagonal), X_t among the even powers.
template <typename R>
__device__ void combo(R* D, int n, R cI, R c2, const R* A2, R c4, const R* A4, R c6, const R* A6) {
  for (int idx = threadIdx.x; idx < n * n; idx += kThreads) {
    const int i = idx / n, j = idx - i * n;
    R v = (i == j) ? cI : R(0);
    if (A2) v += c2 * A2[idx];
    if (A4) v += c4 * A4[idx];
    if (A6) v += c6 * A6[idx];
    D[idx] = v;
  }
}

template <typename R>
__global__ void __launch_bounds__(kThreads) expm_kernel(const R* __restrict__ a, R* __restrict__ out, int n,
                                                        int64_t batch) {
  extern __shared__ unsigned char smem_raw[];
  R* sm = reinterpret_cast<R*>(smem_raw);
  const int nn = n * n;
  R* A = sm;            // scaled A
  R* A2 = A + nn;
  R* A4 = A2 + nn;
  R* A6 = A4 + nn;
  R* T = A6 + nn;       // scratch
  R* U = T + nn;
  R* V = U + nn;
  R* G = V + nn;        // Gauss-Jordan system [n, 2n]
  __shared__ double s_norm;
  __shared__ int s_piv;
  __shared__ R s_fac[BMC_EXPM_MAX_N];

  const int64_t b = blockIdx.x;
  if (b >= batch) return;
  const R* src = a + b * nn;
  R* dst = out + b * nn;

  for (int idx = threadIdx.x; idx < nn; idx += kThreads) A[idx] = src[idx];
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int j = 0; j < n; ++j) {
      double c = 0.0;
      for (int i = 0; i < n; ++i) c += fabs((double)A[i * n + j]);
      m = (c > m || c != c) ? c : m;
    }
    s_norm = m;
  }
  __syncthreads();
  const double l1 = s_norm;
  if (!(l1 <= 1.0e300)) {  // inf or NaN
    for (int idx = threadIdx.x; idx < nn; idx += kThreads) dst[idx] = R(NAN);
    return;
  }
  int s = 0;
  if (l1 > 0.0) {
    const double q = floor(log2(l1 / Thresholds<R>::max_norm));
    s = q > 0.0 ? (int)q : 0;
  }
  if (s > kMaxSquarings) {
    for (int idx = threadIdx.x; idx < nn; idx += kThreads) dst[idx] = R(NAN);
    return;
  }
  int order_idx = Thresholds<R>::n;  // the top order unless l1 falls below a threshold
  for (int i = Thresholds<R>::n - 1; i >= 0; --i)
    if (l1 < Thresholds<R>::at(i)) order_idx = i;
  const R scale = (R)ldexp(1.0, -s);
  for (int idx = threadIdx.x; idx < nn; idx += kThreads) A[idx] *= scale;
  __syncthreads();

  matmul(A2, A, A, n);
  __syncthreads();
  // U = A * (odd part), V = even part
  if (order_idx == 0) {
    combo<R>(T, n, (R)kB3[1], (R)kB3[3], A2, R(0), nullptr, R(0), nullptr);
    combo<R>(V, n, (R)kB3[0], (R)kB3[2], A2, R(0), nullptr, R(0), nullptr);
  } else if (order_idx == 1) {
    matmul(A4, A2, A2, n);
    __syncthreads();
    combo<R>(T, n, (R)kB5[1], (R)kB5[3], A2, (R)kB5[5], A4, R(0), nullptr);
    combo<R>(V, n, (R)kB5[0], (R)kB5[2], A2, (R)kB5[4], A4, R(0), nullptr);
  } else if (order_idx == 2) {
    matmul(A4, A2, A2, n);
    __syncthreads();
    matmul(A6, A4, A2, n);
    __syncthreads();
    combo<R>(T, n, (R)kB7[1], (R)kB7[3], A2, (R)kB7[5], A4, (R)kB7[7], A6);
    combo<R>(V, n, (R)kB7[0], (R)kB7[2], A2, (R)kB7[4], A4, (R)kB7[6], A6);
  } else if (order_idx == 3) {
    // order 9: A8 = A6 A2 enters both parts; it is formed in U first and folded in.
    matmul(A4, A2, A2, n);
    __syncthreads();
    matmul(A6, A4, A2, n);
    __syncthreads();
    matmul(U, A6, A2, n);  // A8
    __syncthreads();
    combo<R>(T, n, (R)kB9[1], (R)kB9[3], A2, (R)kB9[5], A4, (R)kB9[7], A6);
    combo<R>(V, n, (R)kB9[0], (R)kB9[2], A2, (R)kB9[4], A4, (R)kB9[6], A6);
    __syncthreads();
    for (int idx = threadIdx.x; idx < nn; idx += kThreads) {
      T[idx] += (R)kB9[9] * U[idx];
      V[idx] += (R)kB9[8] * U[idx];
    }
  } else {
    // order 13: U = A [A6 (b13 A6 + b11 A4 + b9 A2) + b7 A6 + b5 A4 + b3 A2 + b1 I],
    //           V = A6 (b12 A6 + b10 A4 + b8 A2) + b6 A6 + b4 A4 + b2 A2 + b0 I
    matmul(A4, A2, A2, n);
    __syncthreads();
    matmul(A6, A4, A2, n);
    __syncthreads();
    combo<R>(T, n, R(0), (R)kB13[9], A2, (R)kB13[11], A4, (R)kB13[13], A6);
    __syncthreads();
    matmul(U, A6, T, n);
    __syncthreads();
    for (int idx = threadIdx.x; idx < nn; idx += kThreads) {
      const int i = idx / n, j = idx - i * n;
      U[idx] += (R)kB13[7] * A6[idx] + (R)kB13[5] * A4[idx] + (R)kB13[3] * A2[idx] + (i == j ? (R)kB13[1] : R(0));
    }
    combo<R>(T, n, R(0), (R)kB13[8], A2, (R)kB13[10], A4, (R)kB13[12], A6);
    __syncthreads();
    matmul(V, A6, T, n);
    __syncthreads();
    for (int idx = threadIdx.x; idx < nn; idx += kThreads) {
      const int i = idx / n, j = idx - i * n;
      V[idx] += (R)kB13[6] * A6[idx] + (R)kB13[4] * A4[idx] + (R)kB13[2] * A2[idx] + (i == j ? (R)kB13[0] : R(0));
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < nn; idx += kThreads) T[idx] = U[idx];
  }
  __syncthreads();
  matmul(U, A, T, n);  // U = A * odd part
  __syncthreads();

  // Gauss-Jordan on [V - U | V + U] with partial pivoting
  const int w = 2 * n;
  for (int idx = threadIdx.x; idx < nn; idx += kThreads) {
    const int i = idx / n, j = idx - i * n;
    G[i * w + j] = V[idx] - U[idx];
    G[i * w + n + j] = V[idx] + U[idx];
  }
  __syncthreads();
  for (int c = 0; c < n; ++c) {
    if (threadIdx.x == 0) {
      int p = c;
      R best = fabs(G[c * w + c]);
      for (int r = c + 1; r < n; ++r) {
        const R v = fabs(G[r * w + c]);
        if (v > best) {
          best = v;
          p = r;
        }
      }
      s_piv = p;
    }
    __syncthreads();
    const int p = s_piv;
    if (p != c) {
      for (int j = threadIdx.x; j < w; j += kThreads) {
        const R t = G[c * w + j];
        G[c * w + j] = G[p * w + j];
        G[p * w + j] = t;
      }
    }
    __syncthreads();
    const R inv = R(1) / G[c * w + c];
    __syncthreads();
    for (int j = threadIdx.x; j < w; j += kThreads) G[c * w + j] *= inv;
    __syncthreads();
    for (int r = threadIdx.x; r < n; r += kThreads) s_fac[r] = (r == c) ? R(0) : G[r * w + c];
    __syncthreads();
    for (int idx = threadIdx.x; idx < n * w; idx += kThreads) {
      const int r = idx / w, j = idx - r * w;
      if (r != c) G[idx] -= s_fac[r] * G[c * w + j];
    }
    __syncthreads();
  }
  R* cur = T;
  R* nxt = U;
  for (int idx = threadIdx.x; idx < nn; idx += kThreads) {
    const int i = idx / n, j = idx - i * n;
    cur[idx] = G[i * w + n + j];
  }
  __syncthreads();
  for (int k = 0; k < s; ++k) {
    matmul(nxt, cur, cur, n);
    __syncthreads();
    R* t = cur;
    cur = nxt;
    nxt = t;
  }
  for (int idx = threadIdx.x; idx < nn; idx += kThreads) dst[idx] = cur[idx];
}

template <typename R>
int launch_expm(const void* a, void* out, int n, int64_t batch, hipStream_t stream) {
  const size_t lds = (size_t)(7 * n * n + 2 * n * n) * sizeof(R);
  hipLaunchKernelGGL(expm_kernel<R>, dim3((unsigned)batch), dim3(kThreads), lds, stream, (const R*)a, (R*)out, n,
                     batch);
  BMC_HIP_CHECK(hipGetLastError());
  return BMC_OK;
}

}  // namespace
}  // namespace bmc

using namespace bmc;

extern "C" int bmc_expm(const void* a_dev, void* out_dev, int n, int64_t batch, int dtype, void* stream) {
  if (n < 1 || n > BMC_EXPM_MAX_N) {
    set_error("bmc_expm: n must be 1.." + std::to_string(BMC_EXPM_MAX_N) + ", not " + std::to_string(n));
    return BMC_EINVAL;
  }
  if (batch < 0 || batch > 0x7fffffffLL) {
    set_error("bmc_expm: batch must be 0..2^31-1");
    return BMC_EINVAL;
  }
  if (dtype != BMC_F32 && dtype != BMC_F64) {
    set_error("bmc_expm: dtype must be BMC_F32 or BMC_F64");
    return BMC_EINVAL;
  }
  if (batch == 0) return BMC_OK;
  if (!a_dev || !out_dev) {
    set_error("bmc_expm: null matrix pointer");
    return BMC_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  return dtype == BMC_F64 ? launch_expm<double>(a_dev, out_dev, n, batch, s)
                          : launch_expm<float>(a_dev, out_dev, n, batch, s);
}
