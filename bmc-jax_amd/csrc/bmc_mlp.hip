// The learned Levenberg-Marquardt proposal: a small MLP trained and run with hand-written BF16 MFMA
// GEMMs (FP32 accumulate) on CDNA4. No counterpart in the reference; it supplies the LM start x0.
//
// Layout: activations row-major [batch, features]; weights W_l [out, in] (FP32 master copy), biases [out].
// Hidden layers use GELU (0.5 x (1 + erf(x / sqrt 2))), the last layer is linear; the loss is the mean
// over batch and outputs of the squared error. The input is standardised per feature before the first
// GEMM, (x - shift) * scale, with shift / scale fitted to the training set (bmc_mlp_fit_input_norm).
//
// GEMM: every product is C[M, N] = A[M, K] B^T with both A [M, K] and B [N, K] K-contiguous BF16, so that
// operands are staged with 16-byte loads along K. Each BF16 activation and weight is therefore kept in
// the two orientations the products need: activations [batch, features] (forward, input gradient) and
// [features, batch] (weight gradient), weights [out, in] (forward) and [in, out] (input gradient); the
// transposed copies are written by LDS-tiled epilogue kernels. Leading dimensions are padded to 8
// elements so every 16-byte load is aligned; reads past the logical K are masked to zero.
//
// The kernel: 256 threads (2 x 2 waves) per BM x BN tile (128 x 128 or 64 x 64), K in slabs of 64
// staged through double-buffered LDS, v_mfma_f32_32x32x16_bf16 (lane l: A row l % 32 with k = 8 (l / 32)
// .. + 7, B column l % 32 with the same k, D column l % 32 with rows 8 q + 4 (l / 32) .. + 3 in register
// group q; per flop half the LDS reads of the 16 x 16 x 32 form, which bounded the loop). Products with too
// few tiles to fill 256 CUs (the weight gradients, K = batch) split K over the grid's z axis into FP32
// partials that a second kernel adds in a fixed order (training stays bitwise reproducible).
#include <math.h>

#include <algorithm>
#include <string.h>

#include <string>
#include <vector>

#include "bmc_internal.h"

namespace bmc {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}

inline long pad8(long n) { return (n + 7) & ~7L; }

constexpr int BK = 64;         // K per LDS slab
constexpr int LDS_ROW = BK + 8;  // bf16 per LDS row: 144 bytes, so the 16 rows a fragment read touches
                                 // fall on distinct banks

// Epilogue of one output element (applied by the GEMM when K is not split, else by the reduction).
struct Epi {
  float* C;          // [M, ldc] fp32 or null
  long ldc;
  float alpha;
  int accumulate;    // C += result
  const float* bias; // [N] or null
  int act;           // 0 none, 1 GELU
  float* Cpre;       // pre-activation [M, ldc] or null
  uint16_t* Cbf;     // bf16 [M, ldcb] or null
  long ldcb;
  uint16_t* CbfT;    // bf16 transposed [N, ldct] or null
  long ldct;
};

struct GemmArgs {
  int M, N, K;
  const uint16_t* A;  // [M, lda] bf16, K-contiguous
  long lda;
  const uint16_t* B;  // [N, ldb] bf16, K-contiguous
  long ldb;
  int splits, kchunk; // K split over gridDim.z in chunks of kchunk (a multiple of BK)
  float* part;        // [splits, M, N] fp32 partials when splits > 1
  Epi e;
};

// erfc(z) by the Chebyshev fit of Numerical Recipes (Press et al., routine erfcc): with t = 1 / (1 + |z| / 2),
// erfc(|z|) = t exp(-z^2 + P(t)), P of degree 9, fractional error below 1.2e-7 in exact arithmetic (a few
// float32 ulp more here), and erfc(-z) = 2 - erfc(z). Branch-free: one reciprocal and one exponential
// per value, where a two-range erf diverges within a wave.
__device__ __forceinline__ float erfc_cheb(float z) {
  const float a = fabsf(z);
  const float t = __builtin_amdgcn_rcpf(fmaf(0.5f, a, 1.0f));  // v_rcp_f32, 1 ulp
  float p = 0.17087277f;
  p = fmaf(t, p, -0.82215223f);
  p = fmaf(t, p, 1.48851587f);
  p = fmaf(t, p, -1.13520398f);
  p = fmaf(t, p, 0.27886807f);
  p = fmaf(t, p, -0.18628806f);
  p = fmaf(t, p, 0.09678418f);
  p = fmaf(t, p, 0.37409196f);
  p = fmaf(t, p, 1.00002368f);
  p = fmaf(t, p, -1.26551223f);
  const float e = t * __expf(fmaf(-a, a, p));
  return z >= 0.f ? e : 2.0f - e;
}
// GELU x Phi(x) = x erfc(-x / sqrt 2) / 2 (no cancellation for negative x) and its derivative
// Phi(x) + x phi(x).
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * erfc_cheb(-0.70710678118654752f * x); }
__device__ __forceinline__ float gelu_grad(float x) {
  return fmaf(x * 0.39894228040143268f, __expf(-0.5f * x * x), 0.5f * erfc_cheb(-0.70710678118654752f * x));
}

// v (the accumulated product) -> every output the epilogue asks for, except the transposed bf16 copy.
// b: the bias of column n (read by the caller once per column, ahead of the stores).
__device__ __forceinline__ float finish(const Epi& e, int m, int n, float v, float b) {
  v *= e.alpha;
  const long o = (long)m * e.ldc + n;
  if (e.accumulate) v += e.C[o];
  if (e.bias) v += b;
  if (e.Cpre) e.Cpre[o] = v;
  if (e.act == 1) v = gelu(v);
  if (e.C) e.C[o] = v;
  if (e.Cbf) e.Cbf[(long)m * e.ldcb + n] = f2bf(v);
  return v;
}

// 8 bf16 of row `row` from k, zero past K (and past the last row).
__device__ __forceinline__ uint4 load8(const uint16_t* base, long ld, int row, int rows, int k, int K) {
  if (row >= rows || k >= K) return make_uint4(0u, 0u, 0u, 0u);
  const uint16_t* p = base + (long)row * ld + k;
  if (k + 8 <= K) return *reinterpret_cast<const uint4*>(p);
  uint16_t t[8];
#pragma unroll
  for (int i = 0; i < 8; i++) t[i] = (k + i < K) ? p[i] : (uint16_t)0;
  return make_uint4(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16), t[4] | ((uint32_t)t[5] << 16),
                    t[6] | ((uint32_t)t[7] << 16));
}

// The first of the four consecutive rows that register group q (registers 4 q .. 4 q + 3) of a 32 x 32
// v_mfma_f32_32x32x16_bf16 result holds in lane `lane`, for a fragment starting at row r0: the lane holds
// column lane % 32 at rows 8 q + 4 (lane / 32) + (0..3).
__device__ __forceinline__ int frag_row(int r0, int lane, int q) {
  return r0 + 8 * q + 4 * (lane >> 5);
}

template <int BM, int BN>
__global__ void __launch_bounds__(256) gemm_nt_kernel(GemmArgs g) {
  constexpr int FM = BM / 64, FN = BN / 64;        // 32 x 32 fragments per wave (wave tile BM/2 x BN/2)
  constexpr int CA = BM * BK / 8 / 256, CB = BN * BK / 8 / 256;  // 16-byte chunks per thread per slab
  __shared__ __attribute__((aligned(16))) uint16_t sA[2][BM * LDS_ROW];
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][BN * LDS_ROW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // XCD-aware tile order: workgroups go to the 8 XCDs round-robin by linear id (x fastest), and each XCD
  // has its own L2. In grid order the gridDim.x column tiles of one row block would run on different
  // XCDs, each fetching that row block of A from HBM; remapped, XCD x takes row blocks x, x + 8, ... with
  // all their column tiles in a row, so a row block of A is fetched once. A bijection when gridDim.y is
  // a multiple of 8 (else, and for split-K grids, the grid order is kept).
  int bx = blockIdx.x, by = blockIdx.y;
  if (gridDim.z == 1 && (gridDim.y & 7) == 0) {
    const int b = blockIdx.y * gridDim.x + blockIdx.x, k = b >> 3;
    by = (b & 7) + 8 * (k / (int)gridDim.x);
    bx = k % (int)gridDim.x;
  }
  const int m0 = by * BM, n0 = bx * BN;
  const int wm = (wave >> 1) * (BM / 2), wn = (wave & 1) * (BN / 2);
  const int kbeg = blockIdx.z * g.kchunk;
  const int kend = min(g.K, kbeg + g.kchunk);
  floatx16 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++)
#pragma unroll
      for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;

  uint4 ra[CA], rb[CB];
  // a slab wholly inside both operands (the tile's rows and the slab's K; a uniform test) is read with
  // plain 16-byte loads, all in flight together: load8's element-wise edge path, joined after every load,
  // made the compiler wait for each load before issuing the next
  const bool rows_in = m0 + BM <= g.M && n0 + BN <= g.N;
  auto fetch = [&](int k0) {
    if (rows_in && k0 + BK <= kend) {
#pragma unroll
      for (int c = 0; c < CA; c++) {
        const int e = tid + c * 256, row = e >> 3, kc = (e & 7) * 8;
        ra[c] = *reinterpret_cast<const uint4*>(g.A + (long)(m0 + row) * g.lda + k0 + kc);
      }
#pragma unroll
      for (int c = 0; c < CB; c++) {
        const int e = tid + c * 256, row = e >> 3, kc = (e & 7) * 8;
        rb[c] = *reinterpret_cast<const uint4*>(g.B + (long)(n0 + row) * g.ldb + k0 + kc);
      }
      return;
    }
#pragma unroll
    for (int c = 0; c < CA; c++) {
      const int e = tid + c * 256, row = e >> 3, kc = (e & 7) * 8;
      ra[c] = load8(g.A, g.lda, m0 + row, g.M, k0 + kc, kend);
    }
#pragma unroll
    for (int c = 0; c < CB; c++) {
      const int e = tid + c * 256, row = e >> 3, kc = (e & 7) * 8;
      rb[c] = load8(g.B, g.ldb, n0 + row, g.N, k0 + kc, kend);
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int c = 0; c < CA; c++) {
      const int e = tid + c * 256, row = e >> 3, kc = (e & 7) * 8;
      *reinterpret_cast<uint4*>(&sA[buf][row * LDS_ROW + kc]) = ra[c];
    }
#pragma unroll
    for (int c = 0; c < CB; c++) {
      const int e = tid + c * 256, row = e >> 3, kc = (e & 7) * 8;
      *reinterpret_cast<uint4*>(&sB[buf][row * LDS_ROW + kc]) = rb[c];
    }
  };

  int buf = 0;
  if (kbeg < kend) {
    fetch(kbeg);
    stash(0);
  }
  __syncthreads();
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    const bool more = k0 + BK < kend;
    if (more) fetch(k0 + BK);  // the next slab's global loads overlap this slab's MFMAs
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      const int kk = ks + 8 * (lane >> 5);
      bf16x8 a[FM], b[FN];
#pragma unroll
      for (int i = 0; i < FM; i++)
        a[i] = *reinterpret_cast<const bf16x8*>(&sA[buf][(wm + 32 * i + (lane & 31)) * LDS_ROW + kk]);
#pragma unroll
      for (int j = 0; j < FN; j++)
        b[j] = *reinterpret_cast<const bf16x8*>(&sB[buf][(wn + 32 * j + (lane & 31)) * LDS_ROW + kk]);
#pragma unroll
      for (int i = 0; i < FM; i++)
#pragma unroll
        for (int j = 0; j < FN; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (more) stash(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  if (g.splits > 1) {
    float* P = g.part + (long)blockIdx.z * g.M * g.N;
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
      for (int j = 0; j < FN; j++)
#pragma unroll
        for (int r = 0; r < 16; r++) {
          const int m = frag_row(m0 + wm + 32 * i, lane, r >> 2) + (r & 3), n = n0 + wn + 32 * j + (lane & 31);
          if (m < g.M && n < g.N) P[(long)m * g.N + n] = acc[i][j][r];
        }
    return;
  }
  // each column's bias once, all loads in flight together: read after the first store, it would be read
  // again for every element (a store may alias it), each time waiting out the memory latency
  float bj[FN];
#pragma unroll
  for (int j = 0; j < FN; j++) {
    const int n = n0 + wn + 32 * j + (lane & 31);
    bj[j] = (g.e.bias && n < g.N) ? g.e.bias[n] : 0.f;
  }
  // a hidden layer's forward on a full tile (the bulk of every forward): bias, GELU, the bf16 activation
  // and, in training, the pre-activation and the transposed bf16 copy, with no per-element bounds or flag
  // tests and the row addresses formed once per fragment; the same values as the general form below
  // (alpha 1: v * alpha is v exactly)
  if (m0 + BM <= g.M && n0 + BN <= g.N && g.e.alpha == 1.f && !g.e.accumulate && g.e.bias && g.e.act == 1 &&
      !g.e.C && g.e.Cbf) {
    float* const pre = g.e.Cpre;
    uint16_t* const cbT = g.e.CbfT;
#pragma unroll
    for (int i = 0; i < FM; i++)
#pragma unroll
      for (int j = 0; j < FN; j++)
#pragma unroll
      for (int q4 = 0; q4 < 4; q4++) {
        const int mb = frag_row(m0 + wm + 32 * i, lane, q4), n = n0 + wn + 32 * j + (lane & 31);
        uint16_t* cb = g.e.Cbf + (long)mb * g.e.ldcb + n;
        uint16_t t[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const float v = acc[i][j][4 * q4 + r] + bj[j];
          if (pre) pre[(long)(mb + r) * g.e.ldc + n] = v;
          t[r] = f2bf(gelu(v));
          cb[(long)r * g.e.ldcb] = t[r];
        }
        if (cbT)  // four consecutive rows of column n: one 8-byte store
          *reinterpret_cast<uint2*>(cbT + (long)n * g.e.ldct + mb) =
              make_uint2(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16));
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < FM; i++)
#pragma unroll
    for (int j = 0; j < FN; j++)
#pragma unroll
    for (int q4 = 0; q4 < 4; q4++) {
      const int mb = frag_row(m0 + wm + 32 * i, lane, q4), n = n0 + wn + 32 * j + (lane & 31);
      if (n >= g.N) continue;
      uint16_t t[4];
#pragma unroll
      for (int r = 0; r < 4; r++)
        t[r] = (mb + r < g.M) ? f2bf(finish(g.e, mb + r, n, acc[i][j][4 * q4 + r], bj[j])) : (uint16_t)0;
      if (g.e.CbfT) {
        uint16_t* q = g.e.CbfT + (long)n * g.e.ldct + mb;
        if (mb + 4 <= g.M) {  // four consecutive rows of column n: one 8-byte store
          *reinterpret_cast<uint2*>(q) = make_uint2(t[0] | ((uint32_t)t[1] << 16), t[2] | ((uint32_t)t[3] << 16));
        } else {
          for (int r = 0; r < 4 && mb + r < g.M; r++) q[r] = t[r];
        }
      }
    }
}

// Sum of the split-K partials in a fixed order, then the epilogue.
__global__ void __launch_bounds__(256) gemm_reduce_kernel(GemmArgs g) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)g.M * g.N) return;
  const int m = (int)(i / g.N), n = (int)(i % g.N);
  float v = 0.f;
  for (int s = 0; s < g.splits; s++) v += g.part[(long)s * g.M * g.N + i];
  v = finish(g.e, m, n, v, g.e.bias ? g.e.bias[n] : 0.f);
  if (g.e.CbfT) g.e.CbfT[(long)n * g.e.ldct + m] = f2bf(v);
}

// Elementwise producers of a GEMM operand, over 64 x 64 tiles of an fp32 [rows, cols] matrix, writing
// the bf16 copy [rows, ldo] and its transpose [cols, ldt] (both through LDS, so both stores coalesce):
//   TILE_INPUT  v = (x - shift[c]) * scale[c]    (shift / scale null: v = x; also packs the weights)
//   TILE_GELU   v = x * gelu'(pre), v also written back over x in fp32 (dZ of a hidden layer)
//   TILE_MSE    v = 2 (x - t) / (rows cols), v written to out32; the loss mean((x - t)^2) into *loss
// With colsum set (TILE_GELU, TILE_MSE: v is a layer's dZ), each tile also writes the column sums of its
// 64 rows of v, summed in a fixed order, to colsum[tile row][cols]: the bias gradient's partial sums.
enum { TILE_INPUT = 0, TILE_GELU = 1, TILE_MSE = 2 };

struct TileArgs {
  int mode, rows, cols;
  const float* x;     // [rows, ldx]
  long ldx;
  const float* aux;   // shift [cols] (input), pre [rows, ldx] (gelu), t [rows, ldx] (mse)
  const float* scale; // [cols] (input) or null
  float* out32;       // fp32 result [rows, ldx] or null
  uint16_t* bf;       // [rows, ldo] or null
  long ldo;
  uint16_t* bfT;      // [cols, ldt] or null
  long ldt;
  float* loss;
  float* colsum;      // [ceil(rows / 64), cols] or null
};

__device__ __forceinline__ void tile_body(const TileArgs& a, int tile_r, int tile_c) {
  __shared__ uint16_t t[64][66];
  const int r0 = tile_r * 64, c0 = tile_c * 64;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  float lsum = 0.f, csum = 0.f;
  const float inv = 1.0f / ((float)a.rows * (float)a.cols);
#pragma unroll 4
  for (int k = 0; k < 16; k++) {
    const int r = r0 + 4 * k + ty, c = c0 + tx;
    uint16_t h = 0;
    if (r < a.rows && c < a.cols) {
      const long o = (long)r * a.ldx + c;
      const float x = a.x[o];
      float v;
      if (a.mode == TILE_INPUT) {
        v = a.aux ? (x - a.aux[c]) * a.scale[c] : x;
      } else if (a.mode == TILE_GELU) {
        v = x * gelu_grad(a.aux[o]);
        a.out32[o] = v;
      } else {
        const float d = x - a.aux[o];
        v = 2.0f * d * inv;
        a.out32[o] = v;
        lsum += d * d * inv;
      }
      if (a.mode != TILE_INPUT) csum += v;
      h = f2bf(v);
      if (a.bf) a.bf[(long)r * a.ldo + c] = h;
    }
    t[4 * k + ty][tx] = h;
  }
  if (a.mode == TILE_MSE && a.loss) {
    for (int m = 32; m > 0; m >>= 1) lsum += __shfl_xor(lsum, m, 64);
    if (tx == 0) atomicAdd(a.loss, lsum);
  }
  if (a.colsum) {
    __shared__ float cs[4][64];
    cs[ty][tx] = csum;
    __syncthreads();
    if (ty == 0 && c0 + tx < a.cols)
      a.colsum[(long)tile_r * a.cols + c0 + tx] = (cs[0][tx] + cs[1][tx]) + (cs[2][tx] + cs[3][tx]);
  }
  if (!a.bfT) return;
  __syncthreads();
#pragma unroll 4
  for (int k = 0; k < 16; k++) {
    const int c = c0 + 4 * k + ty, r = r0 + tx;
    if (c < a.cols && r < a.rows) a.bfT[(long)c * a.ldt + r] = t[tx][4 * k + ty];
  }
}

__global__ void __launch_bounds__(256) tile_kernel(TileArgs a) { tile_body(a, blockIdx.y, blockIdx.x); }

// The standardised input alone (inference: no transposed copy): four consecutive features per thread, a
// 16-byte load and an 8-byte store (cols, ldx and ldo multiples of 4, x 16-byte aligned; the host checks).
__global__ void __launch_bounds__(256) input_kernel(const float* x, long ldx, const float* shift, const float* scale,
                                                    int rows, int cols, uint16_t* bf, long ldo) {
  const int q = cols / 4;
  const long n = (long)rows * q;
  for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
    const int r = (int)(e / q), c = (int)(e % q) * 4;
    const float4 v = *reinterpret_cast<const float4*>(x + (long)r * ldx + c);
    float w[4] = {v.x, v.y, v.z, v.w};
    if (shift) {
#pragma unroll
      for (int k = 0; k < 4; k++) w[k] = (w[k] - shift[c + k]) * scale[c + k];
    }
    *reinterpret_cast<uint2*>(bf + (long)r * ldo + c) =
        make_uint2(f2bf(w[0]) | ((uint32_t)f2bf(w[1]) << 16), f2bf(w[2]) | ((uint32_t)f2bf(w[3]) << 16));
  }
}

// Every layer's weight copies (bf16 W and W^T from the fp32 master) in one launch: block b takes tile
// b - first[l] of the layer l whose range holds it.
constexpr int kMaxPackLayers = 8;
struct PackArgs {
  int L;
  int first[kMaxPackLayers + 1];  // tile ranges, first[L] = total
  TileArgs t[kMaxPackLayers];
};

__global__ void __launch_bounds__(256) pack_kernel(PackArgs p) {
  int l = 0;
  while (l + 1 < p.L && (int)blockIdx.x >= p.first[l + 1]) l++;
  const TileArgs& a = p.t[l];
  const int tiles_c = (a.cols + 63) / 64;
  const int i = blockIdx.x - p.first[l];
  tile_body(a, i / tiles_c, i % tiles_c);
}

// Column mean and standard deviation of x [n, d] in float64 (the input standardisation), in two passes
// with a fixed summation order: partial sums over chunks of rows (64 columns x 4 row lanes per block, every
// row read in 256-byte runs), then one thread per column adds the chunks.
constexpr int kStatRows = 2048;

__global__ void __launch_bounds__(256) column_stats_partial_kernel(const float* x, long n, int d, double* part) {
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), q = threadIdx.x >> 6;
  const long r0 = (long)blockIdx.y * kStatRows, r1 = min(n, r0 + kStatRows);
  double s = 0.0, s2 = 0.0;
  if (c < d)
    for (long r = r0 + q; r < r1; r += 4) {
      const double v = x[r * d + c];
      s += v;
      s2 += v * v;
    }
  __shared__ double sh[2][4][64];
  sh[0][q][threadIdx.x & 63] = s;
  sh[1][q][threadIdx.x & 63] = s2;
  __syncthreads();
  if (q == 0 && c < d) {
    const int j = threadIdx.x & 63;
    part[((long)blockIdx.y * 2 + 0) * d + c] = (sh[0][0][j] + sh[0][1][j]) + (sh[0][2][j] + sh[0][3][j]);
    part[((long)blockIdx.y * 2 + 1) * d + c] = (sh[1][0][j] + sh[1][1][j]) + (sh[1][2][j] + sh[1][3][j]);
  }
}

__global__ void column_stats_final_kernel(const double* part, int chunks, long n, int d, float* shift, float* scale) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d) return;
  double S = 0.0, S2 = 0.0;
  for (int k = 0; k < chunks; k++) {
    S += part[((long)k * 2 + 0) * d + c];
    S2 += part[((long)k * 2 + 1) * d + c];
  }
  const double mean = S / (double)n;
  const double var = fmax(S2 / (double)n - mean * mean, 0.0);
  shift[c] = (float)mean;
  scale[c] = (float)(1.0 / fmax(sqrt(var), 1e-6));
}

// db[n] = sum_b dZ[b, n]: the tiles' column sums (tile_kernel's colsum) added in a fixed order, so that
// training stays bitwise reproducible.
__global__ void __launch_bounds__(256) bias_grad_sum_kernel(const float* part, int chunks, int N, float* db) {
  // 64 columns per block, four lanes per column each adding every fourth chunk, then the four in order
  const int n = blockIdx.x * 64 + (threadIdx.x & 63), q = threadIdx.x >> 6;
  float s = 0.f;
  if (n < N)
    for (int k = q; k < chunks; k += 4) s += part[(long)k * N + n];
  __shared__ float sh[4][64];
  sh[q][threadIdx.x & 63] = s;
  __syncthreads();
  if (q == 0 && n < N) {
    const int j = threadIdx.x & 63;
    db[n] = (sh[0][j] + sh[1][j]) + (sh[2][j] + sh[3][j]);
  }
}

// Adam with bias correction (the update of design/optimize.py:95-104, applied to the MLP).
__global__ void adam_kernel(float* w, const float* g, float* m, float* v, long n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

int tiles(const TileArgs& a, hipStream_t s) {
  if (a.rows <= 0 || a.cols <= 0) return BMC_OK;
  hipLaunchKernelGGL(tile_kernel, dim3((a.cols + 63) / 64, (a.rows + 63) / 64), dim3(256), 0, s, a);
  BMC_HIP_CHECK(hipGetLastError());
  return BMC_OK;
}

// Launch shape: 128 x 128 tiles when they alone fill the CUs, else 64 x 64; a product with fewer tiles
// than CUs (the weight gradients, K = batch) splits K until about 512 workgroups run, each keeping at
// least four slabs. A product that already has a tile per CU is not split: its partial sums would cost
// a reduction pass longer than the product itself.
int gemm(GemmArgs g, float* part, size_t part_floats, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0) return BMC_OK;
  const long t128 = (long)((g.M + 127) / 128) * ((g.N + 127) / 128);
  const long t64 = (long)((g.M + 63) / 64) * ((g.N + 63) / 64);
  const bool big = t128 >= 256;
  const long tiles_n = big ? t128 : t64;
  int splits = 1;
  const int slabs = (g.K + BK - 1) / BK;
  while (tiles_n < 256 && tiles_n * splits < 512 && slabs / (2 * splits) >= 4 &&
         (size_t)(2 * splits) * g.M * g.N <= part_floats)
    splits *= 2;
  g.splits = splits;
  g.kchunk = ((slabs + splits - 1) / splits) * BK;
  g.part = part;
  const dim3 grid(big ? (g.N + 127) / 128 : (g.N + 63) / 64, big ? (g.M + 127) / 128 : (g.M + 63) / 64, splits);
  if (big)
    hipLaunchKernelGGL((gemm_nt_kernel<128, 128>), grid, dim3(256), 0, s, g);
  else
    hipLaunchKernelGGL((gemm_nt_kernel<64, 64>), grid, dim3(256), 0, s, g);
  BMC_HIP_CHECK(hipGetLastError());
  if (splits > 1) {
    const long n = (long)g.M * g.N;
    hipLaunchKernelGGL(gemm_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g);
    BMC_HIP_CHECK(hipGetLastError());
  }
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
  long ldb = 0;              // padded batch (leading dimension of every transposed activation)
  std::vector<long> woff, boff;  // offsets of W_l and b_l in the flat parameter vector
  long n_params = 0;
  float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // fp32 [n_params]
  std::vector<uint16_t*> wbf, wbfT;  // W_l as bf16 [out, pad8(in)] and [in, pad8(out)]
  // per layer l (output width dims[l+1]): pre-activation, fp32 output (last layer), gradient of the
  // output / pre-activation (fp32), bf16 copies in both orientations
  std::vector<float*> pre, act, dact;
  std::vector<uint16_t*> act_bf, act_bfT, dz_bf, dz_bfT;
  uint16_t *x_bf = nullptr, *x_bfT = nullptr;  // standardised input [B, pad8(d0)] and [d0, ldb]
  float *shift = nullptr, *scale = nullptr;    // input standardisation [d0] (null until set)
  float* loss = nullptr;
  float* bias_part = nullptr;  // [ceil(max_batch / 64), max width] partial bias gradients (tile column sums)
  float* part = nullptr;       // split-K partials
  size_t part_floats = 0;
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
  for (auto p : h->wbf) F(p);
  for (auto p : h->wbfT) F(p);
  for (auto p : h->pre) F(p);
  for (auto p : h->act) F(p);
  for (auto p : h->dact) F(p);
  for (auto p : h->act_bf) F(p);
  for (auto p : h->act_bfT) F(p);
  for (auto p : h->dz_bf) F(p);
  for (auto p : h->dz_bfT) F(p);
  F(h->x_bf);
  F(h->x_bfT);
  F(h->shift);
  F(h->scale);
  F(h->loss);
  F(h->bias_part);
  F(h->part);
  delete h;
}

// bf16 copies of every W_l in both orientations, from the fp32 master (one launch; layers beyond the
// table's size take launches of their own).
int pack_weights(bmc_mlp* h, hipStream_t s) {
  for (int l0 = 0; l0 < h->L; l0 += kMaxPackLayers) {
    PackArgs p{};
    p.L = std::min(kMaxPackLayers, h->L - l0);
    int tiles = 0;
    for (int j = 0; j < p.L; j++) {
      const int l = l0 + j, K = h->dims[l], N = h->dims[l + 1];
      TileArgs& t = p.t[j];
      t.mode = TILE_INPUT;
      t.rows = N;
      t.cols = K;
      t.x = h->w + h->woff[l];
      t.ldx = K;
      t.bf = h->wbf[l];
      t.ldo = pad8(K);
      t.bfT = h->wbfT[l];
      t.ldt = pad8(N);
      p.first[j] = tiles;
      tiles += ((N + 63) / 64) * ((K + 63) / 64);
    }
    p.first[p.L] = tiles;
    hipLaunchKernelGGL(pack_kernel, dim3(tiles), dim3(256), 0, s, p);
    BMC_HIP_CHECK(hipGetLastError());
  }
  return BMC_OK;
}

// train: also keep what backward reads (the pre-activations and the transposed bf16 copies of the input
// and of every hidden activation); inference writes only each layer's bf16 activations and the output.
int forward(bmc_mlp* h, const float* X, int B, hipStream_t s, bool train) {
  TileArgs t{};
  t.mode = TILE_INPUT;
  t.rows = B;
  t.cols = h->dims[0];
  t.x = X;
  t.ldx = h->dims[0];
  t.aux = h->shift;
  t.scale = h->scale;
  t.bf = h->x_bf;
  t.ldo = pad8(h->dims[0]);
  t.bfT = train ? h->x_bfT : nullptr;
  t.ldt = h->ldb;
  int rc = BMC_OK;
  if (!train && B > 0 && t.cols % 4 == 0 && t.ldx % 4 == 0 && t.ldo % 4 == 0 &&
      reinterpret_cast<uintptr_t>(X) % 16 == 0) {
    const long n = (long)B * (t.cols / 4);
    const int blocks = (int)std::min<long>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(input_kernel, dim3(blocks), dim3(256), 0, s, X, t.ldx, t.aux, t.scale, B, t.cols, t.bf, t.ldo);
    BMC_HIP_CHECK(hipGetLastError());
  } else {
    rc = tiles(t, s);
  }
  if (rc) return rc;
  for (int l = 0; l < h->L; l++) {
    const int K = h->dims[l], N = h->dims[l + 1];
    const bool last = l + 1 == h->L;
    GemmArgs g{};
    g.M = B;
    g.N = N;
    g.K = K;
    g.A = l == 0 ? h->x_bf : h->act_bf[l - 1];
    g.lda = pad8(K);
    g.B = h->wbf[l];
    g.ldb = pad8(K);
    g.e.alpha = 1.f;
    g.e.bias = h->w + h->boff[l];
    g.e.act = last ? 0 : 1;
    g.e.ldc = N;
    g.e.C = last ? h->act[l] : nullptr;  // hidden layers need only the pre-activation and the bf16 copies
    g.e.Cpre = (last || !train) ? nullptr : h->pre[l];
    g.e.Cbf = last ? nullptr : h->act_bf[l];
    g.e.ldcb = pad8(N);
    g.e.CbfT = (last || !train) ? nullptr : h->act_bfT[l];
    g.e.ldct = h->ldb;
    if ((rc = gemm(g, h->part, h->part_floats, s))) return rc;
  }
  return BMC_OK;
}

// Gradients of the MSE loss w.r.t. every parameter into h->g (after forward at the same batch).
int backward(bmc_mlp* h, const float* T, int B, hipStream_t s) {
  const int L = h->L;
  for (int l = L - 1; l >= 0; l--) {
    const int K = h->dims[l], N = h->dims[l + 1];
    // dZ_l: 2 (Y - T) / (B D) for the linear last layer, dact * gelu'(pre) for hidden layers; fp32 in
    // dact[l], bf16 in both orientations
    TileArgs t{};
    t.mode = l + 1 == L ? TILE_MSE : TILE_GELU;
    t.rows = B;
    t.cols = N;
    t.x = l + 1 == L ? h->act[l] : h->dact[l];
    t.ldx = N;
    t.aux = l + 1 == L ? T : h->pre[l];
    t.out32 = h->dact[l];
    t.bf = h->dz_bf[l];
    t.ldo = pad8(N);
    t.bfT = h->dz_bfT[l];
    t.ldt = h->ldb;
    t.loss = h->loss;
    t.colsum = h->bias_part;
    int rc = tiles(t, s);
    if (rc) return rc;
    // dW_l [N, K] = dZ^T [N, B] X_l [B, K]: both operands batch-contiguous
    GemmArgs gw{};
    gw.M = N;
    gw.N = K;
    gw.K = B;
    gw.A = h->dz_bfT[l];
    gw.lda = h->ldb;
    gw.B = l == 0 ? h->x_bfT : h->act_bfT[l - 1];
    gw.ldb = h->ldb;
    gw.e.C = h->g + h->woff[l];
    gw.e.ldc = K;
    gw.e.alpha = 1.f;
    if ((rc = gemm(gw, h->part, h->part_floats, s))) return rc;
    hipLaunchKernelGGL(bias_grad_sum_kernel, dim3((N + 63) / 64), dim3(256), 0, s, h->bias_part, (B + 63) / 64, N,
                       h->g + h->boff[l]);
    BMC_HIP_CHECK(hipGetLastError());
    if (l > 0) {
      // dX_l [B, K] = dZ [B, N] W [N, K]: A = dZ (N-contiguous), B^T = W^T [K, N]
      GemmArgs gx{};
      gx.M = B;
      gx.N = K;
      gx.K = N;
      gx.A = h->dz_bf[l];
      gx.lda = pad8(N);
      gx.B = h->wbfT[l];
      gx.ldb = pad8(N);
      gx.e.C = h->dact[l - 1];
      gx.e.ldc = K;
      gx.e.alpha = 1.f;
      if ((rc = gemm(gx, h->part, h->part_floats, s))) return rc;
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
  h->ldb = pad8(max_batch);
  long widest = 0, biggest_w = 0;
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
    widest = std::max<long>(widest, std::max(dims[l], dims[l + 1]));
    biggest_w = std::max<long>(biggest_w, (long)dims[l] * dims[l + 1]);
  }
  // split-K partials: up to 16 splits of a weight gradient or 2 of a batch-sized product
  h->part_floats = (size_t)std::max(16 * biggest_w, 2L * max_batch * widest);
  auto A = [&](void** p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 16)) == hipSuccess; };
  const long d0 = dims[0];
  bool ok = A((void**)&h->w, 4 * h->n_params) && A((void**)&h->g, 4 * h->n_params) &&
            A((void**)&h->m, 4 * h->n_params) && A((void**)&h->v, 4 * h->n_params) &&
            A((void**)&h->x_bf, 2 * max_batch * pad8(d0)) && A((void**)&h->x_bfT, 2 * d0 * h->ldb) &&
            A((void**)&h->loss, 4) && A((void**)&h->part, 4 * h->part_floats) &&
            A((void**)&h->bias_part, 4L * ((max_batch + 63) / 64) * *std::max_element(dims + 1, dims + n_layers + 1));
  for (int l = 0; ok && l < n_layers; l++) {
    const long K = dims[l], N = dims[l + 1];
    const size_t n = (size_t)max_batch * N;
    float *pre = nullptr, *act = nullptr, *dact = nullptr;
    uint16_t *abf = nullptr, *abfT = nullptr, *zbf = nullptr, *zbfT = nullptr, *wb = nullptr, *wbT = nullptr;
    ok = A((void**)&pre, 4 * n) && A((void**)&act, 4 * n) && A((void**)&dact, 4 * n) &&
         A((void**)&abf, 2 * max_batch * pad8(N)) && A((void**)&abfT, 2 * N * h->ldb) &&
         A((void**)&zbf, 2 * max_batch * pad8(N)) && A((void**)&zbfT, 2 * N * h->ldb) &&
         A((void**)&wb, 2 * N * pad8(K)) && A((void**)&wbT, 2 * K * pad8(N));
    h->pre.push_back(pre);
    h->act.push_back(act);
    h->dact.push_back(dact);
    h->act_bf.push_back(abf);
    h->act_bfT.push_back(abfT);
    h->dz_bf.push_back(zbf);
    h->dz_bfT.push_back(zbfT);
    h->wbf.push_back(wb);
    h->wbfT.push_back(wbT);
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
  int rc = pack_weights(h, 0);
  if (rc) return rc;
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

int bmc_mlp_set_input_norm(bmc_mlp* h, const float* shift_host, const float* scale_host) {
  if (!h || (!shift_host) != (!scale_host)) {
    set_error("bmc_mlp_set_input_norm: give both shift and scale, or neither");
    return BMC_EINVAL;
  }
  const size_t bytes = 4 * (size_t)h->dims[0];
  if (!shift_host) {
    if (h->shift) (void)hipFree(h->shift);
    if (h->scale) (void)hipFree(h->scale);
    h->shift = h->scale = nullptr;
    return BMC_OK;
  }
  if (!h->shift && (hipMalloc(&h->shift, bytes) != hipSuccess || hipMalloc(&h->scale, bytes) != hipSuccess)) {
    set_error("bmc_mlp_set_input_norm: device allocation failed");
    return BMC_ENOMEM;
  }
  BMC_HIP_CHECK(hipMemcpy(h->shift, shift_host, bytes, hipMemcpyHostToDevice));
  BMC_HIP_CHECK(hipMemcpy(h->scale, scale_host, bytes, hipMemcpyHostToDevice));
  return BMC_OK;
}

int bmc_mlp_get_input_norm(const bmc_mlp* h, float* shift_host, float* scale_host) {
  if (!h || !shift_host || !scale_host) {
    set_error("bmc_mlp_get_input_norm: null argument");
    return BMC_EINVAL;
  }
  const size_t bytes = 4 * (size_t)h->dims[0];
  if (!h->shift) {
    for (int i = 0; i < h->dims[0]; i++) {
      shift_host[i] = 0.f;
      scale_host[i] = 1.f;
    }
    return BMC_OK;
  }
  BMC_HIP_CHECK(hipMemcpy(shift_host, h->shift, bytes, hipMemcpyDeviceToHost));
  BMC_HIP_CHECK(hipMemcpy(scale_host, h->scale, bytes, hipMemcpyDeviceToHost));
  return BMC_OK;
}

int bmc_mlp_fit_input_norm(bmc_mlp* h, const float* x_dev, int64_t n, void* stream) {
  if (!h || !x_dev || n < 2) {
    set_error("bmc_mlp_fit_input_norm: need at least two rows");
    return BMC_EINVAL;
  }
  const int d = h->dims[0];
  const size_t bytes = 4 * (size_t)d;
  if (!h->shift && (hipMalloc(&h->shift, bytes) != hipSuccess || hipMalloc(&h->scale, bytes) != hipSuccess)) {
    set_error("bmc_mlp_fit_input_norm: device allocation failed");
    return BMC_ENOMEM;
  }
  const long chunks = (n + kStatRows - 1) / kStatRows;
  double* part = nullptr;
  if (hipMalloc(&part, sizeof(double) * 2 * chunks * d) != hipSuccess) {
    set_error("bmc_mlp_fit_input_norm: device allocation failed");
    return BMC_ENOMEM;
  }
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(column_stats_partial_kernel, dim3((d + 63) / 64, (unsigned)chunks), dim3(256), 0, s, x_dev,
                     (long)n, d, part);
  hipLaunchKernelGGL(column_stats_final_kernel, dim3((d + 255) / 256), dim3(256), 0, s, part, (int)chunks, (long)n, d,
                     h->shift, h->scale);
  hipError_t e = hipGetLastError();
  hipError_t se = hipStreamSynchronize(s);
  (void)hipFree(part);
  if (e != hipSuccess || se != hipSuccess) {
    set_error(std::string("bmc_mlp_fit_input_norm: ") + hipGetErrorString(e != hipSuccess ? e : se));
    return BMC_EHIP;
  }
  return BMC_OK;
}

int bmc_mlp_forward(bmc_mlp* h, const float* x_dev, int batch, float* y_dev, void* stream) {
  if (!h || !x_dev || !y_dev || batch < 0 || batch > h->max_batch) {
    set_error("bmc_mlp_forward: bad arguments (batch above max_batch?)");
    return BMC_EINVAL;
  }
  if (batch == 0) return BMC_OK;
  hipStream_t s = (hipStream_t)stream;
  int rc = forward(h, x_dev, batch, s, false);
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
  int rc = forward(h, x_dev, batch, s, true);
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
  int rc = forward(h, x_dev, batch, s, true);
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
                     h->n_params, lr, b1, b2, 1e-8f, c1, c2);
  BMC_HIP_CHECK(hipGetLastError());
  return pack_weights(h, s);
}

}  // extern "C"
