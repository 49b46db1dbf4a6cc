// Device building blocks of the simulate and adjoint kernels: the sparse generator of one event,
// its tangents, the truncated Taylor series and the column-layout products of a team of lanes.
// Included by each HIP translation unit (no relocatable device code: every unit keeps its own copy).
#pragma once

#include <type_traits>
#include <utility>

#include "bmc_internal.h"

namespace bmc {

// Taylor degree thresholds theta_m: the largest 1-norm for which the truncated series T_m has
// backward error <= 2^-53 (Al-Mohy & Higham 2011, Sec. 3; recomputed for this build).
static __constant__ double kTheta[41] = {0.0,      2.22e-16, 2.581e-08, 1.386e-05, 0.0003397, 0.002401, 0.009066,
                                  0.02384,  0.04991,  0.08958,   0.1442,    0.2142,    0.2996,   0.3998,
                                  0.5139,   0.6411,   0.7803,    0.9305,    1.091,     1.26,     1.438,
                                  1.624,    1.816,    2.015,     2.219,     2.429,     2.643,    2.861,
                                  3.084,    3.31,     3.54,      3.772,     4.008,     4.245,    4.486,
                                  4.728,    4.973,    5.219,     5.468,     5.717,     5.969};
// 1 / k for the Taylor recursion (a table load instead of a float64 division per term)
static __constant__ double kInv[41] = {0.0, 1.0, 1.0 / 2, 1.0 / 3, 1.0 / 4, 1.0 / 5, 1.0 / 6, 1.0 / 7, 1.0 / 8, 1.0 / 9,
                                1.0 / 10, 1.0 / 11, 1.0 / 12, 1.0 / 13, 1.0 / 14, 1.0 / 15, 1.0 / 16, 1.0 / 17,
                                1.0 / 18, 1.0 / 19, 1.0 / 20, 1.0 / 21, 1.0 / 22, 1.0 / 23, 1.0 / 24, 1.0 / 25,
                                1.0 / 26, 1.0 / 27, 1.0 / 28, 1.0 / 29, 1.0 / 30, 1.0 / 31, 1.0 / 32, 1.0 / 33,
                                1.0 / 34, 1.0 / 35, 1.0 / 36, 1.0 / 37, 1.0 / 38, 1.0 / 39, 1.0 / 40};
constexpr int kMMax = 30;  // the most accurate cap for the squaring path (fewest squarings before the rounding of the series takes over)

__device__ __forceinline__ double load_real(const void* p, long i, int dtype) {
  return dtype == BMC_F64 ? static_cast<const double*>(p)[i] : (double)static_cast<const float*>(p)[i];
}
__device__ __forceinline__ void store_real(void* p, long i, double v, int dtype) {
  if (dtype == BMC_F64)
    static_cast<double*>(p)[i] = v;
  else
    static_cast<float*>(p)[i] = (float)v;
}

template <int K, int END, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (K < END) {
    f(std::integral_constant<int, K>{});
    static_for<K + 1, END>(f);
  }
}

// Team geometry. Teams of 8, 16 and 32 lanes tile a wave. 12-lane teams (augmented size 9..12, two CEST
// pools) sit five to a wave, at lanes [0, 12), [12, 24), [24, 36), [36, 48) and [48, 64), the last
// with four more idle lanes: 60 of 64 lanes carry a column instead of 44 of 64 with 16-lane teams.
template <int TEAM>
__host__ __device__ constexpr int teams_per_block() {
  return TEAM == 12 ? 20 : 256 / TEAM;  // 256-thread blocks
}
template <int TEAM>
__host__ __device__ constexpr int teams_per_wave() {
  return TEAM == 12 ? 5 : 64 / TEAM;
}
// Teams of one tangent pass, rounded up to whole waves: numbered this way, every wave's teams share one
// pass, so its tangent kinds are wave-uniform (scalar branches in the event setup).
template <int TEAM>
__host__ __device__ constexpr long pass_teams(long rows) {
  return (rows + teams_per_wave<TEAM>() - 1) / teams_per_wave<TEAM>() * teams_per_wave<TEAM>();
}
template <int TEAM>
__device__ __forceinline__ int team_base_in_wave() {  // the team's first lane within its wave
  if constexpr (TEAM == 12) {
    const int wl = threadIdx.x & 63;
    return (wl >= 48 ? 4 : wl / 12) * 12;
  } else {
    return (threadIdx.x & 63) / TEAM * TEAM;
  }
}
template <int TEAM>
__device__ __forceinline__ int team_lane() {
  if constexpr (TEAM == 12)
    return (threadIdx.x & 63) - team_base_in_wave<12>();
  else
    return threadIdx.x % TEAM;
}
template <int TEAM>
__device__ __forceinline__ int team_index() {  // the team within its 256-thread block
  if constexpr (TEAM == 12)
    return (threadIdx.x >> 6) * 5 + team_base_in_wave<12>() / 12;
  else
    return threadIdx.x / TEAM;
}

// Broadcast lane K of the team to every lane of the team.
template <int TEAM, int K>
__device__ __forceinline__ double bcast(double x) {
  if constexpr (TEAM == 12) {
    return __shfl(x, team_base_in_wave<12>() + K, 64);
  } else if constexpr (TEAM == 16) {
    return __builtin_amdgcn_mov_dpp(x, 0x150 + K, 0xf, 0xf, true);  // row_newbcast:K
  } else if constexpr (TEAM == 8) {
    // two teams share a DPP row: lanes 0-7 take lane K, lanes 8-15 lane 8 + K. The first move fills
    // the row with lane K, the second overwrites lanes 8-15 only (bank_mask 0xc), so no select is needed.
    const double all = __builtin_amdgcn_mov_dpp(x, 0x150 + K, 0xf, 0xf, true);
    return __builtin_amdgcn_update_dpp(all, x, 0x150 + 8 + K, 0xf, 0xc, false);
  } else if constexpr (TEAM == 32 && K < 16) {
    // a 32-lane team spans two DPP rows: lane K to every lane of its row (row_newbcast), then lane 15
    // of the team's first row, which now holds it, to the second row (row_bcast:15, rows 1 and 3 only)
    const double t = __builtin_amdgcn_mov_dpp(x, 0x150 + K, 0xf, 0xf, true);
    return __builtin_amdgcn_update_dpp(t, t, 0x142, 0xa, 0xf, false);
  } else {
    return __shfl(x, K, TEAM);
  }
}

// Lanes of one team exchange values through LDS: no lane reads before every write issued earlier in
// program order has landed (a team never spans two waves, whose LDS accesses complete in order).
__device__ __forceinline__ void team_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// x plus the value of another lane of its DPP row, picked by the DPP control `CTRL`.
template <int CTRL>
__device__ __forceinline__ double dpp_add(double x) {
  return x + __builtin_amdgcn_mov_dpp(x, CTRL, 0xf, 0xf, true);
}

// Sum over the team, on every lane of it. Teams of up to 16 lanes reduce with DPP moves inside their row:
// lane ^ 1 (quad_perm [1,0,3,2]), lane ^ 2 (quad_perm [2,3,0,1]), then the mirrored quad of the half row
// and the mirrored half of the row. 32-lane teams keep ds_bpermute butterflies, which timed faster for
// the four-pool forward kernel.
template <int TEAM>
__device__ __forceinline__ double team_sum(double x) {
  static_assert(TEAM == 8 || TEAM == 12 || TEAM == 16 || TEAM == 32, "team size");
  if constexpr (TEAM == 32) {
#pragma unroll
    for (int m = 1; m < TEAM; m <<= 1) x += __shfl_xor(x, m, TEAM);
    return x;
  } else {
    x = dpp_add<0xb1>(x);
    x = dpp_add<0x4e>(x);
    if constexpr (TEAM == 12) {
      // the team's three quads (team bases are multiples of 4)
      const int b = team_base_in_wave<12>();
      return __shfl(x, b, 64) + __shfl(x, b + 4, 64) + __shfl(x, b + 8, 64);
    } else {
      x = dpp_add<0x141>(x);                            // row_half_mirror
      if constexpr (TEAM == 16) x = dpp_add<0x140>(x);  // row_mirror
      return x;
    }
  }
}

// ---- parameters --------------------------------------------------------------------------------
// LDS copy of one voxel's packed record, plus derived values.
template <int NC>
struct PrmIdx {
  static constexpr int stride = 13 + 5 * NC;
  static constexpr int L = stride, KS = stride + 1, KWMT = stride + 2, DW0 = stride + 3;
  static constexpr int size = stride + 4;
};

// MT saturation lineshape g(off, t2, w0) and its partial derivatives (matrix.py:24-81, 183-213).
struct MtShape {
  double g, g_off, g_t2, g_w0;
};

static __device__ void sl_sum(double d, double t2, double& g, double& g_d, double& g_t2) {
  // matrix.py:34-38: sum_u sqrt(2/pi) t2 / T exp(-2 (d t2 / T)^2) * pi * du, T = max(|3u^2-1|, 1e-6)
  double s = 0.0, sd = 0.0, st = 0.0;
  const double c = 1.2533141373155001;  // sqrt(2/pi) * pi * 0.01 * 100 / 100 -> folded below
  for (int i = 0; i <= 100; i++) {
    double u = i * 0.01;
    double T = fabs(3.0 * u * u - 1.0);
    T = T < 1e-6 ? 1e-6 : T;
    double a = d * t2 / T;
    double e = exp(-2.0 * a * a);
    double base = t2 / T * e;
    s += base;
    sd += base * (-4.0 * a * t2 / T);
    st += e / T * (1.0 - 4.0 * a * a);
  }
  const double k = 0.7978845608028654 * 3.141592653589793 * 0.01;  // sqrt(2/pi) * pi * du
  (void)c;
  g = s * k;
  g_d = sd * k;
  g_t2 = st * k;
}

static __device__ MtShape mt_shape(int ls, double off, double t2, double w0) {
  MtShape r{0.0, 0.0, 0.0, 0.0};
  if (ls == 1) {  // Lorentzian t2 / (1 + (off t2)^2), matrix.py:70-81
    double q = off * t2, den = 1.0 + q * q;
    r.g = t2 / den;
    r.g_off = -2.0 * off * t2 * t2 * t2 / (den * den);
    r.g_t2 = (1.0 - q * q) / (den * den);
  } else if (ls == 2) {
    if (fabs(off) >= w0) {
      sl_sum(off, t2, r.g, r.g_off, r.g_t2);
    } else {  // cubic Hermite bridge, matrix.py:40-60
      double px[4] = {-300.0 - w0, -100.0 - w0, 100.0 + w0, 300.0 + w0};
      double dpx[4] = {-1.0, -1.0, 1.0, 1.0};
      double py[4], pyd[4], pyt[4];
#pragma unroll
      for (int i = 0; i < 4; i++) sl_sum(px[i], t2, py[i], pyd[i], pyt[i]);
      double d0y = 30.0 * (py[1] - py[0]), d1y = 30.0 * (py[3] - py[2]);
      double den = px[2] - px[1] + 1.0;
      double q = (off - px[1] + 1.0) / den;
      double sg = q > 0.0 ? 1.0 : (q < 0.0 ? -1.0 : 0.0);
      double c = fabs(q), c2 = c * c, c3 = c2 * c;
      double h0 = 2 * c3 - 3 * c2 + 1, h1 = -2 * c3 + 3 * c2, h2 = c3 - 2 * c2 + c, h3 = c3 - c2;
      double h0c = 6 * c2 - 6 * c, h1c = -6 * c2 + 6 * c, h2c = 3 * c2 - 4 * c + 1, h3c = 3 * c2 - 2 * c;
      r.g = h0 * py[1] + h1 * py[2] + h2 * d0y + h3 * d1y;
      double g_c = h0c * py[1] + h1c * py[2] + h2c * d0y + h3c * d1y;
      r.g_off = g_c * sg / den;
      r.g_t2 = h0 * pyt[1] + h1 * pyt[2] + h2 * 30.0 * (pyt[1] - pyt[0]) + h3 * 30.0 * (pyt[3] - pyt[2]);
      // w0 moves the four nodes and the interval
      double dq = (-dpx[1] * den - (off - px[1] + 1.0) * (dpx[2] - dpx[1])) / (den * den);
      double pyw[4];
#pragma unroll
      for (int i = 0; i < 4; i++) pyw[i] = pyd[i] * dpx[i];
      r.g_w0 = g_c * sg * dq + h0 * pyw[1] + h1 * pyw[2] + h2 * 30.0 * (pyw[1] - pyw[0]) +
               h3 * 30.0 * (pyw[3] - pyw[2]);
    }
  }
  return r;
}

// ---- the generator -----------------------------------------------------------------------------
// Entries of X = aug * dt * 2^-s (matrix.py:238-293), the non-zero structure made explicit.
template <int NC>
struct Gen {
  double axx, offw, w1x, w1y, azz, cw;
  double b[NC > 0 ? NC : 1], o[NC > 0 ? NC : 1], c[NC > 0 ? NC : 1], k[NC > 0 ? NC : 1],
      kw[NC > 0 ? NC : 1], cc[NC > 0 ? NC : 1];
  double dmt, kwmt, kmt, cmt;
};

// y = X v (v augmented with vaug, the homogeneous coordinate).
template <int NC>
__device__ __forceinline__ void apply_X(const Gen<NC>& G, const double* v, double vaug, double* y) {
  constexpr int NT = 1 + NC;
  constexpr int xw = 0, yw = NT, zw = 2 * NT, zmt = 3 * NT;
  double sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
  for (int i = 0; i < NC; i++) {
    sx = fma(G.k[i], v[1 + i], sx);
    sy = fma(G.k[i], v[NT + 1 + i], sy);
    sz = fma(G.k[i], v[2 * NT + 1 + i], sz);
  }
  y[xw] = fma(G.axx, v[xw], fma(G.offw, v[yw], fma(-G.w1y, v[zw], sx)));
  y[yw] = fma(-G.offw, v[xw], fma(G.axx, v[yw], fma(G.w1x, v[zw], sy)));
  y[zw] = fma(G.w1y, v[xw], fma(-G.w1x, v[yw], fma(G.azz, v[zw], fma(G.kmt, v[zmt], fma(G.cw, vaug, sz)))));
#pragma unroll
  for (int i = 0; i < NC; i++) {
    const int xc = 1 + i, yc = NT + 1 + i, zc = 2 * NT + 1 + i;
    y[xc] = fma(G.kw[i], v[xw], fma(G.b[i], v[xc], fma(-G.o[i], v[yc], -G.w1y * v[zc])));
    y[yc] = fma(G.kw[i], v[yw], fma(G.o[i], v[xc], fma(G.b[i], v[yc], G.w1x * v[zc])));
    y[zc] = fma(G.kw[i], v[zw], fma(G.w1y, v[xc], fma(-G.w1x, v[yc], fma(G.c[i], v[zc], G.cc[i] * vaug))));
  }
  y[zmt] = fma(G.kwmt, v[zw], fma(G.dmt, v[zmt], G.cmt * vaug));
}

// Per-event coefficients of one tangent's generator derivative dX (already times dt 2^-s).
struct TanCo {
  double c0, c1, c2;
};

// y += dX v for tangent kind `kind` (pool `ix`), matrix.py:238-293 differentiated by hand.
template <int NC>
__device__ __forceinline__ void apply_dX(int kind, int ix, const TanCo& t, const double* dwp, const double* v,
                                         double vaug, double* y) {
  constexpr int NT = 1 + NC;
  constexpr int xw = 0, yw = NT, zw = 2 * NT, zmt = 3 * NT;
  switch (kind) {
    case TK_W_R1:
      y[zw] = fma(-t.c0, v[zw], fma(t.c0, vaug, y[zw]));
      break;
    case TK_W_R2:
      y[xw] = fma(-t.c0, v[xw], y[xw]);
      y[yw] = fma(-t.c0, v[yw], y[yw]);
      break;
    case TK_MT_K:
      y[zw] = fma(-t.c0, v[zw], fma(t.c1, v[zmt], y[zw]));
      y[zmt] = fma(t.c0, v[zw], fma(-t.c1, v[zmt], y[zmt]));
      break;
    case TK_MT_F:
      y[zw] = fma(-t.c0, v[zw], y[zw]);
      y[zmt] = fma(t.c0, v[zw], fma(t.c1, vaug, y[zmt]));
      break;
    case TK_MT_R1:
      y[zmt] = fma(-t.c0, v[zmt], fma(t.c1, vaug, y[zmt]));
      break;
    case TK_MT_DW:
    case TK_MT_R2:
      y[zmt] = fma(t.c2, v[zmt], y[zmt]);
      break;
    case TK_B0INH:
      // d off_w = c0, d o_i = -c0, d w_mt folded into c2
      y[xw] = fma(t.c0, v[yw], y[xw]);
      y[yw] = fma(-t.c0, v[xw], y[yw]);
#pragma unroll
      for (int i = 0; i < NC; i++) {
        y[1 + i] = fma(t.c0, v[NT + 1 + i], y[1 + i]);
        y[NT + 1 + i] = fma(-t.c0, v[1 + i], y[NT + 1 + i]);
      }
      y[zmt] = fma(t.c2, v[zmt], y[zmt]);
      break;
    case TK_B0:
    case TK_GAMMA:
      // d off_w = c0, d o_i = dw_i c1 - c0
      y[xw] = fma(t.c0, v[yw], y[xw]);
      y[yw] = fma(-t.c0, v[xw], y[yw]);
#pragma unroll
      for (int i = 0; i < NC; i++) {
        double doi = fma(dwp[i], t.c1, -t.c0);
        y[1 + i] = fma(-doi, v[NT + 1 + i], y[1 + i]);
        y[NT + 1 + i] = fma(doi, v[1 + i], y[NT + 1 + i]);
      }
      y[zmt] = fma(t.c2, v[zmt], y[zmt]);
      break;
    case TK_RELB1:
      // d w1x = c0, d w1y = c1
      y[xw] = fma(-t.c1, v[zw], y[xw]);
      y[yw] = fma(t.c0, v[zw], y[yw]);
      y[zw] = fma(t.c1, v[xw], fma(-t.c0, v[yw], y[zw]));
#pragma unroll
      for (int i = 0; i < NC; i++) {
        const int xc = 1 + i, yc = NT + 1 + i, zc = 2 * NT + 1 + i;
        y[xc] = fma(-t.c1, v[zc], y[xc]);
        y[yc] = fma(t.c0, v[zc], y[yc]);
        y[zc] = fma(t.c1, v[xc], fma(-t.c0, v[yc], y[zc]));
      }
      y[zmt] = fma(t.c2, v[zmt], y[zmt]);
      break;
    default:
      break;
  }
  if (kind >= TK_C_K) {
#pragma unroll
    for (int i = 0; i < NC; i++) {
      if (i != ix) continue;
      const int xc = 1 + i, yc = NT + 1 + i, zc = 2 * NT + 1 + i;
      switch (kind) {
        case TK_C_K:  // d k = 1 (c1), d kw = f (c0), d diag = -1, d water diag = -f
          y[xw] = fma(-t.c0, v[xw], fma(t.c1, v[xc], y[xw]));
          y[yw] = fma(-t.c0, v[yw], fma(t.c1, v[yc], y[yw]));
          y[zw] = fma(-t.c0, v[zw], fma(t.c1, v[zc], y[zw]));
          y[xc] = fma(t.c0, v[xw], fma(-t.c1, v[xc], y[xc]));
          y[yc] = fma(t.c0, v[yw], fma(-t.c1, v[yc], y[yc]));
          y[zc] = fma(t.c0, v[zw], fma(-t.c1, v[zc], y[zc]));
          break;
        case TK_C_F:  // d kw = k (c0), C_zc = r1c f -> d = r1c (c1)
          y[xw] = fma(-t.c0, v[xw], y[xw]);
          y[yw] = fma(-t.c0, v[yw], y[yw]);
          y[zw] = fma(-t.c0, v[zw], y[zw]);
          y[xc] = fma(t.c0, v[xw], y[xc]);
          y[yc] = fma(t.c0, v[yw], y[yc]);
          y[zc] = fma(t.c0, v[zw], fma(t.c1, vaug, y[zc]));
          break;
        case TK_C_DW:  // d o_i = c0
          y[xc] = fma(-t.c0, v[yc], y[xc]);
          y[yc] = fma(t.c0, v[xc], y[yc]);
          break;
        case TK_C_R1:
          y[zc] = fma(-t.c0, v[zc], fma(t.c1, vaug, y[zc]));
          break;
        case TK_C_R2:
          y[xc] = fma(-t.c0, v[xc], y[xc]);
          y[yc] = fma(-t.c0, v[yc], y[yc]);
          break;
        default:
          break;
      }
    }
  }
}

// The generator derivative of one tangent in one of four shapes that together cover every tangent kind,
// with the shape's coefficients q (already times dt 2^-s):
//   WATER  water / MT relaxation, exchange and saturation terms (W_R1, W_R2, MT_*):
//            y[xw] += q0 v[xw], y[yw] += q0 v[yw], y[zw] += q1 v[zw] + q2 v[zmt] + q3 vaug,
//            y[zmt] += q4 v[zw] + q5 v[zmt] + q6 vaug
//   POOL   one CEST pool ix (C_K, C_F, C_DW, C_R1, C_R2), c = its (x, y, z):
//            y[w] += q0 v[w] + q1 v[c] (w = xw, yw, zw and c = xc, yc, zc in turn),
//            y[xc] += q2 v[xw] + q3 v[xc] + q4 v[yc], y[yc] += q2 v[yw] + q3 v[yc] - q4 v[xc],
//            y[zc] += q2 v[zw] + q5 v[zc] + q6 vaug
//   SHIFT  a change of every offset (B0INH, B0, GAMMA): y[x] += g v[y], y[y] -= g v[x] for water
//            (g = q0) and pool i (g = q[1 + i]), y[zmt] += q6 v[zmt]
//   B1     a change of the RF vector (RELB1): y[x] -= q1 v[z], y[y] += q0 v[z], y[z] += q1 v[x] - q0 v[y]
//            for water and every pool, y[zmt] += q2 v[zmt]
// The shape is found once per event (dx_shape); the Taylor loop then applies it with a four-way branch
// instead of dispatching on the tangent kind at every term.
enum DxShape { DXS_NONE = 0, DXS_WATER, DXS_POOL, DXS_SHIFT, DXS_B1 };
struct DxCo {
  int shape, ix;
  double q[7];
};
// Coefficients a kind can make nonzero in its shape (bit j for q[j]), forward and transposed
// (dx_transpose) together: W_R1 q1, q3; W_R2 q0; C_K q0, q1, q2, q3, q5; C_F q0, q2, q6 and q1 transposed.
constexpr int kQmWR1 = 0x0a, kQmWR2 = 0x01, kQmCK = 0x2f, kQmCF = 0x47;

template <int NC>
__device__ __forceinline__ DxCo dx_shape(int kind, int ix, const TanCo& t, const double* dwp) {
  DxCo d;
  d.shape = DXS_NONE;
  d.ix = ix;
#pragma unroll
  for (int i = 0; i < 7; i++) d.q[i] = 0.0;
  switch (kind) {
    case TK_W_R1:
      d.shape = DXS_WATER, d.q[1] = -t.c0, d.q[3] = t.c0;
      break;
    case TK_W_R2:
      d.shape = DXS_WATER, d.q[0] = -t.c0;
      break;
    case TK_MT_K:
      d.shape = DXS_WATER, d.q[1] = -t.c0, d.q[2] = t.c1, d.q[4] = t.c0, d.q[5] = -t.c1;
      break;
    case TK_MT_F:
      d.shape = DXS_WATER, d.q[1] = -t.c0, d.q[4] = t.c0, d.q[6] = t.c1;
      break;
    case TK_MT_R1:
      d.shape = DXS_WATER, d.q[5] = -t.c0, d.q[6] = t.c1;
      break;
    case TK_MT_DW:
    case TK_MT_R2:
      d.shape = DXS_WATER, d.q[5] = t.c2;
      break;
    case TK_B0INH:
    case TK_B0:
    case TK_GAMMA:
      // water: d off_w = c0; pool i: d o_i = dw_i c1 - c0 (c1 = 0 for B0INH)
      d.shape = DXS_SHIFT, d.q[0] = t.c0, d.q[6] = t.c2;
#pragma unroll
      for (int i = 0; i < NC; i++) d.q[1 + i] = (kind == TK_B0INH) ? t.c0 : fma(-dwp[i], t.c1, t.c0);
      break;
    case TK_RELB1:
      d.shape = DXS_B1, d.q[0] = t.c0, d.q[1] = t.c1, d.q[2] = t.c2;
      break;
    case TK_C_K:  // d k = 1 (c1), d kw = f (c0)
      d.shape = DXS_POOL, d.q[0] = -t.c0, d.q[1] = t.c1, d.q[2] = t.c0, d.q[3] = -t.c1, d.q[5] = -t.c1;
      break;
    case TK_C_F:  // d kw = k (c0), d C_zc = r1c (c1)
      d.shape = DXS_POOL, d.q[0] = -t.c0, d.q[2] = t.c0, d.q[6] = t.c1;
      break;
    case TK_C_DW:  // d o_i = c0
      d.shape = DXS_POOL, d.q[4] = -t.c0;
      break;
    case TK_C_R1:
      d.shape = DXS_POOL, d.q[5] = -t.c0, d.q[6] = t.c1;
      break;
    case TK_C_R2:
      d.shape = DXS_POOL, d.q[3] = -t.c0;
      break;
    default:
      break;
  }
  return d;
}

// y += dX v for a tangent in shape form (the same values as apply_dX with the kind it came from). With
// SHAPE >= 0 the shape (and, for DXS_POOL, the pool IX) is fixed at compile time and d's are ignored.
// QM: the coefficients q[j] that can be nonzero (bit j), known at compile time when the tangent's kind
// is; the FMAs of the others are not emitted (qm_of_kind).
template <int QM, int J>
__device__ __forceinline__ double qfma(const double* q, double a, double acc) {
  if constexpr ((QM >> J) & 1)
    return fma(q[J], a, acc);
  else
    return acc;
}
template <int NC, int SHAPE = -1, int IX = -1, int QM = 0x7f>
__device__ __forceinline__ void apply_dx_shape(const DxCo& d, const double* v, double vaug, double* y) {
  constexpr int NT = 1 + NC;
  constexpr int xw = 0, yw = NT, zw = 2 * NT, zmt = 3 * NT;
  const double* q = d.q;
  const int shape = SHAPE >= 0 ? SHAPE : d.shape;
  const int pix = IX >= 0 ? IX : d.ix;
  if (shape == DXS_WATER) {
    y[xw] = qfma<QM, 0>(q, v[xw], y[xw]);
    y[yw] = qfma<QM, 0>(q, v[yw], y[yw]);
    y[zw] = qfma<QM, 1>(q, v[zw], qfma<QM, 2>(q, v[zmt], qfma<QM, 3>(q, vaug, y[zw])));
    y[zmt] = qfma<QM, 4>(q, v[zw], qfma<QM, 5>(q, v[zmt], qfma<QM, 6>(q, vaug, y[zmt])));
  } else if (shape == DXS_POOL) {
#pragma unroll
    for (int i = 0; i < NC; i++) {
      if (i != pix) continue;
      const int xc = 1 + i, yc = NT + 1 + i, zc = 2 * NT + 1 + i;
      y[xw] = qfma<QM, 0>(q, v[xw], qfma<QM, 1>(q, v[xc], y[xw]));
      y[yw] = qfma<QM, 0>(q, v[yw], qfma<QM, 1>(q, v[yc], y[yw]));
      y[zw] = qfma<QM, 0>(q, v[zw], qfma<QM, 1>(q, v[zc], y[zw]));
      const double vx = v[xc], vy = v[yc];
      y[xc] = qfma<QM, 2>(q, v[xw], qfma<QM, 3>(q, vx, qfma<QM, 4>(q, vy, y[xc])));
      if constexpr ((QM >> 4) & 1)
        y[yc] = fma(q[2], v[yw], fma(q[3], vy, fma(-q[4], vx, y[yc])));
      else
        y[yc] = qfma<QM, 2>(q, v[yw], qfma<QM, 3>(q, vy, y[yc]));
      y[zc] = qfma<QM, 2>(q, v[zw], qfma<QM, 5>(q, v[zc], qfma<QM, 6>(q, vaug, y[zc])));
    }
  } else if (shape == DXS_SHIFT) {
    y[xw] = fma(q[0], v[yw], y[xw]);
    y[yw] = fma(-q[0], v[xw], y[yw]);
#pragma unroll
    for (int i = 0; i < NC; i++) {
      const int xc = 1 + i, yc = NT + 1 + i;
      y[xc] = fma(q[1 + i], v[yc], y[xc]);
      y[yc] = fma(-q[1 + i], v[xc], y[yc]);
    }
    y[zmt] = fma(q[6], v[zmt], y[zmt]);
  } else if (shape == DXS_B1) {
#pragma unroll
    for (int i = 0; i < NT; i++) {
      const int xi = i, yi = NT + i, zi = 2 * NT + i;
      const double vx = v[xi], vy = v[yi], vz = v[zi];
      y[xi] = fma(-q[1], vz, y[xi]);
      y[yi] = fma(q[0], vz, y[yi]);
      y[zi] = fma(q[1], vx, fma(-q[0], vy, y[zi]));
    }
    y[zmt] = fma(q[2], v[zmt], y[zmt]);
  }
}

// ---- the palindrome factorisation -------------------------------------------------------------
// A chained run E_{2h-1} ... E_0 whose events read the same backwards (E_{2h-1-i} = E_i) is Q^R Q with
// Q = E_{h-1} ... E_0 and Q^R = E_0 ... E_{h-1}. The Bloch-McConnell matrix A of every event is
// symmetrisable, A = S A^T S^-1 (matrix.py:238-286): at RF phase 0, S = diag(1, f_i, -1, -f_i, 1, f_i, f_mt)
// over (x, y, z of water and pool i, z of MT) makes exchange (k_i against k_i f_i), precession and RF
// (antisymmetric, flipped by the y sign) and relaxation (diagonal) agree; at phase phi, S is conjugated by
// the rotation by phi about z. Hence Q^R's linear part is S Q_A^T S^-1, and its affine column is S w with
// w the affine column of exp(X'_0) ... exp(X'_{h-1}), X' = [[A^T, S^-1 C], [0, 0]]: a chain through the
// same h events in reverse order with transposed generators, which the team's spare lane runs beside
// the others. A pool with f_i = 0 and k_i = 0 (or MT with f = 0 and k = 0) is decoupled and takes s = 1.

// The scale s of one pool's (x, y, z) in S (water: 1): its fraction, or 1 when the pool is decoupled
// (fraction 0 and rate 0). mirror_ok has ruled out a zero fraction with a nonzero rate.
__device__ __forceinline__ double sym_scale(double f) { return f > 0.0 ? f : 1.0; }

// The generator of the transposed chain: A^T (precession, RF and pool offsets change sign, exchange
// k <-> k f swaps) with the affine column S^-1 C (matrix.py:289-293: C = (r1w, r1_i f_i, r1mt f_mt) on
// the z rows, divided by the pools' scales).
template <int NC>
__device__ __forceinline__ void transpose_gen(Gen<NC>& G, const double* prm, double sc) {
  G.offw = -G.offw;
  G.w1x = -G.w1x;
  G.w1y = -G.w1y;
#pragma unroll
  for (int i = 0; i < NC; i++) {
    G.o[i] = -G.o[i];
    const double k = G.k[i];
    G.k[i] = G.kw[i];
    G.kw[i] = k;
    if (prm[13 + NC + i] > 0.0) G.cc[i] = prm[13 + 3 * NC + i] * sc;
  }
  const double kmt = G.kmt;
  G.kmt = G.kwmt;
  G.kwmt = kmt;
  if (prm[4] > 0.0) G.cmt = prm[6] * sc;
}

// The generator derivative of the transposed chain in shape form: dA^T (the off-diagonal pair of every
// coupling swapped, antisymmetric terms negated) and d(S^-1 C), which drops the fraction from the
// C_F / MT_F and C_R1 / MT_R1 terms of a pool with a positive fraction.
template <int NC>
__device__ __forceinline__ DxCo dx_transpose(int kind, int ix, DxCo d, const TanCo& t, const double* prm) {
  double q[7];
#pragma unroll
  for (int i = 0; i < 7; i++) q[i] = d.q[i];
  if (d.shape == DXS_WATER) {
    d.q[2] = q[4];
    d.q[4] = q[2];
    if (prm[4] > 0.0) {
      if (kind == TK_MT_F) d.q[6] = 0.0;
      if (kind == TK_MT_R1) d.q[6] = t.c0;
    }
  } else if (d.shape == DXS_POOL) {
    d.q[1] = q[2];
    d.q[2] = q[1];
    d.q[4] = -q[4];
    double f = 0.0;
#pragma unroll
    for (int i = 0; i < NC; i++)
      if (i == ix) f = prm[13 + NC + i];
    if (f > 0.0) {
      if (kind == TK_C_F) d.q[6] = 0.0;
      if (kind == TK_C_R1) d.q[6] = t.c0;
    }
  } else if (d.shape == DXS_SHIFT) {
#pragma unroll
    for (int i = 0; i <= NC; i++) d.q[i] = -q[i];
  } else if (d.shape == DXS_B1) {
    d.q[0] = -q[0];
    d.q[1] = -q[1];
  }
  return d;
}

// Per-team cache of the MT lineshape, keyed on the event's RF frequency and amplitude presence.
struct MtCache {
  double fr;
  MtShape s;
};

// Builds the generator of one event (times `sc` = dt 2^-s) and, for every tangent, its coefficients.
// Returns the unscaled augmented 1-norm ||aug dt||_1 in l1. mmax caps the Taylor degree (the
// squaring count follows from it).
template <int NC, int PT>
__device__ void build_gen(const double* prm, const Ev& e, MtCache& mc, const int* kind, const int* kidx,
                          Gen<NC>& G, TanCo* tc, double& l1, double& sc_out, double nan_l1, int& m_out,
                          int& s_out, bool& is_nan, int mmax = kMMax) {
  using I = PrmIdx<NC>;
  const double r1w = prm[0], r2w = prm[1], kmt = prm[3], fmt = prm[4], dwmt = prm[5], r1mt = prm[6], r2mt = prm[7];
  const int ls = (int)prm[8];
  const double b0 = prm[9], gam = prm[10], b0i = prm[11], relb1 = prm[12];
  const double L = prm[I::L], ks = prm[I::KS], kwmt = prm[I::KWMT], dw0 = prm[I::DW0];
  const double* kk = prm + 13;
  const double* ff = prm + 13 + NC;
  const double* dw = prm + 13 + 2 * NC;
  const double* r1c = prm + 13 + 3 * NC;
  const double* r2c = prm + 13 + 4 * NC;
  const double twopi = 6.283185307179586;
  const double w1 = twopi * e.amp * relb1;
  const double w1x = w1 * e.cph, w1y = w1 * e.sph;
  const double eff = twopi * e.fr - dw0;
  const double off_mt = eff - dwmt * 1e-6 * L;
  const bool mt_on = fmt > 1e-9 && ls != 0;
  double wmt = 0.0;
  MtShape sh{0.0, 0.0, 0.0, 0.0};
  if (mt_on && w1 != 0.0) {
    if (e.fr != mc.fr) {
      mc.fr = e.fr;
      mc.s = mt_shape(ls, off_mt, 1.0 / r2mt, L * 1e-6);
    }
    sh = mc.s;
    wmt = w1 * w1 * sh.g;
  }
  // the 1-norm of the augmented generator (columns), unscaled by 2^-s
  const double dt = e.dt;
  double axx = -r2w - ks, azz = -r1w - ks - kwmt, dmt = -r1mt - kmt - wmt;
  double skw = 0.0;
#pragma unroll
  for (int i = 0; i < NC; i++) skw += fabs(kk[i] * ff[i]);
  double cx = fabs(axx) + fabs(eff) + fabs(w1y) + skw;
  double cy = fabs(eff) + fabs(axx) + fabs(w1x) + skw;
  double cz = fabs(w1y) + fabs(w1x) + fabs(azz) + skw + fabs(kwmt);
  double cmx = fabs(kmt) + fabs(dmt);
  double caug = fabs(r1w) + fabs(r1mt * fmt);
  double n1 = fmax(fmax(cx, cy), fmax(cz, cmx));
#pragma unroll
  for (int i = 0; i < NC; i++) {
    double oi = dw[i] * 1e-6 * L - eff;
    double bi = -r2c[i] - kk[i], ci = -r1c[i] - kk[i];
    n1 = fmax(n1, fabs(kk[i]) + fabs(bi) + fabs(oi) + fabs(w1y));
    n1 = fmax(n1, fabs(kk[i]) + fabs(oi) + fabs(bi) + fabs(w1x));
    n1 = fmax(n1, fabs(kk[i]) + fabs(w1y) + fabs(w1x) + fabs(ci));
    caug += fabs(r1c[i] * ff[i]);
  }
  n1 = fmax(n1, caug);
  l1 = n1 * dt;
  is_nan = !(l1 < nan_l1);
  // Taylor degree and squarings
  int m, s;
  if (l1 <= kTheta[mmax]) {
    s = 0;
    m = 1;
    while (m < mmax && kTheta[m] < l1) m++;
    if (l1 == 0.0) m = 1;
  } else {
    s = (int)ceil(log2(l1 / kTheta[mmax]));
    if (s < 0) s = 0;
    m = mmax;
    while (s > 0 && l1 * ldexp(1.0, -s) > kTheta[m]) s++;  // guard rounding of log2
  }
  if (is_nan) s = 0;
  m_out = m;
  s_out = s;
  const double sc = dt * ldexp(1.0, -s);
  sc_out = sc;
  G.axx = axx * sc;
  G.offw = eff * sc;
  G.w1x = w1x * sc;
  G.w1y = w1y * sc;
  G.azz = azz * sc;
  G.cw = r1w * sc;
#pragma unroll
  for (int i = 0; i < NC; i++) {
    G.b[i] = (-r2c[i] - kk[i]) * sc;
    G.o[i] = (dw[i] * 1e-6 * L - eff) * sc;
    G.c[i] = (-r1c[i] - kk[i]) * sc;
    G.k[i] = kk[i] * sc;
    G.kw[i] = kk[i] * ff[i] * sc;
    G.cc[i] = r1c[i] * ff[i] * sc;
  }
  G.dmt = dmt * sc;
  G.kwmt = kwmt * sc;
  G.kmt = kmt * sc;
  G.cmt = r1mt * fmt * sc;
  (void)b0;
  (void)gam;
  // tangent coefficients
#pragma unroll
  for (int p = 0; p < PT; p++) {
    TanCo t{0.0, 0.0, 0.0};
    const int ix = kidx[p];
    switch (kind[p]) {
      case TK_W_R1:
      case TK_W_R2:
        t.c0 = sc;
        break;
      case TK_MT_K:
        t.c0 = fmt * sc;
        t.c1 = sc;
        break;
      case TK_MT_F:
        t.c0 = kmt * sc;
        t.c1 = r1mt * sc;
        break;
      case TK_MT_R1:
        t.c0 = sc;
        t.c1 = fmt * sc;
        break;
      case TK_MT_DW:  // d off_mt = -1e-6 L
        if (mt_on) t.c2 = -(w1 * w1 * sh.g_off * (-1e-6 * L)) * sc;
        break;
      case TK_MT_R2:  // t2 = 1 / r2
        if (mt_on) t.c2 = -(w1 * w1 * sh.g_t2 * (-1.0 / (r2mt * r2mt))) * sc;
        break;
      case TK_B0INH:
        t.c0 = -1e-6 * L * sc;
        if (mt_on) t.c2 = -(w1 * w1 * sh.g_off * (-1e-6 * L)) * sc;
        break;
      case TK_B0:
      case TK_GAMMA: {
        const double dL = (kind[p] == TK_B0) ? gam * 1e6 : 1e6 * b0;
        t.c0 = -dL * b0i * 1e-6 * sc;
        t.c1 = 1e-6 * dL * sc;
        if (mt_on) {
          double doff = -1e-6 * dL * (b0i + dwmt);
          t.c2 = -(w1 * w1 * (sh.g_off * doff + sh.g_w0 * 1e-6 * dL)) * sc;
        }
      } break;
      case TK_RELB1: {
        const double dw1 = twopi * e.amp;
        t.c0 = dw1 * e.cph * sc;
        t.c1 = dw1 * e.sph * sc;
        if (mt_on) t.c2 = -(2.0 * w1 * dw1 * sh.g) * sc;
      } break;
      case TK_C_K:
        t.c0 = ff[ix] * sc;
        t.c1 = sc;
        break;
      case TK_C_F:
        t.c0 = kk[ix] * sc;
        t.c1 = r1c[ix] * sc;
        break;
      case TK_C_DW:
        t.c0 = 1e-6 * L * sc;
        break;
      case TK_C_R1:
        t.c0 = sc;
        t.c1 = ff[ix] * sc;
        break;
      case TK_C_R2:
        t.c0 = sc;
        break;
      default:
        break;
    }
    tc[p] = t;
  }
}

// Truncated Taylor series acc = T_m(X) v0 with tangents dacc = d(T_m(X) v0) (v0's own tangent dv0),
// by Horner's rule: w <- v0 + (X w) / k for k = m, ..., 1 from w = v0, one FMA per entry and term
// where the power sum needs a scale and an add. The tangents follow the same recursion,
// dw <- dv0 + (X dw + dX w) / k. The homogeneous coordinate of every iterate stays vaug0 (X's last row
// is zero). acc / dacc may alias v0 / dv0 (in-place use): v0 is copied first.
// The Horner recursion itself, over terms k = m..1, with the tangents' dX in shape form (SHAPE / IX as in
// apply_dx_shape).
template <int NC, int PT, int SHAPE, int IX, int QM0 = 0x7f, int QM1 = 0x7f>
__device__ __forceinline__ void taylor_terms(const Gen<NC>& G, const DxCo* dx, int m, const double* v, double vaug0,
                                             const double (*dv)[3 * (1 + NC) + 1], double* acc,
                                             double (*dacc)[3 * (1 + NC) + 1]) {
  constexpr int N = 3 * (1 + NC) + 1;
  double ik = kInv[m];
  for (int k = m; k >= 1; k--) {
    const double ik_next = kInv[k - 1];  // the next term's 1 / k is in flight during this one
    double y[N];
    apply_X<NC>(G, acc, vaug0, y);
#pragma unroll
    for (int p = 0; p < PT; p++) {
      double dy[N];
      apply_X<NC>(G, dacc[p], 0.0, dy);
      if (p == 0)  // dX at the current iterate
        apply_dx_shape<NC, SHAPE, IX, QM0>(dx[p], acc, vaug0, dy);
      else
        apply_dx_shape<NC, SHAPE, IX, QM1>(dx[p], acc, vaug0, dy);
#pragma unroll
      for (int i = 0; i < N; i++) dacc[p][i] = fma(dy[i], ik, dv[p][i]);
    }
#pragma unroll
    for (int i = 0; i < N; i++) acc[i] = fma(y[i], ik, v[i]);
    ik = ik_next;
  }
}

// Truncated Taylor series acc = T_m(X) v0 with tangents dacc = d(T_m(X) v0) (v0's own tangent dv0),
// by Horner's rule: w <- v0 + (X w) / k for k = m, ..., 1 from w = v0, one FMA per entry and term
// where the power sum needs a scale and an add. The tangents follow the same recursion,
// dw <- dv0 + (X dw + dX w) / k. The homogeneous coordinate of every iterate stays vaug0 (X's last row
// is zero). acc / dacc may alias v0 / dv0 (in-place use): v0 is copied first.
// UNSWITCH: when every tangent has the same dX shape (and pool), run a copy of the recursion with that
// shape fixed at compile time, so no term branches on it (for the hottest call site only: each copy is
// more code).
// tr_prm (the team's parameters) marks the lane of a palindrome factorisation that runs the transposed
// chain: its tangent generators are transposed as well (dx_transpose).
// dv0 is either an array of PT tangent columns or nullptr (zero tangents), told apart by its type: a runtime
// test of a local array's address against null cannot be folded (0 is a valid private address), and it
// kept the caller's array in scratch memory.
template <int NC, int PT, bool UNSWITCH = false, typename DV0>
__device__ __forceinline__ void taylor(const Gen<NC>& G, const TanCo* tc, const int* kind, const int* kidx,
                                       const double* dwp, int m, const double* v0, double vaug0, DV0 dv0,
                                       double* acc, double (*dacc)[3 * (1 + NC) + 1], const double* tr_prm = nullptr) {
  constexpr int N = 3 * (1 + NC) + 1;
  double v[N], dv[PT > 0 ? PT : 1][N];
  DxCo dx[PT > 0 ? PT : 1];
#pragma unroll
  for (int p = 0; p < PT; p++) {
    dx[p] = dx_shape<NC>(kind[p], kidx[p], tc[p], dwp);
    if (tr_prm) dx[p] = dx_transpose<NC>(kind[p], kidx[p], dx[p], tc[p], tr_prm);
  }
#pragma unroll
  for (int i = 0; i < N; i++) v[i] = v0[i];
#pragma unroll
  for (int p = 0; p < PT; p++)
#pragma unroll
    for (int i = 0; i < N; i++) {
      if constexpr (std::is_same_v<DV0, std::nullptr_t>)
        dv[p][i] = 0.0;
      else
        dv[p][i] = dv0[p][i];
    }
#pragma unroll
  for (int i = 0; i < N; i++) acc[i] = v[i];
#pragma unroll
  for (int p = 0; p < PT; p++)
#pragma unroll
    for (int i = 0; i < N; i++) dacc[p][i] = dv[p][i];
  if constexpr (UNSWITCH && PT > 0) {
    bool same = true;
#pragma unroll
    for (int p = 1; p < PT; p++) same = same && dx[p].shape == dx[0].shape && dx[p].ix == dx[0].ix;
    if (same) {
      const int sh = dx[0].shape;
      // the usual pairs of fitted values also at the kind level: only the coefficients the two kinds
      // (forward or transposed) can make nonzero are applied
      if constexpr (PT == 2) {
        if (kind[0] == TK_W_R1 && kind[1] == TK_W_R2)
          return taylor_terms<NC, PT, DXS_WATER, -1, kQmWR1, kQmWR2>(G, dx, m, v, vaug0, dv, acc, dacc);
        if (kind[0] == TK_C_K && kind[1] == TK_C_F) {
          if constexpr (NC > 0)
            if (dx[0].ix == 0) return taylor_terms<NC, PT, DXS_POOL, 0, kQmCK, kQmCF>(G, dx, m, v, vaug0, dv, acc, dacc);
          if constexpr (NC > 1)
            if (dx[0].ix == 1) return taylor_terms<NC, PT, DXS_POOL, 1, kQmCK, kQmCF>(G, dx, m, v, vaug0, dv, acc, dacc);
        }
      }
      if (sh == DXS_WATER) return taylor_terms<NC, PT, DXS_WATER, -1>(G, dx, m, v, vaug0, dv, acc, dacc);
      if (sh == DXS_SHIFT) return taylor_terms<NC, PT, DXS_SHIFT, -1>(G, dx, m, v, vaug0, dv, acc, dacc);
      if (sh == DXS_B1) return taylor_terms<NC, PT, DXS_B1, -1>(G, dx, m, v, vaug0, dv, acc, dacc);
      if (sh == DXS_POOL) {
        if constexpr (NC > 0)
          if (dx[0].ix == 0) return taylor_terms<NC, PT, DXS_POOL, 0>(G, dx, m, v, vaug0, dv, acc, dacc);
        if constexpr (NC > 1)
          if (dx[0].ix == 1) return taylor_terms<NC, PT, DXS_POOL, 1>(G, dx, m, v, vaug0, dv, acc, dacc);
        if constexpr (NC > 2)
          if (dx[0].ix == 2) return taylor_terms<NC, PT, DXS_POOL, 2>(G, dx, m, v, vaug0, dv, acc, dacc);
        if constexpr (NC > 3)
          if (dx[0].ix == 3) return taylor_terms<NC, PT, DXS_POOL, 3>(G, dx, m, v, vaug0, dv, acc, dacc);
      }
    }
  }
  taylor_terms<NC, PT, -1, -1>(G, dx, m, v, vaug0, dv, acc, dacc);
}

// out = A B for affine propagators in column layout (lane j holds column j, rows 0..N-1):
// out[:, j] = sum_k A[:, k] B[k, j] + A[:, N] [j == N]; tangents by the product rule.
// STAGED (12-lane teams, at most one tangent): 12-lane teams do not align with DPP rows, so their
// broadcasts are ds_bpermutes, two per double; staged, the left factor's columns (then the tangent's) are
// written to the team's LDS stage (N * N doubles) once and read back as 16-byte broadcast loads, 3.5x the
// products per second in isolation. With two tangents the staged loop spilled: those kernels keep the
// broadcasts. It raises the register pressure of a loop it is inlined into: it runs in staged_event.
// BD: A and its tangents are block-diagonal, transverse rows [0, 2 NT) apart from the z rows [2 NT, N)
// (an event without RF, rf_free): column k of A is broadcast and multiplied over its own block's rows
// only. The terms left out are exact zeros and the others accumulate in the same order, so the result
// is the same as the dense product's.
template <int NC, int PT, int TEAM, bool STAGED = false, bool BD = false>
__device__ __forceinline__ void mul_cols(const double* A, const double (*dA)[3 * (1 + NC) + 1], double* B,
                                         double (*dB)[3 * (1 + NC) + 1], int j, double* stage = nullptr) {
  constexpr int N = 3 * (1 + NC) + 1, NXY = 2 * (1 + NC);
  double o[N], od[PT > 0 ? PT : 1][N];
  const bool aff = (j == N);
#pragma unroll
  for (int i = 0; i < N; i++) {
    o[i] = aff ? A[i] : 0.0;
#pragma unroll
    for (int p = 0; p < PT; p++) od[p][i] = aff ? dA[p][i] : 0.0;
  }
  if constexpr (STAGED) {
    static_assert(TEAM == 12 && N % 2 == 0 && PT <= 1, "the staged product: 12-lane teams, one tangent");
    volatile double* stw = stage;  // volatile: a copy loop would become a memcpy from the registers
    const double2* st2 = reinterpret_cast<const double2*>(stage);
    static_for<0, 1 + PT>([&](auto rc) {
      constexpr int r = decltype(rc)::value;  // 0: A itself, then the tangent's dA
      team_lds_sync();                        // the previous round's reads have landed
      if (j < N) {
#pragma unroll
        for (int i = 0; i < N; i++) stw[j * N + i] = r == 0 ? A[i] : dA[0][i];
      }
      team_lds_sync();
#pragma unroll
      for (int k = 0; k < N; k++) {
        const double bk = B[k];
        const int i0 = (BD && k >= NXY) ? NXY : 0, i1 = (BD && k < NXY) ? NXY : N;  // NXY is even
#pragma unroll
        for (int i2 = i0 / 2; i2 < i1 / 2; i2++) {
          const double2 a2 = st2[(k * N) / 2 + i2];  // A[2 i2][k], A[2 i2 + 1][k], one load for the team
#pragma unroll
          for (int h = 0; h < 2; h++) {
            const int i = 2 * i2 + h;
            const double a = h == 0 ? a2.x : a2.y;
            if constexpr (r == 0) {
              o[i] = fma(a, bk, o[i]);
              if constexpr (PT > 0) od[0][i] = fma(a, dB[0][k], od[0][i]);
            } else {
              od[0][i] = fma(a, bk, od[0][i]);
            }
          }
        }
        __builtin_amdgcn_sched_barrier(0);  // one staged column's loads in flight at a time
      }
    });
    team_lds_sync();  // the stage is free again once every read of it has landed
  } else {
    static_for<0, N>([&](auto kc) {
      constexpr int k = decltype(kc)::value;
      constexpr int i0 = (BD && k >= NXY) ? NXY : 0, i1 = (BD && k < NXY) ? NXY : N;
      const double bk = B[k];
#pragma unroll
      for (int i = i0; i < i1; i++) {
        const double a = bcast<TEAM, k>(A[i]);
        o[i] = fma(a, bk, o[i]);
#pragma unroll
        for (int p = 0; p < PT; p++) od[p][i] = fma(a, dB[p][k], od[p][i]);
      }
#pragma unroll
      for (int p = 0; p < PT; p++)
#pragma unroll
        for (int i = i0; i < i1; i++) od[p][i] = fma(bcast<TEAM, k>(dA[p][i]), bk, od[p][i]);
    });
  }
#pragma unroll
  for (int i = 0; i < N; i++) {
    B[i] = o[i];
#pragma unroll
    for (int p = 0; p < PT; p++) dB[p][i] = od[p][i];
  }
}

// M <- A M (state uniform over the team, A in column layout), with tangents.
template <int NC, int PT, int TEAM>
__device__ __forceinline__ void apply_cols(const double* A, const double (*dA)[3 * (1 + NC) + 1], double* M,
                                           double (*dM)[3 * (1 + NC) + 1], int j) {
  constexpr int N = 3 * (1 + NC) + 1;
  double mj = (j == N) ? 1.0 : 0.0;
  double dmj[PT > 0 ? PT : 1];
#pragma unroll
  for (int p = 0; p < PT; p++) dmj[p] = 0.0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    if (j == i) {
      mj = M[i];
#pragma unroll
      for (int p = 0; p < PT; p++) dmj[p] = dM[p][i];
    }
  }
#pragma unroll
  for (int p = 0; p < PT; p++)
#pragma unroll
    for (int i = 0; i < N; i++) dM[p][i] = team_sum<TEAM>(fma(dA[p][i], mj, A[i] * dmj[p]));
#pragma unroll
  for (int i = 0; i < N; i++) M[i] = team_sum<TEAM>(A[i] * mj);
}

// An event without RF: its generator and every tangent's derivative of it have no term between the
// transverse rows (x, y of every pool) and the z rows (z of every pool, the MT pool, the affine
// column), so its propagator is block-diagonal in them. Only the RF vector (w1x, w1y) and the B1 tangent,
// whose coefficients are derivatives of it, couple the two.
template <int NC, int PT>
__device__ __forceinline__ bool rf_free(const Gen<NC>& G, const TanCo* tc, const int* kind) {
  bool f = G.w1x == 0.0 && G.w1y == 0.0;
#pragma unroll
  for (int p = 0; p < PT; p++) f = f && (kind[p] != TK_RELB1 || (tc[p].c0 == 0.0 && tc[p].c1 == 0.0));
  return f;
}

// The propagator of one event in column layout: Taylor column + s squarings. Returns rf_free: the
// propagator is then block-diagonal (mul_cols' BD form applies to it).
template <int NC, int PT, int TEAM, bool STAGED = false>
__device__ __forceinline__ bool event_cols(const Gen<NC>& G, const TanCo* tc, const int* kind, const int* kidx,
                                           const double* dwp, int m, int s, bool is_nan, int j, double* R,
                                           double (*dR)[3 * (1 + NC) + 1], double* stage = nullptr) {
  constexpr int N = 3 * (1 + NC) + 1;
#pragma unroll
  for (int i = 0; i < N; i++) R[i] = (i == j) ? 1.0 : 0.0;
  taylor<NC, PT>(G, tc, kind, kidx, dwp, m, R, (j == N) ? 1.0 : 0.0, nullptr, R, dR);  // column j of T_m(X)
  // squarings: mul_cols finishes the product before it writes back, so R may be both factors; an event
  // without RF (the delays) squares its two blocks apart
  const bool bd = rf_free<NC, PT>(G, tc, kind);
  if (bd) {
    for (int q = 0; q < s; q++) mul_cols<NC, PT, TEAM, STAGED, true>(R, dR, R, dR, j, stage);
  } else {
    for (int q = 0; q < s; q++) mul_cols<NC, PT, TEAM, STAGED>(R, dR, R, dR, j, stage);
  }
  if (is_nan) {
#pragma unroll
    for (int i = 0; i < N; i++) {
      R[i] = __builtin_nan("");
#pragma unroll
      for (int p = 0; p < PT; p++) dR[p][i] = __builtin_nan("");
    }
  }
  return bd;
}


// An event that needs squarings, in 12-lane teams with an LDS stage and at most one tangent: its
// propagator columns (Taylor series on the identity columns, then the squarings through the stage),
// multiplied into the chained columns `Pi` when `chain` (else returned as they are). A call of its own
// with its arguments and result passed by value, like mirror_chain, so that the staged products'
// registers stay out of the event loop.
template <int NC, int PT>
struct ColsIn {  // the event's generator, its tangents and the chained columns
  Gen<NC> G;
  TanCo tc[PT > 0 ? PT : 1];
  int kind[PT > 0 ? PT : 1], kidx[PT > 0 ? PT : 1];
  double dwp[NC > 0 ? NC : 1];
  double Pi[3 * (1 + NC) + 1];
  double dPi[PT > 0 ? PT : 1][3 * (1 + NC) + 1];
};
template <int NC, int PT>
struct ColsOut {
  double R[3 * (1 + NC) + 1];
  double dR[PT > 0 ? PT : 1][3 * (1 + NC) + 1];
};

template <int NC, int PT, int TEAM>
__device__ __noinline__ ColsOut<NC, PT> staged_event(ColsIn<NC, PT> in, int m, int s, bool is_nan, int j, bool chain,
                                                     double* stage_g) {
  constexpr int N = 3 * (1 + NC) + 1;
  constexpr int PTA = PT > 0 ? PT : 1;
  double* stage = stage_g;
  __builtin_assume(__builtin_amdgcn_is_shared((const __attribute__((address_space(0))) void*)stage));
  // local copies: the by-value argument and the returned record live in scratch memory
  const Gen<NC> G = in.G;
  TanCo tc[PTA];
  int kind[PTA], kidx[PTA];
  double dwp[NC > 0 ? NC : 1];
#pragma unroll
  for (int p = 0; p < PTA; p++) {
    tc[p] = in.tc[p];
    kind[p] = in.kind[p];
    kidx[p] = in.kidx[p];
  }
#pragma unroll
  for (int i = 0; i < (NC > 0 ? NC : 1); i++) dwp[i] = in.dwp[i];
  double R[N], dR[PTA][N];
  const bool bd = event_cols<NC, PT, TEAM, true>(G, tc, kind, kidx, dwp, m, s, is_nan, j, R, dR, stage);
  if (chain) {
    double Pi[N], dPi[PTA][N];
#pragma unroll
    for (int i = 0; i < N; i++) {
      Pi[i] = in.Pi[i];
#pragma unroll
      for (int p = 0; p < PTA; p++) dPi[p][i] = in.dPi[p][i];
    }
    if (bd)
      mul_cols<NC, PT, TEAM, true, true>(R, dR, Pi, dPi, j, stage);
    else
      mul_cols<NC, PT, TEAM, true>(R, dR, Pi, dPi, j, stage);
#pragma unroll
    for (int i = 0; i < N; i++) {
      R[i] = Pi[i];
#pragma unroll
      for (int p = 0; p < PTA; p++) dR[p][i] = dPi[p][i];
    }
  }
  ColsOut<NC, PT> out;
#pragma unroll
  for (int i = 0; i < N; i++) {
    out.R[i] = R[i];
#pragma unroll
    for (int p = 0; p < PTA; p++) out.dR[p][i] = dR[p][i];
  }
  return out;
}

// M <- P M with P the kept propagator in LDS (lane j's column j at stride N + 1, tangent columns
// after it), with tangents. Entries are read as they are used, so the kept block adds no registers
// beyond the state's.
template <int NC, int PT, int TEAM>
__device__ __forceinline__ void apply_kept(const double* keep, double* M, double (*dM)[3 * (1 + NC) + 1], int j) {
  constexpr int N = 3 * (1 + NC) + 1, NA = N + 1;
  constexpr int PTA = PT > 0 ? PT : 1;
  const bool own = j < NA;
  double mj = (j == N) ? 1.0 : 0.0;
  double dmj[PTA];
#pragma unroll
  for (int p = 0; p < PTA; p++) dmj[p] = 0.0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    if (j == i) {
      mj = M[i];
#pragma unroll
      for (int p = 0; p < PTA; p++) dmj[p] = dM[p][i];
    }
  }
#pragma unroll
  for (int i = 0; i < N; i++) {
    const double a = own ? keep[i * NA + j] : 0.0;
#pragma unroll
    for (int p = 0; p < PTA; p++) {
      const double da = own ? keep[((1 + p) * N + i) * NA + j] : 0.0;
      dM[p][i] = team_sum<TEAM>(fma(da, mj, a * dmj[p]));
    }
    M[i] = team_sum<TEAM>(a * mj);
  }
}

// M <- R M for the rotation R = [[c, -s], [s, c]] of every pool's (x, y) about z, with tangents. The
// generator of an RF event at phase phi + d is R(d) X(phi) R(d)^T: R commutes with the offset
// precession, the transverse relaxation and the exchange terms, and turns the RF vector (w1x, w1y).
template <int NC, int PTA>
__device__ __forceinline__ void rotate_xy(double* M, double (*dM)[3 * (1 + NC) + 1], double c, double s) {
  constexpr int NT = 1 + NC;
#pragma unroll
  for (int i = 0; i < NT; i++) {
    const double x = M[i], y = M[NT + i];
    M[i] = fma(c, x, -s * y);
    M[NT + i] = fma(s, x, c * y);
#pragma unroll
    for (int p = 0; p < PTA; p++) {
      const double dx = dM[p][i], dy = dM[p][NT + i];
      dM[p][i] = fma(c, dx, -s * dy);
      dM[p][NT + i] = fma(s, dx, c * dy);
    }
  }
}

// Whether the team's parameters and tangents admit the palindrome factorisation: every pool (and MT)
// either has a positive fraction or is decoupled (fraction 0, rate 0) with no tangent along its rate.
template <int NC, int PT>
__device__ __forceinline__ bool mirror_ok(const double* prm, const int* kind, const int* kidx) {
  constexpr int PTA = PT > 0 ? PT : 1;
  bool ok = prm[4] > 0.0 || (prm[4] == 0.0 && prm[3] == 0.0);
#pragma unroll
  for (int p = 0; p < PTA; p++) ok = ok && !(p < PT && kind[p] == TK_MT_K && !(prm[4] > 0.0));
#pragma unroll
  for (int i = 0; i < NC; i++) {
    const double f = prm[13 + NC + i];
    ok = ok && (f > 0.0 || (f == 0.0 && prm[13 + i] == 0.0));
#pragma unroll
    for (int p = 0; p < PTA; p++) ok = ok && !(p < PT && kind[p] == TK_C_K && kidx[p] == i && !(f > 0.0));
  }
  return ok;
}

// y = S^e x for S of the palindrome factorisation (e = +1 or -1) at RF phase (c, s) = (cos phi, sin phi):
// S = R(phi) S0 R(-phi), S0 = diag(1, s_i, -1, -s_i, 1, s_i, s_mt).
template <int NC>
__device__ __forceinline__ void apply_sym(const double* prm, double c, double s, bool inverse, const double* x,
                                          double* y) {
  constexpr int NT = 1 + NC, N = 3 * NT + 1;
#pragma unroll
  for (int i = 0; i < N; i++) y[i] = x[i];
#pragma unroll
  for (int i = 0; i < NT; i++) {  // R(-phi)
    const double a = y[i], b = y[NT + i];
    y[i] = fma(c, a, s * b);
    y[NT + i] = fma(-s, a, c * b);
  }
#pragma unroll
  for (int i = 0; i < NT; i++) {
    const double sc = i == 0 ? 1.0 : sym_scale(prm[13 + NC + i - 1]);
    const double m = inverse ? 1.0 / sc : sc;
    y[i] *= m;
    y[NT + i] *= -m;
    y[2 * NT + i] *= m;
  }
  {
    const double sc = sym_scale(prm[4]);
    y[3 * NT] *= inverse ? 1.0 / sc : sc;
  }
#pragma unroll
  for (int i = 0; i < NT; i++) {  // R(phi)
    const double a = y[i], b = y[NT + i];
    y[i] = fma(c, a, -s * b);
    y[NT + i] = fma(s, a, c * b);
  }
}

// y += g (.) x for g = S^-1 dS of one tangent: 1 / f on the (x, y, z) of the pool whose fraction the
// tangent moves (C_F), 1 / f_mt on MT's z (MT_F), when that fraction is positive; no other tangent
// moves S.
// Branch-free, every entry at a compile-time index (a zero g where the tangent does not move S): stores
// under a runtime condition were merged into one store at a runtime index, which sent the arrays it
// writes to scratch memory.
template <int NC>
__device__ __forceinline__ void add_dsym(const double* prm, int kind, int ix, const double* x, double* y) {
  constexpr int NT = 1 + NC;
  {
    const double f = prm[4];
    const double g = (kind == TK_MT_F && f > 0.0) ? 1.0 / f : 0.0;
    y[3 * NT] = fma(g, x[3 * NT], y[3 * NT]);
  }
#pragma unroll
  for (int i = 0; i < NC; i++) {
    const double f = prm[13 + NC + i];
    const double g = (kind == TK_C_F && i == ix && f > 0.0) ? 1.0 / f : 0.0;
    y[1 + i] = fma(g, x[1 + i], y[1 + i]);
    y[NT + 1 + i] = fma(g, x[NT + 1 + i], y[NT + 1 + i]);
    y[2 * NT + 1 + i] = fma(g, x[2 * NT + 1 + i], y[2 * NT + 1 + i]);
  }
}


// The propagator of the palindromic run [e0, e0 + 2h) (with tangents) by its factorisation Q^R Q: lanes
// 0..N chain the columns of Q over events e0 .. e0 + h - 1 as a chained block does, while lane N + 1
// (spare in every team wider than the augmented size) chains the transposed generators over the same
// events in reverse order from the affine unit column, giving w. Then, per lane j <= N,
//   v = S^-1 Q[:, j], dv = S^-1 dQ[:, j] - g v, r = Q_A^T v, dr = Q_A^T dv + dQ_A^T v,
//   Pi[:, j] = S (r + [j = N] w), dPi[:, j] = S (g (r + [j = N] w) + dr + [j = N] dw)
// with g = S^-1 dS. Q_A's columns (then each tangent's dQ_A) and w pass through `stage`, the team's
// LDS area of (N + 1) N doubles. Returns false, with Pi and dPi unspecified, when an event needs
// squarings on some lane of the team (its propagator cannot be chained column by column); the caller
// then chains the run event by event.
// It is a call of its own (not inlined), with its arguments and result passed by value, so that its
// register allocation does not raise the pressure in the caller's event loops.
template <int NC, int PT>
struct MirrorIn {  // a pass's tangents (kinds, pools) and the pools' offsets
  int kind[PT > 0 ? PT : 1], kidx[PT > 0 ? PT : 1];
  double dwp[NC > 0 ? NC : 1];
};
template <int NC, int PT>
struct MirrorOut {  // the run's propagator column and its tangents; ok = false: chain event by event
  double Pi[3 * (1 + NC) + 1];
  double dPi[PT > 0 ? PT : 1][3 * (1 + NC) + 1];
  bool ok;
};

template <int NC, int PT, int TEAM>
__device__ __noinline__ MirrorOut<NC, PT> mirror_chain(const double* prm_g, MirrorIn<NC, PT> in, const Ev* evr, int e0,
                                                       int h, double cph, double sph, double nan_l1, int lane,
                                                       double* stage_g) {
  constexpr int NT = 1 + NC, N = 3 * NT + 1, NA = N + 1;
  constexpr int PTA = PT > 0 ? PT : 1;
  static_assert(TEAM > NA, "the factorisation needs a spare lane");
  // prm and stage are the team's LDS areas: a call of its own receives them as generic pointers, whose
  // loads are flat loads; told that they point into LDS, the compiler emits LDS loads
  const double* prm = prm_g;
  double* stage = stage_g;
  __builtin_assume(__builtin_amdgcn_is_shared((const __attribute__((address_space(0))) void*)prm));
  __builtin_assume(__builtin_amdgcn_is_shared((const __attribute__((address_space(0))) void*)stage));
  MirrorOut<NC, PT> out;
  out.ok = false;
  // the chain works on local copies: the by-value argument and the returned record live in scratch
  // memory, and working in them directly sent every event's series through it
  double Pi[N], dPi[PTA][N], dwp[NC > 0 ? NC : 1];
  int kind[PTA], kidx[PTA];
#pragma unroll
  for (int p = 0; p < PTA; p++) {  // the caller's tangents are wave-uniform (pass_teams)
    kind[p] = __builtin_amdgcn_readfirstlane(in.kind[p]);
    kidx[p] = __builtin_amdgcn_readfirstlane(in.kidx[p]);
  }
#pragma unroll
  for (int i = 0; i < (NC > 0 ? NC : 1); i++) dwp[i] = in.dwp[i];
  MtCache mc;
  mc.fr = __builtin_nan("");
  mc.s = MtShape{0.0, 0.0, 0.0, 0.0};
  const bool rev = lane == NA;
  const double vaug = (lane == N || rev) ? 1.0 : 0.0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    Pi[i] = (i == lane) ? 1.0 : 0.0;
#pragma unroll
    for (int p = 0; p < PTA; p++) dPi[p][i] = 0.0;
  }
  bool bad = false;
  int ei = rev ? e0 + h - 1 : e0;
  const int step = rev ? -1 : 1;
  Ev e_next = evr[ei];
  for (int t = 0; t < h; t++) {
    const Ev e = e_next;
    ei += step;
    if (t + 1 < h) e_next = evr[ei];
    if (e.dt > 1e-9) {
      Gen<NC> G;
      TanCo tc[PTA];
      double l1, sc;
      int m, s;
      bool isn;
      build_gen<NC, PT>(prm, e, mc, kind, kidx, G, tc, l1, sc, nan_l1, m, s, isn);
      bad = bad || s > 0 || isn;
      if (!bad) {
        if (rev) transpose_gen<NC>(G, prm, sc);
        taylor<NC, PT, true>(G, tc, kind, kidx, dwp, m, Pi, vaug, dPi, Pi, dPi, rev ? prm : nullptr);
      }
    }
  }
  if (team_sum<TEAM>(bad ? 1.0 : 0.0) != 0.0) return out;

  // recombination: lanes < N stage their column of Q_A (or dQ_A), the reversed lane w (or dw) as column N
  const int slot = lane < N ? lane : (rev ? N : -1);
  // the columns go to the stage by volatile stores: a plain copy loop became a memcpy from the local
  // arrays, which kept them in scratch memory through the whole chain
  volatile double* stage_w = stage;
  double v[N], dv[PTA][N], r[N], dr[PTA][N];
  apply_sym<NC>(prm, cph, sph, true, Pi, v);
#pragma unroll
  for (int p = 0; p < PTA; p++) {
    apply_sym<NC>(prm, cph, sph, true, dPi[p], dv[p]);
    double gv[N];
#pragma unroll
    for (int i = 0; i < N; i++) gv[i] = 0.0;
    if (p < PT) add_dsym<NC>(prm, kind[p], kidx[p], v, gv);
#pragma unroll
    for (int i = 0; i < N; i++) dv[p][i] -= gv[i];
  }
  const bool aff = lane == N;
  if (slot >= 0) {
#pragma unroll
    for (int b = 0; b < N; b++) stage_w[slot * N + b] = Pi[b];
  }
  team_lds_sync();
#pragma unroll
  for (int a = 0; a < N; a++) {
    double ra = 0.0, dra[PTA];
#pragma unroll
    for (int p = 0; p < PTA; p++) dra[p] = 0.0;
#pragma unroll
    for (int b = 0; b < N; b++) {
      const double qba = stage[a * N + b];
      ra = fma(qba, v[b], ra);
#pragma unroll
      for (int p = 0; p < PT; p++) dra[p] = fma(qba, dv[p][b], dra[p]);
    }
    r[a] = aff ? ra + stage[N * N + a] : ra;
#pragma unroll
    for (int p = 0; p < PTA; p++) dr[p][a] = dra[p];
  }
  // the tangent index is a compile-time constant (a loop holding the LDS syncs was left rolled, and
  // indexing dPi / dr by a runtime p kept them in scratch memory)
  static_for<0, PT>([&](auto pc) {
    constexpr int p = decltype(pc)::value;
    team_lds_sync();
    if (slot >= 0) {
#pragma unroll
      for (int b = 0; b < N; b++) stage_w[slot * N + b] = dPi[p][b];
    }
    team_lds_sync();
#pragma unroll
    for (int a = 0; a < N; a++) {
      double da = 0.0;
#pragma unroll
      for (int b = 0; b < N; b++) da = fma(stage[a * N + b], v[b], da);
      dr[p][a] += aff ? da + stage[N * N + a] : da;
    }
  });
  team_lds_sync();  // the stage is free again once every read of it has landed
  const bool own = lane <= N;
  apply_sym<NC>(prm, cph, sph, false, r, v);
#pragma unroll
  for (int i = 0; i < N; i++) out.Pi[i] = own ? v[i] : 0.0;
#pragma unroll
  for (int p = 0; p < PTA; p++) {
    if (p < PT) add_dsym<NC>(prm, kind[p], kidx[p], r, dr[p]);
    apply_sym<NC>(prm, cph, sph, false, dr[p], v);
#pragma unroll
    for (int i = 0; i < N; i++) out.dPi[p][i] = (own && p < PT) ? v[i] : 0.0;
  }
  out.ok = true;
  return out;
}

// Runs one offset row's program (the event scan, simulate.py:80-132) on the team's state M and its
// tangents dM: SEQ ops event by event, BLOCK ops chained into one propagator applied `count` times.
// prm / st0: the team's parameters and start state (+ seeds) in LDS; st_park: LDS room for the state
// while a block is chained. traj (optional): last_M after every event at traj_base + event * N.
// With KEEP (a plan holding KEEP ops), KEEP ops also store the chained propagator in `keep` (the team's
// LDS slot, (1 + PT) N (N + 1) doubles, lane j's column at stride N + 1) and REUSE ops apply it
// conjugated by the op's rotation (rot[2 op], rot[2 op + 1]) = (cos d, sin d). Without it a KEEP op is
// a BLOCK and the kept-pulse code is not compiled in: it would cost the other plans registers.
// With MIR, mir is the team's LDS stage of (N + 1) N doubles for the palindrome factorisation of chained
// ops with a half (mirror_chain), in teams with a spare lane; without it those ops chain every event.
// With STAGE (12-lane teams, at most one tangent), mir is that stage also for the events that need
// squarings, which then run in staged_event.
template <int NC, int PT, int TEAM, bool KEEP = false, bool MIR = false, bool STAGE = false>
__device__ __forceinline__ void run_row(const double* prm, const double* st0, double* st_park, const int* kind,
                                        const int* kidx, const double* dwp, const Ev* ev_row, const Op* ops, int op0,
                                        int op1, int reset_after_adc, double nan_l1, int lane,
                                        double* M, double (*dM)[3 * (1 + NC) + 1], void* traj, long traj_base,
                                        int traj_dtype, const double* rot = nullptr, double* keep = nullptr,
                                        double* mir = nullptr) {
  constexpr int NT = 1 + NC, N = 3 * NT + 1, NA = N + 1;
  constexpr int PTA = PT > 0 ? PT : 1;
  auto load_start = [&]() {
#pragma unroll
    for (int i = 0; i < N; i++) {
      M[i] = st0[i];
#pragma unroll
      for (int p = 0; p < PTA; p++) dM[p][i] = st0[(1 + p) * N + i];
    }
  };
  load_start();

  MtCache mc;
  mc.fr = __builtin_nan("");
  mc.s = MtShape{0.0, 0.0, 0.0, 0.0};
  const Ev* evr = ev_row;
  bool pending = false;

  for (int oi = op0; oi < op1; oi++) {
    const Op op = ops[oi];
    if (op.kind == OP_SEQ) {
      Ev e_next = evr[op.e0];
      for (int ei = op.e0; ei < op.e0 + op.len; ei++) {
        const Ev e = e_next;
        if (ei + 1 < op.e0 + op.len) e_next = evr[ei + 1];
        if (pending) {
          load_start();
          pending = false;
        }
        if (e.dt > 1e-9) {  // simulate.py:51
          Gen<NC> G;
          TanCo tc[PTA];
          double l1, sc;
          int m, s;
          bool isn;
          build_gen<NC, PT>(prm, e, mc, kind, kidx, G, tc, l1, sc, nan_l1, m, s, isn);
          if (s == 0) {
            taylor<NC, PT>(G, tc, kind, kidx, dwp, m, M, 1.0, dM, M, dM);  // in place
            if (isn) {
#pragma unroll
              for (int i = 0; i < N; i++) {
                M[i] = __builtin_nan("");
#pragma unroll
                for (int p = 0; p < PTA; p++) dM[p][i] = __builtin_nan("");
              }
            }
          } else {
            double R[N], dR[PTA][N];
            if constexpr (STAGE) {
              ColsIn<NC, PT> in{};
              in.G = G;
#pragma unroll
              for (int p = 0; p < PTA; p++) {
                in.tc[p] = tc[p];
                in.kind[p] = kind[p];
                in.kidx[p] = kidx[p];
              }
#pragma unroll
              for (int i = 0; i < (NC > 0 ? NC : 1); i++) in.dwp[i] = NC > 0 ? dwp[i] : 0.0;
              const ColsOut<NC, PT> co = staged_event<NC, PT, TEAM>(in, m, s, isn, lane, false, mir);
#pragma unroll
              for (int i = 0; i < N; i++) {
                R[i] = co.R[i];
#pragma unroll
                for (int p = 0; p < PTA; p++) dR[p][i] = co.dR[p][i];
              }
            } else {
              event_cols<NC, PT, TEAM>(G, tc, kind, kidx, dwp, m, s, isn, lane, R, dR);
            }
            apply_cols<NC, PT, TEAM>(R, dR, M, dM, lane);
          }
        }
        if (e.type == BMC_EV_GRADIENT || e.type == BMC_EV_SPOILER) {  // simulate.py:55-72
#pragma unroll
          for (int i = 0; i < 2 * NT; i++) {
            M[i] = 0.0;
#pragma unroll
            for (int p = 0; p < PTA; p++) dM[p][i] = 0.0;
          }
        }
        if (traj && lane == 0) {  // the trajectory: last_M after every event
          const long base = traj_base + (long)ei * N;
#pragma unroll
          for (int i = 0; i < N; i++) store_real(traj, base + i, M[i], traj_dtype);
        }
        if (e.type == BMC_EV_ADC && reset_after_adc) pending = true;  // simulate.py:105-114
      }
    } else if (KEEP && op.kind == OP_REUSE) {
      // a block equal to the kept one up to an RF phase offset d: M <- R(d) P R(-d) M
      if (pending) {
        load_start();
        pending = false;
      }
      const double c = rot[2 * oi], sn = rot[2 * oi + 1];
      rotate_xy<NC, PTA>(M, dM, c, -sn);
      apply_kept<NC, PT, TEAM>(keep, M, dM, lane);
      rotate_xy<NC, PTA>(M, dM, c, sn);
    } else {
      if (pending) {
        load_start();
        pending = false;
      }
      // the state is not touched while the block is chained: park it in LDS
      double* st = st_park;
      if (lane == 0) {
#pragma unroll
        for (int i = 0; i < N; i++) {
          st[i] = M[i];
#pragma unroll
          for (int p = 0; p < PTA; p++) st[(1 + p) * N + i] = dM[p][i];
        }
      }
      double Pi[N], dPi[PTA][N];
      int ei0 = op.e0;  // the first event chained one by one below
      bool mirrored = false;
      if constexpr (MIR && TEAM > NA) {  // a palindromic opening is built from its first half (mirror_chain)
        if (op.half > 0 && rot && mir && mirror_ok<NC, PT>(prm, kind, kidx)) {
          MirrorIn<NC, PT> in;
#pragma unroll
          for (int p = 0; p < PTA; p++) {
            in.kind[p] = kind[p];
            in.kidx[p] = kidx[p];
          }
#pragma unroll
          for (int i = 0; i < (NC > 0 ? NC : 1); i++) in.dwp[i] = NC > 0 ? dwp[i] : 0.0;
          const MirrorOut<NC, PT> mo =
              mirror_chain<NC, PT, TEAM>(prm, in, evr, op.e0, op.half, rot[2 * oi], rot[2 * oi + 1], nan_l1, lane, mir);
          mirrored = mo.ok;
          if (mirrored) {
            ei0 = op.e0 + 2 * op.half;
#pragma unroll
            for (int i = 0; i < N; i++) {
              Pi[i] = mo.Pi[i];
#pragma unroll
              for (int p = 0; p < PTA; p++) dPi[p][i] = mo.dPi[p][i];
            }
          }
        }
      }
      if (!mirrored) {
#pragma unroll
        for (int i = 0; i < N; i++) {
          Pi[i] = (i == lane) ? 1.0 : 0.0;
#pragma unroll
          for (int p = 0; p < PTA; p++) dPi[p][i] = 0.0;
        }
      }
      Ev e_next = evr[min(ei0, op.e0 + op.len - 1)];
      for (int ei = ei0; ei < op.e0 + op.len; ei++) {
        const Ev e = e_next;
        if (ei + 1 < op.e0 + op.len) e_next = evr[ei + 1];  // the next record is in flight meanwhile
        if (e.dt > 1e-9) {
          Gen<NC> G;
          TanCo tc[PTA];
          double l1, sc;
          int m, s;
          bool isn;
          build_gen<NC, PT>(prm, e, mc, kind, kidx, G, tc, l1, sc, nan_l1, m, s, isn);
          if (s == 0) {
            // exp(X) Pi = T_m(X) Pi: each lane runs the series on its own column of the chained
            // propagator (homogeneous coordinate Pi[N][j] = [j == N]); no dense product is needed
            taylor<NC, PT, true>(G, tc, kind, kidx, dwp, m, Pi, (lane == N) ? 1.0 : 0.0, dPi, Pi, dPi);  // in place
            if (isn) {
#pragma unroll
              for (int i = 0; i < N; i++) {
                Pi[i] = __builtin_nan("");
#pragma unroll
                for (int p = 0; p < PTA; p++) dPi[p][i] = __builtin_nan("");
              }
            }
          } else if constexpr (STAGE) {
            ColsIn<NC, PT> in{};
            in.G = G;
#pragma unroll
            for (int p = 0; p < PTA; p++) {
              in.tc[p] = tc[p];
              in.kind[p] = kind[p];
              in.kidx[p] = kidx[p];
            }
#pragma unroll
            for (int i = 0; i < (NC > 0 ? NC : 1); i++) in.dwp[i] = NC > 0 ? dwp[i] : 0.0;
#pragma unroll
            for (int i = 0; i < N; i++) {
              in.Pi[i] = Pi[i];
#pragma unroll
              for (int p = 0; p < PTA; p++) in.dPi[p][i] = dPi[p][i];
            }
            const ColsOut<NC, PT> co = staged_event<NC, PT, TEAM>(in, m, s, isn, lane, true, mir);
#pragma unroll
            for (int i = 0; i < N; i++) {
              Pi[i] = co.R[i];
#pragma unroll
              for (int p = 0; p < PTA; p++) dPi[p][i] = co.dR[p][i];
            }
          } else {
            double R[N], dR[PTA][N];
            if (event_cols<NC, PT, TEAM>(G, tc, kind, kidx, dwp, m, s, isn, lane, R, dR))
              mul_cols<NC, PT, TEAM, false, true>(R, dR, Pi, dPi, lane);  // a delay: block-diagonal R
            else
              mul_cols<NC, PT, TEAM>(R, dR, Pi, dPi, lane);
          }
        }
        if (e.type == BMC_EV_GRADIENT || e.type == BMC_EV_SPOILER) {
#pragma unroll
          for (int i = 0; i < 2 * NT; i++) {
            Pi[i] = 0.0;
#pragma unroll
            for (int p = 0; p < PTA; p++) dPi[p][i] = 0.0;
          }
        }
      }
      if (KEEP && op.kind == OP_KEEP && lane < NA) {  // each lane keeps its own column (read back by itself)
#pragma unroll
        for (int i = 0; i < N; i++) {
          keep[i * NA + lane] = Pi[i];
#pragma unroll
          for (int p = 0; p < PTA; p++) keep[((1 + p) * N + i) * NA + lane] = dPi[p][i];
        }
      }
#pragma unroll
      for (int i = 0; i < N; i++) {
        M[i] = st[i];
#pragma unroll
        for (int p = 0; p < PTA; p++) dM[p][i] = st[(1 + p) * N + i];
      }
      if (op.count < 4) {
        for (int c = 0; c < op.count; c++) apply_cols<NC, PT, TEAM>(Pi, dPi, M, dM, lane);
      } else {
        // M <- Pi^count M by binary powering: about log2(count) squarings and popcount(count)
        // applications instead of count applications, each of which costs (1 + PT) N team sums
        // (the powers of Pi commute, so the order of the applications does not matter)
        for (int c = op.count;;) {
          if (c & 1) apply_cols<NC, PT, TEAM>(Pi, dPi, M, dM, lane);
          c >>= 1;
          if (c == 0) break;
          mul_cols<NC, PT, TEAM>(Pi, dPi, Pi, dPi, lane);  // Pi <- Pi^2 (the product finishes before it writes)
        }
      }
    }
  }

}

}  // namespace bmc
