// The Jacobian of single-event rows (a CW saturation block and its readout, fitting/lm.py:39 on
// sequences/saturation.py:28-84 rows) by accumulating the squaring chain in reverse.
//
// A row whose only propagating event is E = exp(X) = T^n (T = T_m(X / n) the Taylor polynomial,
// n = 2^s) gives Z = e_z^T E v, v = [M0; 1]. Forward mode (bmc_sim.hip) carries dE_p through every
// squaring, two extra products per tangent and level. Here the derivative is contracted instead:
//   dZ/dp = e_z^T dE v = sum_j e_z^T T^j dT_p T^(n-1-j) v = <dT_p, G_n>,
//   G_n = sum_j (T^T)^j e_z v^T (T^T)^(n-1-j),   G_1 = e_z v^T,   G_2n = G_n E^T + E^T G_n (E = T^n),
// a recursion that runs upwards beside the squarings, so no level needs to be stored. Each level costs
// the squaring plus two products for G (and one for the rows of E that G E^T needs), whatever the
// number of tangents; the tangents enter only through dT_p of the Taylor polynomial (forward mode,
// Horner, a chunk of tangents at a time) and one team reduction each.
//
// Work mapping as in the forward kernel: one team of lanes per (voxel, row); lane j holds column j of
// E, row j of E and column j of G (rows < N; the augmented row of G never reaches dT, whose own
// augmented row is zero). The Taylor degree is capped at 18 instead of the forward kernel's 30 (more
// squarings, shorter series: cheaper once the squarings no longer carry tangents); the state stays
// within 1e-12 of the forward kernel's.
#include <math.h>
#include <stdlib.h>

#include <algorithm>

#include "bmc_cwj.h"

namespace bmc {

#ifndef BMC_CWJ_WAVES_PER_SIMD
#define BMC_CWJ_WAVES_PER_SIMD 2
#endif

// One team's (voxel, row). Every thread of the block calls it (block barriers before the first exit).
template <int NC, int TEAM, bool LISTED>
__device__ __forceinline__ void cwj_team(const SimArgs& a, const int32_t* row_live, long team, long total, int n_vox,
                                         double (*sh_prm)[PrmIdx<NC>::size], double (*sh_m0)[3 * (1 + NC) + 1],
                                         int* sh_tk, int* sh_ti) {
  constexpr int NT = 1 + NC, N = 3 * NT + 1, NA = N + 1;
  using I = PrmIdx<NC>;
  const int lane = threadIdx.x % TEAM;
  const int tib = threadIdx.x / TEAM;
  bool dead = team >= total;
  const long tt = dead ? 0 : team;
  const int o = (int)(tt % a.O);
  const int vi = (int)(tt / a.O);
  const int v = LISTED ? a.vlist[vi] : vi;
  if (!dead && a.active && !a.active[v]) dead = true;
  if (a.active && !__syncthreads_or(!dead)) return;

  if (threadIdx.x < kMaxTan) {
    sh_tk[threadIdx.x] = threadIdx.x < a.P ? a.t_kind[threadIdx.x] : TK_NONE;
    sh_ti[threadIdx.x] = threadIdx.x < a.P ? a.t_idx[threadIdx.x] : 0;
  }
  double* prm = sh_prm[tib];
  for (int i = lane; i < I::stride; i += TEAM) {
    double val = load_real(a.params, (long)v * a.stride + i, a.dtype);
    for (int w = 0; w < a.n_write; w++)
      if (a.w_slot[w] == i && (!a.w_keep || a.w_keep[(long)v * a.n_write + w] != 0.0)) {  // write_params (fitting/params.py:100-112)
        double xv = a.xw[(long)v * a.n_write + w] * a.w_scale[w];
        val = (a.dtype == BMC_F32) ? (double)(float)xv : xv;
      }
    prm[i] = val;
  }
  for (int i = lane; i < N; i += TEAM) sh_m0[tib][i] = load_real(a.m0, (long)v * N + i, a.dtype);
  __syncthreads();
  if (lane == 0) {
    double L = prm[10] * 1e6 * prm[9];
    double ks = 0.0;
    for (int i = 0; i < NC; i++) ks += prm[13 + i] * prm[13 + NC + i];
    prm[I::L] = L;
    prm[I::KS] = ks;
    prm[I::KWMT] = prm[3] * prm[4];
    prm[I::DW0] = L * prm[11] * 1e-6;
  }
  __syncthreads();
  if (__ballot(!dead) == 0) return;  // a team never splits across waves

  double dwp[NC > 0 ? NC : 1];
#pragma unroll
  for (int i = 0; i < NC; i++) dwp[i] = prm[13 + 2 * NC + i];
  const int le = row_live[o];
  const long vo = (long)v * a.O + o;
  cwj_row<NC, TEAM>(
      prm, sh_m0[tib], dwp, le >= 0 ? a.ev + (long)o * a.T + le : nullptr, a.P,
      [&](int p, int& k, int& ix) {
        k = sh_tk[p];
        ix = sh_ti[p];
      },
      a.jc, a.nan_l1, lane, a.jac_out != nullptr,
      [&](const double* M, bool isn) {
        if (dead) return;
        if (a.m_out)
          for (int i = 0; i < N; i++) store_real(a.m_out, vo * N + i, isn ? __builtin_nan("") : M[i], a.dtype);
        if (a.z_out) {
          double z = 0.0;
#pragma unroll
          for (int i = 0; i < N; i++)
            if (i == a.jc) z = M[i];
          a.z_out[vo] = isn ? __builtin_nan("") : z;
        }
      },
      [&](int gp, double val, bool isn) {
        if (!dead) store_real(a.jac_out, vo * a.P + gp, isn ? __builtin_nan("") : val * a.t_scale[gp], a.jac_dtype);
      });
}

template <int NC, int TEAM>
__global__ void __launch_bounds__(256, BMC_CWJ_WAVES_PER_SIMD) cwj_kernel(SimArgs a, const int32_t* row_live) {
  constexpr int N = 3 * (1 + NC) + 1;
  constexpr int TPB = 256 / TEAM;
  __shared__ double sh_prm[TPB][PrmIdx<NC>::size];
  __shared__ double sh_m0[TPB][N];
  __shared__ int sh_tk[kMaxTan], sh_ti[kMaxTan];
  const int tib = threadIdx.x / TEAM;
  cwj_team<NC, TEAM, false>(a, row_live, a.team0 + (long)blockIdx.x * TPB + tib, (long)a.V * a.O, a.V, sh_prm,
                            sh_m0, sh_tk, sh_ti);
}

// The same over a compacted voxel list: a fixed grid strides over the live (voxel, row) teams.
template <int NC, int TEAM>
__global__ void __launch_bounds__(256, BMC_CWJ_WAVES_PER_SIMD) cwj_list_kernel(SimArgs a, const int32_t* row_live) {
  constexpr int N = 3 * (1 + NC) + 1;
  constexpr int TPB = 256 / TEAM;
  __shared__ double sh_prm[TPB][PrmIdx<NC>::size];
  __shared__ double sh_m0[TPB][N];
  __shared__ int sh_tk[kMaxTan], sh_ti[kMaxTan];
  const int tib = threadIdx.x / TEAM;
  const int nv = *a.n_vlist;
  const long total = (long)nv * a.O;
  for (long base = (long)blockIdx.x * TPB; base < total; base += (long)gridDim.x * TPB) {
    cwj_team<NC, TEAM, true>(a, row_live, base + tib, total, nv, sh_prm, sh_m0, sh_tk, sh_ti);
    __syncthreads();
  }
}

template <int NC>
static int launch_cwj_nc(const SimArgs& a, const int32_t* row_live, hipStream_t stream) {
  constexpr int NA = 3 * (1 + NC) + 2;
  // four CEST pools (NA = 17) run on 16-lane teams with the augmented column spread (cwj_row)
  constexpr int TEAM = NA <= 8 ? 8 : (NA <= 17 ? 16 : 32);
  constexpr int TPB = 256 / TEAM;
  const long teams = (long)a.V * a.O;
  if (a.vlist) {
    const long blocks = std::min<long>((teams + TPB - 1) / TPB, 2048);
    if (blocks <= 0) return BMC_OK;
    hipLaunchKernelGGL((cwj_list_kernel<NC, TEAM>), dim3((unsigned)blocks), dim3(256), 0, stream, a, row_live);
    BMC_HIP_CHECK(hipGetLastError());
    return BMC_OK;
  }
  constexpr long kMaxBlocks = 1L << 23;
  for (long t0 = 0; t0 < teams; t0 += kMaxBlocks * TPB) {
    const long blocks = std::min(kMaxBlocks, (teams - t0 + TPB - 1) / TPB);
    SimArgs c = a;
    c.team0 = t0;
    hipLaunchKernelGGL((cwj_kernel<NC, TEAM>), dim3((unsigned)blocks), dim3(256), 0, stream, c, row_live);
    BMC_HIP_CHECK(hipGetLastError());
  }
  return BMC_OK;
}

bool use_cwj(const Plan& p, const SimArgs& a, bool forward_mode) {
  if (forward_mode || !p.cw_rows || !p.d_live || a.dm0 || a.traj || a.P <= 0 || !a.jac_out) return false;
  return true;
}

int launch_cwj(int nc, const SimArgs& a, const int32_t* row_live, hipStream_t stream) {
  switch (nc) {
    case 0: return launch_cwj_nc<0>(a, row_live, stream);
    case 1: return launch_cwj_nc<1>(a, row_live, stream);
    case 2: return launch_cwj_nc<2>(a, row_live, stream);
    case 3: return launch_cwj_nc<3>(a, row_live, stream);
    case 4: return launch_cwj_nc<4>(a, row_live, stream);
    default:
      set_error("simulate: n_cest must be 0.." + std::to_string(BMC_MAX_CEST));
      return BMC_ENOTSUP;
  }
}

}  // namespace bmc
