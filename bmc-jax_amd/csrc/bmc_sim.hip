// Forward simulation with forward-mode tangents: the event scan of BMC-JAX
// (src/bmc_jax/simulation/simulate.py:22-148) and its Jacobian (jax.jacfwd at fitting/lm.py:39),
// written for CDNA4.
//
// Work mapping. A team of lanes simulates one (voxel, offset row, tangent pass): 8 lanes (two teams per
// DPP row) when the augmented state has at most 8 entries (one CEST pool), 12 lanes (five teams per
// wave, broadcasts by ds_bpermute) up to 12 entries (two pools), 16 (one DPP row) up to 16, 32 beyond.
//  * The generator X = [[A, C], [0, 0]] dt (matrix.py:83-295, expm.py:98-106) is sparse and
//    identical on every lane of a team, so each lane rebuilds it from the parameters and applies it
//    with no communication.
//  * A propagator exp(X) is held by columns: lane j owns column j (rows 0..N-1; the augmented last
//    row is implicit). Column j is the truncated Taylor series T_m(X / 2^s) e_j, with m and s chosen
//    from the 1-norm by the Al-Mohy--Higham backward-error bound for float64, followed by s
//    squarings. Dense products (squarings, chaining a pulse) broadcast column k of the left factor
//    from lane k with DPP row_newbcast and multiply it into the local column of the right factor.
//  * The state M and its tangents are identical on every lane. An event applied on its own
//    (s == 0) acts on the state directly with the Taylor series; otherwise its propagator is built
//    by columns and applied with a team reduction.
//  * Runs of identical events (a repeated pulse and its gap) are chained into one propagator and
//    applied `count` times (the row program built on the host). A pulse that recurs later in the row
//    with every sample's RF phase shifted by the same d (the phase accumulated by a scanner's pulse
//    train) is chained once, kept in LDS, and applied as R(d) P R(-d), R a rotation about z. An event that needs no squaring is
//    chained by running its Taylor series on each lane's column of the chained propagator, which
//    needs no dense product at all; only events with squarings form their own propagator first.
// Tangents travel beside every quantity (dual numbers): d(X v) = dX v + X dv, where dX is the
// analytic derivative of the generator w.r.t. one packed parameter slot.
// All arithmetic is float64 (CDNA4 issues FP64 FMA at the FP32 rate).

#include <algorithm>

#include "bmc_device.h"

namespace bmc {

// Waves per SIMD the register allocation is held to (2 halves the registers per lane).
#ifndef BMC_SIM_WAVES_PER_SIMD
#define BMC_SIM_WAVES_PER_SIMD 2
#endif
// Most tangents one pass carries (the launch switch below instantiates 0..this many).
#ifndef BMC_TANGENTS_PER_PASS
#define BMC_TANGENTS_PER_PASS 2
#endif

// Doubles of one team's keep slot: the propagator and its PT tangents, N rows by N + 1 columns.
template <int NC, int PT>
__host__ __device__ constexpr long keep_slot_doubles() {
  return (long)(1 + (PT > 0 ? PT : 1)) * (3 * (1 + NC) + 1) * (3 * (1 + NC) + 2);
}

// Whether a kernel runs its squarings' products through the team's LDS stage (staged_event): 12-lane
// teams with at most one tangent per pass.
template <int TEAM, int PT>
__host__ __device__ constexpr bool staged_products() {
  return TEAM == 12 && PT <= 1;
}
// Doubles of one team's LDS stage: (N + 1) N in the kernels that factorise palindromes (MIR, for
// mirror_chain) and in those with staged products (all plans then run the same products, bit for bit),
// 1 (unused) in the others.
template <int NC, int PT, int TEAM, bool MIR>
__host__ __device__ constexpr int stage_doubles() {
  return (MIR || staged_products<TEAM, PT>()) ? (3 * (1 + NC) + 2) * (3 * (1 + NC) + 1) : 1;
}

// One team's (voxel, offset row, pass) of a launch with `total` teams over n_vox voxels (voxel
// index through a.vlist when a compacted list is given). Each pass holds pass_teams(n_vox O) teams,
// the last few of them idle, so that no wave spans two passes. Every thread of the block calls it: it
// holds block barriers before its first early exit.
template <int NC, int PT, int TEAM, bool LISTED, bool KEEP, bool MIR>
__device__ __forceinline__ void sim_team(const SimArgs& a, long team, long total, int n_vox,
                                         double (*sh_prm)[PrmIdx<NC>::size],
                                         double (*sh_state)[(1 + (PT > 0 ? PT : 1)) * (3 * (1 + NC) + 1)],
                                         double (*sh_start)[(1 + (PT > 0 ? PT : 1)) * (3 * (1 + NC) + 1)],
                                         int* sh_tk, int* sh_ti, double* keep, double* mir) {
  constexpr int NT = 1 + NC, N = 3 * NT + 1;
  using I = PrmIdx<NC>;
  constexpr int PTA = PT > 0 ? PT : 1;
  const int lane = team_lane<TEAM>();
  const int tib = team_index<TEAM>();
  const long rows = (long)n_vox * a.O;
  const long per_pass = pass_teams<TEAM>(rows);
  bool dead = team >= total;
  const long tt = dead ? 0 : team;
  const int q = (int)(tt / per_pass);
  long r = tt % per_pass;
  if (r >= rows) {  // the padding that rounds a pass up to whole waves
    dead = true;
    r = 0;
  }
  const int o = (int)(r % a.O);
  const int vi = (int)(r / a.O);
  const int v = LISTED ? a.vlist[vi] : vi;
  if (!dead && a.active && !a.active[v]) dead = true;
  // a block with no live team leaves before touching its records (stopped LM problems)
  if (a.active && !__syncthreads_or(!dead)) return;

  if (threadIdx.x < kMaxTan) {
    sh_tk[threadIdx.x] = threadIdx.x < a.P ? a.t_kind[threadIdx.x] : TK_NONE;
    sh_ti[threadIdx.x] = threadIdx.x < a.P ? a.t_idx[threadIdx.x] : 0;
  }
  // parameters -> LDS (written slots applied, as write_params does: fitting/params.py:100-112)
  double* prm = sh_prm[tib];
  for (int i = lane; i < I::stride; i += TEAM) {
    double val = load_real(a.params, (long)v * a.stride + i, a.dtype);
    for (int w = 0; w < a.n_write; w++)
      if (a.w_slot[w] == i && (!a.w_keep || a.w_keep[(long)v * a.n_write + w] != 0.0)) {
        double xv = a.xw[(long)v * a.n_write + w] * a.w_scale[w];
        val = (a.dtype == BMC_F32) ? (double)(float)xv : xv;
      }
    prm[i] = val;
  }
  // the start state and its seeds -> LDS (reloaded at every ADC reset)
  double* st0 = sh_start[tib];
  for (int i = lane; i < (1 + PTA) * N; i += TEAM) {
    const int p = i / N - 1, r = i % N;
    const int gp = a.t_off + q * PT + p;
    double val = 0.0;
    if (p < 0)
      val = load_real(a.m0, (long)v * N + r, a.dtype);
    else if (p < PT && gp < a.P && a.dm0)
      val = load_real(a.dm0, ((long)v * a.P + gp) * N + r, a.dtype);
    st0[i] = val;
  }
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
  {
    // a wave whose teams are all dead leaves together (a team never splits across waves)
    unsigned long long live = __ballot(!dead);
    if (live == 0) return;
  }

  // tangents of this pass: the same on every lane of the wave (no wave spans two passes), so they are
  // read into scalar registers and the event setup branches on them as a wave
  int kind[PTA], kidx[PTA];
#pragma unroll
  for (int p = 0; p < PTA; p++) {
    const int gp = a.t_off + q * PT + p;
    kind[p] = __builtin_amdgcn_readfirstlane((p < PT && gp < a.P) ? sh_tk[gp] : TK_NONE);
    kidx[p] = __builtin_amdgcn_readfirstlane((p < PT && gp < a.P) ? sh_ti[gp] : 0);
  }
  double dwp[NC > 0 ? NC : 1];
#pragma unroll
  for (int i = 0; i < NC; i++) dwp[i] = prm[13 + 2 * NC + i];

  // state (identical on every lane of the team); the start state is reloaded from memory on a reset
  double M[N];
  double dM[PTA][N];
  run_row<NC, PT, TEAM, KEEP, MIR, staged_products<TEAM, PT>()>(prm, st0, sh_state[tib], kind, kidx, dwp, a.ev + (long)o * a.T, a.ops, a.row_ptr[o],
                        a.row_ptr[o + 1], a.reset_after_adc, a.nan_l1, lane, M, dM,
                        (a.traj && !dead && q == 0) ? a.traj : nullptr, ((long)v * a.O + o) * a.T * N, a.dtype,
                        a.rot, keep, mir);
  if (dead || lane != 0) return;
  const long vo = (long)v * a.O + o;
  if (q == 0 && a.t_off == 0) {
    if (a.m_out)
      for (int i = 0; i < N; i++) store_real(a.m_out, vo * N + i, M[i], a.dtype);
    if (a.z_out) a.z_out[vo] = M[a.jc];
  }
  if (a.jac_out) {
#pragma unroll
    for (int p = 0; p < PTA; p++) {
      const int gp = a.t_off + q * PT + p;
      if (p < PT && gp < a.P) {
        double val = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++)
          if (i == a.jc) val = dM[p][i];
        store_real(a.jac_out, vo * a.P + gp, val * a.t_scale[gp], a.jac_dtype);
      }
    }
  }
}

template <int NC, int PT, int TEAM, bool KEEP, bool MIR>
__global__ void __launch_bounds__(256, BMC_SIM_WAVES_PER_SIMD) sim_kernel(SimArgs a) {
  constexpr int N = 3 * (1 + NC) + 1;
  constexpr int TPB = teams_per_block<TEAM>();
  constexpr int PTA = PT > 0 ? PT : 1;
  __shared__ double sh_prm[TPB][PrmIdx<NC>::size];
  __shared__ double sh_state[TPB][(1 + PTA) * N];  // the state, parked while a block is chained
  __shared__ double sh_start[TPB][(1 + PTA) * N];  // the start state and its tangent seeds
  __shared__ int sh_tk[kMaxTan], sh_ti[kMaxTan];
  extern __shared__ double dyn_keep[];  // TPB keep slots when the plan holds KEEP ops
  __shared__ __attribute__((aligned(16))) double sh_mir[TPB][stage_doubles<NC, PT, TEAM, MIR>()];
  const int tib = team_index<TEAM>();
  double* keep = KEEP ? dyn_keep + (long)tib * keep_slot_doubles<NC, PT>() : nullptr;
  double* mir = (MIR || staged_products<TEAM, PT>()) ? sh_mir[tib] : nullptr;
  sim_team<NC, PT, TEAM, false, KEEP, MIR>(a, a.team0 + (long)blockIdx.x * TPB + tib, pass_teams<TEAM>((long)a.V * a.O) * a.n_pass, a.V, sh_prm,
                         sh_state, sh_start, sh_tk, sh_ti, keep, mir);
}

// The same over a compacted voxel list (a.vlist, device count a.n_vlist): a fixed grid strides over the
// live teams; the trip count is block-uniform.
template <int NC, int PT, int TEAM, bool KEEP, bool MIR>
__global__ void __launch_bounds__(256, BMC_SIM_WAVES_PER_SIMD) sim_list_kernel(SimArgs a) {
  constexpr int N = 3 * (1 + NC) + 1;
  constexpr int TPB = teams_per_block<TEAM>();
  constexpr int PTA = PT > 0 ? PT : 1;
  __shared__ double sh_prm[TPB][PrmIdx<NC>::size];
  __shared__ double sh_state[TPB][(1 + PTA) * N];
  __shared__ double sh_start[TPB][(1 + PTA) * N];
  __shared__ int sh_tk[kMaxTan], sh_ti[kMaxTan];
  extern __shared__ double dyn_keep[];
  __shared__ __attribute__((aligned(16))) double sh_mir[TPB][stage_doubles<NC, PT, TEAM, MIR>()];
  const int tib = team_index<TEAM>();
  double* keep = KEEP ? dyn_keep + (long)tib * keep_slot_doubles<NC, PT>() : nullptr;
  double* mir = (MIR || staged_products<TEAM, PT>()) ? sh_mir[tib] : nullptr;
  const int nv = *a.n_vlist;
  const long total = pass_teams<TEAM>((long)nv * a.O) * a.n_pass;
  for (long base = (long)blockIdx.x * TPB; base < total; base += (long)gridDim.x * TPB) {
    sim_team<NC, PT, TEAM, true, KEEP, MIR>(a, base + tib, total, nv, sh_prm, sh_state, sh_start, sh_tk, sh_ti, keep, mir);
    __syncthreads();  // the next item overwrites the block's LDS
  }
}

template <int NC, int PT, bool KEEP, bool MIR>
static int launch_nc_pt_k(const SimArgs& a, hipStream_t stream) {
  constexpr int NA = 3 * (1 + NC) + 2;
  constexpr int TEAM = NA <= 8 ? 8 : (NA <= 12 ? 12 : (NA <= 16 ? 16 : 32));  // one lane per propagator column
  constexpr int TPB = teams_per_block<TEAM>();
  const long teams = pass_teams<TEAM>((long)a.V * a.O) * a.n_pass;
  const size_t dyn = KEEP ? (size_t)TPB * keep_slot_doubles<NC, PT>() * sizeof(double) : 0;
  if (dyn > 65536) {  // gfx950 gives a workgroup up to 160 KiB of LDS; dynamic LDS above 64 KiB is opt-in
    BMC_HIP_CHECK(hipFuncSetAttribute((const void*)sim_kernel<NC, PT, TEAM, KEEP, MIR>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
    BMC_HIP_CHECK(hipFuncSetAttribute((const void*)sim_list_kernel<NC, PT, TEAM, KEEP, MIR>,
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
  }
  if (a.vlist) {  // persistent grid over a device-side voxel list of at most V entries
    const long blocks = std::min<long>((teams + TPB - 1) / TPB, 2048);
    if (blocks <= 0) return BMC_OK;
    hipLaunchKernelGGL((sim_list_kernel<NC, PT, TEAM, KEEP, MIR>), dim3((unsigned)blocks), dim3(256), dyn, stream, a);
    BMC_HIP_CHECK(hipGetLastError());
    return BMC_OK;
  }
  // HIP caps a launch at 2^32 - 1 work-items: split the teams into launches of at most 2^23 blocks
  constexpr long kMaxBlocks = 1L << 23;
  for (long t0 = 0; t0 < teams; t0 += kMaxBlocks * TPB) {
    const long blocks = std::min(kMaxBlocks, (teams - t0 + TPB - 1) / TPB);
    SimArgs c = a;
    c.team0 = t0;
    hipLaunchKernelGGL((sim_kernel<NC, PT, TEAM, KEEP, MIR>), dim3((unsigned)blocks), dim3(256), dyn, stream, c);
    BMC_HIP_CHECK(hipGetLastError());
  }
  return BMC_OK;
}

// Plans with kept pulses run the kernels that hold the kept-pulse code and its LDS slots; plans with
// palindromic runs (and no kept pulses) the kernels that factorise them, when a team has a spare lane.
// Each variant is compiled apart: the code of one raised the register pressure of the others' event
// loops (the plain chaining path ran 1.37x slower with the factorisation compiled in).
template <int NC, int PT>
static int launch_nc_pt(const SimArgs& a, hipStream_t stream) {
  constexpr bool kSpare = (3 * (1 + NC) + 2) < (3 * (1 + NC) + 2 <= 8 ? 8 : 3 * (1 + NC) + 2 <= 12 ? 12 : 3 * (1 + NC) + 2 <= 16 ? 16 : 32);
  if (a.keep) return launch_nc_pt_k<NC, PT, true, false>(a, stream);
  if constexpr (kSpare)
    if (a.mirror) return launch_nc_pt_k<NC, PT, false, true>(a, stream);
  return launch_nc_pt_k<NC, PT, false, false>(a, stream);
}

template <int NC>
static int launch_nc(const SimArgs& a, hipStream_t stream) {
  switch (a.PT) {
    case 0:
      return launch_nc_pt<NC, 0>(a, stream);
    case 1:
      return launch_nc_pt<NC, 1>(a, stream);
    case 2:
      return launch_nc_pt<NC, 2>(a, stream);
#if BMC_TANGENTS_PER_PASS > 2
    case 3:
      return launch_nc_pt<NC, 3>(a, stream);
#endif
#if BMC_TANGENTS_PER_PASS > 3
    case 4:
      return launch_nc_pt<NC, 4>(a, stream);
#endif
#if BMC_TANGENTS_PER_PASS > 4
    case 5:
      return launch_nc_pt<NC, 5>(a, stream);
#endif
    default:
      set_error("simulate: tangents per pass must be 0.." + std::to_string(BMC_TANGENTS_PER_PASS));
      return BMC_EINVAL;
  }
}

static int launch_sim_nc(int nc, const SimArgs& a, hipStream_t stream) {
  switch (nc) {
    case 0:
      return launch_nc<0>(a, stream);
    case 1:
      return launch_nc<1>(a, stream);
    case 2:
      return launch_nc<2>(a, stream);
    case 3:
      return launch_nc<3>(a, stream);
    case 4:
      return launch_nc<4>(a, stream);
    default:
      set_error("simulate: n_cest must be 0.." + std::to_string(BMC_MAX_CEST));
      return BMC_ENOTSUP;
  }
}

// P tangents in passes of PT: when the last pass would carry fewer than PT, it runs as its own launch
// with that many tangents per pass, so no lane works on an empty tangent slot (5 tangents: 2 + 2 + 1).
// Only launches whose one pass fills the GPU many times over are split: a small launch (a few tubes)
// is latency-bound, and two launches in a row would take about twice as long as one.
constexpr long kSplitMinTeams = 1L << 17;

int launch_sim(int nc, const SimArgs& a, hipStream_t stream) {
  const int P = a.jac_out ? a.P : 0;
  if (a.PT <= 1 || a.n_pass <= 1 || a.t_off != 0 || P % a.PT == 0 || a.traj || (long)a.V * a.O < kSplitMinTeams)
    return launch_sim_nc(nc, a, stream);
  SimArgs head = a;
  head.n_pass = a.n_pass - 1;
  int rc = launch_sim_nc(nc, head, stream);
  if (rc != BMC_OK) return rc;
  SimArgs tail = a;
  tail.t_off = (a.n_pass - 1) * a.PT;
  tail.PT = P - tail.t_off;
  tail.n_pass = 1;
  return launch_sim_nc(nc, tail, stream);
}

int tan_kind_of_slot(int slot, int nc, int* idx) {
  *idx = 0;
  switch (slot) {
    case 0: return TK_W_R1;
    case 1: return TK_W_R2;
    case 2: return TK_NONE;  // water_f enters only M_init
    case 3: return TK_MT_K;
    case 4: return TK_MT_F;
    case 5: return TK_MT_DW;
    case 6: return TK_MT_R1;
    case 7: return TK_MT_R2;
    case 8: return TK_NONE;  // mt_lineshape is discrete
    case 9: return TK_B0;
    case 10: return TK_GAMMA;
    case 11: return TK_B0INH;
    case 12: return TK_RELB1;
    default: break;
  }
  if (slot < 0) return TK_NONE;
  int r = slot - 13;
  if (r < 0 || r >= 5 * nc) return -1;
  *idx = r % nc;
  return TK_C_K + r / nc;
}

}  // namespace bmc
