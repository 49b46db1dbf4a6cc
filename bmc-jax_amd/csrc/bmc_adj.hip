// Reverse mode of the event scan and of its forward tangents with respect to every event's RF
// amplitude and RF frequency: the gradient that jax.value_and_grad takes through jacfwd of the scan
// in the reference's protocol design (src/bmc_jax/design/optimize.py:80-93, simulation/simulate.py:22-132).
//
// Given cotangents gz on Z = state[jc] and gj on J_p = dstate[jc]/dx_p * scale_p for every (voxel,
// row), the kernel accumulates dL/d rf_amp[o, t] and dL/d rf_freq[o, t], summed over voxels.
//
// Work mapping. A team of lanes (8 for at most one CEST pool, else 16) owns one (row, voxel), voxel
// fastest, so the teams of a wave usually share a row and their per-event gradients are summed in
// registers before one atomic add.
//  * Phase 1 runs the scan forward (state M and its P tangents D_p, all on every lane of the team, as
//    in the simulate kernel) and stores a checkpoint every K = ceil(sqrt(T)) events in the workspace.
//  * Phase 2 walks the segments backwards. It re-runs a segment forward from its checkpoint, storing
//    the state before each event, then walks the segment's events in reverse with the adjoint
//    (lambda for M, mu_p for D_p). For an event with propagator E (and tangents E_p = dE/dx_p):
//        grad_q += lambda^T E_q M + sum_p mu_p^T (E_q D_p + E_pq M)
//        lambda <- E^T lambda + sum_p E_p^T mu_p,    mu_p <- E^T mu_p
//    where q is the event's amplitude or frequency. Lane j builds column j of E, E_p, E_q, E_pq by
//    the same truncated Taylor series and squarings as the forward kernel, carried to second order
//    (the derivative of exactly the forward algorithm at its chosen degree and squaring count).
//    Spoilers zero the transverse adjoint entries; an ADC reset cuts the dependence on earlier
//    events, whose gradients are then zero.
// The tangents must have generator derivatives that do not depend on the RF (relaxation rates,
// exchange rates, fractions, CEST shifts); the mixed second derivative of the generator is then zero.

#include <algorithm>
#include <cmath>
#include <cstring>

#include "bmc_device.h"

namespace bmc {

int fill_sim_args(const bmc_plan* plan, int V, int nc, int dtype, const void* params, const void* m0,
                  const void* dm0, int P, const int32_t* slot, const double* scale, int jc, SimArgs& a);

struct AdjArgs {
  const Ev* ev;  // [O, T]
  int O, T, reset_after_adc;
  int V, dtype, stride;
  const void* params;  // [V, stride]
  const void* m0;      // [V, N]
  const void* dm0;     // [V, p_stride, N] or null
  int P;               // tangents of this launch: entries p_off .. p_off + P - 1 of the caller's p_stride
  int p_stride, p_off;
  int t_kind[kMaxTan], t_idx[kMaxTan];
  double t_scale[kMaxTan];
  int jc;
  const double* gz;  // [V, O] or null
  const double* gj;  // [V, O, P] or null
  double* grad_amp;  // [O, T] accumulated, or null
  double* grad_freq; // [O, T] accumulated, or null
  double* ws;        // workspace, (n_seg + K) records per team
  int K, n_seg;
  long team0;
  double nan_l1;
};

// Columns of T_m(X) and its derivatives in the P tangent directions, in the direction q, and the
// P mixed second derivatives (lane j: column j; the augmented coordinate as in `taylor`).
template <int NC, int P>
__device__ __forceinline__ void taylor2(const Gen<NC>& G, const TanCo* tc, const int* kind, const int* kidx,
                                        const double* dwp, const TanCo& tq, int qkind, int m, int j, double* R,
                                        double (*Rp)[3 * (1 + NC) + 1], double* Rq, double (*Rpq)[3 * (1 + NC) + 1]) {
  constexpr int N = 3 * (1 + NC) + 1;
  constexpr int PA = P > 0 ? P : 1;
  double t[N], tp[PA][N], tqv[N], tpq[PA][N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    t[i] = (i == j) ? 1.0 : 0.0;
    R[i] = t[i];
    tqv[i] = 0.0;
    Rq[i] = 0.0;
#pragma unroll
    for (int p = 0; p < PA; p++) {
      tp[p][i] = 0.0;
      Rp[p][i] = 0.0;
      tpq[p][i] = 0.0;
      Rpq[p][i] = 0.0;
    }
  }
  double taug = (j == N) ? 1.0 : 0.0;
  for (int k = 1; k <= m; k++) {
    const double ik = kInv[k];
    double y[N];
    // mixed terms first: they read the previous first-order terms
#pragma unroll
    for (int p = 0; p < P; p++) {
      apply_X<NC>(G, tpq[p], 0.0, y);
      apply_dX<NC>(kind[p], kidx[p], tc[p], dwp, tqv, 0.0, y);
      apply_dX<NC>(qkind, 0, tq, dwp, tp[p], 0.0, y);
#pragma unroll
      for (int i = 0; i < N; i++) {
        tpq[p][i] = y[i] * ik;
        Rpq[p][i] += tpq[p][i];
      }
    }
    apply_X<NC>(G, tqv, 0.0, y);
    apply_dX<NC>(qkind, 0, tq, dwp, t, taug, y);
#pragma unroll
    for (int i = 0; i < N; i++) {
      tqv[i] = y[i] * ik;
      Rq[i] += tqv[i];
    }
#pragma unroll
    for (int p = 0; p < P; p++) {
      apply_X<NC>(G, tp[p], 0.0, y);
      apply_dX<NC>(kind[p], kidx[p], tc[p], dwp, t, taug, y);
#pragma unroll
      for (int i = 0; i < N; i++) {
        tp[p][i] = y[i] * ik;
        Rp[p][i] += tp[p][i];
      }
    }
    apply_X<NC>(G, t, taug, y);
#pragma unroll
    for (int i = 0; i < N; i++) {
      t[i] = y[i] * ik;
      R[i] += t[i];
    }
    taug = 0.0;
  }
}

// One squaring of (R, R_p, R_q, R_pq) in column layout, to second order by the product rule.
template <int NC, int P, int TEAM>
__device__ __forceinline__ void square2(double* R, double (*Rp)[3 * (1 + NC) + 1], double* Rq,
                                        double (*Rpq)[3 * (1 + NC) + 1], int j) {
  constexpr int N = 3 * (1 + NC) + 1;
  constexpr int PA = P > 0 ? P : 1;
  const bool aff = (j == N);
  double oR[N], oRp[PA][N], oRq[N], oRpq[PA][N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    oR[i] = aff ? R[i] : 0.0;
    oRq[i] = aff ? Rq[i] : 0.0;
#pragma unroll
    for (int p = 0; p < PA; p++) {
      oRp[p][i] = aff ? Rp[p][i] : 0.0;
      oRpq[p][i] = aff ? Rpq[p][i] : 0.0;
    }
  }
  static_for<0, N>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const double bk = R[k], bqk = Rq[k];
    double bpk[PA], bpqk[PA];
#pragma unroll
    for (int p = 0; p < PA; p++) {
      bpk[p] = Rp[p][k];
      bpqk[p] = Rpq[p][k];
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
      const double a = bcast<TEAM, k>(R[i]);
      oR[i] = fma(a, bk, oR[i]);
      oRq[i] = fma(a, bqk, oRq[i]);
      const double aq = bcast<TEAM, k>(Rq[i]);
      oRq[i] = fma(aq, bk, oRq[i]);
#pragma unroll
      for (int p = 0; p < P; p++) {
        oRp[p][i] = fma(a, bpk[p], oRp[p][i]);
        oRpq[p][i] = fma(a, bpqk[p], fma(aq, bpk[p], oRpq[p][i]));
        const double ap = bcast<TEAM, k>(Rp[p][i]);
        oRp[p][i] = fma(ap, bk, oRp[p][i]);
        oRpq[p][i] = fma(ap, bqk, oRpq[p][i]);
        oRpq[p][i] = fma(bcast<TEAM, k>(Rpq[p][i]), bk, oRpq[p][i]);
      }
    }
  });
#pragma unroll
  for (int i = 0; i < N; i++) {
    R[i] = oR[i];
    Rq[i] = oRq[i];
#pragma unroll
    for (int p = 0; p < PA; p++) {
      Rp[p][i] = oRp[p][i];
      Rpq[p][i] = oRpq[p][i];
    }
  }
}

// Coefficients of dX/dq for the event's RF amplitude (q = 0, the rel_b1 pattern of apply_dX) or RF
// frequency (q = 1, the B0-shift pattern), times dt 2^-s.
template <int NC>
__device__ __forceinline__ TanCo rf_tangent(const double* prm, const Ev& e, const MtCache& mc, double sc, int q,
                                            int& qkind) {
  const double fmt = prm[4];
  const int ls = (int)prm[8];
  const double relb1 = prm[12];
  const double twopi = 6.283185307179586;
  const double w1 = twopi * e.amp * relb1;
  const bool mt_on = fmt > 1e-9 && ls != 0 && w1 != 0.0;  // build_gen fills the cache only then
  TanCo t{0.0, 0.0, 0.0};
  if (q == 0) {
    qkind = TK_RELB1;
    const double dw1 = twopi * relb1;  // d w1 / d amp
    t.c0 = dw1 * e.cph * sc;
    t.c1 = dw1 * e.sph * sc;
    if (mt_on) t.c2 = -(2.0 * w1 * dw1 * mc.s.g) * sc;
  } else {
    qkind = TK_B0INH;
    t.c0 = twopi * sc;  // d eff / d freq
    if (mt_on) t.c2 = -(w1 * w1 * mc.s.g_off * twopi) * sc;
  }
  return t;
}

#ifndef BMC_ADJ_WAVES_PER_SIMD
#define BMC_ADJ_WAVES_PER_SIMD 1
#endif

template <int NC, int P, int TEAM>
__global__ void __launch_bounds__(256, BMC_ADJ_WAVES_PER_SIMD) adj_kernel(AdjArgs a) {
  constexpr int NT = 1 + NC, N = 3 * NT + 1;
  constexpr int TPB = 256 / TEAM;
  constexpr int PA = P > 0 ? P : 1;
  constexpr int REC = (1 + P) * N + 1;  // state, tangents, flag
  using I = PrmIdx<NC>;
  __shared__ double sh_prm[TPB][I::size];
  __shared__ double sh_start[TPB][(1 + PA) * N];

  const int lane = threadIdx.x % TEAM;
  const int tib = threadIdx.x / TEAM;
  const long team = a.team0 + (long)blockIdx.x * TPB + tib;  // the workspace is padded to whole blocks
  const long total = (long)a.V * a.O;
  const bool dead = team >= total;
  const long tt = dead ? 0 : team;
  const int v = (int)(tt % a.V);
  const int o = (int)(tt / a.V);

  double* prm = sh_prm[tib];
  for (int i = lane; i < I::stride; i += TEAM) prm[i] = load_real(a.params, (long)v * a.stride + i, a.dtype);
  double* st0 = sh_start[tib];
  for (int i = lane; i < (1 + PA) * N; i += TEAM) {
    const int p = i / N - 1, r = i % N;
    double val = 0.0;
    if (p < 0)
      val = load_real(a.m0, (long)v * N + r, a.dtype);
    else if (p < P && a.dm0)
      val = load_real(a.dm0, ((long)v * a.p_stride + a.p_off + p) * N + r, a.dtype);
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

  int kind[PA], kidx[PA];
#pragma unroll
  for (int p = 0; p < PA; p++) {
    kind[p] = p < P ? a.t_kind[p] : TK_NONE;
    kidx[p] = p < P ? a.t_idx[p] : 0;
  }
  double dwp[NC > 0 ? NC : 1];
#pragma unroll
  for (int i = 0; i < NC; i++) dwp[i] = prm[13 + 2 * NC + i];

  double M[N], D[PA][N];
  auto load_start = [&]() {
#pragma unroll
    for (int i = 0; i < N; i++) {
      M[i] = st0[i];
#pragma unroll
      for (int p = 0; p < PA; p++) D[p][i] = st0[(1 + p) * N + i];
    }
  };
  MtCache mc;
  mc.fr = __builtin_nan("");
  mc.s = MtShape{0.0, 0.0, 0.0, 0.0};
  const Ev* evr = a.ev + (long)o * a.T;
  double* wsT = a.ws + team * (long)(a.n_seg + a.K) * REC;  // this team's checkpoints, then its segment
  double* seg = wsT + (long)a.n_seg * REC;

  // every lane stores the whole (team-uniform) record and reads back only what it stored itself
  auto put = [&](double* r, bool flag) {
#pragma unroll
    for (int i = 0; i < N; i++) {
      r[i] = M[i];
#pragma unroll
      for (int p = 0; p < P; p++) r[(1 + p) * N + i] = D[p][i];
    }
    r[REC - 1] = flag ? 1.0 : 0.0;
  };
  auto get = [&](const double* r) {
#pragma unroll
    for (int i = 0; i < N; i++) {
      M[i] = r[i];
#pragma unroll
      for (int p = 0; p < P; p++) D[p][i] = r[(1 + p) * N + i];
    }
    return r[REC - 1] != 0.0;
  };
  // one event forward (the pending reset is handled by the caller), simulate.py:48-72
  auto forward = [&](const Ev& e) {
    if (e.dt > 1e-9) {
      Gen<NC> G;
      TanCo tc[PA];
      double l1, sc;
      int m, s;
      bool isn;
      build_gen<NC, P>(prm, e, mc, kind, kidx, G, tc, l1, sc, a.nan_l1, m, s, isn);
      if (s == 0) {
        taylor<NC, P>(G, tc, kind, kidx, dwp, m, M, 1.0, D, M, D);
        if (isn) {
#pragma unroll
          for (int i = 0; i < N; i++) {
            M[i] = __builtin_nan("");
#pragma unroll
            for (int p = 0; p < PA; p++) D[p][i] = __builtin_nan("");
          }
        }
      } else {
        double R[N], dR[PA][N];
        event_cols<NC, P, TEAM>(G, tc, kind, kidx, dwp, m, s, isn, lane, R, dR);
        apply_cols<NC, P, TEAM>(R, dR, M, D, lane);
      }
    }
    if (e.type == BMC_EV_GRADIENT || e.type == BMC_EV_SPOILER) {
#pragma unroll
      for (int i = 0; i < 2 * NT; i++) {
        M[i] = 0.0;
#pragma unroll
        for (int p = 0; p < PA; p++) D[p][i] = 0.0;
      }
    }
  };

  // ---- phase 1: forward with checkpoints ----
  if (!dead) {
    load_start();
    bool pending = false;
    for (int t = 0; t < a.T; t++) {
      if (t % a.K == 0) put(wsT + (long)(t / a.K) * REC, pending);
      if (pending) {
        load_start();
        pending = false;
      }
      const Ev e = evr[t];
      forward(e);
      if (e.type == BMC_EV_ADC && a.reset_after_adc) pending = true;
    }
  }

  // ---- phase 2: reverse ----
  double lam[N], mu[PA][N];
  {
    const long vo = (long)v * a.O + o;
#pragma unroll
    for (int i = 0; i < N; i++) {
      lam[i] = (i == a.jc && a.gz && !dead) ? a.gz[vo] : 0.0;
#pragma unroll
      for (int p = 0; p < PA; p++)
        mu[p][i] = (p < P && i == a.jc && a.gj && !dead) ? a.gj[vo * a.p_stride + a.p_off + p] * a.t_scale[p] : 0.0;
    }
  }
  bool live = !dead;
  // every team of this wave is live on the same row: sum the wave, else each team adds its own
  const int o_first = __shfl(o, 0, 64);
  const bool wave_row = __ballot(dead || o != o_first) == 0;
  const int n_q = 2;
  for (int sg = a.n_seg - 1; sg >= 0; sg--) {
    const int t0 = sg * a.K, t1 = min(a.T, t0 + a.K);
    if (live) {
      bool pending = get(wsT + (long)sg * REC);
      for (int t = t0; t < t1; t++) {
        const bool reset_here = pending;
        if (pending) {
          load_start();
          pending = false;
        }
        put(seg + (long)(t - t0) * REC, reset_here);
        const Ev e = evr[t];
        forward(e);
        if (e.type == BMC_EV_ADC && a.reset_after_adc) pending = true;
      }
    }
    for (int t = t1 - 1; t >= t0; t--) {
      double gq[2] = {0.0, 0.0};
      if (live) {
        const bool reset_here = get(seg + (long)(t - t0) * REC);
        const Ev e = evr[t];
        if (e.type == BMC_EV_GRADIENT || e.type == BMC_EV_SPOILER) {
#pragma unroll
          for (int i = 0; i < 2 * NT; i++) {
            lam[i] = 0.0;
#pragma unroll
            for (int p = 0; p < PA; p++) mu[p][i] = 0.0;
          }
        }
        if (e.dt > 1e-9) {
          Gen<NC> G;
          TanCo tc[PA];
          double l1, sc;
          int m, s;
          bool isn;
          build_gen<NC, P>(prm, e, mc, kind, kidx, G, tc, l1, sc, a.nan_l1, m, s, isn);
          // this lane's entries of the state (augmented coordinate 1) and of the tangents (0)
          double Mj = (lane == N) ? 1.0 : 0.0, Dj[PA];
#pragma unroll
          for (int p = 0; p < PA; p++) Dj[p] = 0.0;
#pragma unroll
          for (int i = 0; i < N; i++)
            if (i == lane) {
              Mj = M[i];
#pragma unroll
              for (int p = 0; p < PA; p++) Dj[p] = D[p][i];
            }
          double lam_j = 0.0, mu_j[PA];
#pragma unroll
          for (int p = 0; p < PA; p++) mu_j[p] = 0.0;
          bool updated = false;
          for (int q = 0; q < n_q; q++) {
            double* gout = q == 0 ? a.grad_amp : a.grad_freq;
            if (!gout && (updated || q + 1 < n_q)) continue;  // the last pass still updates the adjoint
            int qkind;
            const TanCo tq = rf_tangent<NC>(prm, e, mc, sc, q, qkind);
            double R[N], Rp[PA][N], Rq[N], Rpq[PA][N];
            taylor2<NC, P>(G, tc, kind, kidx, dwp, tq, qkind, m, lane, R, Rp, Rq, Rpq);
            for (int r = 0; r < s; r++) square2<NC, P, TEAM>(R, Rp, Rq, Rpq, lane);
            if (isn) {
#pragma unroll
              for (int i = 0; i < N; i++) {
                R[i] = Rq[i] = __builtin_nan("");
#pragma unroll
                for (int p = 0; p < PA; p++) Rp[p][i] = Rpq[p][i] = __builtin_nan("");
              }
            }
            if (gout) {
              double lq = 0.0, c = 0.0;
#pragma unroll
              for (int i = 0; i < N; i++) lq = fma(lam[i], Rq[i], lq);
              c = Mj * lq;
#pragma unroll
              for (int p = 0; p < P; p++) {
                double a1 = 0.0, a2 = 0.0;
#pragma unroll
                for (int i = 0; i < N; i++) {
                  a1 = fma(mu[p][i], Rq[i], a1);
                  a2 = fma(mu[p][i], Rpq[p][i], a2);
                }
                c = fma(Dj[p], a1, fma(Mj, a2, c));
              }
              gq[q] = team_sum<TEAM>(c);
            }
            if (!updated) {
              double l = 0.0;
#pragma unroll
              for (int i = 0; i < N; i++) l = fma(lam[i], R[i], l);
#pragma unroll
              for (int p = 0; p < P; p++) {
                double u = 0.0;
#pragma unroll
                for (int i = 0; i < N; i++) {
                  l = fma(mu[p][i], Rp[p][i], l);
                  u = fma(mu[p][i], R[i], u);
                }
                mu_j[p] = u;
              }
              lam_j = l;
              updated = true;
            }
          }
          // the new adjoint, entry j from lane j, made uniform over the team
          static_for<0, N>([&](auto ic) {
            constexpr int i = decltype(ic)::value;
            lam[i] = bcast<TEAM, i>(lam_j);
#pragma unroll
            for (int p = 0; p < P; p++) mu[p][i] = bcast<TEAM, i>(mu_j[p]);
          });
        }
        if (reset_here) {  // the state was reloaded here: nothing earlier reaches the output
          live = false;
        }
      }
      // sum the teams of the wave (same row) and add once per wave
#pragma unroll
      for (int q = 0; q < 2; q++) {
        double* gout = q == 0 ? a.grad_amp : a.grad_freq;
        if (!gout) continue;
        double g = (lane == 0) ? gq[q] : 0.0;
        if (wave_row) {
#pragma unroll
          for (int off = TEAM; off < 64; off <<= 1) g += __shfl_xor(g, off, 64);
          if (threadIdx.x % 64 == 0 && g != 0.0) atomicAdd(gout + (long)o_first * a.T + t, g);
        } else if (lane == 0 && !dead && g != 0.0) {
          atomicAdd(gout + (long)o * a.T + t, g);
        }
      }
    }
  }
}

template <int NC, int P>
static int launch_adj_np(const AdjArgs& a, hipStream_t stream) {
  constexpr int NA = 3 * (1 + NC) + 2;
  // one lane per propagator column (two teams per DPP row at 8, a DPP row at 16, two rows beyond)
  constexpr int TEAM = NA <= 8 ? 8 : (NA <= 16 ? 16 : 32);
  constexpr int TPB = 256 / TEAM;
  const long teams = (long)a.V * a.O;
  constexpr long kMaxBlocks = 1L << 23;
  for (long t0 = 0; t0 < teams; t0 += kMaxBlocks * TPB) {
    const long blocks = std::min(kMaxBlocks, (teams - t0 + TPB - 1) / TPB);
    AdjArgs c = a;
    c.team0 = t0;
    hipLaunchKernelGGL((adj_kernel<NC, P, TEAM>), dim3((unsigned)blocks), dim3(256), 0, stream, c);
    BMC_HIP_CHECK(hipGetLastError());
  }
  return BMC_OK;
}

template <int NC>
static int launch_adj_nc(const AdjArgs& a, hipStream_t stream) {
  switch (a.P) {
    case 0:
      return launch_adj_np<NC, 0>(a, stream);
    case 1:
      return launch_adj_np<NC, 1>(a, stream);
    case 2:
      return launch_adj_np<NC, 2>(a, stream);
    default:
      set_error("adjoint: at most 2 tangents");
      return BMC_ENOTSUP;
  }
}

static void adj_segments(int T, int* K, int* n_seg) {
  int k = (int)ceil(sqrt((double)(T > 0 ? T : 1)));
  if (k < 1) k = 1;
  *K = k;
  *n_seg = (T + k - 1) / k;
}

// Tangents one launch carries (the rest run as further launches over the same workspace).
constexpr int kAdjGroup = 2;

static long adj_record(int nc, int P) {
  const int Pg = P < kAdjGroup ? P : kAdjGroup;
  return (long)(1 + Pg) * (3 * (1 + nc) + 1) + 1;
}

static long adj_teams_padded(int V, int O) {
  const long teams = (long)V * O;
  return (teams + 31) / 32 * 32;  // whole blocks of 8- or 16-lane teams
}

}  // namespace bmc

using namespace bmc;

extern "C" int bmc_adjoint_workspace_size(const bmc_plan* plan, int n_voxels, int n_cest, int n_tangents,
                                          int64_t* bytes) {
  if (!plan || !bytes || n_voxels < 0 || n_cest < 0 || n_tangents < 0) {
    set_error("bmc_adjoint_workspace_size: bad argument");
    return BMC_EINVAL;
  }
  const Plan& p = plan->p;
  int K, n_seg;
  adj_segments(p.T, &K, &n_seg);
  *bytes = adj_teams_padded(n_voxels, p.O) * (long)(K + n_seg) * adj_record(n_cest, n_tangents) * 8;
  return BMC_OK;
}

extern "C" int bmc_simulate_adjoint(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
                                    const void* m0_dev, const void* dm0_dev, int n_tangents,
                                    const int32_t* tangent_slot, const double* tangent_scale, int jac_component,
                                    const double* z_cot_dev, const double* jac_cot_dev, double* grad_amp_dev,
                                    double* grad_freq_dev, void* workspace_dev, int64_t workspace_bytes,
                                    void* stream) {
  SimArgs s;
  int rc = fill_sim_args(plan, n_voxels, n_cest, dtype, params_dev, m0_dev, dm0_dev, n_tangents, tangent_slot,
                         tangent_scale, jac_component, s);
  if (rc != BMC_OK) return rc;
  for (int p = 0; p < n_tangents; p++) {
    const int k = s.t_kind[p];
    if (k == TK_MT_DW || k == TK_MT_R2 || k == TK_B0 || k == TK_GAMMA || k == TK_B0INH || k == TK_RELB1) {
      set_error("bmc_simulate_adjoint: tangent slot " + std::to_string(tangent_slot[p]) +
                " has a generator derivative that depends on the RF (not supported)");
      return BMC_ENOTSUP;
    }
  }
  if (!grad_amp_dev && !grad_freq_dev) {
    set_error("bmc_simulate_adjoint: no gradient output");
    return BMC_EINVAL;
  }
  if (n_voxels == 0) return BMC_OK;
  int64_t need = 0;
  rc = bmc_adjoint_workspace_size(plan, n_voxels, n_cest, n_tangents, &need);
  if (rc != BMC_OK) return rc;
  if (!workspace_dev || workspace_bytes < need) {
    set_error("bmc_simulate_adjoint: workspace smaller than bmc_adjoint_workspace_size");
    return BMC_EINVAL;
  }
  AdjArgs a;
  memset(&a, 0, sizeof(a));
  a.ev = s.ev;
  a.O = s.O;
  a.T = s.T;
  a.reset_after_adc = s.reset_after_adc;
  a.V = s.V;
  a.dtype = s.dtype;
  a.stride = s.stride;
  a.params = s.params;
  a.m0 = s.m0;
  a.dm0 = s.dm0;
  a.p_stride = n_tangents;
  a.jc = jac_component;
  a.gj = jac_cot_dev;
  a.grad_amp = grad_amp_dev;
  a.grad_freq = grad_freq_dev;
  a.ws = static_cast<double*>(workspace_dev);
  adj_segments(a.T, &a.K, &a.n_seg);
  a.nan_l1 = s.nan_l1;
  hipStream_t st = (hipStream_t)stream;
  // L = gz.Z + sum_p gj_p.J_p is linear in the cotangents, so its gradient is the sum of the gradients of
  // tangent groups of at most kAdjGroup (gz taken once, with the first group) added into the outputs
  for (int p0 = 0; p0 == 0 || p0 < n_tangents; p0 += kAdjGroup) {
    AdjArgs g = a;
    g.P = std::min(kAdjGroup, n_tangents - p0);
    g.p_off = p0;
    g.gz = p0 == 0 ? z_cot_dev : nullptr;
    for (int p = 0; p < g.P; p++) {
      g.t_kind[p] = s.t_kind[p0 + p];
      g.t_idx[p] = s.t_idx[p0 + p];
      g.t_scale[p] = s.t_scale[p0 + p];
    }
    switch (n_cest) {
      case 0: rc = launch_adj_nc<0>(g, st); break;
      case 1: rc = launch_adj_nc<1>(g, st); break;
      case 2: rc = launch_adj_nc<2>(g, st); break;
      case 3: rc = launch_adj_nc<3>(g, st); break;
      default: rc = launch_adj_nc<4>(g, st); break;
    }
    if (rc != BMC_OK) return rc;
    if (n_tangents == 0) break;
  }
  return BMC_OK;
}
