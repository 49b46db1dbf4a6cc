// The single-event row and its Jacobian by reverse accumulation through the squarings, shared by the
// simulate / LM kernels (bmc_cwj.hip) and the uneven batch (bmc_uneven.hip). Derivation in bmc_cwj.hip.
#pragma once

#include "bmc_device.h"

namespace bmc {

// Tangents per Taylor pass: two share one primal series; at four CEST pools one, since two spill in the
// Horner loop itself (one tangent per pass halved the four-pool sweep).
template <int NC>
__host__ __device__ constexpr int cwj_chunk() {
  return NC >= 4 ? 1 : 2;
}
// Taylor degree cap of this kernel. With the tangents contracted, a squaring level costs about 1.4
// Taylor terms of the series (primal and every tangent pass), so a lower cap with more squarings is
// cheaper than the forward kernel's 30; 18 keeps the state within 1e-12 of it (1.7 more levels).
constexpr int kCwjMMax = 18;

// One level: G <- G E^T + E^T G, then E <- E E (columns R, rows Er), lane j's share.
template <int NC, int TEAM>
__device__ __forceinline__ void cwj_level(double* R, double* Er, double* Gc, int j) {
  constexpr int N = 3 * (1 + NC) + 1, NA = N + 1;
  double o[N], ner[NA], etg[N], nge[N];
  const bool aff = (j == N);
#pragma unroll
  for (int i = 0; i < N; i++) {
    o[i] = aff ? R[i] : 0.0;  // the E[i][N] E[N][j] term (E[N][j] = [j == N])
    etg[i] = 0.0;
    nge[i] = 0.0;
  }
#pragma unroll
  for (int l = 0; l < NA; l++) ner[l] = 0.0;
  ner[N] = Er[N];  // E[j][N] E[N][N]
  static_for<0, NA>([&](auto lc) {
    constexpr int l = decltype(lc)::value;
#pragma unroll
    for (int i = 0; i < N; i++) {
      const double b = bcast<TEAM, l>(R[i]);  // E[i][l]
      if constexpr (l < N) {
        o[i] = fma(b, R[l], o[i]);        // (E E)[i][j]   += E[i][l] E[l][j]
        etg[l] = fma(b, Gc[i], etg[l]);   // (E^T G)[l][j] += E[i][l] G[i][j]
      }
      ner[l] = fma(Er[i], b, ner[l]);     // (E E)[j][l]   += E[j][i] E[i][l]
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
      const double g = bcast<TEAM, l>(Gc[i]);  // G[i][l]
      nge[i] = fma(g, Er[l], nge[i]);          // (G E^T)[i][j] += G[i][l] E[j][l]
    }
  });
#pragma unroll
  for (int i = 0; i < N; i++) {
    R[i] = o[i];
    Gc[i] = etg[i] + nge[i];
  }
#pragma unroll
  for (int l = 0; l < NA; l++) Er[l] = ner[l];
}

// The same level when the team has one lane per column of E but none for its augmented column N
// (four CEST pools: 17 columns on a 16-lane team). Column N is spread over the team instead: lane j keeps
// E[j][N] as Er[N] (its row's last entry, as in every layout) and G[j][N] as GN, and the terms that
// involve column N broadcast those entries:
//   (E E)[j][N] = E[j][N] + sum_l E[j][l] E[l][N],     (G E^T)[i][j] += G[i][N] E[j][N],
//   G'[j][N] = (G E^T)[j][N] + (E^T G)[j][N] = G[j][N] + sum_l E[l][j] G[l][N]
// (E's augmented row is e_N, and G's augmented row never reaches the rows kept).
template <int NC, int TEAM>
__device__ __forceinline__ void cwj_level_spread(double* R, double* Er, double* Gc, double& GN) {
  constexpr int N = 3 * (1 + NC) + 1, NA = N + 1;
  double o[N], ner[NA], etg[N], nge[N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    o[i] = 0.0;
    etg[i] = 0.0;
    nge[i] = 0.0;
  }
#pragma unroll
  for (int l = 0; l < N; l++) ner[l] = 0.0;
  ner[N] = Er[N];
  double gn = GN;
  static_for<0, N>([&](auto lc) {
    constexpr int l = decltype(lc)::value;
#pragma unroll
    for (int i = 0; i < N; i++) {
      const double b = bcast<TEAM, l>(R[i]);  // E[i][l]
      o[i] = fma(b, R[l], o[i]);              // (E E)[i][j]   += E[i][l] E[l][j]
      etg[l] = fma(b, Gc[i], etg[l]);         // (E^T G)[l][j] += E[i][l] G[i][j]
      ner[l] = fma(Er[i], b, ner[l]);         // (E E)[j][l]   += E[j][i] E[i][l]
    }
#pragma unroll
    for (int i = 0; i < N; i++) {
      const double g = bcast<TEAM, l>(Gc[i]);  // G[i][l]
      nge[i] = fma(g, Er[l], nge[i]);          // (G E^T)[i][j] += G[i][l] E[j][l]
    }
    const double eln = bcast<TEAM, l>(Er[N]);  // E[l][N]
    ner[N] = fma(Er[l], eln, ner[N]);          // (E E)[j][N] += E[j][l] E[l][N]
    const double gln = bcast<TEAM, l>(GN);     // G[l][N]
    nge[l] = fma(gln, Er[N], nge[l]);          // (G E^T)[l][j] += G[l][N] E[j][N]
    gn = fma(R[l], gln, gn);                   // (E^T G)[j][N] += E[l][j] G[l][N]
  });
#pragma unroll
  for (int i = 0; i < N; i++) {
    R[i] = o[i];
    Gc[i] = etg[i] + nge[i];
  }
#pragma unroll
  for (int l = 0; l < NA; l++) Er[l] = ner[l];
  GN = gn;
}

// Entry `lane` of a register array of N entries (zero for lanes >= N).
template <int N>
__device__ __forceinline__ double own_entry(const double* x, int lane) {
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < N; i++)
    if (i == lane) r = x[i];
  return r;
}

// One team's row: the state M = E v (every lane, team reductions) and, for the P tangents of kinds
// kind[] / kidx[] (indexable by tangent), dZ/dp for Z = M[jc]. e_live is the row's propagating event
// (null: none, E = I). store_state(M, isn) and store_jac(p, dZ/dp, isn) are called by lane 0 only.
template <int NC, int TEAM, typename KindAt, typename StoreState, typename StoreJac>
__device__ __forceinline__ void cwj_row(const double* prm, const double* m0, const double* dwp, const Ev* e_live,
                                        int P, KindAt kind_at, int jc, double nan_l1, int lane, bool want_jac,
                                        StoreState&& store_state, StoreJac&& store_jac) {
  constexpr int NT = 1 + NC, N = 3 * NT + 1, NA = N + 1;
  constexpr bool SPREAD = NA > TEAM;  // the augmented column spread over the team (cwj_level_spread)
  static_assert(!SPREAD || TEAM == N, "a spread augmented column needs one lane per other column");
  double vj = (lane == N) ? 1.0 : 0.0;  // lane's entry of the augmented start vector
#pragma unroll
  for (int i = 0; i < N; i++)
    if (lane == i) vj = m0[i];
  double R[N], Er[NA], Gc[N];
  double GN = (SPREAD && lane == jc) ? 1.0 : 0.0;  // spread layout: G[lane][N] (G_1[i][N] = [i == jc])
#pragma unroll
  for (int i = 0; i < N; i++) {
    R[i] = (i == lane) ? 1.0 : 0.0;
    Gc[i] = (i == jc) ? vj : 0.0;  // G_1 = e_z v^T
  }
  int m = 0, s = 0;
  bool isn = false;
  MtCache mc;
  mc.fr = __builtin_nan("");
  mc.s = MtShape{0.0, 0.0, 0.0, 0.0};
  Ev e{};
  if (e_live) {
    e = *e_live;
    Gen<NC> G;
    int k0[1] = {TK_NONE}, i0[1] = {0};
    TanCo tc0[1];
    double l1, sc;
    build_gen<NC, 0>(prm, e, mc, k0, i0, G, tc0, l1, sc, nan_l1, m, s, isn, kCwjMMax);
    double dR0[1][N];
    taylor<NC, 0>(G, tc0, k0, i0, dwp, m, R, (lane == N) ? 1.0 : 0.0, nullptr, R, dR0);  // column lane of T
    // rows of T: Er[l] = T[lane][l]
    static_for<0, (SPREAD ? N : NA)>([&](auto lc) {
      constexpr int l = decltype(lc)::value;
      double x = 0.0;
#pragma unroll
      for (int i = 0; i < N; i++) {
        const double b = bcast<TEAM, l>(R[i]);
        if (i == lane) x = b;
      }
      Er[l] = x;
    });
    if constexpr (SPREAD) {
      // column N of T on every lane (the series of e_N), of which each lane keeps its own entry
      double A[N];
#pragma unroll
      for (int i = 0; i < N; i++) A[i] = 0.0;
      taylor<NC, 0>(G, tc0, k0, i0, dwp, m, A, 1.0, nullptr, A, dR0);
      Er[N] = own_entry<N>(A, lane);
      for (int q = 0; q < s; q++) cwj_level_spread<NC, TEAM>(R, Er, Gc, GN);
    } else {
      if (lane == N) Er[N] = 1.0;
      for (int q = 0; q < s; q++) cwj_level<NC, TEAM>(R, Er, Gc, lane);
    }
  } else if constexpr (SPREAD) {
    Er[N] = 0.0;  // E = I: E[lane][N] = 0
  }
  double M[N];
#pragma unroll
  for (int i = 0; i < N; i++) M[i] = team_sum<TEAM>(R[i] * vj + ((SPREAD && i == lane) ? Er[N] : 0.0));
  if (lane == 0) store_state(M, isn);
  if (!want_jac) return;
  // tangents: dZ/dp = <dT_p, G>, dT_p by forward-mode Horner, kCwjChunk tangents per pass
  constexpr int kCwjChunk = cwj_chunk<NC>();
  for (int q0 = 0; q0 < P; q0 += kCwjChunk) {
    int kind[kCwjChunk], kidx[kCwjChunk];
#pragma unroll
    for (int p = 0; p < kCwjChunk; p++) {
      kind[p] = TK_NONE;
      kidx[p] = 0;
      if (q0 + p < P) kind_at(q0 + p, kind[p], kidx[p]);
    }
    double g[kCwjChunk];
    if (e_live) {
      TanCo tc[kCwjChunk];
      double l1, sc;
      int mq, sq;
      bool isq;
      Gen<NC> Gq;
      build_gen<NC, kCwjChunk>(prm, e, mc, kind, kidx, Gq, tc, l1, sc, nan_l1, mq, sq, isq, kCwjMMax);
      double T[N], dT[kCwjChunk][N];
#pragma unroll
      for (int i = 0; i < N; i++) T[i] = (i == lane) ? 1.0 : 0.0;
      taylor<NC, kCwjChunk>(Gq, tc, kind, kidx, dwp, mq, T, (lane == N) ? 1.0 : 0.0, nullptr, T, dT);
      double d[kCwjChunk];
#pragma unroll
      for (int p = 0; p < kCwjChunk; p++) {
        d[p] = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++) d[p] = fma(dT[p][i], Gc[i], d[p]);
      }
      if constexpr (SPREAD) {
        // the augmented column's share: sum_i dT[i][N] G[i][N], lane i adding its term
#pragma unroll
        for (int i = 0; i < N; i++) T[i] = 0.0;
        taylor<NC, kCwjChunk>(Gq, tc, kind, kidx, dwp, mq, T, 1.0, nullptr, T, dT);
#pragma unroll
        for (int p = 0; p < kCwjChunk; p++) d[p] = fma(own_entry<N>(dT[p], lane), GN, d[p]);
      }
#pragma unroll
      for (int p = 0; p < kCwjChunk; p++) g[p] = team_sum<TEAM>(lane < NA ? d[p] : 0.0);
    } else {
#pragma unroll
      for (int p = 0; p < kCwjChunk; p++) g[p] = 0.0;  // no propagating event: Z = M0[jc]
    }
    if (lane == 0) {
#pragma unroll
      for (int p = 0; p < kCwjChunk; p++)
        if (q0 + p < P) store_jac(q0 + p, g[p], isn);
    }
  }
}

}  // namespace bmc
