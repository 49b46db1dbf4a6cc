// Measurement probe: the dense propagator product of the squarings done two ways on gfx950, to decide
// whether matrix cores pay for it. Not part of the simulate-and-fit ABI (include/bmc_hip.h); built into
// libbmc_probe.so and used by `bench.py --workload squaring` and its GPU test.
//
//   path 0  the library's own column product (mul_cols in bmc_device.h): a team of lanes, lane j holding
//           column j of both factors, every column of the left factor broadcast across the team
//           (ds_bpermute for 12-lane teams, DPP row broadcasts for 16 lanes, DPP + row_bcast for 32);
//   path 1  v_mfma_f64_16x16x4f64: one wave per product, the factors padded to 16 x 16, four MFMAs of
//           K = 4; operands A[i][k] at lane i + 16 k and B[k][j] at lane j + 16 k, the result D[i][j] at
//           lane j + 16 (i mod 4), register i / 4 (measured with one-hot operands); it goes back to the
//           B-operand layout through the wave's LDS tile between products.
// Both repeat E <- P E `reps` times on n_mat matrices E (augmented, NA x NA, row-major; last row
// [0 ... 0 1]) with one shared P, so a launch times reps x n_mat products.
#include <hip/hip_runtime.h>

#include "bmc_device.h"

namespace bmc {
namespace probe {

template <int NC, int TEAM>
__global__ void __launch_bounds__(256) probe_cols_kernel(const double* P, double* E, long n_mat, int reps) {
  constexpr int N = 3 * (1 + NC) + 1, NA = N + 1;
  const int lane = team_lane<TEAM>();
  const long team = (long)blockIdx.x * teams_per_block<TEAM>() + team_index<TEAM>();
  const long t = team < n_mat ? team : 0;
  const int j = lane < NA ? lane : NA - 1;  // idle lanes of the team read a valid column, never stored
  double a[N], b[N], da[1][N], db[1][N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    a[i] = P[i * NA + j];
    b[i] = E[t * NA * NA + i * NA + j];
    da[0][i] = 0.0;
    db[0][i] = 0.0;
  }
  for (int r = 0; r < reps; r++) mul_cols<NC, 0, TEAM>(a, da, b, db, lane);
  if (team < n_mat && lane < NA) {
#pragma unroll
    for (int i = 0; i < N; i++) E[t * NA * NA + i * NA + j] = b[i];
  }
}

// path 2 (12-lane teams): the same column product with the left factor's columns written to the team's
// LDS stage once and read back as 16-byte broadcast loads instead of ds_bpermutes.
template <int NC>
__global__ void __launch_bounds__(256) probe_staged_kernel(const double* P, double* E, long n_mat, int reps) {
  constexpr int N = 3 * (1 + NC) + 1, NA = N + 1, TEAM = 12;
  static_assert(N % 2 == 0, "pairs of rows");
  __shared__ __attribute__((aligned(16))) double stage[teams_per_block<TEAM>()][N * N];
  const int lane = team_lane<TEAM>();
  const int tib = team_index<TEAM>();
  const long team = (long)blockIdx.x * teams_per_block<TEAM>() + tib;
  const long t = team < n_mat ? team : 0;
  const int j = lane < NA ? lane : NA - 1;
  double a[N], b[N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    a[i] = P[i * NA + j];
    b[i] = E[t * NA * NA + i * NA + j];
  }
  volatile double* stw = stage[tib];
  const double2* st2 = reinterpret_cast<const double2*>(stage[tib]);
  for (int r = 0; r < reps; r++) {
    team_lds_sync();
    if (lane < N) {
#pragma unroll
      for (int i = 0; i < N; i++) stw[lane * N + i] = a[i];
    }
    team_lds_sync();
    double o[N];
#pragma unroll
    for (int i = 0; i < N; i++) o[i] = (lane == N) ? a[i] : 0.0;
#pragma unroll
    for (int k = 0; k < N; k++) {
      const double bk = b[k];
#pragma unroll
      for (int i2 = 0; i2 < N / 2; i2++) {
        const double2 a2 = st2[(k * N) / 2 + i2];
        o[2 * i2] = fma(a2.x, bk, o[2 * i2]);
        o[2 * i2 + 1] = fma(a2.y, bk, o[2 * i2 + 1]);
      }
    }
#pragma unroll
    for (int i = 0; i < N; i++) b[i] = o[i];
  }
  if (team < n_mat && lane < NA) {
#pragma unroll
    for (int i = 0; i < N; i++) E[t * NA * NA + i * NA + j] = b[i];
  }
}

typedef double v4d __attribute__((ext_vector_type(4)));

template <int NA>
__global__ void __launch_bounds__(256) probe_mfma_kernel(const double* P, double* E, long n_mat, int reps) {
  static_assert(NA <= 16, "one 16 x 16 tile");
  __shared__ double tile[4][16 * 16];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long m = (long)blockIdx.x * 4 + w;
  const long mm = m < n_mat ? m : 0;
  const int col = l & 15, grp = l >> 4;
  // A operand of step kb: lane i + 16 k holds P[i][4 kb + k]
  double pa[4];
#pragma unroll
  for (int kb = 0; kb < 4; kb++) {
    const int k = 4 * kb + grp;
    pa[kb] = (col < NA && k < NA) ? P[col * NA + k] : 0.0;
  }
  // E in the D layout: lane col + 16 grp, register r holds E[4 r + grp][col]
  v4d e;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int i = 4 * r + grp;
    e[r] = (i < NA && col < NA) ? E[mm * NA * NA + i * NA + col] : 0.0;
  }
  double* t = tile[w];
  for (int rep = 0; rep < reps; rep++) {
#pragma unroll
    for (int r = 0; r < 4; r++) t[(4 * r + grp) * 16 + col] = e[r];
    team_lds_sync();
    v4d acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int kb = 0; kb < 4; kb++) {  // B operand of step kb: lane j + 16 k holds E[4 kb + k][j]
      const double bop = t[(4 * kb + grp) * 16 + col];
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[kb], bop, acc, 0, 0, 0);
    }
    team_lds_sync();  // every lane has read the tile before the next product overwrites it
    e = acc;
  }
  if (m < n_mat) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int i = 4 * r + grp;
      if (i < NA && col < NA) E[m * NA * NA + i * NA + col] = e[r];
    }
  }
}

template <int NC>
static int launch(int path, const double* P, double* E, long n_mat, int reps, hipStream_t s) {
  constexpr int NA = 3 * (1 + NC) + 2;
  constexpr int TEAM = NA <= 8 ? 8 : (NA <= 12 ? 12 : (NA <= 16 ? 16 : 32));
  if (path == 2) {
    if constexpr (TEAM == 12) {
      constexpr int TPB = teams_per_block<TEAM>();
      const long blocks = (n_mat + TPB - 1) / TPB;
      hipLaunchKernelGGL((probe_staged_kernel<NC>), dim3((unsigned)blocks), dim3(256), 0, s, P, E, n_mat, reps);
    } else {
      return -2;  // the staged form is for 12-lane teams only
    }
  } else if (path == 0) {
    constexpr int TPB = teams_per_block<TEAM>();
    const long blocks = (n_mat + TPB - 1) / TPB;
    hipLaunchKernelGGL((probe_cols_kernel<NC, TEAM>), dim3((unsigned)blocks), dim3(256), 0, s, P, E, n_mat, reps);
  } else if constexpr (NA <= 16) {
    const long blocks = (n_mat + 3) / 4;
    hipLaunchKernelGGL((probe_mfma_kernel<NA>), dim3((unsigned)blocks), dim3(256), 0, s, P, E, n_mat, reps);
  } else {
    return -2;  // 17 x 17 does not fit one tile
  }
  return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // namespace probe
}  // namespace bmc

// path 0 (team column product), 1 (FP64 MFMA tile) or 2 (LDS-staged column product, 12-lane teams only);
// nc = CEST pools (augmented size 3 (1 + nc) + 2).
// P: NA x NA, E: n_mat x NA x NA, both row-major float64 device arrays. Returns 0, or a negative code for
// an unsupported (path, nc), a size out of range or a launch error.
extern "C" int bmc_probe_products(int path, int nc, const double* P, double* E, long n_mat, int reps, void* stream) {
  if (n_mat < 1 || n_mat > (1L << 26) || reps < 0 || path < 0 || path > 2) return -1;
  hipStream_t s = (hipStream_t)stream;
  switch (nc) {
    case 1: return bmc::probe::launch<1>(path, P, E, n_mat, reps, s);
    case 2: return bmc::probe::launch<2>(path, P, E, n_mat, reps, s);
    case 3: return bmc::probe::launch<3>(path, P, E, n_mat, reps, s);
    case 4: return bmc::probe::launch<4>(path, P, E, n_mat, reps, s);
    default: return -2;
  }
}
