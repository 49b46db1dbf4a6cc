// Uneven batched Levenberg-Marquardt: many independent fit problems of different sizes -- CEST pool
// count, number of rows (rungs x offsets) and number of free values -- fitted together, every
// iteration on the GPU. The reference fits such problems one `fit` call at a time
// (src/bmc_jax/fitting/independent.py:226-298, lm.py:20-67), since vmap needs equal shapes.
//
// The simulation is one persistent launch per sweep. Its work items are (problem, row, tangent pass)
// triples, grouped into chunks of same-class items (one chunk per wave: 8 teams of 8 lanes for one
// CEST pool, 4 of 16 for two or three, 2 of 32 for four; the reverse Jacobian kernel runs four pools
// as 4 teams of 16). A fixed grid of one-wave workgroups pulls chunks from a global
// counter until the list is exhausted, so long and short problems share the machine without a
// launch per shape. Each team runs the same row program as bmc_simulate (run_row).
// The LM steps (float64 normal equations, damped solve, clip, accept/reject) are ragged versions of
// bmc_lm_fit's, one wave per problem. The iteration can be captured into a hipGraph.
// The item order, tangents per pass and grid size are chosen by the caller (an autotuner times them).

#include <algorithm>
#include <cstring>
#include <memory>
#include <numeric>
#include <vector>

#include "bmc_cwj.h"

namespace bmc {

struct UProb {
  int32_t cls;     // n_cest - 1
  int32_t local;   // index among the problems of its class (params / M0 arrays of the class)
  int32_t P;       // free values
  int32_t row0;    // first row of the plan
  int32_t n_rows;
  int32_t x_off;   // offset into the per-value tables (slots, scales, bounds, x)
  int64_t j_off;   // offset of the Jacobian [n_rows, P]
  int32_t jc;      // water Mz index
  int32_t pad;
};

struct UItem {
  int32_t prob, row, pass, pad;
};

struct USimArgs {
  const Ev* ev;
  const Op* ops;
  const int32_t* row_ptr;
  int T, reset_after_adc, dtype, PT;
  const void* prm[4];
  const void* m0[4];
  const UProb* probs;
  const int32_t* t_kind;
  const int32_t* t_idx;
  const int32_t* t_slot;
  const double* t_scale;
  const double* xw;  // [sum P] scaled values written into the slots
  const UItem* items;
  const int2* chunks;  // (first item, count)
  int n_chunks;
  int* counter;
  double* z_out;    // [R]
  double* jac_out;  // [sum n_rows P] or null
  const int32_t* active;
  double nan_l1;
  const int32_t* row_live;  // [R] or null: the Jacobian sweep of single-event rows by reverse accumulation
};

constexpr int kUMaxN = 3 * (1 + 4) + 1;

template <int NC, int PT, int TEAM>
__device__ __forceinline__ void u_chunk(const USimArgs& a, int2 ch, double (*sh_prm)[PrmIdx<4>::size],
                                        double (*sh_state)[(1 + 3) * kUMaxN], double (*sh_start)[(1 + 3) * kUMaxN]) {
  constexpr int NT = 1 + NC, N = 3 * NT + 1;
  constexpr int PTA = PT > 0 ? PT : 1;
  using I = PrmIdx<NC>;
  const int lane = threadIdx.x % TEAM;
  const int tib = threadIdx.x / TEAM;
  const bool has = tib < ch.y;
  const UItem it = a.items[ch.x + (has ? tib : 0)];
  const UProb pr = a.probs[it.prob];
  const bool dead = !has || (a.active && !a.active[it.prob]);
  const int q = it.pass;

  double* prm = sh_prm[tib];
  for (int i = lane; i < I::stride; i += TEAM) {
    double val = load_real(a.prm[NC - 1], (long)pr.local * I::stride + i, a.dtype);
    for (int w = 0; w < pr.P; w++)
      if (a.t_slot[pr.x_off + w] == i) {  // write_params (fitting/params.py:100-112)
        const double xv = a.xw[pr.x_off + w] * a.t_scale[pr.x_off + w];
        val = (a.dtype == BMC_F32) ? (double)(float)xv : xv;
      }
    prm[i] = val;
  }
  double* st0 = sh_start[tib];
  for (int i = lane; i < (1 + PTA) * N; i += TEAM)
    st0[i] = (i < N) ? load_real(a.m0[NC - 1], (long)pr.local * N + i, a.dtype) : 0.0;
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
  if (__ballot(!dead) == 0) return;  // wave-uniform

  int kind[PTA], kidx[PTA];
#pragma unroll
  for (int p = 0; p < PTA; p++) {
    const int gp = q * PT + p;
    const bool on = p < PT && gp < pr.P;
    kind[p] = on ? a.t_kind[pr.x_off + gp] : TK_NONE;
    kidx[p] = on ? a.t_idx[pr.x_off + gp] : 0;
  }
  double dwp[NC];
#pragma unroll
  for (int i = 0; i < NC; i++) dwp[i] = prm[13 + 2 * NC + i];
  double M[N], dM[PTA][N];
  run_row<NC, PT, TEAM>(prm, st0, sh_state[tib], kind, kidx, dwp, a.ev + (long)it.row * a.T, a.ops,
                        a.row_ptr[it.row], a.row_ptr[it.row + 1], a.reset_after_adc, a.nan_l1, lane, M, dM, nullptr,
                        0, 0);
  if (dead || lane != 0) return;
  if (q == 0) a.z_out[it.row] = M[pr.jc];
  if (a.jac_out) {
#pragma unroll
    for (int p = 0; p < PT; p++) {
      const int gp = q * PT + p;
      if (gp < pr.P) {
        double val = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++)
          if (i == pr.jc) val = dM[p][i];
        a.jac_out[pr.j_off + (long)(it.row - pr.row0) * pr.P + gp] = val * a.t_scale[pr.x_off + gp];
      }
    }
  }
}

// One chunk of (problem, row) items of single-event rows: state and Jacobian by reverse accumulation
// through the squarings (cwj_row, bmc_cwj.h), every tangent of the problem in one item.
template <int NC, int TEAM>
__device__ __forceinline__ void u_chunk_cwj(const USimArgs& a, int2 ch, double (*sh_prm)[PrmIdx<4>::size],
                                            double (*sh_start)[(1 + 3) * kUMaxN]) {
  constexpr int NT = 1 + NC, N = 3 * NT + 1;
  using I = PrmIdx<NC>;
  const int lane = threadIdx.x % TEAM;
  const int tib = threadIdx.x / TEAM;
  const bool has = tib < ch.y;
  const UItem it = a.items[ch.x + (has ? tib : 0)];
  const UProb pr = a.probs[it.prob];
  const bool dead = !has || (a.active && !a.active[it.prob]);
  double* prm = sh_prm[tib];
  for (int i = lane; i < I::stride; i += TEAM) {
    double val = load_real(a.prm[NC - 1], (long)pr.local * I::stride + i, a.dtype);
    for (int w = 0; w < pr.P; w++)
      if (a.t_slot[pr.x_off + w] == i) {  // write_params (fitting/params.py:100-112)
        const double xv = a.xw[pr.x_off + w] * a.t_scale[pr.x_off + w];
        val = (a.dtype == BMC_F32) ? (double)(float)xv : xv;
      }
    prm[i] = val;
  }
  double* st0 = sh_start[tib];
  for (int i = lane; i < N; i += TEAM) st0[i] = load_real(a.m0[NC - 1], (long)pr.local * N + i, a.dtype);
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
  if (__ballot(!dead) == 0) return;  // wave-uniform
  double dwp[NC];
#pragma unroll
  for (int i = 0; i < NC; i++) dwp[i] = prm[13 + 2 * NC + i];
  const int le = a.row_live[it.row];
  cwj_row<NC, TEAM>(
      prm, st0, dwp, le >= 0 ? a.ev + (long)it.row * a.T + le : nullptr, pr.P,
      [&](int p, int& k, int& ix) {
        k = a.t_kind[pr.x_off + p];
        ix = a.t_idx[pr.x_off + p];
      },
      pr.jc, a.nan_l1, lane, a.jac_out != nullptr,
      [&](const double* M, bool isn) {
        if (dead) return;
        double z = 0.0;
#pragma unroll
        for (int i = 0; i < N; i++)
          if (i == pr.jc) z = M[i];
        a.z_out[it.row] = isn ? __builtin_nan("") : z;
      },
      [&](int gp, double val, bool isn) {
        if (!dead)
          a.jac_out[pr.j_off + (long)(it.row - pr.row0) * pr.P + gp] =
              isn ? __builtin_nan("") : val * a.t_scale[pr.x_off + gp];
      });
}

#ifndef BMC_USIM_WAVES_PER_SIMD
#define BMC_USIM_WAVES_PER_SIMD 1
#endif

template <int PT>
__global__ void __launch_bounds__(64, BMC_USIM_WAVES_PER_SIMD) usim_kernel(USimArgs a) {
  constexpr int PTA = PT > 0 ? PT : 1;
  static_assert(PTA <= 3, "tangents per pass");
  __shared__ double sh_prm[8][PrmIdx<4>::size];  // up to 8 teams per wave (8-lane teams)
  __shared__ double sh_state[8][(1 + 3) * kUMaxN];
  __shared__ double sh_start[8][(1 + 3) * kUMaxN];
  __shared__ int sh_chunk;
  for (;;) {
    if (threadIdx.x == 0) sh_chunk = atomicAdd(a.counter, 1);
    __syncthreads();
    const int c = sh_chunk;
    __syncthreads();
    if (c >= a.n_chunks) break;  // every wave reaches this: the grid drains when the list is done
    const int2 ch = a.chunks[c];
    const int cls = a.probs[a.items[ch.x].prob].cls;
    switch (cls) {
      case 0:
        u_chunk<1, PT, 8>(a, ch, sh_prm, sh_state, sh_start);
        break;
      case 1:
        u_chunk<2, PT, 16>(a, ch, sh_prm, sh_state, sh_start);
        break;
      case 2:
        u_chunk<3, PT, 16>(a, ch, sh_prm, sh_state, sh_start);
        break;
      default:
        u_chunk<4, PT, 32>(a, ch, sh_prm, sh_state, sh_start);
        break;
    }
    __syncthreads();
  }
}

#ifndef BMC_UCWJ_WAVES_PER_SIMD
#define BMC_UCWJ_WAVES_PER_SIMD BMC_USIM_WAVES_PER_SIMD
#endif

// The Jacobian sweep of single-event rows: persistent over the one-item-per-row chunks.
__global__ void __launch_bounds__(64, BMC_UCWJ_WAVES_PER_SIMD) ucwj_kernel(USimArgs a) {
  __shared__ double sh_prm[8][PrmIdx<4>::size];
  __shared__ double sh_start[8][(1 + 3) * kUMaxN];
  __shared__ int sh_chunk;
  for (;;) {
    if (threadIdx.x == 0) sh_chunk = atomicAdd(a.counter, 1);
    __syncthreads();
    const int c = sh_chunk;
    __syncthreads();
    if (c >= a.n_chunks) break;  // every wave reaches this: the grid drains when the list is done
    const int2 ch = a.chunks[c];
    const int cls = a.probs[a.items[ch.x].prob].cls;
    switch (cls) {
      case 0:
        u_chunk_cwj<1, 8>(a, ch, sh_prm, sh_start);
        break;
      case 1:
        u_chunk_cwj<2, 16>(a, ch, sh_prm, sh_start);
        break;
      case 2:
        u_chunk_cwj<3, 16>(a, ch, sh_prm, sh_start);
        break;
      default:
        u_chunk_cwj<4, 16>(a, ch, sh_prm, sh_start);  // column N spread over the team (cwj_row)
        break;
    }
    __syncthreads();
  }
}

__global__ void set_int_kernel(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__device__ __forceinline__ double u_ldr(const void* p, long i, int dtype) {
  return dtype == BMC_F64 ? static_cast<const double*>(p)[i] : (double)static_cast<const float*>(p)[i];
}

__device__ __forceinline__ double u_wave_sum(double v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}

struct ULm {
  const UProb* probs;
  int n_prob, dtype, n_iter, early;
  const double* z;    // [R] at x
  const double* J;    // ragged
  const double* zt;   // [R] at the trial
  const void* zmeas;  // [R]
  const double* lo;   // [sum P]
  const double* hi;
  double* x;          // [sum P]
  double* xt;         // [sum P]
  double* lam;        // [n_prob]
  double* cost;       // [n_prob]
  int32_t* active;    // [n_prob]
  int32_t* it;        // [1]
  int32_t* n_done;    // [n_prob] or null
  double ftol, rms_tol;
};

// One wave per problem: J^T J and J^T r over its rows (float64), the damped solve, clip (lm.py:43-57).
__global__ void __launch_bounds__(64) u_normal_eq_kernel(ULm a) {
  const int v = blockIdx.x;
  const int lane = threadIdx.x;
  if (v >= a.n_prob || !a.active[v]) return;
  const UProb pr = a.probs[v];
  const int P = pr.P, R = pr.n_rows;
  __shared__ double H[kMaxTan][kMaxTan + 1];
  __shared__ double g[kMaxTan];
  const double* Jv = a.J + pr.j_off;
  double c = 0.0;
  for (int o = lane; o < R; o += 64) {
    const double r = a.z[pr.row0 + o] - u_ldr(a.zmeas, pr.row0 + o, a.dtype);
    c = fma(r, r, c);
  }
  c = u_wave_sum(c);
  if (a.early && sqrt(c / (double)R) <= a.rms_tol) {
    // already within the tolerance where its Jacobian was just taken: no step, stopped
    if (lane == 0) {
      a.cost[v] = c;
      a.active[v] = 0;
    }
    return;
  }
  for (int p = 0; p < P; p++) {
    double acc = 0.0;
    for (int o = lane; o < R; o += 64) {
      const double r = a.z[pr.row0 + o] - u_ldr(a.zmeas, pr.row0 + o, a.dtype);
      acc = fma(Jv[(long)o * P + p], r, acc);
    }
    acc = u_wave_sum(acc);
    if (lane == 0) g[p] = acc;
  }
  const int npair = P * (P + 1) / 2;
  for (int pi = lane; pi < npair; pi += 64) {
    int p = 0;
    while ((p + 1) * (p + 2) / 2 <= pi) p++;
    const int q = pi - p * (p + 1) / 2;
    double acc = 0.0;
    for (int o = 0; o < R; o++) acc = fma(Jv[(long)o * P + p], Jv[(long)o * P + q], acc);
    H[p][q] = acc;
    H[q][p] = acc;
  }
  __syncthreads();
  if (lane != 0) return;
  const double lam = a.lam[v];
  a.cost[v] = c;
  for (int p = 0; p < P; p++) {
    H[p][p] += lam * (H[p][p] + 1e-9);
    H[p][P] = -g[p];
  }
  for (int cI = 0; cI < P; cI++) {
    int piv = cI;
    double best = fabs(H[cI][cI]);
    for (int r = cI + 1; r < P; r++)
      if (fabs(H[r][cI]) > best) {
        best = fabs(H[r][cI]);
        piv = r;
      }
    if (piv != cI)
      for (int q = 0; q <= P; q++) {
        const double t = H[cI][q];
        H[cI][q] = H[piv][q];
        H[piv][q] = t;
      }
    for (int r = cI + 1; r < P; r++) {
      const double m = H[r][cI] / H[cI][cI];
      for (int q = cI; q <= P; q++) H[r][q] -= m * H[cI][q];
    }
  }
  double* d = g;
  for (int r = P - 1; r >= 0; r--) {
    double s = H[r][P];
    for (int q = r + 1; q < P; q++) s -= H[r][q] * d[q];
    d[r] = s / H[r][r];
  }
  for (int p = 0; p < P; p++) {
    const int k = pr.x_off + p;
    a.xt[k] = fmin(fmax(a.x[k] + d[p], a.lo[k]), a.hi[k]);
  }
}

// Accept or reject the trial of every problem (lm.py:58-61), with optional early stopping.
__global__ void __launch_bounds__(64) u_accept_kernel(ULm a) {
  const int v = blockIdx.x;
  const int lane = threadIdx.x;
  if (v >= a.n_prob || !a.active[v]) return;
  const UProb pr = a.probs[v];
  double c = 0.0;
  for (int o = lane; o < pr.n_rows; o += 64) {
    const double r = a.zt[pr.row0 + o] - u_ldr(a.zmeas, pr.row0 + o, a.dtype);
    c = fma(r, r, c);
  }
  c = u_wave_sum(c);
  const double c0 = a.cost[v];
  const bool better = c < c0;
  if (better && lane < pr.P) a.x[pr.x_off + lane] = a.xt[pr.x_off + lane];
  if (lane != 0) return;
  const double lam = a.lam[v];
  a.lam[v] = better ? lam * 0.5 : lam * 3.0;
  const double cout = better ? c : c0;
  if (better) a.cost[v] = c;
  if (a.n_done) a.n_done[v] = *a.it + 1;
  if (a.early) {
    const bool small = sqrt(cout / (double)pr.n_rows) <= a.rms_tol;
    const bool flat = better && (c0 - c) <= a.ftol * c0;
    const bool stuck = a.lam[v] > 1e12;
    if (small || flat || stuck) a.active[v] = 0;
  }
}

__global__ void u_init_kernel(ULm a, double lam0) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v == 0) *a.it = 0;
  if (v < a.n_prob) {
    a.lam[v] = lam0;
    a.active[v] = 1;
    a.cost[v] = 0.0;
    if (a.n_done) a.n_done[v] = 0;
  }
}

__global__ void u_bump_kernel(int32_t* it) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *it += 1;
}

struct DevMem {
  std::vector<void*> ptrs;
  ~DevMem() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  int alloc(T** out, size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      set_error("uneven batch: device allocation failed");
      return BMC_ENOMEM;
    }
    ptrs.push_back(p);
    *out = static_cast<T*>(p);
    return BMC_OK;
  }
};

}  // namespace bmc

using namespace bmc;

struct bmc_ubatch {
  const bmc_plan* plan = nullptr;
  int n_prob = 0, dtype = 0, R = 0, sumP = 0;
  int64_t sumJ = 0;
  std::vector<UProb> h_probs;
  DevMem mem;
  UProb* d_probs = nullptr;
  int32_t *d_kind = nullptr, *d_idx = nullptr, *d_slot = nullptr;
  double *d_scale = nullptr, *d_lo = nullptr, *d_hi = nullptr;
  const void* prm[4] = {nullptr, nullptr, nullptr, nullptr};
  const void* m0[4] = {nullptr, nullptr, nullptr, nullptr};
  // configuration (items and chunks are rebuilt by bmc_ubatch_configure)
  int PT = 1, grid = 1024, order = 1;
  UItem *d_items_j = nullptr, *d_items_p = nullptr;
  int2 *d_chunks_j = nullptr, *d_chunks_p = nullptr;
  int2* d_chunks_r = nullptr;  // the primal items, chunked for the reverse Jacobian kernel's teams
  int n_chunks_j = 0, n_chunks_p = 0, n_chunks_r = 0;
  std::vector<void*> cfg_ptrs;
  int* d_counter = nullptr;
  ~bmc_ubatch() {
    for (void* p : cfg_ptrs) (void)hipFree(p);
  }
};

namespace {

// Items of every (problem, row, pass) with `passes(P)` passes, grouped into same-class chunks of
// 64 / TEAM items. order 0: problem order; 1: classes with more pools first, then more rows first.
// per4: teams per wave of four-pool items (2 for 32-lane teams, 4 for the reverse kernel's 16-lane ones).
void build_items(const bmc_ubatch& b, int PT, int order, bool tangents, std::vector<UItem>& items,
                 std::vector<int2>& chunks, int per4 = 2) {
  std::vector<int> ids(b.n_prob);
  std::iota(ids.begin(), ids.end(), 0);
  if (order == 1)
    std::stable_sort(ids.begin(), ids.end(), [&](int x, int y) {
      const UProb &a = b.h_probs[x], &c = b.h_probs[y];
      if (a.cls != c.cls) return a.cls > c.cls;
      return (long)a.n_rows * a.P > (long)c.n_rows * c.P;
    });
  items.clear();
  chunks.clear();
  for (int cls = 3; cls >= 0; cls--) {
    const int per = (cls == 3) ? per4 : (cls == 0 ? 8 : 4);  // teams per wave: 32, 16 or 8 lanes
    std::vector<UItem> mine;
    for (int id : ids) {
      const UProb& p = b.h_probs[id];
      if (p.cls != cls) continue;
      const int n_pass = tangents ? std::max(1, (p.P + PT - 1) / PT) : 1;
      for (int q = 0; q < n_pass; q++)
        for (int r = 0; r < p.n_rows; r++) mine.push_back(UItem{id, p.row0 + r, q, 0});
    }
    for (size_t s = 0; s < mine.size(); s += per) {
      const int cnt = (int)std::min<size_t>(per, mine.size() - s);
      chunks.push_back(make_int2((int)(items.size() + s), cnt));
    }
    items.insert(items.end(), mine.begin(), mine.end());
  }
}

int upload(bmc_ubatch* b, const std::vector<UItem>& items, const std::vector<int2>& chunks, UItem** di, int2** dc) {
  void *pi = nullptr, *pc = nullptr;
  if (hipMalloc(&pi, std::max<size_t>(items.size(), 1) * sizeof(UItem)) != hipSuccess ||
      hipMalloc(&pc, std::max<size_t>(chunks.size(), 1) * sizeof(int2)) != hipSuccess) {
    if (pi) (void)hipFree(pi);
    set_error("uneven batch: device allocation failed");
    return BMC_ENOMEM;
  }
  b->cfg_ptrs.push_back(pi);
  b->cfg_ptrs.push_back(pc);
  BMC_HIP_CHECK(hipMemcpy(pi, items.data(), items.size() * sizeof(UItem), hipMemcpyHostToDevice));
  BMC_HIP_CHECK(hipMemcpy(pc, chunks.data(), chunks.size() * sizeof(int2), hipMemcpyHostToDevice));
  *di = static_cast<UItem*>(pi);
  *dc = static_cast<int2*>(pc);
  return BMC_OK;
}

int upload_chunks(bmc_ubatch* b, const std::vector<int2>& chunks, int2** dc) {
  void* pc = nullptr;
  if (hipMalloc(&pc, std::max<size_t>(chunks.size(), 1) * sizeof(int2)) != hipSuccess) {
    set_error("uneven batch: device allocation failed");
    return BMC_ENOMEM;
  }
  b->cfg_ptrs.push_back(pc);
  BMC_HIP_CHECK(hipMemcpy(pc, chunks.data(), chunks.size() * sizeof(int2), hipMemcpyHostToDevice));
  *dc = static_cast<int2*>(pc);
  return BMC_OK;
}

USimArgs sim_args(const bmc_ubatch* b, bool tangents, const double* xw, double* z_out, double* jac_out,
                  const int32_t* active, bool forward_mode) {
  USimArgs a;
  memset(&a, 0, sizeof(a));
  const Plan& p = b->plan->p;
  a.ev = p.d_ev;
  a.ops = p.d_ops;
  a.row_ptr = p.d_row_ptr;
  a.T = p.T;
  // single-event rows: the Jacobian sweep takes one item per row and every tangent at once
  const bool reverse = tangents && p.cw_rows && p.d_live && !forward_mode;
  a.reset_after_adc = p.reset_after_adc;
  a.dtype = b->dtype;
  a.PT = tangents ? b->PT : 0;
  for (int c = 0; c < 4; c++) {
    a.prm[c] = b->prm[c];
    a.m0[c] = b->m0[c];
  }
  a.probs = b->d_probs;
  a.t_kind = b->d_kind;
  a.t_idx = b->d_idx;
  a.t_slot = b->d_slot;
  a.t_scale = b->d_scale;
  a.xw = xw;
  a.items = tangents && !reverse ? b->d_items_j : b->d_items_p;
  a.chunks = reverse ? b->d_chunks_r : (tangents ? b->d_chunks_j : b->d_chunks_p);
  a.n_chunks = reverse ? b->n_chunks_r : (tangents ? b->n_chunks_j : b->n_chunks_p);
  a.row_live = reverse ? p.d_live : nullptr;
  a.counter = b->d_counter;
  a.z_out = z_out;
  a.jac_out = tangents ? jac_out : nullptr;
  a.active = active;
  a.nan_l1 = (b->dtype == BMC_F32 ? 3.925724783138660 : 5.371920351148152) * 131072.0;
  return a;
}

int launch_usim(const bmc_ubatch* b, const USimArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, st, a.counter, 0);
  BMC_HIP_CHECK(hipGetLastError());
  const dim3 grid(std::max(1, std::min(b->grid, std::max(a.n_chunks, 1)))), block(64);
  if (a.row_live) {
    hipLaunchKernelGGL(ucwj_kernel, grid, block, 0, st, a);
    BMC_HIP_CHECK(hipGetLastError());
    return BMC_OK;
  }
  switch (a.PT) {
    case 0:
      hipLaunchKernelGGL(usim_kernel<0>, grid, block, 0, st, a);
      break;
    case 1:
      hipLaunchKernelGGL(usim_kernel<1>, grid, block, 0, st, a);
      break;
    default:
      hipLaunchKernelGGL(usim_kernel<2>, grid, block, 0, st, a);
      break;
  }
  BMC_HIP_CHECK(hipGetLastError());
  return BMC_OK;
}

}  // namespace

extern "C" int bmc_ubatch_create(const bmc_plan* plan, int n_problems, int dtype, const int32_t* n_cest,
                                 const int32_t* n_params, const int32_t* row0, const int32_t* n_rows,
                                 const int32_t* slots, const double* scales, const double* lo, const double* hi,
                                 const void* const* params_dev, const void* const* m0_dev, bmc_ubatch** out) {
  if (!plan || n_problems <= 0 || !n_cest || !n_params || !row0 || !n_rows || !slots || !scales || !lo || !hi ||
      !params_dev || !m0_dev || !out || (dtype != BMC_F32 && dtype != BMC_F64)) {
    set_error("bmc_ubatch_create: bad argument");
    return BMC_EINVAL;
  }
  if (plan->p.has_keep) {  // the one-wave persistent workgroups hold no LDS slot for a kept pulse
    set_error("bmc_ubatch_create: the plan keeps pulses for reuse (compression level 2); create it with level <= 1");
    return BMC_ENOTSUP;
  }
  auto b = std::make_unique<bmc_ubatch>();
  b->plan = plan;
  b->n_prob = n_problems;
  b->dtype = dtype;
  b->R = plan->p.O;
  int local[4] = {0, 0, 0, 0};
  std::vector<int32_t> kind, idx;
  int x_off = 0;
  int64_t j_off = 0;
  for (int i = 0; i < n_problems; i++) {
    const int nc = n_cest[i], P = n_params[i];
    if (nc < 1 || nc > 4 || P < 1 || P > BMC_MAX_TANGENTS || n_rows[i] < 1 || row0[i] < 0 ||
        row0[i] + n_rows[i] > plan->p.O) {
      set_error("bmc_ubatch_create: problem " + std::to_string(i) + " has n_cest outside 1..4, P outside 1.." +
                std::to_string(BMC_MAX_TANGENTS) + " or rows outside the plan");
      return BMC_EINVAL;
    }
    if (!params_dev[nc - 1] || !m0_dev[nc - 1]) {
      set_error("bmc_ubatch_create: no parameter array for n_cest " + std::to_string(nc));
      return BMC_EINVAL;
    }
    UProb p;
    p.cls = nc - 1;
    p.local = local[nc - 1]++;
    p.P = P;
    p.row0 = row0[i];
    p.n_rows = n_rows[i];
    p.x_off = x_off;
    p.j_off = j_off;
    p.jc = 2 * (1 + nc);
    p.pad = 0;
    for (int k = 0; k < P; k++) {
      int ix = 0;
      const int kd = tan_kind_of_slot(slots[x_off + k], nc, &ix);
      if (kd < 0) {
        set_error("bmc_ubatch_create: slot outside the packed record of problem " + std::to_string(i));
        return BMC_EINVAL;
      }
      kind.push_back(kd);
      idx.push_back(ix);
    }
    b->h_probs.push_back(p);
    x_off += P;
    j_off += (int64_t)n_rows[i] * P;
  }
  b->sumP = x_off;
  b->sumJ = j_off;
  for (int c = 0; c < 4; c++) {
    b->prm[c] = params_dev[c];
    b->m0[c] = m0_dev[c];
  }
  int rc;
  if ((rc = b->mem.alloc(&b->d_probs, n_problems)) || (rc = b->mem.alloc(&b->d_kind, x_off)) ||
      (rc = b->mem.alloc(&b->d_idx, x_off)) || (rc = b->mem.alloc(&b->d_slot, x_off)) ||
      (rc = b->mem.alloc(&b->d_scale, x_off)) || (rc = b->mem.alloc(&b->d_lo, x_off)) ||
      (rc = b->mem.alloc(&b->d_hi, x_off)) || (rc = b->mem.alloc(&b->d_counter, 1)))
    return rc;
  BMC_HIP_CHECK(hipMemcpy(b->d_probs, b->h_probs.data(), n_problems * sizeof(UProb), hipMemcpyHostToDevice));
  BMC_HIP_CHECK(hipMemcpy(b->d_kind, kind.data(), x_off * sizeof(int32_t), hipMemcpyHostToDevice));
  BMC_HIP_CHECK(hipMemcpy(b->d_idx, idx.data(), x_off * sizeof(int32_t), hipMemcpyHostToDevice));
  BMC_HIP_CHECK(hipMemcpy(b->d_slot, slots, x_off * sizeof(int32_t), hipMemcpyHostToDevice));
  BMC_HIP_CHECK(hipMemcpy(b->d_scale, scales, x_off * sizeof(double), hipMemcpyHostToDevice));
  BMC_HIP_CHECK(hipMemcpy(b->d_lo, lo, x_off * sizeof(double), hipMemcpyHostToDevice));
  BMC_HIP_CHECK(hipMemcpy(b->d_hi, hi, x_off * sizeof(double), hipMemcpyHostToDevice));
  *out = b.release();
  if ((rc = bmc_ubatch_configure(*out, 1, 1024, 1)) != BMC_OK) {
    delete *out;
    *out = nullptr;
    return rc;
  }
  return BMC_OK;
}

extern "C" int bmc_ubatch_configure(bmc_ubatch* b, int tangents_per_pass, int grid, int order) {
  if (!b || tangents_per_pass < 1 || tangents_per_pass > 2 || grid < 1 || order < 0 || order > 1) {
    set_error("bmc_ubatch_configure: tangents_per_pass 1..2, grid >= 1, order 0..1");
    return BMC_EINVAL;
  }
  for (void* p : b->cfg_ptrs) (void)hipFree(p);
  b->cfg_ptrs.clear();
  b->PT = tangents_per_pass;
  b->grid = grid;
  b->order = order;
  std::vector<UItem> items;
  std::vector<int2> chunks;
  build_items(*b, b->PT, order, true, items, chunks);
  int rc = upload(b, items, chunks, &b->d_items_j, &b->d_chunks_j);
  if (rc != BMC_OK) return rc;
  b->n_chunks_j = (int)chunks.size();
  build_items(*b, b->PT, order, false, items, chunks);
  rc = upload(b, items, chunks, &b->d_items_p, &b->d_chunks_p);
  if (rc != BMC_OK) return rc;
  b->n_chunks_p = (int)chunks.size();
  build_items(*b, b->PT, order, false, items, chunks, 4);  // the same items, four 16-lane teams per wave
  rc = upload_chunks(b, chunks, &b->d_chunks_r);
  if (rc != BMC_OK) return rc;
  b->n_chunks_r = (int)chunks.size();
  return BMC_OK;
}

extern "C" int bmc_ubatch_destroy(bmc_ubatch* b) {
  delete b;
  return BMC_OK;
}

extern "C" int bmc_ubatch_sizes(const bmc_ubatch* b, int* n_rows, int* n_values, int64_t* n_jacobian) {
  if (!b) {
    set_error("bmc_ubatch_sizes: null batch");
    return BMC_EINVAL;
  }
  if (n_rows) *n_rows = b->R;
  if (n_values) *n_values = b->sumP;
  if (n_jacobian) *n_jacobian = b->sumJ;
  return BMC_OK;
}

extern "C" int bmc_ubatch_simulate_ex(bmc_ubatch* b, const double* x_dev, double* z_out_dev, double* jac_out_dev,
                                      unsigned flags, void* stream) {
  if (!b || !x_dev || !z_out_dev || (flags & ~BMC_SIM_FORWARD_JACOBIAN)) {
    set_error("bmc_ubatch_simulate: null argument or unknown flag bits");
    return BMC_EINVAL;
  }
  const USimArgs a = sim_args(b, jac_out_dev != nullptr, x_dev, z_out_dev, jac_out_dev, nullptr,
                              (flags & BMC_SIM_FORWARD_JACOBIAN) != 0);
  return launch_usim(b, a, (hipStream_t)stream);
}

extern "C" int bmc_ubatch_simulate(bmc_ubatch* b, const double* x_dev, double* z_out_dev, double* jac_out_dev,
                                   void* stream) {
  return bmc_ubatch_simulate_ex(b, x_dev, z_out_dev, jac_out_dev, 0u, stream);
}

extern "C" int bmc_ubatch_lm_fit(bmc_ubatch* b, const void* z_meas_dev, double* x_dev, int n_iter, double lam0,
                                 unsigned flags, double early_ftol, double early_rms, double* cost_dev,
                                 int32_t* n_iter_done_dev, void* stream) {
  if (!b || !z_meas_dev || !x_dev || n_iter < 0) {
    set_error("bmc_ubatch_lm_fit: bad argument");
    return BMC_EINVAL;
  }
  if (n_iter == 0) return BMC_OK;
  DevMem m;
  double *z, *J, *zt, *xt, *lam, *cost;
  int32_t *active, *it;
  int rc;
  if ((rc = m.alloc(&z, b->R)) || (rc = m.alloc(&zt, b->R)) || (rc = m.alloc(&J, b->sumJ)) ||
      (rc = m.alloc(&xt, b->sumP)) || (rc = m.alloc(&lam, b->n_prob)) || (rc = m.alloc(&cost, b->n_prob)) ||
      (rc = m.alloc(&active, b->n_prob)) || (rc = m.alloc(&it, 1)))
    return rc;
  ULm L;
  memset(&L, 0, sizeof(L));
  L.probs = b->d_probs;
  L.n_prob = b->n_prob;
  L.dtype = b->dtype;
  L.n_iter = n_iter;
  L.early = (flags & BMC_LM_EARLY_EXIT) ? 1 : 0;
  L.z = z;
  L.J = J;
  L.zt = zt;
  L.zmeas = z_meas_dev;
  L.lo = b->d_lo;
  L.hi = b->d_hi;
  L.x = x_dev;
  L.xt = xt;
  L.lam = lam;
  L.cost = cost;
  L.active = active;
  L.it = it;
  L.n_done = n_iter_done_dev;
  L.ftol = early_ftol;
  L.rms_tol = early_rms;
  const bool fwd = (flags & BMC_LM_FORWARD_JACOBIAN) != 0;
  const USimArgs sj = sim_args(b, true, x_dev, z, J, active, fwd);
  const USimArgs sp = sim_args(b, false, xt, zt, nullptr, active, fwd);

  hipStream_t st = (hipStream_t)stream;
  const bool use_graph = (flags & BMC_LM_GRAPH) != 0;
  hipStream_t own = nullptr;
  if (use_graph && st == nullptr) {  // the null stream cannot be captured
    if ((rc = stream_after_null(&own)) != BMC_OK) return rc;
    st = own;
  }
  struct OwnGuard {
    hipStream_t s;
    ~OwnGuard() {
      if (s) (void)hipStreamDestroy(s);
    }
  } own_guard{own};
  hipLaunchKernelGGL(u_init_kernel, dim3((b->n_prob + 255) / 256), dim3(256), 0, st, L, lam0);
  BMC_HIP_CHECK(hipGetLastError());
  auto iteration = [&]() -> int {
    int r;
    if ((r = launch_usim(b, sj, st)) != BMC_OK) return r;
    hipLaunchKernelGGL(u_normal_eq_kernel, dim3(b->n_prob), dim3(64), 0, st, L);
    BMC_HIP_CHECK(hipGetLastError());
    if ((r = launch_usim(b, sp, st)) != BMC_OK) return r;
    hipLaunchKernelGGL(u_accept_kernel, dim3(b->n_prob), dim3(64), 0, st, L);
    BMC_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(u_bump_kernel, dim3(1), dim3(1), 0, st, it);
    BMC_HIP_CHECK(hipGetLastError());
    return BMC_OK;
  };
  int result = BMC_OK;
  if (use_graph) {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    BMC_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int r = iteration();
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (r != BMC_OK) return r;
    if (e != hipSuccess) {
      set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      return BMC_EHIP;
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e == hipSuccess) {
      for (int i = 0; i < n_iter && e == hipSuccess; i++) e = hipGraphLaunch(exec, st);
    }
    if (e != hipSuccess) {
      set_error(std::string("hipGraph: ") + hipGetErrorString(e));
      result = BMC_EHIP;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
  } else {
    for (int i = 0; i < n_iter && result == BMC_OK; i++) result = iteration();
  }
  if (result == BMC_OK && cost_dev) {
    BMC_HIP_CHECK(hipMemcpyAsync(cost_dev, cost, b->n_prob * sizeof(double), hipMemcpyDeviceToDevice, st));
  }
  BMC_HIP_CHECK(hipStreamSynchronize(st));
  return result;
}
