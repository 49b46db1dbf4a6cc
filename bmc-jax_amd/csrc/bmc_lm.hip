// Batched bounded Levenberg-Marquardt on the GPU: BMC-JAX's levenberg_marquardt
// (src/bmc_jax/fitting/lm.py:20-67) vmapped over voxels by fit_batch (fitting/independent.py:118-128).
//
// One iteration is four launches on one stream, every buffer fixed, so the iteration can be captured
// once as a hipGraph and replayed:
//   1. sim_kernel at x with one tangent per free value  -> z [V, O], J [V, O, P]      (float64)
//   2. normal_eq_kernel: r = z - z_meas, c = r.r, g = J^T r, H = J^T J in float64, A = H + lam (diag H
//      + 1e-9 I) (+ diag(1 - free)), delta = A^-1 (-g) by LU with partial pivoting, x_t = clip(x + delta)
//   3. sim_kernel at x_t, primal only                    -> z_t [V, O]
//   4. accept_kernel: c_t = |z_t - z_meas|^2; accept iff c_t < c (lam * 0.5) else reject (lam * 3)
// The iteration counter lives on the device so a replayed graph advances it. With BMC_LM_START_CHECK a
// primal sweep at the start values first stops the problems whose start already meets the rms tolerance.
#include <math.h>
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "bmc_internal.h"

namespace bmc {

int fill_sim_args(const bmc_plan* plan, int V, int nc, int dtype, const void* params, const void* m0,
                  const void* dm0, int P, const int32_t* slot, const double* scale, int jc, SimArgs& a);

namespace {

constexpr int kMaxP = BMC_MAX_TANGENTS;

struct LmState {
  double* x;        // [V, P] current (scaled)
  double* xt;       // [V, P] trial
  double* lam;      // [V]
  double* cost;     // [V] cost at x (set by normal_eq)
  int32_t* active;  // [V]
  int32_t* jac_live;  // [V] active and moved by the last step: its J and z must be recomputed
  int32_t* it;      // [1] iteration counter
  int32_t* n_done;  // [V] or null
  double* ctraj;    // [V, n_iter] or null
  double* xtraj;    // [V, n_iter, P] or null
};

__device__ __forceinline__ double ldr(const void* p, long i, int dtype) {
  return dtype == BMC_F64 ? static_cast<const double*>(p)[i] : (double)static_cast<const float*>(p)[i];
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) v += __shfl_xor(v, m, 64);
  return v;
}

struct NeArgs {
  int V, O, P, dtype, has_free, early;
  double rms_tol;
  const double* z;
  const double* J;
  const void* zmeas;
  const double* free_;
  double lo[kMaxP], hi[kMaxP];
  LmState s;
};

// One wave per voxel. Lanes accumulate J^T J entries (pairs) and J^T r over the offsets.
__global__ void __launch_bounds__(64) normal_eq_kernel(NeArgs a) {
  const int v = blockIdx.x;
  const int lane = threadIdx.x;
  if (v >= a.V || !a.s.active[v]) return;
  const int P = a.P, O = a.O;
  __shared__ double H[kMaxP][kMaxP + 1];
  __shared__ double g[kMaxP];
  __shared__ double fr[kMaxP];
  const double* Jv = a.J + (long)v * O * P;
  const double* zv = a.z + (long)v * O;
  if (lane < P) fr[lane] = a.has_free ? a.free_[(long)v * P + lane] : 1.0;
  __syncthreads();
  // cost and J^T r
  double c = 0.0;
  for (int o = lane; o < O; o += 64) {
    double r = zv[o] - ldr(a.zmeas, (long)v * O + o, a.dtype);
    c = fma(r, r, c);
  }
  c = wave_sum(c);
  if (a.early && sqrt(c / (double)O) <= a.rms_tol) {
    // with early exit, a problem already within the tolerance at the point its Jacobian was just taken
    // at (its start, in the first iteration) takes no step and stops
    if (lane == 0) {
      a.s.cost[v] = c;
      a.s.active[v] = 0;
      a.s.jac_live[v] = 0;
    }
    return;
  }
  for (int p = 0; p < P; p++) {
    double acc = 0.0;
    for (int o = lane; o < O; o += 64) {
      double r = zv[o] - ldr(a.zmeas, (long)v * O + o, a.dtype);
      acc = fma(Jv[(long)o * P + p] * fr[p], r, acc);
    }
    acc = wave_sum(acc);
    if (lane == 0) g[p] = acc;
  }
  // J^T J, one lane per (p, q) pair, q <= p
  const int npair = P * (P + 1) / 2;
  for (int pi = lane; pi < npair; pi += 64) {
    int p = 0;
    while ((p + 1) * (p + 2) / 2 <= pi) p++;
    const int q = pi - p * (p + 1) / 2;
    double acc = 0.0;
    for (int o = 0; o < O; o++) acc = fma(Jv[(long)o * P + p], Jv[(long)o * P + q], acc);
    acc *= fr[p] * fr[q];
    H[p][q] = acc;
    H[q][p] = acc;
  }
  __syncthreads();
  if (lane != 0) return;
  const double lam = a.s.lam[v];
  a.s.cost[v] = c;
  // the damped system in place (lm.py:53-55), then Gaussian elimination with partial pivoting
  // (jnp.linalg.solve, lm.py:56)
  for (int p = 0; p < P; p++) {
    H[p][p] += lam * (H[p][p] + 1e-9);
    if (a.has_free) H[p][p] += 1.0 - fr[p];
    H[p][P] = -g[p];
  }
  for (int cI = 0; cI < P; cI++) {
    int pr = cI;
    double best = fabs(H[cI][cI]);
    for (int r = cI + 1; r < P; r++)
      if (fabs(H[r][cI]) > best) {
        best = fabs(H[r][cI]);
        pr = r;
      }
    if (pr != cI)
      for (int q = 0; q <= P; q++) {
        double t = H[cI][q];
        H[cI][q] = H[pr][q];
        H[pr][q] = t;
      }
    for (int r = cI + 1; r < P; r++) {
      double m = H[r][cI] / H[cI][cI];
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
    double xn = a.s.x[(long)v * P + p] + d[p];
    xn = fmin(fmax(xn, a.lo[p]), a.hi[p]);  // jnp.clip (lm.py:57)
    a.s.xt[(long)v * P + p] = xn;
  }
}

struct AcArgs {
  int V, O, P, dtype, n_iter, early;
  double ftol, rms_tol;
  const double* zt;
  const void* zmeas;
  LmState s;
};

__global__ void __launch_bounds__(64) accept_kernel(AcArgs a) {
  const int v = blockIdx.x;
  const int lane = threadIdx.x;
  if (v >= a.V) return;
  const int it = *a.s.it;
  if (!a.s.active[v]) {
    // a stopped problem keeps its last cost and x in the trajectory
    if (lane == 0 && a.s.ctraj && it < a.n_iter) a.s.ctraj[(long)v * a.n_iter + it] = a.s.cost[v];
    if (lane < a.P && a.s.xtraj && it < a.n_iter)
      a.s.xtraj[((long)v * a.n_iter + it) * a.P + lane] = a.s.x[(long)v * a.P + lane];
    return;
  }
  double c = 0.0;
  const double* zv = a.zt + (long)v * a.O;
  for (int o = lane; o < a.O; o += 64) {
    double r = zv[o] - ldr(a.zmeas, (long)v * a.O + o, a.dtype);
    c = fma(r, r, c);
  }
  c = wave_sum(c);
  const double c0 = a.s.cost[v];
  const bool better = c < c0;  // lm.py:58 (a NaN trial is never better)
  if (better && lane < a.P) a.s.x[(long)v * a.P + lane] = a.s.xt[(long)v * a.P + lane];
  if (lane < a.P && a.s.xtraj)
    a.s.xtraj[((long)v * a.n_iter + it) * a.P + lane] = better ? a.s.xt[(long)v * a.P + lane] : a.s.x[(long)v * a.P + lane];
  if (lane != 0) return;
  const double lam = a.s.lam[v];
  a.s.lam[v] = better ? lam * 0.5 : lam * 3.0;  // lm.py:59-60
  // a rejected step leaves x, hence z and J, unchanged: the next iteration reuses them
  a.s.jac_live[v] = better ? 1 : 0;
  const double cout = better ? c : c0;
  if (better) a.s.cost[v] = c;
  if (a.s.ctraj) a.s.ctraj[(long)v * a.n_iter + it] = cout;
  if (a.s.n_done) a.s.n_done[v] = it + 1;
  if (a.early) {
    const bool small = sqrt(cout / (double)a.O) <= a.rms_tol;
    const bool flat = better && (c0 - c) <= a.ftol * c0;
    const bool stuck = a.s.lam[v] > 1e12;
    if (small || flat || stuck) {
      a.s.active[v] = 0;
      a.s.jac_live[v] = 0;
    }
  }
}

// The start check (BMC_LM_START_CHECK): cost at the start values from a primal sweep; a problem whose
// rms is already at most rms_tol is stopped before its first Jacobian (it keeps x0, 0 iterations).
__global__ void __launch_bounds__(64) start_kernel(AcArgs a) {
  const int v = blockIdx.x;
  const int lane = threadIdx.x;
  if (v >= a.V || !a.s.active[v]) return;
  double c = 0.0;
  const double* zv = a.zt + (long)v * a.O;
  for (int o = lane; o < a.O; o += 64) {
    double r = zv[o] - ldr(a.zmeas, (long)v * a.O + o, a.dtype);
    c = fma(r, r, c);
  }
  c = wave_sum(c);
  if (lane != 0) return;
  a.s.cost[v] = c;
  if (sqrt(c / (double)a.O) <= a.rms_tol) {
    a.s.active[v] = 0;
    a.s.jac_live[v] = 0;
  }
}

// list <- indices v with flag[v] != 0 (any order), count <- their number (count cleared beforehand).
// One atomic per wave: the wave's hits are ranked by ballot.
__global__ void __launch_bounds__(256) compact_kernel(const int32_t* flag, int V, int32_t* list, int32_t* count) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  const bool f = v < V && flag[v] != 0;
  const unsigned long long m = __ballot(f);
  const int lane = threadIdx.x % 64;
  int base = 0;
  if (lane == 0 && m) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, 0, 64);
  if (f) list[base + __popcll(m & ((1ull << lane) - 1ull))] = v;
}

__global__ void clear2_kernel(int32_t* c) {
  if (threadIdx.x < 2 && blockIdx.x == 0) c[threadIdx.x] = 0;
}

__global__ void bump_iter_kernel(int32_t* it) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *it += 1;
}

__global__ void init_kernel(LmState s, int V, double lam0) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v == 0) *s.it = 0;
  if (v < V) {
    s.lam[v] = lam0;
    s.active[v] = 1;
    s.jac_live[v] = 1;
    s.cost[v] = 0.0;
    if (s.n_done) s.n_done[v] = 0;
  }
}

// The work buffers of one fit, carved from one block of device memory (each part 256-byte aligned).
struct LmWork {
  double *z, *J, *zt, *xt, *lam, *cost;
  int32_t *active, *jac_live, *it, *list_j, *list_a, *cnt;
};

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

size_t lm_work_bytes(int V, int O, int P) {
  const size_t v = (size_t)V, o = (size_t)O, p = (size_t)P;
  return align256(8 * v * o) + align256(8 * v * o * p) + align256(8 * v * o) + align256(8 * v * p) +
         2 * align256(8 * v) + 4 * align256(4 * v) + align256(4) + align256(8);
}

void lm_carve(char* base, int V, int O, int P, LmWork& w) {
  const size_t v = (size_t)V, o = (size_t)O, p = (size_t)P;
  auto take = [&](size_t n) {
    char* q = base;
    base += align256(n);
    return q;
  };
  w.z = (double*)take(8 * v * o);
  w.J = (double*)take(8 * v * o * p);
  w.zt = (double*)take(8 * v * o);
  w.xt = (double*)take(8 * v * p);
  w.lam = (double*)take(8 * v);
  w.cost = (double*)take(8 * v);
  w.active = (int32_t*)take(4 * v);
  w.jac_live = (int32_t*)take(4 * v);
  w.list_j = (int32_t*)take(4 * v);
  w.list_a = (int32_t*)take(4 * v);
  w.it = (int32_t*)take(4);
  w.cnt = (int32_t*)take(8);
}

// One fit on a carved workspace: every launch on `stream`, in stream order, no allocation and no
// synchronisation (the caller owns the workspace's lifetime). Every buffer a kernel reads is written
// earlier in the same call (init_kernel, clear2_kernel, the sweeps), so the workspace's contents on
// entry do not matter.
int lm_run(const bmc_plan* plan, int V, int n_cest, int dtype, const void* params_dev, const void* m0_dev, int P,
           const int32_t* slot, const double* scale, const double* lo, const double* hi, int jac_component,
           const void* z_meas_dev, double* x_dev, int n_iter, double lam0, const double* free_dev, unsigned flags,
           double early_ftol, double early_rms, double* cost_traj_dev, double* x_traj_dev, int32_t* n_iter_done_dev,
           const LmWork& w, hipStream_t stream) {
  SimArgs sj;
  int rc = fill_sim_args(plan, V, n_cest, dtype, params_dev, m0_dev, nullptr, P, slot, scale, jac_component, sj);
  if (rc != BMC_OK) return rc;
  const int O = plan->p.O;
  LmState st;
  st.x = x_dev;
  st.xt = w.xt;
  st.lam = w.lam;
  st.cost = w.cost;
  st.active = w.active;
  st.jac_live = w.jac_live;
  st.it = w.it;
  st.n_done = n_iter_done_dev;
  st.ctraj = cost_traj_dev;
  st.xtraj = x_traj_dev;

  // launch 1: sim with tangents at x
  sj.xw = x_dev;
  sj.w_keep = free_dev;  // a held value is not written: the record's own number stays
  sj.n_write = P;
  for (int p = 0; p < P; p++) {
    sj.w_slot[p] = slot[p];
    sj.w_scale[p] = scale[p];
  }
  sj.z_out = w.z;
  sj.jac_out = w.J;
  sj.jac_dtype = BMC_F64;
  // each sweep runs over a compacted list of the voxels it has to simulate
  int32_t* cnt = w.cnt;
  sj.active = nullptr;
  sj.vlist = w.list_j;
  sj.n_vlist = cnt;
  sj.PT = tangents_per_pass(P);
  sj.n_pass = (P + sj.PT - 1) / sj.PT;
  // launch 3: primal at the trial point
  SimArgs sp = sj;
  sp.vlist = w.list_a;
  sp.n_vlist = cnt + 1;
  sp.xw = st.xt;
  sp.z_out = w.zt;
  sp.jac_out = nullptr;
  sp.PT = 0;
  sp.n_pass = 1;

  NeArgs ne;
  memset(&ne, 0, sizeof(ne));
  ne.V = V;
  ne.O = O;
  ne.P = P;
  ne.dtype = dtype;
  ne.has_free = free_dev != nullptr;
  ne.early = (flags & BMC_LM_EARLY_EXIT) ? 1 : 0;
  ne.rms_tol = early_rms;
  ne.z = w.z;
  ne.J = w.J;
  ne.zmeas = z_meas_dev;
  ne.free_ = free_dev;
  for (int p = 0; p < P; p++) {
    ne.lo[p] = lo[p];
    ne.hi[p] = hi[p];
  }
  ne.s = st;
  AcArgs ac;
  memset(&ac, 0, sizeof(ac));
  ac.V = V;
  ac.O = O;
  ac.P = P;
  ac.dtype = dtype;
  ac.n_iter = n_iter;
  ac.early = (flags & BMC_LM_EARLY_EXIT) ? 1 : 0;
  ac.ftol = early_ftol;
  ac.rms_tol = early_rms;
  ac.zt = w.zt;
  ac.zmeas = z_meas_dev;
  ac.s = st;

  // single-event rows take the Jacobian by reverse accumulation through the squarings
  const bool cwj = use_cwj(plan->p, sj, (flags & BMC_LM_FORWARD_JACOBIAN) != 0);
  hipLaunchKernelGGL(init_kernel, dim3((V + 255) / 256), dim3(256), 0, stream, st, V, lam0);
  BMC_HIP_CHECK(hipGetLastError());
  if ((flags & BMC_LM_START_CHECK) && (flags & BMC_LM_EARLY_EXIT)) {
    SimArgs s0 = sp;  // primal at the start values of every problem
    s0.xw = x_dev;
    hipLaunchKernelGGL(clear2_kernel, dim3(1), dim3(64), 0, stream, cnt);
    hipLaunchKernelGGL(compact_kernel, dim3((V + 255) / 256), dim3(256), 0, stream, (const int32_t*)st.active, V,
                       w.list_a, cnt + 1);
    BMC_HIP_CHECK(hipGetLastError());
    if ((rc = launch_sim(n_cest, s0, stream)) != BMC_OK) return rc;
    hipLaunchKernelGGL(start_kernel, dim3(V), dim3(64), 0, stream, ac);
    BMC_HIP_CHECK(hipGetLastError());
  }

  auto iteration = [&]() -> int {
    int r;
    hipLaunchKernelGGL(clear2_kernel, dim3(1), dim3(64), 0, stream, cnt);
    hipLaunchKernelGGL(compact_kernel, dim3((V + 255) / 256), dim3(256), 0, stream, (const int32_t*)st.jac_live, V,
                       w.list_j, cnt);
    BMC_HIP_CHECK(hipGetLastError());
    if ((r = cwj ? launch_cwj(n_cest, sj, plan->p.d_live, stream) : launch_sim(n_cest, sj, stream)) != BMC_OK)
      return r;
    hipLaunchKernelGGL(normal_eq_kernel, dim3(V), dim3(64), 0, stream, ne);
    BMC_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(compact_kernel, dim3((V + 255) / 256), dim3(256), 0, stream, (const int32_t*)st.active, V,
                       w.list_a, cnt + 1);
    BMC_HIP_CHECK(hipGetLastError());
    if ((r = launch_sim(n_cest, sp, stream)) != BMC_OK) return r;
    hipLaunchKernelGGL(accept_kernel, dim3(V), dim3(64), 0, stream, ac);
    BMC_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(bump_iter_kernel, dim3(1), dim3(64), 0, stream, st.it);
    BMC_HIP_CHECK(hipGetLastError());
    return BMC_OK;
  };

  if (!(flags & BMC_LM_GRAPH)) {
    for (int it = 0; it < n_iter; it++)
      if ((rc = iteration()) != BMC_OK) return rc;
    return BMC_OK;
  }
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  BMC_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
  rc = iteration();
  hipError_t ce = hipStreamEndCapture(stream, &graph);
  if (rc != BMC_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (ce != hipSuccess) {
    set_error(std::string("bmc_lm_fit: graph capture: ") + hipGetErrorString(ce));
    return BMC_EHIP;
  }
  hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(graph);
    set_error(std::string("bmc_lm_fit: graph instantiate: ") + hipGetErrorString(ie));
    return BMC_EHIP;
  }
  hipError_t le = hipSuccess;
  for (int it = 0; it < n_iter && le == hipSuccess; it++) le = hipGraphLaunch(exec, stream);
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  if (le != hipSuccess) {
    set_error(std::string("bmc_lm_fit: graph launch: ") + hipGetErrorString(le));
    return BMC_EHIP;
  }
  return BMC_OK;
}

int lm_check(int P, const double* lo, const double* hi, const double* x_dev, const void* z_meas_dev, int n_iter,
             const char* who) {
  if (P <= 0 || P > kMaxP || !lo || !hi || !x_dev || !z_meas_dev || n_iter < 0) {
    set_error(std::string(who) + ": need 1.." + std::to_string(kMaxP) + " free values, bounds, x and z_meas");
    return BMC_EINVAL;
  }
  return BMC_OK;
}

}  // namespace

}  // namespace bmc

using namespace bmc;

extern "C" int bmc_lm_fit(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
                          const void* m0_dev, int n_params, const int32_t* slot, const double* scale, const double* lo,
                          const double* hi, int jac_component, const void* z_meas_dev, double* x_dev, int n_iter,
                          double lam0, const double* free_dev, unsigned flags, double early_ftol, double early_rms,
                          double* cost_traj_dev, double* x_traj_dev, int32_t* n_iter_done_dev, void* stream_) {
  const int V = n_voxels, P = n_params;
  int rc = lm_check(P, lo, hi, x_dev, z_meas_dev, n_iter, "bmc_lm_fit");
  if (rc != BMC_OK) return rc;
  SimArgs probe;  // argument checks before anything is allocated
  if ((rc = fill_sim_args(plan, V, n_cest, dtype, params_dev, m0_dev, nullptr, P, slot, scale, jac_component,
                          probe)) != BMC_OK)
    return rc;
  if (V == 0 || n_iter == 0) return BMC_OK;
  hipStream_t stream = (hipStream_t)stream_;
  hipStream_t own = nullptr;
  if ((flags & BMC_LM_GRAPH) && stream == nullptr) {  // the null stream cannot be captured
    if ((rc = stream_after_null(&own)) != BMC_OK) return rc;
    stream = own;
  }
  struct Guard {
    hipStream_t s = nullptr;
    void* mem = nullptr;
    ~Guard() {
      if (mem) (void)hipFree(mem);
      if (s) (void)hipStreamDestroy(s);
    }
  } guard;
  guard.s = own;
  const size_t bytes = lm_work_bytes(V, plan->p.O, P);
  if (hipMalloc(&guard.mem, bytes) != hipSuccess) {
    guard.mem = nullptr;
    set_error("bmc_lm_fit: device allocation of " + std::to_string(bytes) + " bytes failed");
    return BMC_ENOMEM;
  }
  LmWork w;
  lm_carve((char*)guard.mem, V, plan->p.O, P, w);
  rc = lm_run(plan, V, n_cest, dtype, params_dev, m0_dev, P, slot, scale, lo, hi, jac_component, z_meas_dev, x_dev,
              n_iter, lam0, free_dev, flags, early_ftol, early_rms, cost_traj_dev, x_traj_dev, n_iter_done_dev, w,
              stream);
  // the workspace is freed on return: finish the work that uses it first
  hipError_t se = hipStreamSynchronize(stream);
  if (rc != BMC_OK) return rc;
  if (se != hipSuccess) {
    set_error(std::string("bmc_lm_fit: hipStreamSynchronize: ") + hipGetErrorString(se));
    return BMC_EHIP;
  }
  return BMC_OK;
}

// ---- the allocation-free variant -----------------------------------------------------------------
// Its workspace is this file-scope device array. Each device holds its own copy from the moment the
// library's code object is loaded on it until the process ends. Host code never passes `scratch`
// itself (on the host that name is only a stand-in for the symbol): it takes the device address from
// hipGetSymbolAddress once per device and passes that.
static __device__ __align__(256) unsigned char scratch[BMC_LM_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call on that device.
int scratch_of_current_device(void** out) {
  int dev = -1;
  BMC_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("bmc_lm_fit_noalloc: device index " + std::to_string(dev) + " beyond the library's table");
    return BMC_ENOTSUP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("bmc_lm_fit_noalloc: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return BMC_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return BMC_OK;
}

}  // namespace

extern "C" int bmc_lm_fit_noalloc(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
                                  const void* m0_dev, int n_params, const int32_t* slot, const double* scale,
                                  const double* lo, const double* hi, int jac_component, const void* z_meas_dev,
                                  double* x_dev, int n_iter, double lam0, const double* free_dev, unsigned flags,
                                  double early_ftol, double early_rms, double* cost_traj_dev, double* x_traj_dev,
                                  int32_t* n_iter_done_dev, void* stream) {
  const int V = n_voxels, P = n_params;
  int rc = lm_check(P, lo, hi, x_dev, z_meas_dev, n_iter, "bmc_lm_fit_noalloc");
  if (rc != BMC_OK) return rc;
  if (flags & BMC_LM_GRAPH) {
    set_error("bmc_lm_fit_noalloc: BMC_LM_GRAPH is refused (instantiating a graph allocates); run without it");
    return BMC_EINVAL;
  }
  SimArgs probe;
  if ((rc = fill_sim_args(plan, V, n_cest, dtype, params_dev, m0_dev, nullptr, P, slot, scale, jac_component,
                          probe)) != BMC_OK)
    return rc;
  const size_t bytes = lm_work_bytes(V, plan->p.O, P);
  if (bytes > (size_t)BMC_LM_NOALLOC_SCRATCH_BYTES) {
    set_error("bmc_lm_fit_noalloc: " + std::to_string(V) + " voxels x " + std::to_string(plan->p.O) + " offsets x " +
              std::to_string(P) + " values need " + std::to_string(bytes) + " bytes of workspace, above the fixed " +
              std::to_string((size_t)BMC_LM_NOALLOC_SCRATCH_BYTES) + "; split the voxels or call bmc_lm_fit");
    return BMC_EINVAL;
  }
  if (V == 0 || n_iter == 0) return BMC_OK;
  void* base = nullptr;
  if ((rc = scratch_of_current_device(&base)) != BMC_OK) return rc;
  LmWork w;
  lm_carve((char*)base, V, plan->p.O, P, w);
  return lm_run(plan, V, n_cest, dtype, params_dev, m0_dev, P, slot, scale, lo, hi, jac_component, z_meas_dev, x_dev,
                n_iter, lam0, free_dev, flags, early_ftol, early_rms, cost_traj_dev, x_traj_dev, n_iter_done_dev, w,
                (hipStream_t)stream);
}
