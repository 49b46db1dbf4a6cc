// Internal declarations shared by the HIP translation units of libbmc_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "bmc_hip.h"

namespace bmc {

// One event, widened to float64 on upload. cos/sin of the phase are taken on the host once.
struct Ev {
  double dt, amp, cph, sph, fr;
  int32_t type, pad;
};

// One instruction of a row program.
//   kind 0 (SEQ):   events [e0, e0 + len) applied one by one to the state.
//   kind 1 (BLOCK): the propagator of events [e0, e0 + len) is chained once, then applied `count`
//                   times; the block holds no ADC.
//   kind 2 (KEEP):  as BLOCK, and the chained propagator is also kept (the team's LDS slot) for later
//                   REUSE ops of the same row.
//   kind 3 (REUSE): events [e0, e0 + len) equal the kept block up to a constant RF phase offset d
//                   (every sample's phase shifted by d): their propagator is R(d) P R(d)^T, with P the
//                   kept one and R(d) the rotation by d about z of every pool's (x, y). Applied once
//                   as M <- R(d) P R(-d) M; the op's (cos d, sin d) is in the plan's rotation table.
// `half` (BLOCK and KEEP ops): when h > 0, events [e0, e0 + 2h) read the same backwards as forwards
// (bit for bit), are RF or delay events only, and every RF event among them with a nonzero amplitude
// has the same phase phi, whose (cos phi, sin phi) is the op's entry in the rotation table. Kernels
// with a spare lane per team may then chain only the first h events (the run's palindrome
// factorisation); the others chain all of them as usual.
struct Op {
  int32_t kind, e0, len, count, half;
};

enum { OP_SEQ = 0, OP_BLOCK = 1, OP_KEEP = 2, OP_REUSE = 3 };

// Tangent kinds: how the generator depends on a packed slot (matrix.py:83-295 differentiated).
enum TanKind {
  TK_NONE = 0,
  TK_W_R1,
  TK_W_R2,
  TK_MT_K,
  TK_MT_F,
  TK_MT_R1,
  TK_MT_DW,
  TK_MT_R2,
  TK_B0,
  TK_GAMMA,
  TK_B0INH,
  TK_RELB1,
  TK_C_K,
  TK_C_F,
  TK_C_DW,
  TK_C_R1,
  TK_C_R2
};

constexpr int kMaxTan = BMC_MAX_TANGENTS;

struct SimArgs {
  const Ev* ev;           // [O, T]
  const Op* ops;          // program
  const int32_t* row_ptr; // [O + 1]
  int O, T, reset_after_adc;
  int V, dtype, stride;
  const void* params;     // [V, stride] dtype
  const void* m0;         // [V, N] dtype
  const void* dm0;        // [V, P, N] dtype or null
  // values written into slots before simulating (fit_batch's write_params): [V, n_write] float64
  const double* xw;
  // [V, n_write] float64 or null: where 0, the slot keeps the record's own value (a held entry of a
  // partly free fit, whose model uses the carried value exactly: fitting/joint.py:292-294)
  const double* w_keep;
  int n_write;
  int w_slot[kMaxTan];
  double w_scale[kMaxTan];
  // tangents
  int P, PT, n_pass;
  int t_kind[kMaxTan], t_idx[kMaxTan];
  double t_scale[kMaxTan];
  int jc;
  void* m_out;            // [V, O, N] dtype or null
  double* z_out;          // [V, O] float64 or null: state[jc]
  void* jac_out;          // [V, O, P] (jac_dtype) or null
  int jac_dtype;
  const int32_t* active;  // [V] or null: skip voxels whose flag is 0
  double nan_l1;          // reference NaN threshold on the augmented 1-norm
  void* traj;             // [V, O, T, N] dtype or null: state after every event (SEQ-only plans)
  long team0;             // first team of this launch (a launch covers at most 2^31 threads)
  const int32_t* vlist;   // [<= V] or null: simulate only these voxels (compacted list)
  const int32_t* n_vlist; // device count of vlist
  int t_off;              // first tangent of this launch's pass 0 (a launch may carry the last pass alone)
  const double* rot;      // [n_ops, 2] (cos d, sin d) of REUSE ops
  int keep;               // the plan holds KEEP ops: the launch gives every team an LDS slot for one propagator
  int mirror;             // the plan holds chained ops with a palindromic half (and no KEEP ops)
};

struct Plan {
  int O = 0, T = 0, reset_after_adc = 1;
  Ev* d_ev = nullptr;
  Op* d_ops = nullptr;
  int32_t* d_row_ptr = nullptr;
  double* d_rot = nullptr;  // [n_ops, 2]
  int32_t* d_live = nullptr;  // [O] the one propagating event of a single-event row (-1: none)
  int64_t n_ops = 0, n_built = 0, n_applied = 0;
  int compressed = 1;
  int has_keep = 0;
  int has_mirror = 0;  // some BLOCK / KEEP op has a palindromic half
  int cw_rows = 0;  // every row propagates at most one event (a CW block) and ends at its readout
  std::vector<Op> h_ops;          // host copy of the program
  std::vector<int32_t> h_row_ptr; // host copy of the row offsets
};

void set_error(const std::string& msg);
int tan_kind_of_slot(int slot, int nc, int* idx);
int launch_sim(int nc, const SimArgs& a, hipStream_t stream);
// The Jacobian of single-event rows by reverse accumulation through the squarings (bmc_cwj.hip):
// use_cwj says whether a tangent launch of this plan can take it (no M0 seeds, no trajectory, and the
// caller did not ask for forward mode with BMC_SIM_FORWARD_JACOBIAN / BMC_LM_FORWARD_JACOBIAN),
// launch_cwj runs it.
bool use_cwj(const Plan& p, const SimArgs& a, bool forward_mode);
int launch_cwj(int nc, const SimArgs& a, const int32_t* row_live, hipStream_t stream);
// Tangents carried per simulate pass for P tangents (register budget of the kernel).
int tangents_per_pass(int P);
// A stream of the library's own for work a caller issued on the null stream but that has to run on a
// capturable stream (a graph-captured LM iteration): non-blocking, and made to wait for everything
// already queued on the null stream, so that the caller's earlier writes to its buffers (and any
// earlier use of memory its allocator hands out again) complete first. The caller synchronises it
// before returning.
int stream_after_null(hipStream_t* out);

}  // namespace bmc

struct bmc_plan {
  bmc::Plan p;
};

#define BMC_HIP_CHECK(expr)                                                        \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess) {                                                        \
      bmc::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));          \
      return BMC_EHIP;                                                             \
    }                                                                              \
  } while (0)
