// C ABI of libbmc_hip.so: sequence plans (host-compiled row programs) and the simulate entry point.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "bmc_internal.h"

namespace bmc {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

namespace {

struct EvKey {
  int32_t type;
  double dt, amp, ph, fr;
  bool operator==(const EvKey& o) const {
    return type == o.type && memcmp(&dt, &o.dt, 8) == 0 && memcmp(&amp, &o.amp, 8) == 0 &&
           memcmp(&ph, &o.ph, 8) == 0 && memcmp(&fr, &o.fr, 8) == 0;
  }
};

struct EvKeyHash {
  size_t operator()(const EvKey& k) const {
    uint64_t h = 1469598103934665603ull ^ (uint64_t)k.type;
    const double* d[4] = {&k.dt, &k.amp, &k.ph, &k.fr};
    for (int i = 0; i < 4; i++) {
      uint64_t b;
      memcpy(&b, d[i], 8);
      h = (h ^ b) * 1099511628211ull;
    }
    return (size_t)h;
  }
};

// Pulses of one row that recur with their RF phase shifted: maximal runs of RF events, grouped when
// they agree event by event in type, dt, amplitude and frequency (bit-exact) and their phases differ
// by one constant d on every sample with a nonzero amplitude (to kPhaseTol rad). The propagator of a
// member is then R(d) P R(d)^T with P that of the group's first member (R: rotation about z of every
// pool's (x, y)), which the kernel applies without rebuilding it.
constexpr double kPhaseTol = 1e-12;

struct PulseInfo {
  int a = 0, b = 0;       // events [a, b)
  int ref = 0;            // first sample with a nonzero amplitude (a if none)
  int group = -1;
  bool has_later = false; // a later member of the group starts after this pulse
};

std::vector<PulseInfo> find_pulses(const std::vector<EvKey>& row) {
  const int T = (int)row.size();
  std::vector<PulseInfo> pulses;
  for (int e = 0; e < T;) {
    if (row[e].type != BMC_EV_RF) {
      e++;
      continue;
    }
    PulseInfo p;
    p.a = e;
    while (e < T && row[e].type == BMC_EV_RF) e++;
    p.b = e;
    p.ref = p.a;
    for (int t = p.a; t < p.b; t++)
      if (row[t].amp != 0.0) {
        p.ref = t;
        break;
      }
    pulses.push_back(p);
  }
  // bucket by a hash of the phase-free contents, then compare phases within a bucket
  auto same_shape = [&](const PulseInfo& x, const PulseInfo& y) {
    if (x.b - x.a != y.b - y.a || x.ref - x.a != y.ref - y.a) return false;
    for (int t = 0; t < x.b - x.a; t++) {
      const EvKey &u = row[x.a + t], &v = row[y.a + t];
      if (u.type != v.type || memcmp(&u.dt, &v.dt, 8) || memcmp(&u.amp, &v.amp, 8) || memcmp(&u.fr, &v.fr, 8))
        return false;
    }
    const double d = row[x.ref].ph - row[y.ref].ph;
    for (int t = 0; t < x.b - x.a; t++) {
      const EvKey &u = row[x.a + t], &v = row[y.a + t];
      if (u.amp != 0.0 && !(fabs((u.ph - v.ph) - d) <= kPhaseTol)) return false;
    }
    return true;
  };
  std::unordered_map<uint64_t, std::vector<int>> buckets;  // hash -> group leaders (pulse indices)
  int n_groups = 0;
  std::vector<int> leader_group;
  for (int i = 0; i < (int)pulses.size(); i++) {
    uint64_t h = 1469598103934665603ull ^ (uint64_t)(pulses[i].b - pulses[i].a);
    for (int t = pulses[i].a; t < pulses[i].b; t++) {
      uint64_t w[3];
      memcpy(&w[0], &row[t].dt, 8);
      memcpy(&w[1], &row[t].amp, 8);
      memcpy(&w[2], &row[t].fr, 8);
      for (int k = 0; k < 3; k++) h = (h ^ w[k]) * 1099511628211ull;
    }
    std::vector<int>& lead = buckets[h];
    for (int l : lead)
      if (same_shape(pulses[i], pulses[l])) {
        pulses[i].group = pulses[l].group;
        break;
      }
    if (pulses[i].group < 0) {
      pulses[i].group = n_groups++;
      lead.push_back(i);
    }
  }
  std::vector<int> last_start(n_groups, -1);
  for (const PulseInfo& p : pulses) last_start[p.group] = std::max(last_start[p.group], p.a);
  for (PulseInfo& p : pulses) p.has_later = last_start[p.group] > p.a;
  return pulses;
}

// Shortest half worth factorising: the kernel's recombination costs about one chained event.
constexpr int kMinMirrorHalf = 4;
// Mirrored RF samples may differ in amplitude by this relative amount (a symmetric shape sampled in
// floating point, e.g. exp(-((i + 0.5) dt - T / 2)^2 ...), is symmetric to an ulp or two, not bit for
// bit); the first half's values are used for both. The effect on a state is of the same relative size.
constexpr double kMirrorAmpTol = 1e-13;

bool near_mirror(const EvKey& a, const EvKey& b) {
  return a.type == b.type && memcmp(&a.dt, &b.dt, 8) == 0 && memcmp(&a.ph, &b.ph, 8) == 0 &&
         memcmp(&a.fr, &b.fr, 8) == 0 && fabs(a.amp - b.amp) <= kMirrorAmpTol * fmax(fabs(a.amp), fabs(b.amp));
}

// The longest even palindromic prefix [e0, e0 + 2h) of the chained run [e0, e0 + len): RF and delay
// events only, each equal to its mirror image (near_mirror), every RF event with a nonzero amplitude at
// one phase (bit for bit), which goes to *phase. Returns h, or 0 below kMinMirrorHalf.
int mirror_half(const std::vector<EvKey>& row, int e0, int len, double* phase) {
  int lim = 0;  // the prefix of allowed events at one phase
  bool have_ph = false;
  double ph = 0.0;
  for (; lim < len; lim++) {
    const EvKey& k = row[e0 + lim];
    if (k.type != BMC_EV_RF && k.type != BMC_EV_DELAY) break;
    if (k.type == BMC_EV_RF && k.amp != 0.0) {
      if (have_ph && memcmp(&k.ph, &ph, 8) != 0) break;
      ph = k.ph;
      have_ph = true;
    }
  }
  for (int L = lim & ~1; L >= 2 * kMinMirrorHalf; L -= 2) {
    if (!near_mirror(row[e0 + L - 1], row[e0])) continue;
    bool pal = true;
    for (int t = 1; t < L / 2 && pal; t++) pal = near_mirror(row[e0 + t], row[e0 + L - 1 - t]);
    if (pal) {
      *phase = ph;
      return L / 2;
    }
  }
  return 0;
}

// Greedy run detection for one row: a block [i, i + p) repeated c >= 2 times back to back, with no
// ADC inside, becomes one BLOCK op. With level >= 2, a pulse that is not part of such a run and
// recurs later with a shifted phase becomes KEEP (chained, kept), and its later members REUSE ops
// with their rotation (cos d, sin d) appended to `rot`. Everything else is SEQ. Candidate periods come
// from the next few occurrences of event i's exact values.
void compile_row(const std::vector<EvKey>& row, int level, std::vector<Op>& ops, std::vector<double>& rot,
                 int64_t& built, int64_t& applied) {
  const int T = (int)row.size();
  std::unordered_map<EvKey, std::vector<int>, EvKeyHash> pos;
  for (int e = 0; e < T; e++) pos[row[e]].push_back(e);
  std::vector<int> next_adc(T + 1, T);
  for (int e = T - 1; e >= 0; e--) next_adc[e] = (row[e].type == BMC_EV_ADC) ? e : next_adc[e + 1];
  std::vector<PulseInfo> pulses;
  std::vector<int> pulse_at(T, -1);  // pulse index starting at event e
  if (level >= 2) {
    pulses = find_pulses(row);
    for (int i = 0; i < (int)pulses.size(); i++) pulse_at[pulses[i].a] = i;
  }
  int kept_group = -1;
  double kept_phase = 0.0;
  auto push = [&](Op op, double c, double sn) {
    ops.push_back(op);
    rot.push_back(c);
    rot.push_back(sn);
  };
  // with level 3, a chained op whose events open with a palindrome records its half and phase (cos, sin)
  auto push_chained = [&](Op op) {
    double ph = 0.0;
    op.half = level >= 3 ? mirror_half(row, op.e0, op.len, &ph) : 0;
    push(op, op.half ? cos(ph) : 1.0, op.half ? sin(ph) : 0.0);
  };
  auto flush_seq = [&](int s0, int s1) {
    if (s1 > s0) {
      push(Op{OP_SEQ, s0, s1 - s0, 1, 0}, 1.0, 0.0);
      built += s1 - s0;
      applied += s1 - s0;
    }
  };
  int seq_start = 0, i = 0;
  while (i < T) {
    int best_p = 0, best_c = 0;
    const std::vector<int>& occ = pos[row[i]];
    auto it = std::upper_bound(occ.begin(), occ.end(), i);
    for (int tries = 0; it != occ.end() && tries < 8; ++it, ++tries) {
      const int p = *it - i;
      if (next_adc[i] < i + p) break;  // the block would hold an ADC
      int c = 1;
      while (i + (c + 1) * p <= T) {
        bool same = true;
        for (int t = 0; t < p && same; t++) same = row[i + c * p + t] == row[i + t];
        if (!same) break;
        c++;
      }
      if (c >= 2 && (long)c * p > (long)best_c * best_p) {
        best_p = p;
        best_c = c;
      }
    }
    if (best_c >= 2) {
      flush_seq(seq_start, i);
      push_chained(Op{OP_BLOCK, i, best_p, best_c, 0});
      built += best_p;
      applied += best_c;
      i += best_p * best_c;
      seq_start = i;
      continue;
    }
    const int pi = pulse_at[i];
    if (pi >= 0) {
      const PulseInfo& pu = pulses[pi];
      if (pu.group == kept_group) {
        flush_seq(seq_start, i);
        const double d = row[pu.ref].ph - kept_phase;
        push(Op{OP_REUSE, pu.a, pu.b - pu.a, 1, 0}, cos(d), sin(d));
        applied += 1;
        i = seq_start = pu.b;
        continue;
      }
      if (pu.has_later) {
        flush_seq(seq_start, i);
        push_chained(Op{OP_KEEP, pu.a, pu.b - pu.a, 1, 0});
        kept_group = pu.group;
        kept_phase = row[pu.ref].ph;
        built += pu.b - pu.a;
        applied += 1;
        i = seq_start = pu.b;
        continue;
      }
    }
    i++;
  }
  flush_seq(seq_start, T);
}

}  // namespace
}  // namespace bmc

using namespace bmc;

extern "C" {

int bmc_abi_version(void) { return BMC_ABI_VERSION; }

const char* bmc_last_error(void) { return g_err.c_str(); }

int bmc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int bmc_plan_create(int n_offsets, int n_events, const int32_t* type, const double* dt, const double* rf_amp,
                    const double* rf_phase, const double* rf_freq, int reset_after_adc, bmc_plan** out) {
  return bmc_plan_create_ex(n_offsets, n_events, type, dt, rf_amp, rf_phase, rf_freq, reset_after_adc, 3, out);
}

}  // extern "C"

namespace bmc {
namespace {

// The row programs of an event table (host only). Returns BMC_OK or BMC_EINVAL.
int compile_plan(int O, int T, const int32_t* type, const double* dt, const double* rf_amp, const double* rf_phase,
                 const double* rf_freq, int level, std::vector<Ev>* ev, std::vector<Op>& ops, std::vector<double>& rot,
                 std::vector<int32_t>& row_ptr, int64_t& built, int64_t& applied) {
  if (O <= 0 || T < 0 || (T > 0 && (!type || !dt || !rf_amp || !rf_phase || !rf_freq)) || level < 0 || level > 3) {
    set_error("bmc_plan_create: bad arguments");
    return BMC_EINVAL;
  }
  if (ev) ev->assign((size_t)O * T, Ev{});
  ops.clear();
  rot.clear();
  row_ptr.assign(O + 1, 0);
  built = applied = 0;
  std::vector<EvKey> row(T);
  for (int o = 0; o < O; o++) {
    for (int e = 0; e < T; e++) {
      const size_t k = (size_t)o * T + e;
      if (type[k] < 0 || type[k] > 4 || !(dt[k] >= 0.0) || !isfinite(dt[k])) {
        set_error("bmc_plan_create: event type must be 0..4 and dt finite and >= 0");
        return BMC_EINVAL;
      }
      row[e] = EvKey{type[k], dt[k], rf_amp[k], rf_phase[k], rf_freq[k]};
      if (ev) (*ev)[k] = Ev{dt[k], rf_amp[k], cos(rf_phase[k]), sin(rf_phase[k]), rf_freq[k], type[k], 0};
    }
    row_ptr[o] = (int32_t)ops.size();
    if (level > 0) {
      compile_row(row, level, ops, rot, built, applied);
    } else if (T > 0) {
      ops.push_back(Op{OP_SEQ, 0, T, 1, 0});
      rot.push_back(1.0);
      rot.push_back(0.0);
      built += T;
      applied += T;
    }
  }
  row_ptr[O] = (int32_t)ops.size();
  return BMC_OK;
}

}  // namespace
}  // namespace bmc

extern "C" {

int bmc_plan_compile(int n_offsets, int n_events, const int32_t* type, const double* dt, const double* rf_amp,
                     const double* rf_phase, const double* rf_freq, int compress, int64_t capacity, int32_t* ops_out,
                     double* rot_out, int32_t* row_ptr_out, int64_t* n_ops_out) {
  std::vector<Op> ops;
  std::vector<double> rot;
  std::vector<int32_t> row_ptr;
  int64_t built, applied;
  int rc = compile_plan(n_offsets, n_events, type, dt, rf_amp, rf_phase, rf_freq, compress, nullptr, ops, rot, row_ptr,
                        built, applied);
  if (rc != BMC_OK) return rc;
  if (n_ops_out) *n_ops_out = (int64_t)ops.size();
  if (!ops_out && !rot_out && !row_ptr_out) return BMC_OK;  // size query
  if (!ops_out || !rot_out || !row_ptr_out || capacity < (int64_t)ops.size()) {
    set_error("bmc_plan_compile: null output or capacity below the op count");
    return BMC_EINVAL;
  }
  for (size_t i = 0; i < ops.size(); i++) {
    ops_out[5 * i + 0] = ops[i].kind;
    ops_out[5 * i + 1] = ops[i].e0;
    ops_out[5 * i + 2] = ops[i].len;
    ops_out[5 * i + 3] = ops[i].count;
    ops_out[5 * i + 4] = ops[i].half;
    rot_out[2 * i] = rot[2 * i];
    rot_out[2 * i + 1] = rot[2 * i + 1];
  }
  for (size_t i = 0; i < row_ptr.size(); i++) row_ptr_out[i] = row_ptr[i];
  return BMC_OK;
}

int bmc_plan_create_ex(int n_offsets, int n_events, const int32_t* type, const double* dt, const double* rf_amp,
                       const double* rf_phase, const double* rf_freq, int reset_after_adc, int compress,
                       bmc_plan** out) {
  if (!out) {
    set_error("bmc_plan_create: bad arguments");
    return BMC_EINVAL;
  }
  *out = nullptr;
  const int O = n_offsets, T = n_events;
  std::vector<Ev> ev;
  std::vector<Op> ops;
  std::vector<double> rot;
  std::vector<int32_t> row_ptr;
  int64_t built = 0, applied = 0;
  int rc = compile_plan(O, T, type, dt, rf_amp, rf_phase, rf_freq, compress, &ev, ops, rot, row_ptr, built, applied);
  if (rc != BMC_OK) return rc;
  bmc_plan* pl = new bmc_plan();
  Plan& p = pl->p;
  p.O = O;
  p.T = T;
  p.reset_after_adc = reset_after_adc ? 1 : 0;
  p.compressed = compress ? 1 : 0;
  p.has_keep = std::any_of(ops.begin(), ops.end(), [](const Op& op) { return op.kind == OP_KEEP; }) ? 1 : 0;
  p.has_mirror = std::any_of(ops.begin(), ops.end(), [](const Op& op) { return op.half > 0; }) ? 1 : 0;
  // single-event rows: at most one event with dt > 1e-9 (simulate.py:51), a delay or an RF block; no
  // spoiler or gradient anywhere; after that event only zero-length events, any readout being the
  // row's last event (so the row's last_M is the state the event leaves, simulate.py:105-132)
  std::vector<int32_t> live(O, -1);
  bool cw = T > 0;
  for (int o = 0; o < O && cw; o++) {
    for (int e = 0; e < T && cw; e++) {
      const size_t k = (size_t)o * T + e;
      if (type[k] == BMC_EV_GRADIENT || type[k] == BMC_EV_SPOILER) cw = false;
      if (dt[k] > 1e-9) {
        if (live[o] >= 0 || (type[k] != BMC_EV_DELAY && type[k] != BMC_EV_RF)) cw = false;
        live[o] = e;
      }
      if (type[k] == BMC_EV_ADC && live[o] >= 0 && e != T - 1) cw = false;
    }
  }
  p.cw_rows = cw ? 1 : 0;
  p.n_ops = (int64_t)ops.size();
  p.n_built = built;
  p.n_applied = applied;
  p.h_ops = ops;
  p.h_row_ptr = row_ptr;
  auto fail = [&](hipError_t e) {
    set_error(std::string("bmc_plan_create: ") + hipGetErrorString(e));
    bmc_plan_destroy(pl);
    return e == hipErrorOutOfMemory ? BMC_ENOMEM : BMC_EHIP;
  };
  hipError_t e;
  if ((e = hipMalloc(&p.d_ev, sizeof(Ev) * std::max<size_t>(ev.size(), 1))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&p.d_ops, sizeof(Op) * std::max<size_t>(ops.size(), 1))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&p.d_rot, sizeof(double) * std::max<size_t>(rot.size(), 2))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&p.d_row_ptr, sizeof(int32_t) * row_ptr.size())) != hipSuccess) return fail(e);
  if (p.cw_rows) {
    if ((e = hipMalloc(&p.d_live, sizeof(int32_t) * O)) != hipSuccess) return fail(e);
    if ((e = hipMemcpy(p.d_live, live.data(), sizeof(int32_t) * O, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
  }
  if (!ev.empty() && (e = hipMemcpy(p.d_ev, ev.data(), sizeof(Ev) * ev.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e);
  if (!ops.empty() && (e = hipMemcpy(p.d_ops, ops.data(), sizeof(Op) * ops.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(e);
  if (!rot.empty() && (e = hipMemcpy(p.d_rot, rot.data(), sizeof(double) * rot.size(), hipMemcpyHostToDevice)) !=
                          hipSuccess)
    return fail(e);
  if ((e = hipMemcpy(p.d_row_ptr, row_ptr.data(), sizeof(int32_t) * row_ptr.size(), hipMemcpyHostToDevice)) !=
      hipSuccess)
    return fail(e);
  *out = pl;
  return BMC_OK;
}

int bmc_plan_destroy(bmc_plan* plan) {
  if (!plan) return BMC_OK;
  Plan& p = plan->p;
  if (p.d_ev) (void)hipFree(p.d_ev);
  if (p.d_ops) (void)hipFree(p.d_ops);
  if (p.d_row_ptr) (void)hipFree(p.d_row_ptr);
  if (p.d_rot) (void)hipFree(p.d_rot);
  if (p.d_live) (void)hipFree(p.d_live);
  delete plan;
  return BMC_OK;
}

int bmc_plan_stats(const bmc_plan* plan, int64_t* n_ops, int64_t* n_built, int64_t* n_applied) {
  if (!plan) {
    set_error("bmc_plan_stats: null plan");
    return BMC_EINVAL;
  }
  if (n_ops) *n_ops = plan->p.n_ops;
  if (n_built) *n_built = plan->p.n_built;
  if (n_applied) *n_applied = plan->p.n_applied;
  return BMC_OK;
}

int bmc_plan_program(const bmc_plan* plan, int64_t capacity, int32_t* ops_out, int32_t* row_ptr_out) {
  if (!plan || !ops_out || !row_ptr_out || capacity < plan->p.n_ops) {
    set_error("bmc_plan_program: null argument or capacity below n_ops");
    return BMC_EINVAL;
  }
  const Plan& p = plan->p;
  for (size_t i = 0; i < p.h_ops.size(); i++) {
    ops_out[5 * i + 0] = p.h_ops[i].kind;
    ops_out[5 * i + 1] = p.h_ops[i].e0;
    ops_out[5 * i + 2] = p.h_ops[i].len;
    ops_out[5 * i + 3] = p.h_ops[i].count;
    ops_out[5 * i + 4] = p.h_ops[i].half;
  }
  for (size_t i = 0; i < p.h_row_ptr.size(); i++) row_ptr_out[i] = p.h_row_ptr[i];
  return BMC_OK;
}

}  // extern "C"

namespace bmc {

// Fills the parts of SimArgs common to simulate and the LM loop. Returns BMC_OK or an error.
int fill_sim_args(const bmc_plan* plan, int V, int nc, int dtype, const void* params, const void* m0,
                  const void* dm0, int P, const int32_t* slot, const double* scale, int jc, SimArgs& a) {
  if (!plan) {
    set_error("null plan");
    return BMC_EINVAL;
  }
  if (V < 0 || nc < 0 || (dtype != BMC_F32 && dtype != BMC_F64) || !params || !m0) {
    set_error("bad sizes, dtype or null input");
    return BMC_EINVAL;
  }
  if (nc > BMC_MAX_CEST) {
    set_error("n_cest " + std::to_string(nc) + " exceeds the compiled maximum " + std::to_string(BMC_MAX_CEST));
    return BMC_ENOTSUP;
  }
  const int N = 3 * (1 + nc) + 1;
  if (P < 0 || P > BMC_MAX_TANGENTS || (P > 0 && (!slot || !scale))) {
    set_error("tangent count must be 0.." + std::to_string(BMC_MAX_TANGENTS));
    return BMC_EINVAL;
  }
  if (jc < 0 || jc >= N) {
    set_error("jac_component out of range");
    return BMC_EINVAL;
  }
  memset(&a, 0, sizeof(a));
  const Plan& p = plan->p;
  a.ev = p.d_ev;
  a.ops = p.d_ops;
  a.row_ptr = p.d_row_ptr;
  a.rot = p.d_rot;
  a.keep = p.has_keep;
  a.mirror = p.has_mirror && !p.has_keep;
  a.O = p.O;
  a.T = p.T;
  a.reset_after_adc = p.reset_after_adc;
  a.V = V;
  a.dtype = dtype;
  a.stride = 13 + 5 * nc;
  a.params = params;
  a.m0 = m0;
  a.dm0 = dm0;
  a.P = P;
  for (int i = 0; i < P; i++) {
    int idx = 0;
    int k = tan_kind_of_slot(slot[i], nc, &idx);
    if (k < 0) {
      set_error("tangent slot " + std::to_string(slot[i]) + " outside the packed record");
      return BMC_EINVAL;
    }
    a.t_kind[i] = k;
    a.t_idx[i] = idx;
    a.t_scale[i] = scale[i];
  }
  a.jc = jc;
  // reference NaN rule (expm.py:73-83): floor(log2(l1 / theta)) > 16  <=>  l1 >= theta 2^17
  a.nan_l1 = (dtype == BMC_F32 ? 3.925724783138660 : 5.371920351148152) * 131072.0;
  return BMC_OK;
}

// Tangents per pass for a tangent count (register budget of the kernel).
#ifndef BMC_TANGENTS_PER_PASS
#define BMC_TANGENTS_PER_PASS 2
#endif
#if BMC_TANGENTS_PER_PASS < 1 || BMC_TANGENTS_PER_PASS > 5
#error "the simulate kernels are instantiated for 1 to 5 tangents per pass"
#endif
int tangents_per_pass(int P) { return P <= 0 ? 0 : (P < BMC_TANGENTS_PER_PASS ? P : BMC_TANGENTS_PER_PASS); }

int stream_after_null(hipStream_t* out) {
  *out = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t ready = nullptr;
  BMC_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  hipError_t e = hipEventCreateWithFlags(&ready, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(ready, nullptr);
  if (e == hipSuccess) e = hipStreamWaitEvent(s, ready, 0);
  if (ready) (void)hipEventDestroy(ready);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(s);
    set_error(std::string("ordering a stream after the null stream: ") + hipGetErrorString(e));
    return BMC_EHIP;
  }
  *out = s;
  return BMC_OK;
}

}  // namespace bmc

extern "C" int bmc_tangents_per_pass(int n_tangents) { return tangents_per_pass(n_tangents); }

extern "C" int bmc_simulate_ex(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
                               const void* m0_dev, const void* dm0_dev, int n_tangents, const int32_t* tangent_slot,
                               const double* tangent_scale, int jac_component, void* m_out_dev, void* jac_out_dev,
                               unsigned flags, void* stream) {
  if (flags & ~BMC_SIM_FORWARD_JACOBIAN) {
    set_error("bmc_simulate_ex: unknown flag bits");
    return BMC_EINVAL;
  }
  SimArgs a;
  int rc = fill_sim_args(plan, n_voxels, n_cest, dtype, params_dev, m0_dev, dm0_dev, n_tangents, tangent_slot,
                         tangent_scale, jac_component, a);
  if (rc != BMC_OK) return rc;
  if (n_voxels == 0) return BMC_OK;
  a.m_out = m_out_dev;
  a.jac_out = jac_out_dev;
  a.jac_dtype = dtype;
  const int want_tan = (jac_out_dev && n_tangents > 0) ? n_tangents : 0;
  a.PT = tangents_per_pass(want_tan);
  a.n_pass = want_tan ? (want_tan + a.PT - 1) / a.PT : 1;
  if (!want_tan) a.P = n_tangents;  // seeds unused
  if (want_tan && use_cwj(plan->p, a, (flags & BMC_SIM_FORWARD_JACOBIAN) != 0))
    return launch_cwj(n_cest, a, plan->p.d_live, (hipStream_t)stream);
  return launch_sim(n_cest, a, (hipStream_t)stream);
}

extern "C" int bmc_simulate(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
                            const void* m0_dev, const void* dm0_dev, int n_tangents, const int32_t* tangent_slot,
                            const double* tangent_scale, int jac_component, void* m_out_dev, void* jac_out_dev,
                            void* stream) {
  return bmc_simulate_ex(plan, n_voxels, n_cest, dtype, params_dev, m0_dev, dm0_dev, n_tangents, tangent_slot,
                         tangent_scale, jac_component, m_out_dev, jac_out_dev, 0u, stream);
}

extern "C" int bmc_simulate_trajectory(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
                                       const void* m0_dev, void* traj_out_dev, void* stream) {
  SimArgs a;
  int rc = fill_sim_args(plan, n_voxels, n_cest, dtype, params_dev, m0_dev, nullptr, 0, nullptr, nullptr, 0, a);
  if (rc != BMC_OK) return rc;
  if (plan->p.compressed) {
    set_error("bmc_simulate_trajectory: the plan must be created with compress = 0");
    return BMC_EINVAL;
  }
  if (!traj_out_dev) {
    set_error("bmc_simulate_trajectory: null output");
    return BMC_EINVAL;
  }
  if (n_voxels == 0) return BMC_OK;
  a.traj = traj_out_dev;
  a.PT = 0;
  a.n_pass = 1;
  return launch_sim(n_cest, a, (hipStream_t)stream);
}
