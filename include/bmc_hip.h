/*
 * bmc_hip.h -- C ABI of the MI355X simulate-and-fit engine (libbmc_hip.so).
 *
 * The reference (MPR-UKD/BMC-JAX) is pure Python/JAX and has no FFI; its drop-in boundary is the Python
 * API. Each entry point below replaces the JAX computation behind one reference function (cited per
 * entry, paths relative to the BMC-JAX tree). A thin Python layer (bmc-jax_amd/bmc_jax) binds these
 * with ctypes and keeps the reference's Python signatures.
 *
 * Conventions
 *   - Every entry point returns 0 on success and a negative BMC_E* code on failure; bmc_last_error()
 *     then returns a message (thread-local).
 *   - Pointers named *_dev are device pointers (hipMalloc / torch CUDA tensors); the caller owns all
 *     buffers and the library never frees caller memory. `stream` is a hipStream_t (NULL = default).
 *   - `dtype` is BMC_F32 or BMC_F64 and applies to every real-valued buffer of the call. The kernels
 *     compute in float64 internally (CDNA4 runs FP64 FMA at the FP32 rate); float32 mode rounds only
 *     inputs and outputs.
 *
 * Packed parameter record, one per voxel, `13 + 5 * n_cest` reals (simulation/params.py:18-49):
 *   [0] water_r1 [1] water_r2 [2] water_f [3] mt_k [4] mt_f [5] mt_dw [6] mt_r1 [7] mt_r2
 *   [8] mt_lineshape (0 none, 1 Lorentzian, 2 SuperLorentzian) [9] b0 [10] gamma [11] b0_inhom
 *   [12] rel_b1, then cest_k[n_cest], cest_f[n_cest], cest_dw[n_cest], cest_r1[n_cest], cest_r2[n_cest].
 * State vector (length N = 3 (1 + n_cest) + 1): x of water and pools, y of water and pools, z of water
 * and pools, MT z (simulation/matrix.py:224-234). Water Mz sits at index 2 (1 + n_cest).
 */
#ifndef BMC_HIP_H
#define BMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BMC_ABI_VERSION 3

#define BMC_F32 0
#define BMC_F64 1

#define BMC_OK 0
#define BMC_EINVAL -1  /* bad shape, size or argument */
#define BMC_ENOMEM -2  /* device allocation failed */
#define BMC_EHIP -3    /* a HIP runtime call failed */
#define BMC_ENOTSUP -4 /* configuration outside what the kernels support */

/* Event types (simulation/events.py:12-18). */
#define BMC_EV_DELAY 0
#define BMC_EV_RF 1
#define BMC_EV_GRADIENT 2
#define BMC_EV_ADC 3
#define BMC_EV_SPOILER 4

/* Largest number of CEST pools the kernels are compiled for (state length <= 16, augmented <= 17). */
#define BMC_MAX_CEST 4
/* Largest number of free parameters (tangents) per call. */
#define BMC_MAX_TANGENTS 16

int bmc_abi_version(void);
const char* bmc_last_error(void);
/* Tangents bmc_simulate and bmc_lm_fit carry per pass of the kernel for n_tangents tangents (the
 * kernel runs ceil(n_tangents / that) passes, each with the primal state). Informational. */
int bmc_tangents_per_pass(int n_tangents);

/* Number of HIP devices visible (0 if none). */
int bmc_device_count(void);

/*
 * A compiled saturation sequence: the event table (simulation/events.py:20-32) uploaded to the device
 * together with a per-offset program that marks runs of identical events (a repeated pulse and its
 * gap) so their propagator is built once and applied many times. Replaces nothing in the reference
 * directly: it is the device-resident form of the `events` argument of simulate_spectrum
 * (simulation/simulate.py:134-148).
 *
 * Host arrays, each [n_offsets, n_events] row-major; rf_* and dt in float64 (callers holding float32
 * events widen them, which is exact). reset_after_adc mirrors simulate_sequence's argument
 * (simulate.py:80); the reference always simulates with True.
 */
typedef struct bmc_plan bmc_plan;
int bmc_plan_create(int n_offsets, int n_events, const int32_t* type, const double* dt, const double* rf_amp,
                    const double* rf_phase, const double* rf_freq, int reset_after_adc, bmc_plan** out);
/* As bmc_plan_create with a compression level: 0 keeps every event in order as its own step (no chained
 * runs), which bmc_simulate_trajectory needs; 1 chains runs of identical events repeated back to back;
 * 2 also keeps a pulse that recurs later in its row with every sample's RF phase shifted by one
 * constant d (a scanner train's accumulated phase, sequences/pulseq.py:60-66) and applies it again
 * rotated by d about z instead of rebuilding it; 3 (bmc_plan_create's level) also marks chained runs
 * that open with a palindrome (a symmetric pulse shape at one RF phase: mirrored samples equal in type,
 * dt, phase and frequency and in amplitude to 1e-13 relative), whose propagator the kernels build from
 * the first half (the same result to rounding). */
int bmc_plan_create_ex(int n_offsets, int n_events, const int32_t* type, const double* dt, const double* rf_amp,
                       const double* rf_phase, const double* rf_freq, int reset_after_adc, int compress,
                       bmc_plan** out);
/* The row programs bmc_plan_create_ex would build, on the host only (no device call): ops_out
 * [n_ops, 5] and rot_out [n_ops, 2] as bmc_plan_program describes them, row_ptr_out [n_offsets + 1].
 * With all three outputs NULL it only stores the op count in n_ops_out (a size query). Replaces
 * nothing in the reference; it lets a caller inspect a compilation without a GPU. */
int bmc_plan_compile(int n_offsets, int n_events, const int32_t* type, const double* dt, const double* rf_amp,
                     const double* rf_phase, const double* rf_freq, int compress, int64_t capacity, int32_t* ops_out,
                     double* rot_out, int32_t* row_ptr_out, int64_t* n_ops_out);
int bmc_plan_destroy(bmc_plan* plan);
/* Statistics of a plan: ops and the number of events whose propagator is built per offset row. */
int bmc_plan_stats(const bmc_plan* plan, int64_t* n_ops, int64_t* n_built, int64_t* n_applied);
/* The row programs: ops_out [n_ops, 5] = (kind, first event, length, repeat count, half), row_ptr_out
 * [n_offsets + 1]. Kinds: 0 SEQ (events one by one), 1 BLOCK (chained once, applied `count` times),
 * 2 KEEP (as BLOCK, and kept), 3 REUSE (the kept block rotated by the op's phase offset d: (cos d,
 * sin d) in bmc_plan_compile's rot_out). half h > 0 on a BLOCK or KEEP op: its first 2h events read
 * the same backwards (RF and delay events, one RF phase phi, whose (cos phi, sin phi) is the op's
 * rot_out entry), so the kernel builds their propagator from the first h (palindrome factorisation).
 * capacity is the number of ops ops_out can hold. */
int bmc_plan_program(const bmc_plan* plan, int64_t capacity, int32_t* ops_out, int32_t* row_ptr_out);

/*
 * Forward simulation with forward-mode tangents: for every voxel v and offset row o, the final state
 * `last_M` of the event scan and, for each tangent p, its derivative w.r.t. packed parameter slot
 * tangent_slot[p], times tangent_scale[p].
 *
 * Replaces: simulate_spectrum / simulate_voxels (simulation/simulate.py:134-148, voxels.py:89-129) for
 * the state, and jax.jacfwd of the residual (fitting/lm.py:39, fitting/joint.py:300) for the tangents.
 *
 *   params_dev   [V, 13 + 5 n_cest]          packed records
 *   m0_dev       [V, N]                      initial state per voxel (M_init)
 *   dm0_dev      [V, P, N] or NULL           derivative of M_init per tangent (NULL = zero)
 *   tangent_slot [P] host                    packed slot index, or -1 for "no generator dependence"
 *   tangent_scale[P] host                    multiplier of tangent p (the FitSpec scale)
 *   m_out_dev    [V, O, N] or NULL           final states
 *   jac_out_dev  [V, O, P] or NULL           d state[jac_component] / d x_p
 * NaN semantics follow the reference: an event whose reference squaring count exceeds 16
 * (simulation/expm.py:83) makes the state NaN until the next ADC reset.
 * When every row of the plan propagates at most one event (CW saturation rows: no spoiler or gradient,
 * a readout only as a row's last event) and dm0_dev is NULL, the tangents come from the reverse-
 * accumulation kernel instead (same outputs, to rounding). bmc_lm_fit takes the same route for its
 * Jacobian sweeps. The library reads no environment: the mode is chosen per call (bmc_simulate_ex).
 */
int bmc_simulate(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
                 const void* m0_dev, const void* dm0_dev, int n_tangents, const int32_t* tangent_slot,
                 const double* tangent_scale, int jac_component, void* m_out_dev, void* jac_out_dev,
                 void* stream);
/* bmc_simulate with a flags word. BMC_SIM_FORWARD_JACOBIAN: take the tangents in forward mode (what
 * jax.jacfwd computes, fitting/lm.py:39) even where the reverse-accumulation kernel applies. Same
 * outputs to rounding; the bit exists to compare the two routes and to time them. Unknown bits:
 * BMC_EINVAL. bmc_simulate is bmc_simulate_ex with flags 0. */
#define BMC_SIM_FORWARD_JACOBIAN 1u
int bmc_simulate_ex(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
                    const void* m0_dev, const void* dm0_dev, int n_tangents, const int32_t* tangent_slot,
                    const double* tangent_scale, int jac_component, void* m_out_dev, void* jac_out_dev,
                    unsigned flags, void* stream);

/*
 * The state after every event: simulate_sequence's return_full_trajectory=True (simulate.py:115-120,
 * 128-129): traj_out_dev [V, O, T, N] in dtype. The plan must come from bmc_plan_create_ex(compress=0).
 */
int bmc_simulate_trajectory(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
                            const void* m0_dev, void* traj_out_dev, void* stream);

/*
 * Batched bounded Levenberg-Marquardt, one independent problem per voxel, every iteration on the GPU:
 * simulate with tangents at x, float64 normal equations J^T J, J^T r with the damped solve
 * A = H + lam (diag H + 1e-9 I) (free mask: J * free, A += diag(1 - free)), clip to [lo, hi], a primal
 * simulation at the trial point, accept iff strictly better (lam x 0.5) else reject (lam x 3).
 *
 * Replaces: levenberg_marquardt vmapped by fit_batch (fitting/lm.py:20-67, independent.py:118-128) and
 * fit_joint's non-coupled path (joint.py:359-374) when every slot is a plain value of a packed field
 * (held entries through free_dev; factor, offset, start-state and B0-spread slots have no kernel form).
 *
 *   x_dev        [V, P] float64 in/out       scaled start values in, fitted scaled values out
 *   lo, hi       [P] host float64            bounds, scaled units
 *   z_meas_dev   [V, O] dtype                measured Z (residual r = Z_sim[:, jac_component] - z_meas)
 *   free_dev     [V, P] float64 or NULL      1 free / 0 held (NULL = all free, arithmetic unchanged); a
 *                                            held value is never written into the record, so the
 *                                            model keeps the record's own number for it exactly
 *   cost_traj_dev[V, n_iter] float64 or NULL cost after every iteration (the reference's C_traj)
 *   x_traj_dev   [V, n_iter, P] float64 or NULL
 *   flags        BMC_LM_EARLY_EXIT: stop a problem once it has converged (never the reference's
 *                behaviour, which runs every iteration): its rms at most early_rms (checked after
 *                every accepted step and, from the Jacobian sweep's own state, before the first step),
 *                a relative cost decrease of at most early_ftol, or lam above 1e12; BMC_LM_GRAPH: replay the iteration as a
 *                captured hipGraph; BMC_LM_START_CHECK (with BMC_LM_EARLY_EXIT): simulate every start
 *                first and stop, before any Jacobian, the problems whose start rms is at most
 *                early_rms (they return x0 after 0 iterations); BMC_LM_FORWARD_JACOBIAN: Jacobian sweeps
 *                in forward mode, as BMC_SIM_FORWARD_JACOBIAN.
 *   n_iter_done_dev [V] int32 or NULL        iterations each problem ran
 */
#define BMC_LM_EARLY_EXIT 1u
#define BMC_LM_GRAPH 2u
#define BMC_LM_START_CHECK 4u
#define BMC_LM_FORWARD_JACOBIAN 8u
int bmc_lm_fit(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
               const void* m0_dev, int n_params, const int32_t* slot, const double* scale, const double* lo,
               const double* hi, int jac_component, const void* z_meas_dev, double* x_dev, int n_iter,
               double lam0, const double* free_dev, unsigned flags, double early_ftol, double early_rms,
               double* cost_traj_dev, double* x_traj_dev, int32_t* n_iter_done_dev, void* stream);

/*
 * bmc_lm_fit without any device allocation. No counterpart in the reference: it is for callers that
 * cannot call hipMalloc (work issued under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation). Arguments, arithmetic and results are bmc_lm_fit's (the same kernels); the work buffers
 * come from a fixed device array of BMC_LM_NOALLOC_SCRATCH_BYTES that the library holds on every device
 * it runs on until the process ends. A call needs about 8 V (O (P + 2) + P + 2) + 16 V bytes (each of
 * its twelve buffers rounded up to 256 B); a larger call returns BMC_EINVAL with the size it needed.
 * BMC_LM_GRAPH is refused (instantiating a graph allocates). The call does not synchronise `stream`.
 * Every call on one device shares the one array: calls issued in order on one stream are ordered, but
 * calls that overlap on one GPU (other streams, other threads) race and corrupt each other.
 */
#define BMC_LM_NOALLOC_SCRATCH_BYTES (64u << 20)
int bmc_lm_fit_noalloc(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
                       const void* m0_dev, int n_params, const int32_t* slot, const double* scale, const double* lo,
                       const double* hi, int jac_component, const void* z_meas_dev, double* x_dev, int n_iter,
                       double lam0, const double* free_dev, unsigned flags, double early_ftol, double early_rms,
                       double* cost_traj_dev, double* x_traj_dev, int32_t* n_iter_done_dev, void* stream);

/*
 * Reverse mode of bmc_simulate with respect to the events' RF amplitudes and RF frequencies: given
 * cotangents on Z = state[jac_component] and on the Jacobian J[v, o, p] that bmc_simulate returns with
 * the same tangents, adds dL/d rf_amp[o, t] and dL/d rf_freq[o, t], summed over voxels, into the
 * outputs (the caller clears them).
 *
 * Replaces: jax.value_and_grad through jacfwd of the scan in optimize_design
 * (design/optimize.py:80-93, over design/crlb.py:38-58 and simulation/simulate.py:22-132).
 *
 * The forward scan is stored at every K = ceil(sqrt(T))-th event and each segment is recomputed in the
 * reverse sweep (sqrt(T) checkpoints); bmc_adjoint_workspace_size gives the workspace it needs
 * (device memory the caller provides; its contents on entry do not matter).
 *   z_cot_dev    [V, O] float64 or NULL      cotangent of Z
 *   jac_cot_dev  [V, O, P] float64 or NULL   cotangent of J (J includes tangent_scale)
 *   grad_amp_dev [O, T] float64 or NULL      += dL / d rf_amp   (Hz)
 *   grad_freq_dev[O, T] float64 or NULL      += dL / d rf_freq  (Hz)
 * Supported: every pool count the library is compiled for and up to BMC_MAX_TANGENTS tangents (run as
 * launches of two tangents each over the one workspace, their gradients added: L is linear in the
 * cotangents), for tangents whose generator derivative does not depend on the RF (water, CEST and MT
 * rates and fractions, CEST shifts); B0, gamma, b0_inhom, rel_b1 and the MT shift and T2 return
 * BMC_ENOTSUP. dm0_dev as in bmc_simulate.
 */
int bmc_adjoint_workspace_size(const bmc_plan* plan, int n_voxels, int n_cest, int n_tangents, int64_t* bytes);
int bmc_simulate_adjoint(const bmc_plan* plan, int n_voxels, int n_cest, int dtype, const void* params_dev,
                         const void* m0_dev, const void* dm0_dev, int n_tangents, const int32_t* tangent_slot,
                         const double* tangent_scale, int jac_component, const double* z_cot_dev,
                         const double* jac_cot_dev, double* grad_amp_dev, double* grad_freq_dev,
                         void* workspace_dev, int64_t workspace_bytes, void* stream);

/*
 * Uneven batched LM: independent fit problems of different sizes (CEST pool count 1..4, number of rows,
 * number of free values) fitted together. Problem i owns rows [row0[i], row0[i] + n_rows[i]) of the plan
 * (its rungs x offsets, front-padded to the plan's event count) and n_params[i] consecutive entries of
 * the per-value tables (slots, scales, bounds in scaled units, x). params_dev[c] / m0_dev[c] hold the
 * packed records [n_c, 13 + 5 (c + 1)] and start states of the problems with c + 1 CEST pools, in the
 * order they appear. Every simulate sweep is one persistent launch over (problem, row, tangent pass)
 * items; every iteration is that sweep, the float64 normal equations and damped solve per problem, a
 * primal sweep at the trial values and the accept/reject step, as bmc_lm_fit.
 *
 * Replaces: one `fit` per problem (fitting/independent.py:226-298 over lm.py:20-67); vmap cannot batch
 * problems of different shapes.
 *
 * bmc_ubatch_configure sets the launch mapping: tangents per pass (1..2), persistent grid size in
 * one-wave workgroups, item order (0 problem order, 1 largest first). The result does not depend on
 * it. bmc_ubatch_simulate: z_out [R] float64 and, if jac_out is not NULL, the ragged Jacobian
 * [sum n_rows[i] n_params[i]] float64 at x [sum n_params] (scaled). bmc_ubatch_lm_fit: x in/out;
 * z_meas [R] in dtype; cost_dev [n] float64 (final cost) and n_iter_done_dev [n] optional; flags as
 * bmc_lm_fit.
 */
typedef struct bmc_ubatch bmc_ubatch;
int bmc_ubatch_create(const bmc_plan* plan, int n_problems, int dtype, const int32_t* n_cest, const int32_t* n_params,
                      const int32_t* row0, const int32_t* n_rows, const int32_t* slots, const double* scales,
                      const double* lo, const double* hi, const void* const* params_dev, const void* const* m0_dev,
                      bmc_ubatch** out);
int bmc_ubatch_configure(bmc_ubatch* batch, int tangents_per_pass, int grid, int order);
int bmc_ubatch_sizes(const bmc_ubatch* batch, int* n_rows, int* n_values, int64_t* n_jacobian);
int bmc_ubatch_simulate(bmc_ubatch* batch, const double* x_dev, double* z_out_dev, double* jac_out_dev, void* stream);
/* bmc_ubatch_simulate with a flags word (BMC_SIM_FORWARD_JACOBIAN as for bmc_simulate_ex). */
int bmc_ubatch_simulate_ex(bmc_ubatch* batch, const double* x_dev, double* z_out_dev, double* jac_out_dev,
                           unsigned flags, void* stream);
int bmc_ubatch_lm_fit(bmc_ubatch* batch, const void* z_meas_dev, double* x_dev, int n_iter, double lam0,
                      unsigned flags, double early_ftol, double early_rms, double* cost_dev, int32_t* n_iter_done_dev,
                      void* stream);
int bmc_ubatch_destroy(bmc_ubatch* batch);

/*
 * The learned LM proposal: an MLP trained and evaluated with hand-written BF16 MFMA GEMMs (FP32
 * accumulate), GELU hidden layers, a linear output, mean-squared-error loss and Adam. No counterpart
 * in the reference; its output is the LM start x0 (plain x0 remains the fallback).
 *   dims [n_layers + 1] layer widths; weights W_l [dims[l+1], dims[l]] then b_l, flat FP32 parameters.
 *   x_dev [batch, dims[0]] FP32, t_dev / y_dev [batch, dims[n_layers]] FP32 (device).
 * bmc_mlp_train_step adds the batch loss into *loss_dev (device float; the caller clears it).
 */
typedef struct bmc_mlp bmc_mlp;
int bmc_mlp_create(int n_layers, const int32_t* dims, int max_batch, uint64_t seed, bmc_mlp** out);
int bmc_mlp_destroy(bmc_mlp* mlp);
int64_t bmc_mlp_n_params(const bmc_mlp* mlp);
int bmc_mlp_set_params(bmc_mlp* mlp, const float* host_params);
int bmc_mlp_get_params(const bmc_mlp* mlp, float* host_params);
int bmc_mlp_forward(bmc_mlp* mlp, const float* x_dev, int batch, float* y_dev, void* stream);
int bmc_mlp_gradients(bmc_mlp* mlp, const float* x_dev, const float* t_dev, int batch, float* grad_host,
                      float* loss_host, void* stream);
int bmc_mlp_train_step(bmc_mlp* mlp, const float* x_dev, const float* t_dev, int batch, float lr, float* loss_dev,
                       void* stream);
/* Input standardisation applied before the first layer, x -> (x - shift) * scale per input feature
 * (without one, the identity). bmc_mlp_fit_input_norm sets shift to the column means and scale to the
 * inverse column standard deviations of x_dev [n, dims[0]] (float64 sums in a fixed order; a standard
 * deviation below 1e-6 counts as 1e-6); set / get move both vectors [dims[0]] from / to the host (set
 * with both NULL removes it). A measured Z-spectrum varies by a few percent around a mean near 1, below
 * BF16's resolution there: standardised, the features use the GEMM's BF16 operands fully. */
int bmc_mlp_fit_input_norm(bmc_mlp* mlp, const float* x_dev, int64_t n, void* stream);
int bmc_mlp_set_input_norm(bmc_mlp* mlp, const float* shift_host, const float* scale_host);
int bmc_mlp_get_input_norm(const bmc_mlp* mlp, float* shift_host, float* scale_host);

/* Dense matrix exponential exp(A) of `batch` row-major n x n matrices (device, float32 or float64 by
 * dtype), for callers that hold the matrix themselves: replaces expm_library / expm_plain
 * (src/bmc_jax/simulation/expm.py:26-83), on which propagate (:86-119) steps. Pade approximation with
 * scaling and squaring in the reference's rule: s = max(0, floor(log2(||A||_1 / theta))), the Pade order
 * from ||A||_1 (float64: 3, 5, 7, 9, 13; float32: 3, 5, 7, computed in float32), Gauss-Jordan solve with
 * partial pivoting, s squarings; more than 16 squarings, or a non-finite norm, gives a NaN matrix.
 * One workgroup per matrix; n <= BMC_EXPM_MAX_N. */
#define BMC_EXPM_MAX_N 24
int bmc_expm(const void* a_dev, void* out_dev, int n, int64_t batch, int dtype, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BMC_HIP_H */
