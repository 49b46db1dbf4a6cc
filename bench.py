"""Benchmark of the simulate-and-fit hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--voxels V]

Workload: the pulsed-train forward + Jacobian configuration -- 4 pools (water, amide +3.5 ppm, amine
+2.0 ppm, SuperLorentzian MT at -2.5 ppm), 50 Gaussian pulses x 200 steps of 0.1 ms with 20 ms gaps,
a spoiler and a 3 s recovery (10,053 events per offset), 61 offsets, 65,536 voxels with seeded
uniform / log-uniform parameters, 5 free values (k_amide, f_amide, water R1, water R2, B0 shift),
float32 inputs and outputs. One step = the final state and the 5-column Jacobian of every voxel and
offset (bmc_simulate). The metric is Z-spectra (forward + Jacobian) per second.

Prints ONE JSON line. `roofline` is the dominant kernel's executed float64 flops (bmc_jax._flops,
which replays the kernel's per-event Taylor degree and squaring count on a sample of voxels) over its
measured launch time, against MI355X's FP64 vector peak (the kernel computes in float64).
`cpu_baseline` times the reference-equivalent C oracle (one float64 Pade exponential per event, no
reuse, forward-mode tangents, OpenMP over offsets) on this host's cores on one voxel's spectrum.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "bmc-jax_amd")]

FREE = [("cest_k", 0), ("cest_f", 0), ("water_r1", None), ("water_r2", None), ("b0_inhom", None)]
PEAK_FP64_TFLOPS = 78.6
# HBM bytes per voxel of one step of this workload, from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
# runs scaled by a 1 GiB copy kernel's readings (profiles/sim_kernel_pulsed_train.md): 1.060e12 B
# per 65,536-voxel step (its two dispatches: 307 GB fetched, 753 GB written), almost all
# register-spill and call-frame traffic (627 GB/s over the step, about 8 % of the HBM bandwidth).
TRAFFIC_BYTES_PER_VOXEL = 1.060e12 / 65536


def workload(n_vox, seed=2):
    """Packed float32 parameters [V, 23], M0 [V, 10], events of the pulsed train."""
    from bmc_jax.sequences import pulsed_train
    rng = np.random.default_rng(seed)
    nc = 2
    V = n_vox
    prm = np.zeros((V, 13 + 5 * nc))
    k = np.exp(rng.uniform(np.log(5.0), np.log(5000.0), (V, 2)))
    f = rng.uniform(1e-5, 0.02, (V, 2))
    prm[:, 0] = 1.0 / rng.uniform(0.8, 3.0, V)        # water R1
    prm[:, 1] = 1.0 / rng.uniform(0.04, 0.15, V)      # water R2
    prm[:, 2] = 1.0                                   # water f
    prm[:, 3:9] = [30.0, 0.1, -2.5, 1.0, 1e5, 2]      # MT k, f, dw, r1, r2 (T2 10 us), SuperLorentzian
    prm[:, 9], prm[:, 10] = 3.0, 267.5153             # B0, gamma
    prm[:, 11] = rng.uniform(-0.3, 0.3, V)            # B0 shift [ppm]
    prm[:, 12] = 1.0
    prm[:, 13:15] = k
    prm[:, 15:17] = f
    prm[:, 17:19] = [3.5, 2.0]                        # dw
    prm[:, 19:21] = [1.0 / 1.3, 1.0 / 1.3]            # r1
    prm[:, 21:23] = [10.0, 20.0]                      # r2 (T2 0.1 s, 0.05 s)
    prm = prm.astype(np.float32)
    N = 3 * (1 + nc) + 1
    m0 = np.zeros((V, N), np.float32)
    m0[:, 6] = prm[:, 2]
    m0[:, 7:9] = prm[:, 15:17]
    m0[:, 9] = prm[:, 4]
    ev = pulsed_train(np.linspace(-6.0, 6.0, 61), b1_peak_uT=2.0, n_pulses=50, pulse_s=0.02, n_samples=200,
                      gap_s=0.02, recovery_s=3.0, b0_T=3.0)
    return prm, m0, ev, nc


def design_workload(n_truths=1024, seed=4):
    """The adjoint protocol-gradient problem: 3 pools (water, amide +3.5 ppm, SuperLorentzian MT at
    -2.5 ppm), a shaped pulse of 256 samples repeated 40 times (20 ms pulses, 10 ms gaps) after a 1 s
    recovery, 31 offsets, CRLB cost of (k, f) averaged over n_truths seeded parameter sets."""
    from bmc_jax.fitting import FitSpec, FreeParam
    from bmc_jax.fitting.experiment import Experiment
    from bmc_jax.sequences import pulsed_train
    from bmc_jax.sequences.saturation import hz_per_ppm
    from bmc_jax.simulation.params import SimParams
    rng = np.random.default_rng(seed)
    f32 = lambda v: np.float32(v)
    sp = SimParams(water_r1=f32(1 / 1.3), water_r2=f32(1 / 0.075), water_f=f32(1.0), cest_k=np.array([50.0], np.float32),
                   cest_f=np.array([2e-3], np.float32), cest_dw=np.array([3.5], np.float32),
                   cest_r1=np.array([1 / 1.3], np.float32), cest_r2=np.array([10.0], np.float32), mt_k=f32(30.0),
                   mt_f=f32(0.1), mt_dw=f32(-2.5), mt_r1=f32(1.0), mt_r2=f32(1e5), mt_lineshape=np.int32(2),
                   b0=f32(3.0), gamma=f32(267.5153), b0_inhom=f32(0.0), rel_b1=f32(1.0))
    m0 = np.zeros(7, np.float32)
    m0[4], m0[5], m0[6] = 1.0, 2e-3, 0.1
    ppm = np.linspace(-5.0, 5.0, 31)
    ev = pulsed_train(ppm, b1_peak_uT=2.0, n_pulses=40, pulse_s=0.02, n_samples=256, gap_s=0.01, recovery_s=1.0,
                      b0_T=3.0, dtype=np.float64)
    typ = np.asarray(ev.type)
    isrf = typ[0] == 1
    sidx = np.where(isrf, (np.cumsum(isrf) - 1) % 256, 0)
    shape0 = np.asarray(ev.rf_amp)[0, np.flatnonzero(isrf)[:256]]
    d0 = np.concatenate([shape0, ppm])
    spec = FitSpec([FreeParam("k", "cest_k", 0, 100.0, 5.0, 20000.0, 100.0),
                    FreeParam("f", "cest_f", 0, 0.001, 1e-6, 0.05, 2e-3)])
    truths = np.stack([np.exp(rng.uniform(np.log(5.0), np.log(5000.0), n_truths)),
                       rng.uniform(1e-3, 0.02, n_truths)], axis=1)

    def build(d):
        import torch
        shape, offs = d[:256], d[256:]
        isrf_t = torch.as_tensor(isrf)
        amp = torch.where(isrf_t, shape[torch.as_tensor(sidx)], torch.zeros((), dtype=d.dtype))
        amp = amp[None, :].expand(len(ppm), -1)
        freq = torch.where(isrf_t[None, :], offs[:, None] * hz_per_ppm(3.0), torch.zeros((), dtype=d.dtype))
        events = ev._replace(rf_amp=amp, rf_freq=freq)
        return Experiment(sp, events, m0, 4)

    return build, d0, spec, truths, ev


def bench_design(args):
    import torch
    from bmc_jax import _flops
    from bmc_jax.design.adjoint import design_cost
    build, d0, spec, truths, ev = design_workload(args.truths)
    d = torch.tensor(d0, dtype=torch.float64, requires_grad=True)

    def step():
        d.grad = None
        c = design_cost(build(d), spec, 0.002, truths, np.array([0, 1]), dtype=np.float32)
        c.backward()
        return float(c.detach()), d.grad

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        c, g = step()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    assert np.isfinite(c) and torch.isfinite(g).all().item()
    # the adjoint launch alone, timed on its stream
    from bmc_jax import _native
    from bmc_jax.design.adjoint import scan_jacobian
    Z, J = scan_jacobian(build(d.detach()), spec, truths, dtype=np.float32)
    gz, gj = torch.zeros_like(Z, dtype=torch.float64), torch.ones_like(J, dtype=torch.float64)
    from bmc_jax.design.adjoint import ScanProblem  # noqa: F401
    ev_np = {k: np.asarray(getattr(ev, k)) for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    plan = _native.Plan(ev_np["type"], ev_np["dt"], ev_np["rf_amp"], ev_np["rf_phase"], ev_np["rf_freq"])
    from bmc_jax.simulation.params import pack
    from bmc_jax.fitting.params import write_params
    from bmc_jax.simulation.voxels import tile_params
    exp = build(d.detach())
    pb = write_params(tile_params(exp.params, truths.shape[0]), truths / spec.scales, spec)
    prm = _native.device_array(pack(pb, truths.shape[0], dtype=np.float32), torch.float32)
    m0 = _native.device_array(np.broadcast_to(exp.m0, (truths.shape[0], 7)).astype(np.float32), torch.float32)
    slots = [_native.slot_index("cest_k", 0, 1), _native.slot_index("cest_f", 0, 1)]
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    _native.simulate_adjoint(plan, prm, 1, _native.BMC_F32, m0, slots=slots, scales=[1.0, 1.0], jc=4,
                             z_cot=gz, jac_cot=gj)
    e1.record(stream)
    torch.cuda.synchronize()
    adj_ms = e0.elapsed_time(e1)
    sample = np.random.default_rng(0).choice(truths.shape[0], size=min(8, truths.shape[0]), replace=False)
    flops = float(_flops.adjoint_flops(pack(pb, truths.shape[0])[sample].astype(np.float64), 1, ev_np, slots).mean()) \
        * truths.shape[0]
    achieved = flops / (adj_ms * 1e-3) / 1e12
    ms = 1e3 * float(np.mean(times))
    out = {"metric": "protocol-design gradients/s (CRLB of k, f; adjoint)", "value": 1e3 / ms,
           "unit": "gradients/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 (float32 I/O)",
           "data": "synthetic (seeded parameter sets)",
           "config": {"workload": "adjoint protocol gradient: 3 pools (water, amide, SL MT), 256-sample shaped pulse "
                                  "x 40 + gaps + 1 s recovery (10,283 events), 31 offsets, "
                                  f"{truths.shape[0]} parameter sets, d = 256 amplitudes + 31 offsets",
                      "truths": int(truths.shape[0]), "offsets": 31, "events_per_offset": int(ev_np["dt"].shape[1])},
           "roofline": {"bound": "valu", "bound_detail": "compute-bound in float64 vector FMA (adjoint kernel)",
                        "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None, "flops_per_launch": flops,
                        "avg_launch_ms": adj_ms},
           "cost": c}
    print(json.dumps(out), flush=True)


def uneven_workload(n_problems=4096, seed=3, tsat=2.0):
    """Heterogeneous QUESP fit problems: 1-8 rungs from {0.5, 1, 2, 3, 4, 5, 6, 8} uT (2 s CW), 31-121 offsets
    per rung, water plus 1-4 CEST pools at 1-4 ppm and an optional MT pool, 2-8 free values, seeded."""
    from bmc_jax.fitting import FitSpec, FreeParam
    from bmc_jax.fitting.uneven import UnevenProblem
    from bmc_jax.sequences import cw_saturation
    from bmc_jax.simulation.events import concat_events
    from bmc_jax.simulation.params import SimParams
    rng = np.random.default_rng(seed)
    probs, truths = [], []
    for _ in range(n_problems):
        nc = int(rng.integers(1, 5))
        mt = bool(rng.integers(0, 2))
        n_rungs = int(rng.integers(1, 9))
        n_off = int(rng.integers(31, 122))
        k = np.exp(rng.uniform(np.log(20.0), np.log(3000.0), nc))
        f = rng.uniform(5e-4, 1e-2, nc)
        dw = rng.uniform(1.0, 4.0, nc)
        t1w, t2w, b0i = rng.uniform(0.9, 2.5), rng.uniform(0.04, 0.12), rng.uniform(-0.2, 0.2)
        a = lambda v: np.asarray(v, np.float64)
        sp = SimParams(water_r1=a(1 / t1w), water_r2=a(1 / t2w), water_f=a(1.0), cest_k=a(k), cest_f=a(f),
                       cest_dw=a(dw), cest_r1=a(np.full(nc, 1 / 1.3)), cest_r2=a(np.full(nc, 20.0)),
                       mt_k=a(20.0 if mt else 0.0), mt_f=a(0.05 if mt else 0.0), mt_dw=a(-2.5 if mt else 0.0),
                       mt_r1=a(1.0 if mt else 0.0), mt_r2=a(1e5 if mt else 1.0), mt_lineshape=np.int32(1 if mt else 0),
                       b0=a(3.0), gamma=a(267.5153), b0_inhom=a(b0i), rel_b1=a(1.0))
        m0 = np.zeros(3 * (1 + nc) + 1)
        m0[2 * (1 + nc)] = 1.0
        m0[2 * (1 + nc) + 1:3 * (1 + nc)] = f
        m0[-1] = 0.05 if mt else 0.0
        b1s = rng.choice([0.5, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 8.0], size=n_rungs, replace=False)
        ev = concat_events([cw_saturation(np.linspace(-6, 6, n_off), b1_uT=float(b), tsat_s=tsat, b0_T=3.0)
                            for b in b1s])
        cands = []
        for i in range(nc):
            cands += [FreeParam(f"k{i}", "cest_k", i, 100.0, 1.0, 1e4, k[i]),
                      FreeParam(f"f{i}", "cest_f", i, 0.001, 1e-5, 0.05, f[i])]
        cands += [FreeParam("r1w", "water_r1", None, 1.0, 0.1, 3.0, 1 / t1w),
                  FreeParam("r2w", "water_r2", None, 10.0, 1.0, 60.0, 1 / t2w),
                  FreeParam("db0", "b0_inhom", None, 0.1, -0.5, 0.5, b0i)]
        if mt:
            cands += [FreeParam("kmt", "mt_k", None, 10.0, 1.0, 100.0, 20.0),
                      FreeParam("fmt", "mt_f", None, 0.01, 1e-3, 0.3, 0.05)]
        P = int(min(rng.integers(2, 9), len(cands)))
        spec = FitSpec([cands[j] for j in sorted(rng.choice(len(cands), size=P, replace=False))])
        truth = spec.x0_phys.copy()
        x0 = np.clip(truth * np.exp(rng.normal(0.0, 0.2, P)), spec.lo_phys, spec.hi_phys)
        x0 = np.where(np.abs(truth) < 1e-12, truth + 0.02, x0)
        probs.append(UnevenProblem(params=sp, events=ev, m0=m0, spec=spec, z_meas=np.zeros(ev.dt.shape[0]),
                                   x0_phys=x0))
        truths.append(truth)
    return probs, truths


def bench_uneven(args):
    import torch
    from bmc_jax import _flops, _native
    from bmc_jax.fitting.independent import _lm_batched
    from bmc_jax.fitting.uneven import UnevenBatch, autotune
    from bmc_jax.simulation.params import n_cest_of, pack
    from bmc_jax.simulation.voxels import tile_params
    from bmc_jax.fitting.residual import spec_slots
    t_build = time.perf_counter()
    probs, truths = uneven_workload(args.problems)
    b = UnevenBatch(probs, dtype=np.float32)
    # measured Z: every problem at its truth (the batch's own sweep), plus sigma 1e-3 Gaussian noise
    x_true = np.concatenate([p.spec.to_scaled(t) for p, t in zip(probs, truths)])
    z, _ = b.simulate(x_true, jacobian=False)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(9)
    b.z_meas = (z + 1e-3 * torch.randn(z.shape, generator=gen, device="cuda", dtype=torch.float64)).to(torch.float32)
    t_build = time.perf_counter() - t_build
    best, table = autotune(b, repeats=2, log=lambda m: print(m, file=sys.stderr, flush=True))
    n_iter = 30
    for _ in range(args.warmup):
        b.lm_fit(n_iter=n_iter)
    times = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, cost, nd = b.lm_fit(n_iter=n_iter)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    wall = float(np.mean(times))
    rms = np.sqrt(cost / b.n_rows)
    # the tangent sweep alone, and its flops from a sample of problems
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    xs = b.x0_scaled()
    b.simulate(xs)
    e0.record()
    b.simulate(xs)
    e1.record()
    torch.cuda.synchronize()
    sweep_ms = e0.elapsed_time(e1)
    ops, rp = b.plan.program()
    ev_all = {k: np.asarray(getattr(b.plan_events, k)) for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    # single-event rows take the reverse-accumulation sweep (csrc/bmc_cwj.hip): count its own flops
    live = _flops.single_event_rows(ev_all)
    rng = np.random.default_rng(0)
    sample = rng.choice(b.n, size=min(64, b.n), replace=False)
    fl = 0.0
    for i in sample:
        p = probs[i]
        nc = n_cest_of(p.params)
        r0, nr = int(b.row0[i]), int(b.n_rows[i])
        ev_i = {k: v[r0:r0 + nr] for k, v in ev_all.items()}
        from bmc_jax.fitting.params import write_params
        pw = write_params(p.params, p.spec.to_scaled(p.x0_phys), p.spec)
        if live is not None:
            fl += float(_flops.reverse_flops(pack(pw, 1), nc, ev_i, spec_slots(p.spec, nc), live[r0:r0 + nr])[0])
        else:
            fl += float(_flops.executed_flops(pack(pw, 1), nc, ev_i, ops, rp[r0:r0 + nr + 1], spec_slots(p.spec, nc),
                                              pt=best[0], mirror=False)[0])
    flops_sweep = fl / len(sample) * b.n
    achieved = flops_sweep / (sweep_ms * 1e-3) / 1e12
    # one bmc_lm_fit per problem (what a per-shape loop of `fit` calls launches), on a sample
    n_one = min(64, b.n)
    zm = b.z_meas.cpu().numpy()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(n_one):
        p = probs[i]
        nc = n_cest_of(p.params)
        r0, nr = int(b.row0[i]), int(b.n_rows[i])
        _lm_batched(tile_params(p.params, 1), p.events, np.asarray(p.m0)[None], 2 * (1 + nc), p.spec,
                    zm[None, r0:r0 + nr], p.spec.to_scaled(p.x0_phys)[None], n_iter, np.float32)
    torch.cuda.synchronize()
    per_problem_s = (time.perf_counter() - t0) / n_one
    out = {"metric": "LM-fitted problems/s (30 iterations, uneven batch)", "value": b.n / wall,
           "unit": "problems/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * wall,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64 (float32 I/O)",
           "data": "synthetic (seeded problems, sigma 1e-3 noise)",
           "config": {"workload": "uneven batched LM: 4096 QUESP problems, 1-8 rungs x 31-121 offsets (2 s CW), "
                                  "1-4 CEST pools + optional MT, 2-8 free values, 30 LM iterations in one hipGraph",
                      "problems": b.n, "rows": int(b.R), "values": int(b.n_values),
                      "autotuned": {"tangents_per_pass": best[0], "grid": best[1], "order": best[2]}},
           "roofline": {"bound": "valu", "bound_detail": "compute-bound in float64 vector FMA (persistent "
                        + ("reverse-accumulation Jacobian sweep)" if live is not None else "sweep)"),
                        "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None, "flops_per_launch": flops_sweep,
                        "avg_launch_ms": sweep_ms},
           "autotune_ms": {f"{k[0]},{k[1]},{k[2]}": v for k, v in table.items()},
           "per_problem_lm_fit_s": per_problem_s, "speedup_vs_per_problem_launches": per_problem_s * b.n / wall,
           "median_rms": float(np.median(rms)), "setup_s": t_build}
    print(json.dumps(out), flush=True)


QUESP_B1_UT = (0.15, 0.3, 0.5, 0.8, 1.0, 1.3, 1.5, 1.8)
# (pool offset ppm, k /s, f, pool T2 s) of six tubes: two of each agent, as in the reference's phantom
TUBES = ((1.0, 1500.0, 5.4e-3, 0.010), (1.0, 1300.0, 2.7e-3, 0.012), (1.9, 900.0, 3.2e-3, 0.020),
         (1.9, 250.0, 3.2e-3, 0.020), (3.2, 120.0, 4.5e-4, 0.015), (3.2, 600.0, 4.5e-4, 0.015))
REF_PHANTOM_FIT_S = 900.0   # examples/README.md:60-64 of the reference: about 15 min on one RTX 3070


def scanner_workload(n_vox=6, seed=8, noise=1e-3):
    """The reference's phantom QUESP fit (examples/08_quesp.py) at its shape, on synthetic data: 8 rungs
    (0.15-1.8 uT) of a converted scanner sequence (20 Gaussian pulses x 200 samples of 0.5 ms, each at
    its accumulated RF phase, spoiler + delay gaps, 4 s lead: 4041 events) at B0 2.8936 T, the 82
    offsets of linspace(-5, 5, 101) with |ppm| >= 1, stacked into one 656-row experiment. Voxels cycle
    through six tubes (one CEST pool each), water R1 drawn per voxel, water R2 0.5605 /s, start
    magnetisation 0.557; three free values (k, f, pool R2) from the tutorial's start (300 /s, 0.008,
    100 /s); Z is the float64 forward model at the truth plus N(0, noise)."""
    from bmc_jax.fitting import Experiment, FitSpec, FreeParam
    from bmc_jax.sequences import scanner_train
    from bmc_jax.simulation.params import SimParams, initial_state
    rng = np.random.default_rng(seed)
    ppm = np.linspace(-5.0, 5.0, 101)
    keep = np.abs(ppm) >= 1.0
    # float32 event values, as the reference's cache loader casts them
    rungs = [scanner_train(ppm[keep], b1, b0_T=2.8936, dtype=np.float32) for b1 in QUESP_B1_UT]
    tube = np.arange(n_vox) % len(TUBES)
    dw, k, f, t2p = (np.array([TUBES[t][j] for t in tube]) for j in range(4))
    a = lambda v: np.asarray(v, np.float64)
    sp = SimParams(water_r1=a(rng.uniform(0.30, 0.45, n_vox)), water_r2=a(np.full(n_vox, 0.5605)),
                   water_f=a(np.ones(n_vox)), cest_k=a(k[:, None]), cest_f=a(f[:, None]), cest_dw=a(dw[:, None]),
                   cest_r1=a(np.ones((n_vox, 1))), cest_r2=a(1.0 / t2p[:, None]), mt_k=a(np.zeros(n_vox)),
                   mt_f=a(np.zeros(n_vox)), mt_dw=a(np.zeros(n_vox)), mt_r1=a(np.zeros(n_vox)),
                   mt_r2=a(np.ones(n_vox)), mt_lineshape=np.zeros(n_vox, np.int32), b0=a(np.full(n_vox, 2.8936)),
                   gamma=a(np.full(n_vox, 267.5153)), b0_inhom=a(np.zeros(n_vox)),
                   rel_b1=a(rng.uniform(0.95, 1.05, n_vox)))
    m0 = np.stack([initial_state(SimParams(*[np.asarray(v)[i] for v in sp]), m_start=0.557) for i in range(n_vox)])
    spec = FitSpec([FreeParam("k", "cest_k", 0, 1000.0, 5.0, 2000.0, 300.0),
                    FreeParam("f", "cest_f", 0, 0.01, 1e-5, 0.03, 0.008),
                    FreeParam("r2p", "cest_r2", 0, 100.0, 10.0, 1000.0, 100.0)])
    ladder = Experiment.stack([Experiment(sp, ev, m0, 4) for ev in rungs])
    truth = np.stack([k, f, 1.0 / t2p], axis=1)
    return ladder, spec, truth, rng


def bench_scanner(args):
    """Wall time of the drop-in fit() of the phantom QUESP workload, 60 LM iterations, against the
    reference's published time for its tutorial 08; then the same protocol over an image of voxels."""
    import torch
    from bmc_jax import _engine
    from bmc_jax.fitting import fit
    out = {}
    for n_vox in (6, args.image_voxels):
        ladder, spec, truth, rng = scanner_workload(n_vox)
        z_true = _engine.simulate_states(ladder.params, ladder.events, ladder.m0, n_voxels=n_vox,
                                         dtype=np.float64)[..., ladder.mz_loc]   # float64 model at the truth
        z = z_true + 1e-3 * rng.standard_normal(z_true.shape)
        fit(ladder, spec, z, n_iter=2)                          # warm: plan, graph, allocator
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = fit(ladder, spec, z, n_iter=60, n_per_block=82)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        rel = np.abs(np.asarray(res.x) / truth - 1.0)
        plan = _engine.plan_for(ladder.events)
        ops, _ = plan.program()
        out[n_vox] = dict(seconds=t, voxels_per_s=n_vox / t, median_rms=float(np.median(res.rms)),
                          median_rel_err_k_f_r2p=np.median(rel, axis=0).tolist(),
                          rows=int(ladder.n_offsets), events_per_row=int(ladder.n_events),
                          kept_pulses_per_row=float(np.sum(ops[:, 0] == 2) / ladder.n_offsets),
                          reused_pulses_per_row=float(np.sum(ops[:, 0] == 3) / ladder.n_offsets))
        print(f"{n_vox} voxels: {t:.3f} s", file=sys.stderr, flush=True)
    six = out[6]
    line = {"metric": "phantom QUESP fit wall time (reference tutorial 08 shape: 6 tubes x 8 rungs x 82 offsets x "
                      "4041 events, 3 free values, 60 LM iterations, drop-in fit())",
            "value": six["seconds"], "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": 1e3 * six["seconds"], "higher_is_better": False, "scaling": "weak",
            "vs_baseline": six["seconds"] / REF_PHANTOM_FIT_S,
            "vs_baseline_note": "reference's published ~15 min on one RTX 3070 (examples/README.md:60-64)",
            "dtype": "f64 simulate (f32 I/O), f64 normal equations",
            "data": "synthetic Z (float64 forward model at seeded truths + N(0, 1e-3)) on synthetic events of the "
                    "converted scanner sequences' layout",
            "config": {"workload": "scanner QUESP fit", "rungs_uT": list(QUESP_B1_UT)},
            "six_tubes": six, "image": dict(voxels=args.image_voxels, **out[args.image_voxels])}
    print(json.dumps(line), flush=True)


PHANTOM_DIR = os.path.join(ROOT, "tests", "golden", "phantom")
# the fits of record drop offsets within 1 ppm of water; the CEST pool's T2 starts at 10 ms
PHANTOM_EXCLUDE_PPM, PHANTOM_POOL_T2 = 1.0, 0.010
# pair checks: fitted f ratios that a shared calibration error cannot move (GAG 100 / 50 mg/mL; one creatine
# and one carnosine solution each at two pH values)
PHANTOM_PAIRS = (("GAG 100 / 50 mg/mL", "GAG100_pH6.31", "GAG50_pH6.20", 2.0),
                 ("creatine 7.51 / 6.31", "creatine_pH7.51", "creatine_pH6.31", 1.0),
                 ("carnosine 7.44 / 6.59", "carnosine_pH7.44", "carnosine_pH6.59", 1.0))


def phantom_experiment(kind, tubes="pool"):
    """The measured seven-tube phantom (2026-09-27, 2.8936 T) as the reference's tutorials 08 (QUESP, eight
    B1 rungs) and 09 (QUEST, six saturation times) fit it: the six tubes with a CEST pool, each with its
    measured water R1, B0 offset and relative B1, water R2 from the PBS tube, start magnetisation 0.557, the
    scanner sequences converted to events (float32, reference row dropped), offsets within 1 ppm of water
    left out of both data and model, the rungs stacked. tubes="all" adds the PBS tube (water only: a pool of
    size and rate 0). Returns (experiment, spec, Z, noise, names, n_per_block)."""
    from bmc_jax.fitting import Experiment, FitSpec, FreeParam
    from bmc_jax.sequences import load_events
    from bmc_jax.simulation.params import SimParams, initial_state
    with np.load(os.path.join(PHANTOM_DIR, "phantom_7tubes.npz"), allow_pickle=False) as d:
        data = {k: d[k] for k in d.files}
    has_pool = np.isfinite(data["vial_dw"])
    tubes = np.flatnonzero(has_pool) if tubes == "pool" else np.arange(has_pool.size)
    pool = has_pool[tubes]
    ppm = data["ppm"]
    keep = np.abs(ppm) >= PHANTOM_EXCLUDE_PPM
    n = tubes.size
    one = np.ones(n, np.float32)
    f32 = lambda v: np.asarray(v, np.float32)
    sp = SimParams(water_r1=f32(data["R1"][tubes]), water_r2=f32(float(data["water_r2"]) * one), water_f=f32(one),
                   cest_k=f32(np.where(pool, 900.0, 0.0)[:, None]), cest_f=f32(np.where(pool, 0.008, 0.0)[:, None]),
                   cest_dw=f32(np.where(pool, data["vial_dw"][tubes], 0.0)[:, None]), cest_r1=f32(np.ones((n, 1))),
                   cest_r2=f32(np.full((n, 1), 1.0 / PHANTOM_POOL_T2)), mt_k=f32(0 * one), mt_f=f32(0 * one),
                   mt_dw=f32(0 * one), mt_r1=f32(0 * one), mt_r2=f32(one), mt_lineshape=np.zeros(n, np.int32),
                   b0=f32(float(data["b0_T"]) * one), gamma=f32(267.5153 * one),
                   b0_inhom=f32(data["b0_hz"][tubes] / (42.5764 * float(data["b0_T"]))),
                   rel_b1=f32(data["rel_b1"][tubes]))
    m_start = float(data["m_start"])
    m0 = np.stack([initial_state(SimParams(*[np.asarray(v)[i] for v in sp]), m_start=m_start) for i in range(n)])
    if kind == "quesp":
        stems = [f"quesp_sim_b1_{('%g' % b).replace('.', 'p')}uT" for b in data["b1_levels_uT"]]
        Z, noise = data["Z_quesp"], data["noise_quesp"]
    else:
        stems = [f"quest_sim_tsat_{('%g' % t).replace('.', 'p')}s" for t in data["tsat_s"]]
        Z, noise = data["Z_quest"], data["noise_quest"]
    rungs = [Experiment(sp, load_events(os.path.join(PHANTOM_DIR, f"{st}_events.npz"), keep=keep), m0, 4)
             for st in stems]
    spec = FitSpec([FreeParam("k", "cest_k", 0, scale=1000.0, lo=5.0, hi=2000.0, x0=300.0),
                    FreeParam("f", "cest_f", 0, scale=0.01, lo=1e-5, hi=0.03, x0=0.008),
                    FreeParam("r2p", "cest_r2", 0, scale=100.0, lo=10.0, hi=1000.0, x0=1.0 / PHANTOM_POOL_T2)])
    names = [str(v) for v in data["vial_names"][tubes]]
    return (Experiment.stack(rungs), spec, Z[tubes][:, :, keep].reshape(n, -1), noise[tubes], names,
            int(keep.sum()))


def phantom_pairs(f, names):
    return [dict(pair=label, f_ratio=float(f[names.index(a)] / f[names.index(b)]), expected=want)
            for label, a, b, want in PHANTOM_PAIRS]


def bench_phantom(args, kind):
    """The reference's tutorial 08 (QUESP) or 09 (QUEST) on the shipped phantom measurement: the drop-in
    fit() with 60 LM iterations, wall time against the reference's published ~15 min (one RTX 3070)."""
    import torch
    from bmc_jax.fitting import fit
    ladder, spec, Z, noise, names, npb = phantom_experiment(kind)
    fit(ladder, spec, Z, n_iter=2, n_per_block=npb)               # warm: plan, graph, allocator
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = fit(ladder, spec, Z, n_iter=60, n_per_block=npb)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    tubes = [dict(tube=nm, k=float(res["k"][i]), f=float(res["f"][i]), t2_pool_ms=1e3 / float(res["r2p"][i]),
                  rms_over_noise=float(res.rms[i] / noise[i]),
                  at_bounds=[s.name for s, b in zip(spec.slots, res.at_bounds[i]) if b]) for i, nm in enumerate(names)]
    tutorial = "08 (QUESP, 8 B1 rungs)" if kind == "quesp" else "09 (QUEST, 6 saturation times)"
    line = {"metric": f"measured-phantom fit wall time, reference tutorial {tutorial}: 6 tubes, 3 free values, "
                      "60 LM iterations, drop-in fit()",
            "value": t, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1e3 * t,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": t / REF_PHANTOM_FIT_S,
            "vs_baseline_note": "reference's published ~15 min on one RTX 3070 (examples/README.md:60-64)",
            "dtype": "f64 simulate (f32 I/O), f64 normal equations",
            "data": "the reference's shipped 2026-09-27 phantom scan and converted scanner sequences",
            "config": {"workload": f"phantom {kind}", "rows": int(ladder.n_offsets),
                       "events_per_row": int(ladder.n_events)},
            "tubes": tubes, "pair_checks": phantom_pairs(np.asarray(res["f"]), names)}
    print(json.dumps(line), flush=True)
    return line


REF_JOINT_FIT_S = 7200.0    # examples/README.md:65 of the reference: about 2 h on one RTX 3070
REF_PER_RUNG_FIT_S = 12600.0  # examples/README.md:66: about 3.5 h on one RTX 3070


def bench_phantom_joint(args, per_rung):
    """The reference's tutorial 10 (examples/10_shared_water.py) and, with per_rung, 11 (11_per_rung.py)
    on the shipped phantom measurement: k, f, pool T2 of the six pool tubes (fit, 60 iterations), then
    fit_joint over all seven tubes with the water R2 factor, M_start and a 15-node B0 spread shared (12
    steps), then (11) the same with one B1 factor per rung shared by the tubes (12 steps)."""
    import torch
    from dataclasses import replace
    from bmc_jax.fitting import FitSpec, FreeParam, fit, fit_joint
    from bmc_jax.simulation import B0Spread
    ladder6, spec3, Z6, _, names6, npb = phantom_experiment("quesp", "pool")
    ladder, _, Z, noise, names, _ = phantom_experiment("quesp", "all")
    pool = [names.index(nm) for nm in names6]
    K, F, T2 = spec3.slots
    R2F = FreeParam("r2w", "water_r2", None, scale=1.0, lo=0.5, hi=2.0, x0=1.0, scope="shared", mode="factor")
    MS = FreeParam("m_start", "m_start", None, scale=1.0, lo=0.2, hi=1.0, x0=0.557, scope="shared")
    SIG = FreeParam("sigma", "b0_spread", None, scale=10.0, lo=0.0, hi=8.0, x0=5.0, scope="shared")
    B1R = FreeParam("b1", "rel_b1", None, scale=1.0, lo=0.7, hi=1.3, x0=1.0, scope="block", mode="factor")
    spread = B0Spread(nodes=15)
    Kp, Fp, T2p = (replace(sl, samples=tuple(pool)) for sl in (K, F, T2))
    spec10 = FitSpec([Kp, Fp, T2p, R2F, MS, SIG])
    fit(ladder6, spec3, Z6, n_iter=2, n_per_block=npb)            # warm: plans, graph, allocator
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain = fit(ladder6, spec3, Z6, n_iter=60, n_per_block=npb)
    x0 = np.tile([sl.x0 for sl in spec10.slots], (len(names), 1))
    x0[pool, :3] = plain.x
    water = fit_joint(ladder, spec10, Z, n_iter=12, n_per_block=npb, x0_phys=x0, spread=spread, max_chunk=3)
    torch.cuda.synchronize()
    t10 = time.perf_counter() - t0
    out = {"tutorial10": {"seconds": t10, "shared": {"water_r2_factor": float(water["r2w"][0]),
                                                     "m_start": float(water["m_start"][0]),
                                                     "b0_spread_hz": float(water["sigma"][0])},
                          "cost_first_last": [float(water.cost[0]), float(water.cost[-1])],
                          "rms_over_noise": (water.rms / noise).tolist(),
                          "pair_checks": phantom_pairs(np.asarray(water["f"]), names)}}
    ref, secs, label = REF_JOINT_FIT_S, t10, "10 (shared water)"
    if per_rung:
        t1 = time.perf_counter()
        x1 = np.column_stack([water.x, np.ones(len(names))])
        rung = fit_joint(ladder, FitSpec([Kp, Fp, T2p, R2F, MS, SIG, B1R]), Z, n_iter=12, n_per_block=npb,
                         x0_phys=x1, spread=spread, max_chunk=3, block_jacobian="split")
        torch.cuda.synchronize()
        t11 = time.perf_counter() - t1
        b1 = rung.blocks["b1"]
        out["tutorial11"] = {"seconds_after_10": t11, "b1_per_rung": b1["x"].tolist(), "b1_sd": b1["sd"].tolist(),
                             "shared": {"water_r2_factor": float(rung["r2w"][0]), "m_start": float(rung["m_start"][0]),
                                        "b0_spread_hz": float(rung["sigma"][0])},
                             "pair_checks": phantom_pairs(np.asarray(rung["f"]), names)}
        ref, secs, label = REF_PER_RUNG_FIT_S, t10 + t11, "11 (shared water, then B1 per rung)"
    line = {"metric": f"measured-phantom joint fit wall time, reference tutorial {label}: 7 tubes, 15-node B0 "
                      "spread, drop-in fit() and fit_joint()",
            "value": secs, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1e3 * secs,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": secs / ref,
            "vs_baseline_note": f"reference's published ~{ref / 3600:.1f} h on one RTX 3070 (examples/README.md:60-66)",
            "dtype": "f64 simulate (f32 I/O), f64 LM steps",
            "data": "the reference's shipped 2026-09-27 phantom scan and converted scanner sequences",
            "config": {"workload": "phantom joint", "tubes": len(names), "spread_nodes": 15, "joint_steps": 12}}
    line.update(out)
    print(json.dumps(line), flush=True)
    return line


def bench_joint(args):
    """The reference's tutorial 10 (examples/10_shared_water.py) at its shape, on synthetic data: seven
    tubes (a water-only tube plus the six tubes of the scanner workload) on the same 656-row QUESP
    ladder. Step 1 fits k, f and pool R2 of the six pool tubes with the water held (fit, 60 iterations);
    step 2 is fit_joint over all seven, 12 LM steps, with the water R2 factor, M_start and the B0 spread
    (Gauss-Hermite, 15 nodes) shared by every tube. Z is the float64 model at truths with those shared
    values (R2 x 1.1, M_start 0.60, spread 3 Hz) plus N(0, 1e-3)."""
    import torch
    from dataclasses import replace
    from bmc_jax import _engine
    from bmc_jax.fitting import Experiment, FitSpec, fit, fit_joint, FreeParam
    from bmc_jax.simulation import B0Spread, SimParams, initial_state
    from bmc_jax.simulation.spread import node_params
    ladder6, spec6, truth6, rng = scanner_workload(6, seed=10)
    # tube 0 holds water only: pool of size 0 and rate 0 (its water is exactly that of no pool)
    cat = lambda a, v: np.concatenate([np.asarray(v, np.asarray(a).dtype).reshape((1,) + np.shape(a)[1:]), a])
    p6 = ladder6.params
    p7 = SimParams(*[cat(getattr(p6, f), v) for f, v in zip(SimParams._fields, (
        0.35, 0.5605, 1.0, [0.0], [0.0], [0.0], [1.0], [100.0], 0.0, 0.0, 0.0, 0.0, 1.0, 0, 2.8936, 267.5153, 0.0,
        1.0))])
    pool = tuple(range(1, 7))
    events = ladder6.events
    r2_factor, m_start, sigma = 1.1, 0.60, 3.0
    truth_p = p7._replace(water_r2=p7.water_r2 * r2_factor)
    spread = B0Spread(nodes=15)
    _, w = spread.nodes_weights()
    z_true = np.zeros((7, ladder6.n_offsets))
    for i in range(7):
        one = SimParams(*[np.asarray(f)[i] for f in truth_p])
        nodes = node_params(one, sigma, spread)
        m0 = initial_state(one, m_start=m_start)
        m = _engine.simulate_states(nodes, events, m0, n_voxels=w.size, dtype=np.float64)
        z_true[i] = np.tensordot(w, m[..., ladder6.mz_loc], axes=(0, 0))
    z = z_true + 1e-3 * rng.standard_normal(z_true.shape)
    m0_held = np.stack([initial_state(SimParams(*[np.asarray(f)[i] for f in p7]), m_start=0.557) for i in range(7)])
    ladder7 = Experiment(p7, events, m0_held, ladder6.mz_loc)
    lad6 = Experiment(p6, events, m0_held[1:], ladder6.mz_loc)
    K, F, T2 = spec6.slots
    Kp, Fp, T2p = (replace(q, samples=pool) for q in (K, F, T2))
    R2F = FreeParam("r2w", "water_r2", None, 1.0, 0.5, 2.0, 1.0, scope="shared", mode="factor")
    MS = FreeParam("m_start", "m_start", None, 1.0, 0.2, 1.0, 0.557, scope="shared")
    SIG = FreeParam("sigma", "b0_spread", None, 10.0, 0.0, 8.0, 5.0, scope="shared")
    spec = FitSpec([Kp, Fp, T2p, R2F, MS, SIG])
    fit(lad6, spec6, z[1:], n_iter=2)                                  # warm: plan, graph
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain = fit(lad6, spec6, z[1:], n_iter=60, n_per_block=82)
    x0 = np.tile([q.x0 for q in spec.slots], (7, 1))
    x0[list(pool), :3] = plain.x
    costs = []
    joint = fit_joint(ladder7, spec, z, n_iter=12, n_per_block=82, x0_phys=x0, spread=spread, max_chunk=3,
                      callback=lambda it, c: costs.append(float(c)))
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    shared = [float(joint["r2w"][0]), float(joint["m_start"][0]), float(joint["sigma"][0])]
    rel = np.abs(np.asarray(joint.x)[list(pool), :3] / np.asarray(truth6) - 1.0)
    line = {"metric": "joint phantom fit wall time (reference tutorial 10 shape: 6-tube fit, 60 iterations, then "
                      "fit_joint of 7 tubes with shared water R2 factor, M_start and 15-node B0 spread, 12 steps)",
            "value": t, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1e3 * t,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": t / REF_JOINT_FIT_S,
            "vs_baseline_note": "reference's published ~2 h on one RTX 3070 (examples/README.md:65)",
            "dtype": "f64 simulate (f32 I/O), f64 host LM step",
            "data": "synthetic Z (float64 model at seeded truths, shared R2 x1.1, M_start 0.60, 3 Hz spread, "
                    "+ N(0, 1e-3)) on synthetic events of the converted scanner sequences' layout",
            "config": {"workload": "joint QUESP fit", "tubes": 7, "spread_nodes": 15, "joint_steps": 12},
            "shared_fitted": {"r2w_factor": shared[0], "m_start": shared[1], "sigma_hz": shared[2]},
            "shared_truth": {"r2w_factor": r2_factor, "m_start": m_start, "sigma_hz": sigma},
            "median_rel_err_k_f_r2p": np.median(rel, axis=0).tolist(), "cost_per_step": costs,
            "median_rms": float(np.median(joint.rms))}
    print(json.dumps(line), flush=True)


def host_threads():
    """Threads for the CPU baselines: the host's cores, at most the 16 a one-GPU box grants a job."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prm, m0, ev, nc, slots, budget_s=15.0):
    """The C oracle on this host's cores: whole spectra of the first voxels, as many as fit the budget."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_c as OC
    threads = host_threads()
    ev64 = {k: np.asarray(getattr(ev, k)) for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    O = ev64["dt"].shape[0]
    t0 = time.perf_counter()
    OC.simulate(prm[:1].astype(np.float64), nc, ev64, m0[:1].astype(np.float64), slots=slots, threads=threads)
    one = time.perf_counter() - t0
    n_vox = int(min(64, max(1, budget_s / max(one, 1e-9))))
    t0 = time.perf_counter()
    OC.simulate(prm[1:1 + n_vox].astype(np.float64), nc, ev64, m0[1:1 + n_vox].astype(np.float64), slots=slots,
                threads=threads)
    wall = time.perf_counter() - t0
    return dict(value=n_vox / wall, unit="Z-spectra/s (fwd+Jac)", cores=threads, kind="port", cpu_model=cpu_model(),
                sample=f"{n_vox} voxels x {O} offsets x {ev64['dt'].shape[1]:,} events, 5 tangents, float64 "
                       f"Pade-13 per event (no reuse), {wall:.1f} s wall on {threads} threads")


PEAK_BF16_TFLOPS = 2500.0


def cpu_lm_baseline(Z, x0, n_iter=30, budget_s=15.0, rms_tol=1e-4):
    """The reference's LM (lm.py:43-61: every voxel its own damping, all n_iter iterations, no early exit)
    with the float64 C oracle as model and Jacobian, on this host's cores, over the first voxels of the
    held-out ladder set: fitted voxels per second (rms at most rms_tol), extrapolated from the sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_c as OC
    from bmc_jax import learned
    from bmc_jax.simulation.params import initial_state, pack
    threads = host_threads()
    spec = learned.fit_spec()
    base = learned.base_params(1, dtype=np.float64)
    rec = pack(base, 1, dtype=np.float64)[0]
    m0 = np.asarray(initial_state(base), np.float64).reshape(-1)
    slots = [_slot(f, i) for f, i in (("cest_k", 0), ("cest_f", 0), ("water_r1", None), ("water_r2", None),
                                      ("b0_inhom", None))]
    ev = learned.ladder_events()
    ev64 = {k: np.asarray(getattr(ev, k)) for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    sc = spec.scales
    lo, hi = spec.lo_phys / sc, spec.hi_phys / sc

    def model(x, want_j=True):
        prm = np.repeat(rec[None], x.shape[0], axis=0)
        prm[:, slots] = x * sc
        m, dm = OC.simulate(prm, 1, ev64, m0, slots=slots if want_j else (), threads=threads)
        return m[:, :, 4], (dm[:, :, :, 4] * sc if want_j else None)

    t0 = time.perf_counter()
    model(x0[:1] / sc)
    one = time.perf_counter() - t0
    V = int(min(len(Z), max(8, budget_s / max(2 * n_iter * one, 1e-9))))
    z = np.asarray(Z[:V], np.float64)
    x = np.asarray(x0[:V], np.float64) / sc
    lam = np.full(V, 1e-2)
    eye = np.eye(len(slots))
    t0 = time.perf_counter()
    for _ in range(n_iter):
        zm, J = model(x)
        r = zm - z
        c = np.einsum("vo,vo->v", r, r)
        H = np.einsum("vop,voq->vpq", J, J)
        g = np.einsum("vop,vo->vp", J, r)
        A = H + lam[:, None, None] * (H * eye + 1e-9 * eye)
        xt = np.clip(x + np.linalg.solve(A, -g[..., None])[..., 0], lo, hi)
        zt, _ = model(xt, want_j=False)
        ct = np.einsum("vo,vo->v", zt - z, zt - z)
        ok = ct < c
        x = np.where(ok[:, None], xt, x)
        lam = np.where(ok, lam * 0.5, lam * 3.0)
    wall = time.perf_counter() - t0
    rms = np.sqrt(np.minimum(ct, c) / z.shape[1])
    fitted = int(np.sum(rms <= rms_tol))
    return dict(value=fitted / wall, unit="fitted voxels/s (plain-x0 LM, all 30 iterations)", cores=threads,
                kind="port", cpu_model=cpu_model(),
                sample=f"{V} held-out voxels x 244 CW rows, 5 free values, 30 iterations of the reference LM with "
                       f"the float64 C oracle (Pade-13, forward-mode tangents) as model, {wall:.1f} s on {threads} "
                       f"threads, {fitted} of {V} at rms <= {rms_tol:g}")


def _slot(field, index):
    from bmc_jax import _native
    return _native.slot_index(field, index, 1)


def nn_record(args, log=None):
    """The learned-proposal comparison (NN-started LM against plain-x0 LM, both early-exit and all 30
    iterations) on the 262,144-voxel held-out ladder set, with the MLP's own roofline and a CPU baseline."""
    import torch
    from bmc_jax import _native, learned
    r = learned.benchmark(n_train=args.nn_train, n_test=args.nn_test, epochs=args.nn_epochs, log=log)
    prop = r.pop("proposal")
    nf = r["noise_free"]
    # the MLP's GEMMs at their roofline: inference over the held-out spectra, timed on the stream
    Zt = r.pop("Z_noise_free")
    stream = torch.cuda.current_stream()
    prop.mlp.forward(Zt)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    prop.mlp.forward(Zt)
    e1.record(stream)
    torch.cuda.synchronize()
    t_inf = 1e-3 * e0.elapsed_time(e1)
    dims = learned.DIMS
    mac = sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1))
    inf_tflops = 2.0 * mac * Zt.shape[0] / t_inf / 1e12
    train_flops = 3 * 2.0 * mac * (args.nn_train // 4096) * 4096 * args.nn_epochs
    out = {
        "metric": "fitted voxels/s (NN-initialised LM vs plain LM, rms <= 1e-4, noise-free held-out ladders)",
        "value": nf["nn"]["fitted_per_s"], "unit": "fitted voxels/s",
        "speedup_vs_plain_all_iterations": nf["speedup_vs_fixed"], "speedup_vs_plain_early_exit": nf["speedup"],
        "nn_lm": nf["nn"], "plain_lm_all_iterations": nf["plain_fixed"], "plain_lm_early_exit": nf["plain"],
        "noisy_sigma_1e-3_rms_le_1.15_sigma": r["noisy"],
        "config": {"workload": "learned LM proposal: 2-pool CW QUESP ladder (0.5, 1, 2, 4 uT, 2 s), 61 offsets, "
                               "MLP 244-512-512-256-5 GELU, input standardised", "n_train": args.nn_train,
                   "epochs": args.nn_epochs, "n_test": args.nn_test},
        "train_seconds": r["train_seconds"], "final_train_loss": r["final_train_loss"],
        "roofline": {"bound": "mfma", "kernel": "gemm_nt_kernel (bf16 MFMA 32x32x16, fp32 accumulate)",
                     "achieved": inf_tflops, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": inf_tflops / PEAK_BF16_TFLOPS,
                     "measured_on": f"MLP inference over {Zt.shape[0]} spectra (batches of {prop.mlp.max_batch}), "
                                    f"{1e3 * t_inf:.2f} ms including the input standardisation tiles",
                     "training_tflops": train_flops / max(r["train_seconds"], 1e-9) / 1e12},
    }
    if not args.no_cpu:
        rng = np.random.default_rng(55)
        x_true, _ = learned.draw_params(64, rng)
        out["cpu_baseline"] = cpu_lm_baseline(learned.simulate_ladder(x_true),
                                              np.broadcast_to(learned.fit_spec().x0_phys, (64, 5)))
    return out


def bench_mlp(args):
    """The learned proposal's network alone: BF16 MFMA inference of 244-512-512-256-5 over 65,536
    synthetic Z-spectrum-like inputs per launch (random weights, inputs standardised), timed with HIP
    events on the launch stream over `steps` forwards."""
    import torch
    from bmc_jax import learned
    n = 65536
    mlp = learned.MLP(dims=learned.DIMS, max_batch=n, seed=1)
    g = torch.Generator(device="cuda").manual_seed(0)
    profile = 0.9 - 0.3 * torch.sin(torch.linspace(0, 3, learned.DIMS[0], device="cuda"))
    X = profile + 0.03 * torch.randn(n, learned.DIMS[0], device="cuda", generator=g)
    mlp.fit_input_norm(X)
    for _ in range(max(args.warmup, 1)):
        mlp.forward(X)
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        mlp.forward(X)
    e1.record(stream)
    torch.cuda.synchronize()
    t = 1e-3 * e0.elapsed_time(e1) / args.steps
    dims = learned.DIMS
    flops = 2.0 * n * sum(dims[i] * dims[i + 1] for i in range(len(dims) - 1))
    print(json.dumps({"metric": "MLP inference spectra/s (244-512-512-256-5, BF16 MFMA)", "value": n / t,
                      "unit": "spectra/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": 1e3 * t, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "bf16 / fp32 accumulate", "data": "synthetic (random weights and inputs)",
                      "config": {"workload": "MLP inference, 65,536 spectra per launch"},
                      "roofline": {"bound": "mfma", "achieved": flops / t / 1e12, "peak": 2500.0,
                                   "unit": "TFLOP/s", "frac": flops / t / 1e12 / 2500.0, "traffic": None}}),
          flush=True)


def bench_squaring(args):
    """The squarings' dense product on one GPU, two ways (bmc_jax._probe): the library's team column product
    (mul_cols: DPP or ds_bpermute broadcasts of the left factor's columns) against one FP64 MFMA tile
    (v_mfma_f64_16x16x4f64, factors padded to 16 x 16). reps products E <- P E on n matrices per launch."""
    import torch
    from bmc_jax import _probe
    n, reps = 65536, 200
    rng = np.random.default_rng(11)
    recs = []
    for nc in (1, 2, 3, 4):
        na = 3 * (1 + nc) + 2
        N = na - 1
        # an orthogonal linear part keeps P^r E bounded over the repetitions
        Q, _ = np.linalg.qr(rng.standard_normal((N, N)))
        P = np.eye(na)
        P[:N, :N] = Q
        P[:N, N] = 0.01 * rng.standard_normal(N)
        E = np.zeros((n, na, na))
        E[:, :N, :] = rng.standard_normal((n, N, na))
        E[:, N, N] = 1.0
        Pd = torch.tensor(P, device="cuda")
        outs = {}
        for path, name in ((_probe.PATH_COLS, "team_columns"), (_probe.PATH_MFMA, "mfma_f64_16x16x4"),
                           (_probe.PATH_STAGED, "team_columns_lds_staged")):
            if (path == _probe.PATH_MFMA and na > 16) or (path == _probe.PATH_STAGED and nc != 2):
                continue
            Ed = torch.tensor(E, device="cuda")
            _probe.products(path, nc, Pd, Ed, 2)                     # warm-up
            Ed.copy_(torch.tensor(E, device="cuda"))
            stream = torch.cuda.current_stream()
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record(stream)
            _probe.products(path, nc, Pd, Ed, reps)
            t1.record(stream)
            torch.cuda.synchronize()
            ms = t0.elapsed_time(t1)
            outs[name] = Ed.cpu().numpy()
            rate = n * reps / (ms * 1e-3)
            useful = 2.0 * N * N * na                                 # the augmented product's flops
            team = 8 if na <= 8 else 12 if na <= 12 else 16 if na <= 16 else 32
            recs.append({"n_cest": nc, "augmented": na, "path": name,
                         "lanes": team if path == _probe.PATH_COLS else 64,
                         "ms": ms, "products_per_s": rate, "useful_tflops": rate * useful / 1e12,
                         "executed_tflops": rate * (useful if path == _probe.PATH_COLS else 2.0 * 16 ** 3) / 1e12})
        for r in recs:
            if r["n_cest"] == nc and r["path"] != "team_columns" and r["path"] in outs:
                a, b = outs["team_columns"], outs[r["path"]]
                r["max_rel_diff_vs_team_columns"] = float(np.max(np.abs(a - b)) / np.max(np.abs(a)))
    best = {}
    for r in recs:
        best.setdefault(r["n_cest"], {})[r["path"]] = r["products_per_s"]
    out = {"metric": "augmented propagator products/s (the squarings' dense product), team columns vs FP64 MFMA",
           "value": max(r["products_per_s"] for r in recs if r["n_cest"] == 2), "unit": "products/s",
           "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded matrices)",
           "config": {"workload": f"E <- P E, {reps} repetitions on {n} augmented matrices per launch, 1-4 CEST pools"},
           "records": recs,
           "mfma_over_team_columns": {nc: v.get("mfma_f64_16x16x4", float("nan")) / v["team_columns"]
                                      for nc, v in best.items()}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--voxels", type=int, default=65536)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--flop-sample", type=int, default=16)
    ap.add_argument("--truths", type=int, default=1024)
    ap.add_argument("--problems", type=int, default=4096)
    ap.add_argument("--image-voxels", type=int, default=1024)
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 1 GiB device buffer (the reference point for scaling HBM counters)")
    ap.add_argument("--workload", choices=("pulsed", "nn", "design", "uneven", "scanner", "joint", "tutorial08",
                                           "tutorial09", "tutorial10", "tutorial11", "squaring", "mlp"), default="pulsed",
                    help="pulsed: forward + Jacobian Z-spectra/s (default); nn: fitted voxels/s, NN-started vs plain LM")
    ap.add_argument("--nn-train", type=int, default=2_000_000)
    ap.add_argument("--nn-test", type=int, default=262_144)
    ap.add_argument("--nn-epochs", type=int, default=30)
    ap.add_argument("--no-nn", action="store_true", help="leave the learned-proposal sub-record out of the line")
    ap.add_argument("--plan-level", type=int, default=3, choices=(2, 3),
                    help="pulsed workload's plan: 3 factorises the symmetric pulses (default), 2 chains every event")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (voxels are independent: "
              "N GPUs would each run 1/N of the batch)", file=sys.stderr)
        sys.exit(2)

    import torch
    from bmc_jax import _engine, _flops, _native
    from bmc_jax.simulation.events import SimEvent

    if args.calibrate_copy:
        _native.torch_cuda()
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": src.numel() * 4}), flush=True)
        return

    if args.workload == "uneven":
        _native.torch_cuda()
        bench_uneven(args)
        return
    if args.workload == "squaring":
        _native.torch_cuda()
        bench_squaring(args)
        return
    if args.workload == "mlp":
        _native.torch_cuda()
        bench_mlp(args)
        return
    if args.workload == "scanner":
        _native.torch_cuda()
        bench_scanner(args)
        return
    if args.workload in ("tutorial10", "tutorial11"):
        _native.torch_cuda()
        bench_phantom_joint(args, args.workload == "tutorial11")
        return
    if args.workload in ("tutorial08", "tutorial09"):
        _native.torch_cuda()
        bench_phantom(args, "quesp" if args.workload == "tutorial08" else "quest")
        return
    if args.workload == "joint":
        _native.torch_cuda()
        bench_joint(args)
        return
    if args.workload == "design":
        _native.torch_cuda()
        bench_design(args)
        return
    if args.workload == "nn":
        _native.torch_cuda()
        rec = nn_record(args, log=lambda m: print(m, file=sys.stderr, flush=True))
        out = {"metric": rec.pop("metric"), "value": rec.pop("value"), "unit": rec.pop("unit"), "n_gpus": 1,
               "steps": 1, "warmup": 0, "ms_per_step": 1e3 * rec["nn_lm"]["seconds"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64 simulate / bf16-fp32 MLP",
               "data": "synthetic (simulator-generated, seeded)"}
        out.update(rec)
        print(json.dumps(out), flush=True)
        return

    _native.torch_cuda()
    prm, m0, ev, nc = workload(args.voxels)
    slots = [_native.slot_index(f, i, nc) for f, i in FREE]
    scales = [100.0, 0.001, 1.0, 10.0, 0.1]
    plan = _engine.plan_for(SimEvent(*ev), compress=args.plan_level)
    prm_d = _native.device_array(prm, torch.float32)
    m0_d = _native.device_array(m0, torch.float32)
    jc = 2 * (1 + nc)
    stream = torch.cuda.current_stream()

    def step():
        return _native.simulate(plan, prm_d, nc, _native.BMC_F32, m0_d, slots=slots, scales=scales, jc=jc)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        m, J = step()
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    ms = 1e3 * wall / args.steps
    V = args.voxels
    spectra_per_s = V / (wall / args.steps)
    assert torch.isfinite(J).all().item() and torch.isfinite(m).all().item()

    # roofline: executed flops of one launch from a sample of voxels
    ops, rp = plan.program()
    rng = np.random.default_rng(0)
    sample = rng.choice(V, size=min(args.flop_sample, V), replace=False)
    ev_np = {k: np.asarray(getattr(ev, k)) for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    pt = int(_native.lib().bmc_tangents_per_pass(len(slots)))
    per_vox = _flops.executed_flops(prm[sample].astype(np.float64), nc, ev_np, ops, rp, slots, pt=pt)
    flops_launch = float(per_vox.mean()) * V
    # the same split by launch: passes of pt tangents, then the remainder alone (when P % pt != 0)
    n_head = (len(slots) // pt) * pt if len(slots) % pt else len(slots)
    by_launch = [float(_flops.executed_flops(prm[sample].astype(np.float64), nc, ev_np, ops, rp, slots[:n_head],
                                             pt=pt).mean()) * V]
    if n_head < len(slots):
        by_launch.append(float(_flops.executed_flops(prm[sample].astype(np.float64), nc, ev_np, ops, rp,
                                                     slots[n_head:], pt=len(slots) - n_head).mean()) * V)
    avg_launch_s = 1e-3 * float(np.mean(launch_ms))
    achieved = flops_launch / avg_launch_s / 1e12
    # reference-equivalent flops: every event its own Pade exponential with 5 tangents, as the reference
    # computes it (the "effective" rate beside the executed one)
    ref_vox = _flops.reference_equivalent_flops(prm[sample[:4]].astype(np.float64), nc, ev_np, len(slots))
    ref_step = float(ref_vox.mean()) * V
    stats = plan.stats()

    out = {
        "metric": "Z-spectra/s (forward + 5-column Jacobian), pulsed train",
        "value": spectra_per_s,
        "unit": "Z-spectra/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded parameter draws; no measured data)",
        "config": {"workload": "pulsed train fwd+Jac: 4 pools (water, amide, amine, SL MT), 50 Gaussian pulses x 200 "
                               "steps + gaps, spoiler, 3 s recovery (10,053 events), 61 offsets, "
                               f"{V} voxels, 5 free values, float32 I/O",
                   "voxels": V, "offsets": 61, "events_per_offset": int(ev_np['dt'].shape[1]), "free_values": 5,
                   "io_dtype": "float32", "plan_built_events_per_offset": stats["n_built"] / 61},
        "roofline": {"bound": "valu", "bound_detail": "compute-bound in float64 vector FMA (MI355X's FP64 vector "
                     "and FP64 matrix peaks are both 78.6 TFLOP/s)", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": TRAFFIC_BYTES_PER_VOXEL * V,
                     "traffic_unit": "bytes per step (measured, scaled by voxels)",
                     "launches_per_step": "2 (tangent passes 2 + 2, then the fifth tangent alone); achieved = "
                                          "executed flops of both / their time, HIP events on the launch stream",
                     "flops_per_step": flops_launch, "flops_by_launch": by_launch, "avg_step_ms": 1e3 * avg_launch_s,
                     "reference_equivalent_flops_per_spectrum": ref_step / V,
                     "effective_tflops": ref_step / avg_launch_s / 1e12},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(prm, m0, ev, nc, slots)
    if not args.no_nn:
        del m, J
        torch.cuda.empty_cache()
        out["nn"] = nn_record(args, log=lambda msg: print(msg, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
