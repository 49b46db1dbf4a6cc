"""Shared builders for the tests: parameter records, sequences, packing (numpy only)."""
import numpy as np

import bmc_oracle as O
import oracle_c as OC

B0_CFG1 = 3.0


def cfg1_params():
    """Configuration 1: water T1 1.3 s, T2 75 ms; amide +3.5 ppm, k 50, f 9.01e-4, T1 1.3 s, T2 0.1 s."""
    return O.make_params(1.3, 0.075, pools=[dict(f=9.01e-4, k=50.0, dw=3.5, t1=1.3, t2=0.1)], b0=B0_CFG1)


def cfg1_events():
    return O.cw_saturation(np.linspace(-6, 6, 61), b1_uT=1.0, tsat_s=2.0, b0_T=B0_CFG1)


def four_pool_params(k_am=50.0, f_am=1e-3, k_amine=800.0, f_amine=4e-3, t1w=1.3, t2w=0.07, b0i=0.0,
                     lineshape="superlorentzian", rel_b1=1.0):
    """Water, amide +3.5, amine +2.0, semi-solid MT at -2.5 ppm (configuration 2's pool model)."""
    return O.make_params(t1w, t2w,
                         pools=[dict(f=f_am, k=k_am, dw=3.5, t1=1.3, t2=0.1),
                                dict(f=f_amine, k=k_amine, dw=2.0, t1=1.3, t2=0.05)],
                         mt=dict(f=0.1, k=30.0, dw=-2.5, t1=1.0, t2=1e-5, lineshape=lineshape),
                         b0=3.0, b0_inhom=b0i, rel_b1=rel_b1)


def small_train(ppm, n_pulses=3, n_samples=16, b0_T=3.0):
    return O.pulsed_train(ppm, b1_peak_uT=2.0, n_pulses=n_pulses, pulse_s=0.02, n_samples=n_samples, gap_s=0.02,
                          recovery_s=1.0, b0_T=b0_T, dtype=np.float64)


def as_dtype(ev, dtype):
    return {k: (v.astype(dtype) if v.dtype.kind == "f" else v) for k, v in ev.items()}


def pack_many(records):
    return np.stack([OC.pack(p) for p in records])


def rel_err(a, b, floor=1e-30):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), floor))
