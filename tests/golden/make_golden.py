"""Writes the float64 golden fixtures of this directory from the C oracle (oracle/bmc_oracle.c).

    python tests/golden/make_golden.py

cfg1_f64.npz   configuration 1 (2-pool CW, 61 offsets): final water Mz and d/d(k, f), every offset.
cfg2_slice.npz configuration 2's full pulsed train (50 Gaussian pulses x 200 steps, 10,053 events) on a
               thin slice: 3 voxels x 6 offsets, with the 5 free values (k_amide, f_amide, water_r1,
               water_r2, b0_inhom). The packed parameters and offsets are stored alongside.
scanner_rows.npz  two rows (-4.0 and +2.0 ppm) of a converted scanner sequence shipped with the
               reference (examples/data/events/quesp_sim_b1_1uT_events.npz: 20 shaped pulses x 200 samples,
               each pulse at its own accumulated RF phase, 4041 events), as float32 values the way the
               reference's loader casts them, and two tube-like voxels (one CEST pool, start magnetisation
               0.557): final states and d/d(k, f, pool R2). Written only where the reference is mounted;
               the fixture itself is data (event values and oracle outputs).
The oracle itself is pinned by tests/test_oracle.py (closed forms, scipy, independent routes).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(HERE, "..", "..", "oracle"), os.path.join(HERE, "..")]

import bmc_oracle as O  # noqa: E402
import oracle_c as OC  # noqa: E402
from helpers import cfg1_events, cfg1_params, four_pool_params  # noqa: E402

CFG2_FREE = [("cest_k", 0), ("cest_f", 0), ("water_r1", None), ("water_r2", None), ("b0_inhom", None)]


def cfg2_voxels(n, seed=2):
    """Configuration 2's per-voxel draws (k log-U(5, 5000), f U(1e-5, 0.02) per CEST pool, water T1 U(0.8, 3),
    T2 U(0.04, 0.15), b0_inhom U(-0.3, 0.3) ppm)."""
    rng = np.random.default_rng(seed)
    recs = []
    for _ in range(n):
        k = np.exp(rng.uniform(np.log(5.0), np.log(5000.0), 2))
        f = rng.uniform(1e-5, 0.02, 2)
        recs.append(four_pool_params(k_am=k[0], f_am=f[0], k_amine=k[1], f_amine=f[1], t1w=rng.uniform(0.8, 3.0),
                                     t2w=rng.uniform(0.04, 0.15), b0i=rng.uniform(-0.3, 0.3)))
    return recs


def main():
    p = cfg1_params()
    ev = cfg1_events()
    idx = [OC.slot_index("cest_k", 0, 1), OC.slot_index("cest_f", 0, 1)]
    m, dm = OC.simulate(OC.pack(p)[None], 1, ev, O.initial_state(p), slots=idx)
    np.savez(os.path.join(HERE, "cfg1_f64.npz"), z=m[0, :, 4], jac=dm[0, :, :, 4], m=m[0])

    recs = cfg2_voxels(3)
    ppm = np.array([-6.0, -2.5, 0.4, 2.0, 3.5, 6.0])
    ev2 = O.pulsed_train(ppm, b1_peak_uT=2.0, n_pulses=50, pulse_s=0.02, n_samples=200, gap_s=0.02, recovery_s=3.0,
                         b0_T=3.0, dtype=np.float32)
    prm = np.stack([OC.pack(r) for r in recs])
    m0 = np.stack([O.initial_state(r) for r in recs])
    idx2 = [OC.slot_index(f, i, 2) for f, i in CFG2_FREE]
    m2, dm2 = OC.simulate(prm, 2, ev2, m0, slots=idx2)
    np.savez(os.path.join(HERE, "cfg2_slice.npz"), params=prm, ppm=ppm, m0=m0, z=m2[..., 6], jac=dm2[..., 6],
             slots=np.array(idx2))
    ref = "/root/reference/examples/data/events/quesp_sim_b1_1uT_events.npz"
    if os.path.exists(ref):
        with np.load(ref) as d:
            offsets = np.asarray(d["offsets_ppm"])
            rows = [int(np.argmin(np.abs(offsets - v))) for v in (-4.0, 2.0)]
            evs = {k: np.asarray(d[k])[rows].astype(np.float32).astype(np.float64)
                   for k in ("dt", "rf_amp", "rf_phase", "rf_freq")}
            evs["type"] = np.asarray(d["type"])[rows].astype(np.int32)
        tubes = [O.make_params(1 / 0.35, 1 / 0.5605, pools=[dict(f=5e-3, k=900.0, dw=1.0, t1=1.0, t2=0.01)],
                               b0=2.8936, rel_b1=0.97),
                 O.make_params(1 / 0.4, 1 / 0.5605, pools=[dict(f=3.2e-3, k=350.0, dw=1.9, t1=1.0, t2=0.02)],
                               b0=2.8936, b0_inhom=0.02, rel_b1=1.03)]
        prm = np.stack([OC.pack(t) for t in tubes])
        m0 = np.stack([O.initial_state(t, m_start=0.557) for t in tubes])
        idx = [OC.slot_index(f, 0, 1) for f in ("cest_k", "cest_f", "cest_r2")]
        m3, dm3 = OC.simulate(prm, 1, evs, m0, slots=idx)
        np.savez_compressed(os.path.join(HERE, "scanner_rows.npz"), params=prm, m0=m0, slots=np.array(idx),
                            offsets_ppm=offsets[rows], m=m3, jac=dm3[..., 4], **evs)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
