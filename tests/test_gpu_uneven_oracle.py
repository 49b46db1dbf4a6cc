"""The uneven batch (bmc_ubatch_*) against the float64 C oracle at the shapes of the 4096-problem
benchmark: problems drawn from bench.py's uneven_workload generator (1-8 CW rungs from {0.5 .. 8} uT,
31-121 offsets per rung, water plus 1-4 CEST pools with or without MT, 2-8 free values).

Each problem's Z and Jacobian at its start values, from both Jacobian sweeps (reverse accumulation,
the default for CW rows, and forward mode), must match the oracle's simulation of that problem alone:
Z within 1e-11, every Jacobian column within 1e-7 of its largest entry. A few LM fits must take the
oracle LM's steps (lm.py:43-61 restated, float64)."""
import numpy as np
import pytest

import bmc_oracle as O
import oracle_c as OC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def drawn():
    from bmc_jax import _native
    _native.torch_cuda()
    import bench
    probs, truths = bench.uneven_workload(n_problems=64, seed=3)
    return probs, truths


def _packed(p, x_phys):
    """The problem's packed float64 record with its free values written (write_params)."""
    from bmc_jax.fitting.residual import spec_slots
    from bmc_jax.simulation.params import n_cest_of, pack
    nc = n_cest_of(p.params)
    rec = pack(p.params, 1, dtype=np.float64)[0].copy()
    rec[spec_slots(p.spec, nc)] = x_phys
    return rec, nc, spec_slots(p.spec, nc)


def _oracle(p, x_phys, rel_h=1e-3):
    """The oracle's Z and Jacobian of one problem. The C oracle differences the generator; with four CEST
    pools and fast exchange the Jacobian amplifies that step's rounding to ~1e-7 at its default step, so
    here the difference is Richardson-extrapolated from steps rel_h and rel_h / 2 (stable to ~1e-9)."""
    rec, nc, slots = _packed(p, x_phys)
    ev = {k: np.asarray(getattr(p.events, k)) for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    m0 = np.asarray(p.m0, np.float64)
    m, d1 = OC.simulate(rec[None], nc, ev, m0, slots=slots, rel_h=rel_h)
    _, d2 = OC.simulate(rec[None], nc, ev, m0, slots=slots, rel_h=rel_h / 2, want_m=False)
    jc = 2 * (1 + nc)
    return m[0, :, jc], (4 * d2[0, :, :, jc] - d1[0, :, :, jc]) / 3 * p.spec.scales


def test_drawn_shapes_cover_the_benchmark(drawn):
    probs, _ = drawn
    rungs = [p.events.dt.shape[0] for p in probs]
    ncs = [len(p.params.cest_k) for p in probs]
    assert set(ncs) == {1, 2, 3, 4}
    assert max(rungs) >= 8 * 31 and min(rungs) <= 121
    assert {p.spec.n for p in probs} >= {2, 5, 8}
    assert any(float(p.params.mt_f) > 0 for p in probs) and any(float(p.params.mt_f) == 0 for p in probs)


@pytest.mark.parametrize("forward", [False, True])
def test_z_and_jacobian_match_the_oracle(drawn, forward):
    from bmc_jax.fitting.uneven import UnevenBatch
    probs, _ = drawn
    b = UnevenBatch(probs, dtype=np.float64)
    b.configure(2, 512, 1)
    x = b.x0_scaled()
    z, J = b.simulate(x, forward_jacobian=forward)
    z = z.cpu().numpy()
    worst_z, worst_j = 0.0, 0.0
    for i, p in enumerate(probs):
        xp = x[b.x_off[i]:b.x_off[i + 1]] * p.spec.scales
        zo, Jo = _oracle(p, xp)
        zi = z[b.row0[i]:b.row0[i] + b.n_rows[i]]
        Ji = b.jacobian_of(J, i)
        worst_z = max(worst_z, float(np.max(np.abs(zi - zo))))
        col = np.maximum(np.abs(Jo).max(axis=0), 1e-30)
        worst_j = max(worst_j, float(np.max(np.abs(Ji - Jo) / col)))
        assert np.max(np.abs(zi - zo)) < 1e-11, (i, np.max(np.abs(zi - zo)))
        assert np.max(np.abs(Ji - Jo) / col) < 1e-7, (i, np.max(np.abs(Ji - Jo) / col, axis=0))
    print(f"uneven vs oracle ({'forward' if forward else 'reverse'}): |dZ| {worst_z:.1e}, |dJ|/|J| {worst_j:.1e}")


def test_lm_follows_the_oracle_lm(drawn):
    """Three small problems of different pool counts, fitted from their start values on their own
    noise-free data: the batch's LM and the oracle LM (C-oracle residual and Jacobian) agree step for
    step in float64."""
    from bmc_jax.fitting.uneven import UnevenBatch
    probs, truths = drawn
    pick = sorted(range(len(probs)), key=lambda i: probs[i].events.dt.shape[0] * probs[i].spec.n)
    chosen, seen = [], set()
    for i in pick:
        nc = len(probs[i].params.cest_k)
        if nc not in seen:
            chosen.append(i)
            seen.add(nc)
        if len(chosen) == 3:
            break
    sub = [probs[i] for i in chosen]
    for p, i in zip(sub, chosen):
        p.z_meas = _oracle(p, truths[i])[0]
    b = UnevenBatch(sub, dtype=np.float64)
    n_iter = 6
    x, cost, nd = b.lm_fit(n_iter=n_iter, graph=False)
    xs = b.split(x)
    for k, p in enumerate(sub):
        sc = p.spec.scales

        def resid(xq, p=p):
            return _oracle(p, xq * sc)[0] - p.z_meas

        def jac(xq, p=p):
            return _oracle(p, xq * sc)[1]

        xo = O.levenberg_marquardt(resid, jac, p.spec.to_scaled(p.x0_phys), p.spec.lo_phys / sc, p.spec.hi_phys / sc,
                                   n_iter=n_iter)
        np.testing.assert_allclose(xs[k] / sc, xo, rtol=1e-7, atol=1e-12)
