"""fit_joint on the GPU: the behaviours the reference pins in its tests/test_joint.py, each checked against
a truth the test builds itself (data simulated by the kernel, then fitted back).

  * a plain spec reproduces fit's numbers;
  * shared, per-block and partly-free values, factors, offsets, the start state and a B0 spread are
    recovered;
  * keep_jacobian / stop_lam change no step; the coupled-only options are refused without coupling;
  * the assembled joint Jacobian equals central differences of the joint residual (float64).
"""
import numpy as np
import pytest

import bmc_oracle as O

pytestmark = pytest.mark.gpu

PPM = np.linspace(-6.0, 6.0, 41)
LADDER = (0.5, 1.0, 1.5, 2.0)
K_RTOL = 5e-2


@pytest.fixture(scope="module")
def api():
    from bmc_jax import _native
    _native.torch_cuda()
    import bmc_jax.fitting as F
    import bmc_jax.simulation as S
    from bmc_jax.io import example_config
    from bmc_jax.sequences import cw_saturation
    return dict(F=F, S=S, example_config=example_config, cw=cw_saturation)


def _ladder(api, p):
    from bmc_jax.simulation.events import concat_events
    return concat_events([api["cw"](PPM, b1_uT=b, tsat_s=2.0, b0_T=p.system.b0) for b in LADDER])


def _batch(jp, n, **fields):
    base = type(jp)(*[np.broadcast_to(np.asarray(a), (n,) + np.shape(a)).copy() for a in jp])
    dt = np.asarray(base.water_r1).dtype
    return base._replace(**{k: np.asarray(v, dt) for k, v in fields.items()})


def _one(b, i):
    return type(b)(*[np.asarray(a)[i] for a in b])


def _data(api, base, events, mz, m_scale=None):
    S = api["S"]
    n = np.shape(base.water_r1)[0]
    m_scale = np.ones(n) if m_scale is None else np.asarray(m_scale)
    m0 = np.stack([S.initial_state(_one(base, i)) * m_scale[i] for i in range(n)])
    z = np.stack([np.asarray(S.simulate_spectrum(_one(base, i), events, m0[i], return_full_trajectory=False))[:, mz]
                  for i in range(n)])
    return m0, z


@pytest.fixture(scope="module")
def specs(api):
    FP = api["F"].FreeParam
    return dict(
        K=FP("k", "cest_k", 0, scale=100.0, lo=5.0, hi=2000.0, x0=150.0),
        F=FP("f", "cest_f", 0, scale=0.001, lo=1e-5, hi=0.01, x0=0.0005),
        R2W=FP("r2w", "water_r2", None, scale=10.0, lo=1.0, hi=100.0, x0=14.0, scope="shared"),
        F_NOE=FP("f_noe", "cest_f", 1, scale=0.001, lo=0.0, hi=0.03, x0=0.002, samples=(0, 1)),
        M_SHARED=FP("m0", "m_start", None, scale=1.0, lo=0.1, hi=1.0, x0=0.9, scope="shared"),
        R2_FACTOR=FP("r2f", "water_r2", None, scale=1.0, lo=0.5, hi=1.5, x0=1.0, scope="shared", mode="factor"),
        DB0=FP("db0", "b0_inhom", None, scale=0.01, lo=-0.1, hi=0.1, x0=0.0, scope="block", mode="add"),
        SIGMA=FP("sig", "b0_spread", None, scale=10.0, lo=0.0, hi=60.0, x0=10.0, scope="shared"),
        B1F=FP("b1f", "rel_b1", None, scale=1.0, lo=0.7, hi=1.3, x0=1.0, scope="block", mode="factor"),
        M_SOME=FP("m0", "m_start", None, scale=1.0, lo=0.1, hi=1.0, x0=0.9, samples=(0, 1)),
    )


@pytest.fixture(scope="module")
def amide(api):
    p = api["example_config"]("amide")
    jp = api["S"].to_sim_params(p)
    events = _ladder(api, p)
    true = _batch(jp, 3, cest_k=[[40.0], [120.0], [300.0]])
    m0, z = _data(api, true, events, p.mz_loc)
    return p, jp, true, events, m0, z


@pytest.fixture(scope="module")
def noe(api):
    p = api["example_config"]("amide_noe")
    jp = api["S"].to_sim_params(p)
    events = _ladder(api, p)
    true = _batch(jp, 3, cest_k=[[50.0, 20.0], [150.0, 20.0], [300.0, 20.0]],
                  cest_f=[[0.0009, 0.0045], [0.0009, 0.009], [0.0009, 0.0]])
    m0, z = _data(api, true, events, p.mz_loc)
    return p, jp, true, events, m0, z


def _exp(api, params, events, m0, mz):
    return api["F"].Experiment(params, events, m0, mz)


def test_a_plain_spec_gives_fits_numbers(api, specs, amide):
    p, jp, true, events, m0, z = amide
    F = api["F"]
    exp = _exp(api, true, events, m0, p.mz_loc)
    spec = F.FitSpec([specs["K"], specs["F"]])
    a, b = F.fit(exp, spec, z, n_iter=20), F.fit_joint(exp, spec, z, n_iter=20)
    np.testing.assert_allclose(b.x, a.x, rtol=1e-3, atol=0)
    np.testing.assert_allclose(b.z_fit, a.z_fit, rtol=0, atol=1e-5)
    assert b.free.all() and np.isfinite(b.sd).all()


def _shared_r2(api, specs, amide, **kw):
    p, jp, true, events, m0, z = amide
    start = true._replace(water_r2=np.full(3, 14.0, np.asarray(true.water_r1).dtype))
    spec = api["F"].FitSpec([specs["K"], specs["F"], specs["R2W"]])
    return api["F"].fit_joint(_exp(api, start, events, m0, p.mz_loc), spec, z, **kw)


def test_a_shared_value_is_recovered(api, specs, amide):
    res = _shared_r2(api, specs, amide, n_iter=30)
    assert np.all(res["r2w"] == res["r2w"][0])
    assert res["r2w"][0] == pytest.approx(20.0, rel=5e-3)
    np.testing.assert_allclose(res["k"], [40.0, 120.0, 300.0], rtol=K_RTOL)
    assert np.all(res.sd[:, 2] == res.sd[0, 2]) and np.isfinite(res.sd).all()
    assert res.cost[-1] < res.cost[0] and res.nan_steps == 0


def test_keeping_the_jacobian_and_stop_lam_change_no_step(api, specs, amide):
    a = _shared_r2(api, specs, amide, n_iter=40)
    b = _shared_r2(api, specs, amide, n_iter=40, keep_jacobian=True)
    np.testing.assert_allclose(b.x, a.x, rtol=1e-6)
    np.testing.assert_allclose(b.cost, a.cost, rtol=1e-6)
    stopped = _shared_r2(api, specs, amide, n_iter=100, stop_lam=1e3)
    assert len(stopped.cost) < 100
    same = _shared_r2(api, specs, amide, n_iter=len(stopped.cost))
    np.testing.assert_allclose(stopped.x, same.x, rtol=1e-6)


def test_the_lm_rules_need_a_coupled_fit(api, specs, amide):
    p, jp, true, events, m0, z = amide
    F = api["F"]
    for kw in (dict(stop_lam=1e7), dict(active_bounds=True), dict(keep_jacobian=True)):
        with pytest.raises(ValueError, match="coupled"):
            F.fit_joint(_exp(api, true, events, m0, p.mz_loc), F.FitSpec([specs["K"], specs["F"]]), z, n_iter=1, **kw)


@pytest.mark.parametrize("shared", [False, True], ids=["independent", "coupled"])
def test_a_pool_free_in_some_samples_only(api, specs, noe, shared):
    p, jp, true, events, m0, z = noe
    F = api["F"]
    start = true._replace(water_r2=np.full(3, 14.0, np.asarray(true.water_r1).dtype)) if shared else true
    spec = F.FitSpec([specs["K"], specs["F"], specs["F_NOE"]] + ([specs["R2W"]] if shared else []))
    res = F.fit_joint(_exp(api, start, events, m0, p.mz_loc), spec, z, n_iter=30)
    np.testing.assert_allclose(res["f_noe"][:2], [0.0045, 0.009], rtol=1e-2)
    assert res["f_noe"][2] == 0.0
    assert not res.free[2, 2] and np.isnan(res.sd[2, 2]) and np.isfinite(res.sd[:2, 2]).all()
    assert not res.at_bounds[2, 2]
    if shared:
        assert res["r2w"][0] == pytest.approx(20.0, rel=5e-3)


def test_misuse_is_refused(api, specs, amide):
    p, jp, true, events, m0, z = amide
    F = api["F"]
    exp = _exp(api, true, events, m0, p.mz_loc)
    for extra in (specs["R2W"], specs["R2_FACTOR"], specs["M_SHARED"], specs["DB0"], specs["SIGMA"]):
        with pytest.raises(ValueError, match="fit_joint"):
            F.fit(exp, F.FitSpec([specs["K"], extra]), z)
    bad = F.FreeParam("f2", "cest_f", 0, 0.001, 0.0, 0.03, 0.002, samples=(0, 7))
    with pytest.raises(ValueError, match="do not exist"):
        F.fit_joint(exp, F.FitSpec([specs["K"], bad]), z, n_iter=1)
    with pytest.raises(ValueError, match="spread"):
        F.fit_joint(exp, F.FitSpec([specs["K"], specs["SIGMA"]]), z, n_iter=1)


def test_a_shared_factor_and_start_state_are_recovered(api, specs, amide):
    p, jp, true, events, m0, z = amide
    F, S = api["F"], api["S"]
    r2_true = np.asarray([16.0, 20.0, 24.0], np.asarray(true.water_r1).dtype)
    true2 = true._replace(water_r2=r2_true)
    _, z2 = _data(api, true2, events, p.mz_loc, m_scale=[0.6, 0.6, 0.6])
    carried = true2._replace(water_r2=(r2_true / 0.8).astype(r2_true.dtype))
    m_eq = np.stack([S.initial_state(_one(carried, i)) for i in range(3)])
    spec = F.FitSpec([specs["K"], specs["F"], specs["R2_FACTOR"], specs["M_SHARED"]])
    res = F.fit_joint(_exp(api, carried, events, m_eq, p.mz_loc), spec, z2, n_iter=40)
    assert res["r2f"][0] == pytest.approx(0.8, rel=5e-3)
    assert res["m0"][0] == pytest.approx(0.6, rel=5e-3)
    np.testing.assert_allclose(res["k"], [40.0, 120.0, 300.0], rtol=K_RTOL)


def test_a_start_state_free_in_some_samples_only(api, specs, amide):
    p, jp, true, events, m0, z = amide
    F = api["F"]
    _, z2 = _data(api, true, events, p.mz_loc, m_scale=[0.7, 0.5, 1.0])
    res = F.fit_joint(_exp(api, true, events, m0, p.mz_loc), F.FitSpec([specs["K"], specs["F"], specs["M_SOME"]]),
                      z2, n_iter=40)
    np.testing.assert_allclose(res["m0"][:2], [0.7, 0.5], rtol=5e-3)
    assert res["m0"][2] == pytest.approx(1.0) and not res.free[2, 2]
    np.testing.assert_allclose(res["k"], [40.0, 120.0, 300.0], rtol=K_RTOL)


def test_one_offset_per_experiment_is_recovered(api, specs, amide):
    p, jp, true, events, m0, z = amide
    F, S = api["F"], api["S"]
    dt = np.asarray(true.water_r1).dtype
    own = np.asarray([0.0, 0.01, -0.01], dt)
    drift = [0.0, 0.03, -0.02, 0.01]
    n_per = PPM.size
    blocks = []
    for b, d in enumerate(drift):
        ev = type(events)(*[np.asarray(a)[b * n_per:(b + 1) * n_per] for a in events])
        blocks.append(np.stack([np.asarray(S.simulate_spectrum(_one(true, i)._replace(b0_inhom=dt.type(own[i] + d)), ev,
                                                               m0[i], return_full_trajectory=False))[:, p.mz_loc]
                                for i in range(3)]))
    z2 = np.concatenate(blocks, axis=1)
    res = F.fit_joint(_exp(api, true._replace(b0_inhom=own), events, m0, p.mz_loc),
                      F.FitSpec([specs["K"], specs["F"], specs["DB0"]]), z2, n_iter=40, n_per_block=n_per)
    np.testing.assert_allclose(res["db0"], drift, atol=2e-3)
    assert np.isnan(res.x[:, 2]).all() and res.blocks["db0"]["sd"].shape == (4,)
    np.testing.assert_allclose(res["k"], [40.0, 120.0, 300.0], rtol=K_RTOL)


@pytest.mark.parametrize("shape", ["gauss", "lorentz"])
def test_a_spread_of_zero_width_is_one_isochromat(api, amide, shape):
    p, jp, true, events, m0, z = amide
    S = api["S"]
    a = np.asarray(S.simulate_spectrum(_one(true, 0), events, m0[0], return_full_trajectory=False))
    b = np.asarray(S.simulate_spectrum_spread(_one(true, 0), events, m0[0], 0.0, S.B0Spread(shape=shape, nodes=5)))
    np.testing.assert_allclose(b, a, rtol=0, atol=2e-6)


def test_a_node_is_the_sample_shifted_by_sigma(api, amide):
    p, jp, true, events, m0, z = amide
    S = api["S"]
    one = _one(true, 0)
    hz = float(one.gamma * one.b0 / (2 * np.pi))
    assert hz == pytest.approx(42.577 * float(one.b0), rel=1e-3)
    s = 12.0
    dt = np.asarray(one.water_r1).dtype
    sh = lambda d: np.asarray(S.simulate_spectrum(one._replace(b0_inhom=dt.type(one.b0_inhom + d)), events, m0[0],
                                                  return_full_trajectory=False))
    two = np.asarray(S.simulate_spectrum_spread(one, events, m0[0], s, S.B0Spread(nodes=2)))
    np.testing.assert_allclose(two, 0.5 * (sh(s / hz) + sh(-s / hz)), rtol=0, atol=2e-6)


def test_a_shared_spread_is_recovered(api, specs, amide):
    p, jp, true, events, m0, z = amide
    F, S = api["F"], api["S"]
    sp = S.B0Spread(nodes=5)
    z2 = np.stack([np.asarray(S.simulate_spectrum_spread(_one(true, i), events, m0[i], 25.0, sp))[:, p.mz_loc]
                   for i in range(3)])
    res = F.fit_joint(_exp(api, true, events, m0, p.mz_loc), F.FitSpec([specs["K"], specs["F"], specs["SIGMA"]]), z2,
                      n_iter=30, spread=sp)
    assert res["sig"][0] == pytest.approx(25.0, rel=2e-2)
    np.testing.assert_allclose(res["k"], [40.0, 120.0, 300.0], rtol=K_RTOL)


JAC_CASES = {
    "plain": (["K", "F"], None, None),
    "shared": (["K", "F", "R2W"], None, None),
    "factor + start state": (["K", "F", "R2_FACTOR", "M_SHARED"], None, None),
    "some samples": (["K", "F", "M_SOME"], None, None),
    "block offset": (["K", "F", "DB0"], None, "full"),
    "block factor + shared, split": (["K", "F", "B1F", "R2W"], None, "split"),
    "spread": (["K", "F", "SIGMA"], ("gauss", 3), None),
    "block + spread": (["K", "F", "B1F", "SIGMA"], ("lorentz", 3), "split"),
}


@pytest.mark.parametrize("case", list(JAC_CASES))
def test_the_joint_jacobian_is_the_derivative(api, specs, amide, case):
    """The scattered per-sample Jacobian against central differences of the joint residual, float64."""
    from bmc_jax import settings
    p, jp, true, events, m0, z = amide
    F, S = api["F"], api["S"]
    names, sp, blocks = JAC_CASES[case]
    settings.use_x64(True)
    try:
        t64 = type(true)(*[np.asarray(a, np.float64) if np.asarray(a).dtype.kind == "f" else np.asarray(a)
                           for a in true])
        spread = None if sp is None else S.B0Spread(shape=sp[0], nodes=sp[1])
        d = F.fit_joint(_exp(api, t64, events, m0.astype(np.float64), p.mz_loc), F.FitSpec([specs[k] for k in names]),
                        z.astype(np.float64), spread=spread, n_per_block=PPM.size if blocks else None,
                        block_jacobian=blocks or "full", _internals=True)
        x = d["x0"] + 0.02 * np.linspace(-1.0, 1.0, d["x0"].size)
        J = np.asarray(d["jac"](x))
        Jn = np.zeros_like(J)
        for c in range(x.size):
            # a 2 s CW event takes ~12 squarings: the forward carries ~1e-12 of rounding noise, so the step
            # is large enough for that noise to stay below the truncation error of the difference
            h = 1e-3 * max(1.0, abs(x[c]))
            e = np.zeros_like(x)
            e[c] = h
            d1 = (np.asarray(d["resid"](x + e)) - np.asarray(d["resid"](x - e))) / (2 * h)
            d2 = (np.asarray(d["resid"](x + e / 2)) - np.asarray(d["resid"](x - e / 2))) / h
            Jn[:, c] = (4 * d2 - d1) / 3          # Richardson: the h^2 term cancels
    finally:
        settings.use_x64(False)
    rel = np.linalg.norm(J - Jn, axis=0) / np.maximum(np.linalg.norm(Jn, axis=0), 1e-30)
    assert rel.max() < 1e-5, f"{case}: columns {np.flatnonzero(rel >= 1e-5)} off by {rel.max():.1e}"


MODEL_CASES = {
    "factor + start state value": (["K", "F", "R2_FACTOR", "M_SHARED"], None, False),
    "start state in some samples": (["K", "F", "M_SOME"], None, False),
    "gaussian spread": (["K", "F", "SIGMA"], ("gauss", 3), False),
    "per-block factor and offset": (["K", "F", "B1F", "DB0"], None, True),
    "lorentzian spread + block factor": (["K", "F", "B1F", "SIGMA"], ("lorentz", 3), True),
}


def _record(sp, i):
    """Sample i of a SimParams batch as an oracle record (float64 values, int lineshape)."""
    out = {}
    for f in sp._fields:
        a = np.asarray(getattr(sp, f))
        v = a[i] if a.ndim else a
        out[f] = int(v) if f == "mt_lineshape" else np.array(v, np.float64)
    return out


def _spread_nodes(shape, nodes, cut=20.0):
    """Probabilists' Gauss-Hermite, or Gauss-Legendre mapped by tan onto a Lorentzian cut at +-cut
    (spread.py:67-74), from scipy's root tables, weights normalised to 1."""
    from scipy.special import roots_hermitenorm, roots_legendre
    if shape == "gauss":
        u, w = roots_hermitenorm(nodes)
    else:
        t, w = roots_legendre(nodes)
        u = np.tan(t * np.arctan(cut))
    return u, w / w.sum()


@pytest.mark.parametrize("case", list(MODEL_CASES))
def test_the_joint_model_is_the_oracle(api, specs, amide, case):
    """The joint model (every slot mode applied per sample and block, the start state rebuilt or scaled,
    the B0 spread as weighted isochromats) against the float64 C oracle run on parameters this test
    builds itself from the joint vector (joint.py:120-138, spread.py:77-106)."""
    import oracle_c as OC
    from bmc_jax import settings
    p, jp, true, events, m0, z = amide
    F, S = api["F"], api["S"]
    names, sp, blocked = MODEL_CASES[case]
    n, n_pts = z.shape
    npb = PPM.size
    settings.use_x64(True)
    try:
        t64 = type(true)(*[np.asarray(a, np.float64) if np.asarray(a).dtype.kind == "f" else np.asarray(a)
                           for a in true])
        spread = None if sp is None else S.B0Spread(sigma_hz=7.0, shape=sp[0], nodes=sp[1])
        spec = F.FitSpec([specs[k] for k in names])
        d = F.fit_joint(_exp(api, t64, events, m0.astype(np.float64), p.mz_loc), spec, z.astype(np.float64),
                        spread=spread, n_per_block=npb if blocked else None, _internals=True)
        x = d["x0"] * (1.0 + 0.03 * np.cos(np.arange(d["x0"].size)))
        Zm = np.asarray(d["resid"](x)).reshape(n, n_pts) + z.astype(np.float64)
    finally:
        settings.use_x64(False)
    ev = {k: np.asarray(getattr(events, k)) for k in ("type", "dt", "rf_amp", "rf_phase", "rf_freq")}
    nb = n_pts // npb if blocked else 1
    Zo = np.zeros((n, n_pts))
    for i in range(n):
        for b in range(nb):
            rec = _record(t64, i)
            m_scale, sigma = None, 7.0
            for p_, (l, members, blk) in enumerate(d["cols"]):
                s = spec.slots[l]
                if i not in members or (blk >= 0 and blk != b):
                    continue
                v = x[p_] * s.scale
                if s.field == "m_start":
                    m_scale = v
                elif s.field == "b0_spread":
                    sigma = v
                else:
                    cur = rec[s.field] if s.index is None else rec[s.field][s.index]
                    new = v if s.mode == "value" else (cur * v if s.mode == "factor" else cur + v)
                    if s.index is None:
                        rec[s.field] = np.array(new)
                    else:
                        rec[s.field][s.index] = new
            start = m0[i].astype(np.float64) if m_scale is None else O.initial_state(rec) * m_scale
            rows = slice(b * npb, (b + 1) * npb) if blocked else slice(None)
            evb = {k: v[rows] for k, v in ev.items()}
            if spread is None:
                nodes, weights = np.zeros(1), np.ones(1)
            else:
                nodes, weights = _spread_nodes(sp[0], sp[1])
            hz = float(rec["gamma"]) * float(rec["b0"]) / (2 * np.pi)
            acc = 0.0
            for u, w in zip(nodes, weights):
                q = dict(rec, b0_inhom=np.array(float(rec["b0_inhom"]) + sigma * u / hz))
                mm, _ = OC.simulate(OC.pack(q)[None], len(q["cest_k"]), evb, start)
                acc = acc + w * mm[0, :, p.mz_loc]
            Zo[i, rows] = acc
    assert np.max(np.abs(Zm - Zo)) < 1e-11, (case, np.max(np.abs(Zm - Zo)))
