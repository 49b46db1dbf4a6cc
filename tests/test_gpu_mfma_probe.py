"""The product probe (csrc/probe/bmc_mfma_probe.hip): the team column product of the squarings
(mul_cols) and the FP64 MFMA tile product both give P^r E of float64 numpy, which also pins the
v_mfma_f64_16x16x4f64 operand and result layouts the MFMA path assumes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _aug(rng, n, na):
    M = np.zeros((n, na, na))
    M[:, :na - 1, :] = rng.standard_normal((n, na - 1, na)) / np.sqrt(na)
    M[:, na - 1, na - 1] = 1.0
    return M


@pytest.mark.parametrize("nc", [1, 2, 3, 4])
@pytest.mark.parametrize("path", [0, 1, 2])
def test_products_match_float64_numpy(nc, path):
    import torch
    from bmc_jax import _native, _probe
    _native.torch_cuda()
    if path == 1 and nc == 4:
        pytest.skip("17 x 17 does not fit one 16 x 16 MFMA tile (the probe returns an error for it)")
    if path == 2 and nc != 2:
        pytest.skip("the LDS-staged form is for 12-lane teams (two CEST pools) only")
    na = 3 * (1 + nc) + 2
    rng = np.random.default_rng(nc)
    P = _aug(rng, 1, na)[0]
    E = _aug(rng, 37, na)             # a ragged count: the last team / wave is partly idle
    reps = 3
    want = E.copy()
    for _ in range(reps):
        want = P @ want
    Ed = torch.tensor(E, device="cuda")
    _probe.products(path, nc, torch.tensor(P, device="cuda"), Ed, reps)
    got = Ed.cpu().numpy()
    assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(want))
    assert np.all(got[:, na - 1, :na - 1] == 0.0) and np.all(got[:, na - 1, na - 1] == 1.0)


def test_probe_refuses_a_tile_it_cannot_hold():
    import torch
    from bmc_jax import _native, _probe
    _native.torch_cuda()
    na = 17
    with pytest.raises(_native.NativeError):
        _probe.products(1, 4, torch.eye(na, dtype=torch.float64, device="cuda"),
                        torch.zeros((2, na, na), dtype=torch.float64, device="cuda"), 1)
