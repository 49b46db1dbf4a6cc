"""The learned proposal's BF16 MFMA kernels (bmc_mlp_*) against a float64 NumPy MLP with the same BF16
rounding points, and what the trained network gains over the plain start."""
import numpy as np
import pytest

import mlp_oracle as MO

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    from bmc_jax import _native
    return _native.torch_cuda()


def test_forward_and_gradients_match_the_oracle(torch_mod):
    """Odd sizes exercise every tile edge of the GEMM (M, N, K not multiples of 64 or 32)."""
    torch = torch_mod
    from bmc_jax.learned import MLP
    dims = (37, 70, 45, 3)
    mlp = MLP(dims=dims, max_batch=130, seed=3)
    rng = np.random.default_rng(0)
    X = rng.normal(size=(130, 37)).astype(np.float32)
    T = rng.normal(size=(130, 3)).astype(np.float32)
    w = mlp.get_params().astype(np.float64)
    Y = mlp.forward(torch.as_tensor(X, device="cuda")).cpu().numpy()
    Yo, _, _ = MO.forward(w, dims, X)
    assert np.max(np.abs(Y - Yo)) < 2e-3 * max(1.0, np.abs(Yo).max())
    g, loss = mlp.gradients(torch.as_tensor(X, device="cuda"), torch.as_tensor(T, device="cuda"))
    go, lo = MO.gradients(w, dims, X, T)
    assert loss == pytest.approx(lo, rel=1e-3)
    cos = float(g @ go / (np.linalg.norm(g) * np.linalg.norm(go)))
    assert cos > 0.9999, cos
    assert np.linalg.norm(g - go) < 5e-3 * np.linalg.norm(go)


@pytest.mark.parametrize("batch", [4096, 1000])
def test_the_proposal_network_at_its_size(torch_mod, batch):
    """The proposal network itself, 244-512-512-256-5, at the training batch of 4096 (and a ragged 1000):
    outputs and every layer's weight and bias gradients against the oracle."""
    torch = torch_mod
    from bmc_jax.learned import MLP
    dims = (244, 512, 512, 256, 5)
    mlp = MLP(dims=dims, max_batch=4096, seed=11)
    rng = np.random.default_rng(2)
    # Z-spectrum-like inputs: a mean near 1 per feature with a few percent of spread, standardised
    X = (0.9 - 0.3 * np.sin(np.linspace(0, 3, 244)) + 0.03 * rng.normal(size=(batch, 244))).astype(np.float32)
    T = rng.normal(size=(batch, 5)).astype(np.float32)
    Xd = torch.as_tensor(X, device="cuda")
    mlp.fit_input_norm(Xd)
    shift, scale = mlp.input_norm()
    np.testing.assert_allclose(shift, X.astype(np.float64).mean(axis=0), rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(1.0 / scale, X.astype(np.float64).std(axis=0), rtol=1e-5)
    w = mlp.get_params().astype(np.float64)
    Y = mlp.forward(Xd).cpu().numpy()
    Yo, _, _ = MO.forward(w, dims, X, shift, scale)
    assert np.max(np.abs(Y - Yo)) < 2e-3 * max(1.0, np.abs(Yo).max())
    g, loss = mlp.gradients(Xd, torch.as_tensor(T, device="cuda"))
    go, lo = MO.gradients(w, dims, X, T, shift, scale)
    assert loss == pytest.approx(lo, rel=1e-3)
    off = 0
    for l in range(len(dims) - 1):
        for n in (dims[l] * dims[l + 1], dims[l + 1]):
            a, b = g[off:off + n], go[off:off + n]
            assert np.linalg.norm(a - b) < 1e-2 * np.linalg.norm(b), (l, n, np.linalg.norm(a - b) / np.linalg.norm(b))
            off += n
    assert off == g.size == go.size


def test_the_network_learns_a_smooth_map(torch_mod):
    torch = torch_mod
    from bmc_jax.learned import MLP
    rng = np.random.default_rng(1)
    X = rng.uniform(-1, 1, (16384, 8)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0] * 2) + X[:, 1] ** 2, X[:, 2] * X[:, 3]], axis=1).astype(np.float32)
    mlp = MLP(dims=(8, 64, 64, 2), max_batch=1024, seed=1)
    hist = mlp.train(torch.as_tensor(X, device="cuda"), torch.as_tensor(T, device="cuda"), epochs=30, batch=1024,
                     lr=3e-3)
    assert hist[-1] < 0.05 * hist[0], hist


def test_the_trained_proposal_beats_the_plain_start(torch_mod):
    """Train on 200k simulated ladders (sigma 1e-3 noise), then fit 512 held-out noise-free voxels for 5 LM
    iterations from the network's x0 and from the plain x0: the network's start is closer and its fits end
    at a lower residual."""
    from bmc_jax import learned
    from bmc_jax.fitting import fit_batch
    from bmc_jax.simulation.params import initial_state
    prop = learned.train_proposal(n_train=200_000, epochs=12, batch=2048, seed=5)
    assert prop.loss[-1] < 0.25 * prop.loss[0], prop.loss
    rng = np.random.default_rng(55)
    x_true, u_true = learned.draw_params(512, rng)
    Z = learned.simulate_ladder(x_true)
    x_nn = learned.propose(prop, Z)
    spec = learned.fit_spec()
    x_plain = np.broadcast_to(spec.x0_phys, x_true.shape)
    err = lambda x: np.median(np.abs(learned.normalise(x) - u_true))
    assert err(x_nn) < 0.5 * err(x_plain)
    base = learned.base_params(512)
    m0 = initial_state(base)
    ev = learned.ladder_events()
    rms = {}
    for name, x0 in (("nn", x_nn), ("plain", x_plain)):
        X = fit_batch(base, ev, m0, 4, spec, Z, X0_phys=x0, n_iter=5, skip_first=False, base_batched=True)
        p = learned.with_values(base, X)
        zf = learned.simulate_ladder(X)
        rms[name] = np.median(np.sqrt(np.mean((zf - Z) ** 2, axis=1)))
    assert rms["nn"] < rms["plain"], rms


def test_an_inference_copy_gives_the_same_outputs(torch_mod):
    """copy_for_inference (weights and input standardisation in a handle sized for larger batches, the
    trained proposal's inference handle) against the training handle, which runs the same spectra in
    chunks of its own batch size, and both against the oracle."""
    torch = torch_mod
    from bmc_jax.learned import DIMS, MLP
    mlp = MLP(dims=DIMS, max_batch=512, seed=7)
    rng = np.random.default_rng(1)
    # Z-spectrum-like inputs (as in the network's own test): a mean near 1 per feature, a few percent spread
    X = torch.as_tensor((0.9 - 0.3 * np.sin(np.linspace(0, 3, DIMS[0])) + 0.03 * rng.normal(size=(3000, DIMS[0])))
                        .astype(np.float32), device="cuda")
    mlp.fit_input_norm(X)
    big = mlp.copy_for_inference(4096)
    assert big.max_batch == 4096 and big.n_params == mlp.n_params
    np.testing.assert_array_equal(big.get_params(), mlp.get_params())
    for a, b in zip(big.input_norm(), mlp.input_norm()):
        np.testing.assert_array_equal(a, b)
    # the hidden activations are rounded to BF16 between layers, so a different GEMM split (the batch
    # sizes differ) may round a sum to the neighbouring BF16 value: most outputs agree bit for bit, and
    # the copy is no further from the float64 oracle (same rounding points) than the original
    y_small = mlp.forward(X).cpu().numpy()
    y_big = big.forward(X).cpu().numpy()
    assert np.mean(y_big == y_small) > 0.5
    shift, scale = mlp.input_norm()
    Yo, _, _ = MO.forward(mlp.get_params().astype(np.float64), DIMS, X.cpu().numpy(), shift, scale)
    e_small, e_big = np.abs(y_small - Yo), np.abs(y_big - Yo)
    scale_y = max(1.0, np.abs(Yo).max())
    assert e_big.max() <= max(1.5 * e_small.max(), 2e-3 * scale_y)
    assert np.median(e_big) < 1e-4 * scale_y


@pytest.mark.parametrize("batch", [8192, 8200])
def test_large_batch_forward_matches_the_oracle(torch_mod, batch):
    """Batches whose 512-wide layers have at least 256 tiles of 128 x 128 run the large-tile GEMM (32 x 32
    fragments; at 8,192 rows, 64 row blocks, in the XCD-aware tile order; at 8,200 rows, 65 row blocks
    with a ragged last one, in grid order) and the full-tile epilogue: the outputs against the oracle."""
    torch = torch_mod
    from bmc_jax.learned import DIMS, MLP
    mlp = MLP(dims=DIMS, max_batch=8200, seed=13)
    rng = np.random.default_rng(4)
    X = (0.9 - 0.3 * np.sin(np.linspace(0, 3, DIMS[0])) + 0.03 * rng.normal(size=(batch, DIMS[0]))).astype(np.float32)
    Xd = torch.as_tensor(X, device="cuda")
    mlp.fit_input_norm(Xd)
    shift, scale = mlp.input_norm()
    Y = mlp.forward(Xd).cpu().numpy()
    Yo, _, _ = MO.forward(mlp.get_params().astype(np.float64), DIMS, X, shift, scale)
    scale_y = max(1.0, np.abs(Yo).max())
    assert np.max(np.abs(Y - Yo)) < 2e-3 * scale_y
    assert np.median(np.abs(Y - Yo)) < 1e-4 * scale_y
