"""Float64 NumPy restatement of the proposal MLP, with the GPU kernels' BF16 rounding points.

TEST INFRASTRUCTURE ONLY (checks bmc_mlp_* gradients and outputs). The kernels round the GEMM operands
to BF16 (inputs, weights, post-activations, output gradients) and accumulate in FP32; this oracle
rounds at the same points and computes everything else in float64.
"""
import numpy as np
from scipy.special import erf


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2 * np.pi)


def unflatten(w, dims):
    Ws, bs, o = [], [], 0
    for l in range(len(dims) - 1):
        n = dims[l] * dims[l + 1]
        Ws.append(np.asarray(w[o:o + n], np.float64).reshape(dims[l + 1], dims[l]))
        o += n
        bs.append(np.asarray(w[o:o + dims[l + 1]], np.float64))
        o += dims[l + 1]
    return Ws, bs


def forward(w, dims, X, shift=None, scale=None):
    """Output and the saved activations; the input is standardised first, (X - shift) * scale in float32
    as the kernel does, when shift / scale are given."""
    Ws, bs = unflatten(w, dims)
    Xs = np.asarray(X, np.float32)
    if shift is not None:
        Xs = (Xs - np.asarray(shift, np.float32)) * np.asarray(scale, np.float32)
    a = bf16(Xs)
    pres, acts = [], [a]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = a @ bf16(W).T + b
        pres.append(z)
        if l + 1 < len(Ws):
            a = bf16(gelu(z))
            acts.append(a)
        else:
            a = z
    return a, pres, acts


def gradients(w, dims, X, T, shift=None, scale=None):
    """Flat gradient of mean((Y - T)^2) and the loss."""
    Ws, bs = unflatten(w, dims)
    Y, pres, acts = forward(w, dims, X, shift, scale)
    n = Y.size
    loss = float(np.sum((Y - T) ** 2) / n)
    dA = 2.0 * (Y - T) / n
    grads = []
    for l in range(len(Ws) - 1, -1, -1):
        dZ = dA if l == len(Ws) - 1 else dA * gelu_grad(pres[l])
        dZb = bf16(dZ)
        gW = dZb.T @ acts[l]
        gb = dZ.sum(axis=0)
        grads.append((gW, gb))
        if l > 0:
            dA = dZb @ bf16(Ws[l])
    flat = []
    for gW, gb in reversed(grads):
        flat += [gW.reshape(-1), gb]
    return np.concatenate(flat), loss
