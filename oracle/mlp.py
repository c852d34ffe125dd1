"""Oracle of the learned warm start's network: a NumPy multilayer perceptron with GELU, its backward pass and Adam.

TEST INFRASTRUCTURE ONLY: the checker of the HIP training and inference kernels; the product package never imports it.

The network has no reference counterpart (the reference has no learned component). It is pinned by a central
finite-difference check of its own FP64 gradients (tests/test_oracle_pins.py) and then used to check the HIP
kernels, which run with BF16 operands and FP32 accumulation: `bf16` reproduces their rounding (round to nearest
even on the FP32 bit pattern) at the same points (weights, and each layer's input activations).

Layers: h_0 = x, z_i = h_{i-1} W_i^T + b_i, h_i = gelu(z_i) for hidden layers, y = z_last (linear output).
gelu(z) = z Phi(z) = z (1 + erf(z / sqrt 2)) / 2 (the exact form, not the tanh approximation).
Loss: mean over samples and outputs of w (y - t)^2 with optional per-output weights w.
Adam (Kingma and Ba 2015) with bias correction.
"""

import math

import numpy as np
from scipy.special import erf


def bf16(x):
    """Round FP32 values to BF16 (round to nearest, ties to even) and return them as FP32."""
    a = np.ascontiguousarray(np.asarray(x, np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isfinite(a), out, a)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(params, x, round_bf16=False):
    """params: list of (W [out, in], b [out]). Returns (y, cache) with cache = (inputs h_{i-1}, pre-activations z_i)."""
    q = bf16 if round_bf16 else (lambda a: a)
    h = np.asarray(x, np.float64)
    hs, zs = [], []
    for i, (W, b) in enumerate(params):
        hin = q(h).astype(np.float64)
        z = hin @ q(W).astype(np.float64).T + b
        hs.append(hin)
        zs.append(z)
        h = gelu(z) if i < len(params) - 1 else z
    return h, (hs, zs)


def loss(y, t, weight=None):
    d = y - t
    w = 1.0 if weight is None else weight
    return float(np.mean(w * d * d))


def backward(params, cache, y, t, weight=None, round_bf16=False):
    """Gradients (dW_i, db_i) of `loss` for the forward pass that produced `cache` and `y`."""
    q = bf16 if round_bf16 else (lambda a: a)
    hs, zs = cache
    n = y.size
    w = 1.0 if weight is None else weight
    dz = 2.0 * w * (y - t) / n
    grads = [None] * len(params)
    for i in range(len(params) - 1, -1, -1):
        W, _ = params[i]
        dq = q(dz).astype(np.float64)  # the products take BF16 operands; the bias sums the FP32 values
        grads[i] = (dq.T @ hs[i], dz.sum(0))
        if i:
            dh = dq @ q(W).astype(np.float64)
            dz = dh * gelu_grad(zs[i - 1])
    return grads


def adam(param, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step in place on float64 arrays; step counts from 1."""
    m *= b1
    m += (1 - b1) * grad
    v *= b2
    v += (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    param -= lr * mh / (np.sqrt(vh) + eps)
