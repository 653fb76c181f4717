"""FP64 CPU reference of the learned-SGS MLP -- TEST INFRASTRUCTURE ONLY.

Forward and backward pass of the 12 -> 64 -> 64 -> 64 -> 2 GELU network in
float64 numpy, in two flavours: `exact` (no rounding) for the finite-difference
check of the backprop formulas, and `bf16` where weights, activations and
output gradients are rounded to bfloat16 exactly where the device kernels
round them, so device gradients can be compared operation-for-operation.
Parameters use the compact layout of les_mlp_layout (normalisers excluded).
"""
import numpy as np

SIZES = [(64, 12), (64,), (64, 64), (64,), (64, 64), (64,), (2, 64), (2,)]
NPAR = sum(int(np.prod(s)) for s in SIZES)  # 9282
C = 0.7978845608028654


def unpack(v):
    out, o = [], 0
    for s in SIZES:
        n = int(np.prod(s))
        out.append(np.asarray(v[o:o + n], dtype=np.float64).reshape(s))
        o += n
    return out


def pack(parts):
    return np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in parts])


def bf16(x):
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + np.tanh(C * (x + 0.044715 * x ** 3)))


def gelu_d(x):
    t = np.tanh(C * (x + 0.044715 * x ** 3))
    return 0.5 * (1 + t) + 0.5 * x * (1 - t * t) * C * (1 + 3 * 0.044715 * x * x)


def loss_and_grad(v, X, T, mode="exact"):
    """X: (n, 12) inputs, T: (n, 2) targets; loss = 0.5 mean_n |y - t|^2."""
    r = bf16 if mode == "bf16" else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(v)
    W1r, W2r, W3r, W4r = r(W1), r(W2), r(W3), r(W4)
    n = X.shape[0]
    x = r(X)
    Z1 = x @ W1r.T + b1
    H1 = r(gelu(Z1))
    Z2 = H1 @ W2r.T + b2
    H2 = r(gelu(Z2))
    Z3 = H2 @ W3r.T + b3
    H3 = r(gelu(Z3))
    Y = H3 @ W4r.T + b4
    E = Y - T
    loss = 0.5 * np.sum(E * E) / n
    dY = E / n
    dYr = r(dY)
    gW4 = dYr.T @ H3
    gb4 = dY.sum(0)
    dH = dYr @ W4r
    grads = [None] * 8
    grads[6], grads[7] = gW4, gb4
    for Z, Hin, W, wi, bi in ((Z3, H2, W3r, 4, 5), (Z2, H1, W2r, 2, 3), (Z1, x, W1r, 0, 1)):
        dZ = dH * gelu_d(Z)
        dZr = r(dZ)
        grads[wi] = dZr.T @ Hin
        grads[bi] = dZ.sum(0)
        dH = dZr @ W
    return loss, pack(grads), Y


def forward(v, X, mode="bf16", chunk=1 << 16):
    """The network's outputs (n, 2) for inputs X (n, 12), in chunks of
    points (the forward half of loss_and_grad)."""
    r = bf16 if mode == "bf16" else (lambda a: np.asarray(a, dtype=np.float64))
    W1, b1, W2, b2, W3, b3, W4, b4 = unpack(v)
    W1r, W2r, W3r, W4r = r(W1), r(W2), r(W3), r(W4)
    out = np.empty((X.shape[0], 2))
    for p0 in range(0, X.shape[0], chunk):
        x = r(X[p0:p0 + chunk])
        H1 = r(gelu(x @ W1r.T + b1))
        H2 = r(gelu(H1 @ W2r.T + b2))
        H3 = r(gelu(H2 @ W3r.T + b3))
        out[p0:p0 + chunk] = H3 @ W4r.T + b4
    return out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for s in SIZES:
        if len(s) == 2:
            parts.append(rng.standard_normal(s) * np.sqrt(2.0 / s[1]))
        else:
            parts.append(rng.standard_normal(s) * 0.1)
    return pack(parts)
