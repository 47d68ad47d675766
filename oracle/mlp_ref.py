"""TEST INFRASTRUCTURE: CPU oracle of the surrogate's multilayer perceptron (numpy).

Used only by tests/ to check the GPU ensemble (mithra_mlp_*).  A network of widths
dims[0] -> ... -> dims[-1], SiLU after every layer but the last, mean-squared-error loss over all
outputs; parameters laid out per member as W_0 (dims[0] x dims[1], row-major), b_0, W_1, b_1, ...

`bf16=True` reproduces where the device rounds to bfloat16 (round to nearest even): the inputs,
the weights, every hidden activation and every output-side gradient before it enters a product;
sums are then exact in float64, so the device differs only by its FP32 accumulation order.
`bf16=False` is the plain float64 network, whose gradient the finite-difference test pins.
"""
import numpy as np


def bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    r = ((u >> 16) & 1) + 0x7FFF
    return (((u + r) >> 16) << 16).astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, dims):
    out, o = [], 0
    for a, b in zip(dims[:-1], dims[1:]):
        W = np.asarray(params[o:o + a * b], dtype=np.float64).reshape(a, b)
        o += a * b
        bias = np.asarray(params[o:o + b], dtype=np.float64)
        o += b
        out.append((W, bias))
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, dims, x, bf16_products=True):
    r = bf16 if bf16_products else (lambda a: np.asarray(a, dtype=np.float64))
    layers = unpack(params, dims)
    X = [r(x)]
    Z = []
    for l, (W, b) in enumerate(layers):
        z = X[-1] @ r(W) + b
        Z.append(z)
        if l + 1 < len(layers):
            X.append(r(silu(z)))
    return Z[-1], (X, Z, layers)


def gradients(params, dims, x, t, bf16_products=True):
    """(gradient in the parameter layout, loss)."""
    r = bf16 if bf16_products else (lambda a: np.asarray(a, dtype=np.float64))
    y, (X, Z, layers) = forward(params, dims, x, bf16_products)
    count = y.size
    e = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(e * e) / count)
    dZ = 2.0 * e / count
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        dZr = r(dZ)
        grads[l] = (X[l].T @ dZr, dZ.sum(axis=0))
        if l == 0:
            break
        dA = dZr @ r(W).T
        dZ = dA * silu_grad(Z[l - 1])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return flat, loss
