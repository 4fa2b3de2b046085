"""CPU reference of the surrogate MLP (test infrastructure only: the product
path is the HIP code in cfdwavemaker_amd/csrc/mlp.hip).

The network has no counterpart in the reference; this restates the library's
own definition: L linear layers (W: dout x din row-major, then b), GELU in
erf form after every layer but the last, standardised inputs / outputs, loss
= mean squared error of the standardised outputs.  mode "f64" evaluates in
float64 throughout (pinned by central finite differences in the tests);
mode "bf16" rounds both operands of every product to bfloat16 (round to
nearest even, from their float32 values) as the matrix cores do, and sums in
float64.
"""
import math

import numpy as np

_erf = np.vectorize(math.erf)


def bf16_round(a):
    """float32 value -> nearest bfloat16 (ties to even), returned as float64."""
    a32 = np.ascontiguousarray(np.asarray(a, np.float32))
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_d(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def layer_dims(nin, width, nhidden, nout):
    dims, d = [], nin
    for l in range(nhidden + 1):
        o = width if l < nhidden else nout
        dims.append((d, o))
        d = o
    return dims


def unpack(p, dims):
    out, q = [], 0
    for din, dout in dims:
        W = np.asarray(p[q:q + din * dout], np.float64).reshape(dout, din)
        q += din * dout
        b = np.asarray(p[q:q + dout], np.float64)
        q += dout
        out.append((W, b))
    return out


def _mm(A, B, mode):
    if mode == "bf16":
        return bf16_round(A) @ bf16_round(B)
    return A @ B


def forward(p, dims, Xs, mode="f64"):
    """Xs: standardised inputs.  Returns output and the per-layer (H, Z)."""
    layers = unpack(p, dims)
    H, cache = np.asarray(Xs, np.float64), []
    for l, (W, b) in enumerate(layers):
        Z = _mm(H, W.T, mode) + b
        cache.append((H, Z))
        if l + 1 < len(layers):
            H = gelu(Z)
            if mode == "bf16":
                H = H.astype(np.float32).astype(np.float64)
        else:
            H = Z
    return H, cache


def loss_grad(p, dims, X, Y, norm, mode="f64"):
    """norm = (in_mean, in_std, out_mean, out_std).  Returns (loss, flat grad)."""
    mi, si, mo, so = (np.asarray(a, np.float64) for a in norm)
    if mode == "bf16":
        Xs = ((np.asarray(X, np.float32) - mi.astype(np.float32)) / si.astype(np.float32)).astype(np.float64)
        Ys = ((np.asarray(Y, np.float32) - mo.astype(np.float32)) / so.astype(np.float32)).astype(np.float64)
    else:
        Xs, Ys = (X - mi) / si, (Y - mo) / so
    out, cache = forward(p, dims, Xs, mode)
    n, d = Ys.shape
    err = out - Ys
    loss = float(np.sum(err * err) / (n * d))
    dZ = 2.0 * err / (n * d)
    layers = unpack(p, dims)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        H, Z = cache[l]
        W, _ = layers[l]
        gW = _mm(dZ.T, H, mode)
        gb = dZ.sum(axis=0)
        grads[l] = (gW, gb)
        if l == 0:
            break
        dH = _mm(dZ, W, mode)
        dZ = dH * gelu_d(cache[l - 1][1])
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def predict(p, dims, X, norm, mode="f64"):
    mi, si, mo, so = (np.asarray(a, np.float64) for a in norm)
    out, _ = forward(p, dims, (np.asarray(X, np.float64) - mi) / si, mode)
    return out * so + mo
