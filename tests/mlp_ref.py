"""NumPy reference of the surrogate network (test infrastructure): forward
pass, weighted mean-square loss and its gradient by backpropagation, in
float64, with the parameter layout of include/pnamp_mlp.h (W_l [out][in]
row-major, then b_l, layer by layer)."""
import numpy as np


def unpack(p, sizes):
    out, o = [], 0
    for l in range(len(sizes) - 1):
        ni, no = sizes[l], sizes[l + 1]
        W = p[o:o + ni * no].reshape(no, ni)
        o += ni * no
        b = p[o:o + no]
        o += no
        out.append((W, b))
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, sizes, X):
    a = np.asarray(X, dtype=np.float64)
    layers = unpack(np.asarray(p, dtype=np.float64), sizes)
    for l, (W, b) in enumerate(layers):
        z = a @ W.T + b
        a = silu(z) if l + 1 < len(layers) else z
    return a


def loss_grad(p, sizes, X, Y, wout=None):
    p = np.asarray(p, dtype=np.float64)
    layers = unpack(p, sizes)
    B, nout = Y.shape
    w = np.ones(nout) if wout is None else np.asarray(wout, dtype=np.float64)
    acts, zs = [np.asarray(X, dtype=np.float64)], []
    for l, (W, b) in enumerate(layers):
        z = acts[-1] @ W.T + b
        zs.append(z)
        acts.append(silu(z) if l + 1 < len(layers) else z)
    r = acts[-1] - Y
    loss = float((w * r * r).sum() / (B * nout))
    d = 2.0 * w * r / (B * nout)
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W, b = layers[l]
        grads.append((d.T @ acts[l], d.sum(0)))
        if l > 0:
            d = (d @ W) * dsilu(zs[l - 1])
    g = []
    for gW, gb in reversed(grads):
        g.append(gW.ravel())
        g.append(gb)
    return loss, np.concatenate(g)
