'''CPU oracle of the learned-guess network kernels -- TEST INFRASTRUCTURE ONLY.

Only tests/ use it; the product never imports it.

Forward, backward and AdamW of the pair MLP in numpy (float64 arithmetic),
rounding to bfloat16 at exactly the points where the HIP kernels store
bfloat16 (inputs, weights, hidden activations, dL/dZ), so the kernels' FP32
results are compared with an exact-arithmetic restatement of the same
computation.
'''
import numpy as np


def bf16(x):
    '''Round float32 values to bfloat16 (nearest even), returned as float.'''
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7fff + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_bits(x):
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    return ((u + 0x7fff + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def silu(z):
    return z / (1. + np.exp(-z))


def exact(x):
    '''No rounding: with rnd=exact the oracle is the float64 network, which
    the tests compare with torch autograd.'''
    return np.asarray(x, dtype=np.float64)


def forward(x, Ws, bs, rnd=bf16):
    a = rnd(x)
    zs, acts = [], [a]
    for l, (W, b) in enumerate(zip(Ws, bs)):
        z = acts[-1] @ rnd(W) + b
        zs.append(z)
        if l < len(Ws) - 1:
            acts.append(rnd(silu(z)))
    return zs[-1], zs, acts


def backward(dy, zs, acts, Ws, rnd=bf16):
    '''dy: dL/dY (float); returns (dWs, dbs).'''
    dWs = [None] * len(Ws)
    dbs = [None] * len(Ws)
    d = dy
    d16 = rnd(dy)
    for l in range(len(Ws) - 1, -1, -1):
        if l == len(Ws) - 1:
            dbs[l] = d.sum(axis=0)
        dWs[l] = acts[l].T @ d16
        if l == 0:
            break
        dA = d16 @ rnd(Ws[l]).T
        z = zs[l - 1]
        s = 1. / (1. + np.exp(-z))
        dz = dA * s * (1. + z * (1. - s))
        dbs[l - 1] = dz.sum(axis=0)
        d16 = rnd(dz)
    return dWs, dbs


def mse_grad(y, t, mask, scale):
    '''Masked MSE: loss = scale sum mask (y - t)^2 and dloss/dy.'''
    r = mask * (y - t)
    return scale * float(np.sum(r * (y - t))), 2. * scale * r


def adamw(w, m, v, g, step, lr, b1=.9, b2=.999, eps=1e-8, wd=1e-4):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    w = w * (1 - lr * wd) - lr * (m / (1 - b1**step)) / (np.sqrt(v / (1 - b2**step)) + eps)
    return w, m, v
