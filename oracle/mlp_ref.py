"""CPU reference of the learned proposal (TEST INFRASTRUCTURE).

numpy restatement of the model the library trains and runs on the GPU, with
the same bf16 rounding points: inputs, weights and post-ReLU activations and
the backpropagated deltas are bf16; every product and sum is fp32.  Used by
tests/ to check the GPU features, forward pass and gradients.
"""
import math

import numpy as np


def bf16(x):
    """Round fp32 to bf16 (nearest even) and return it as fp32."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def shape(Z, hidden=256):
    Z2 = (Z + 1) // 2
    n_in = 4 + 25 * Z
    k_in = (n_in + 31) // 32 * 32
    n_out = (Z - 1) * Z2
    n_op = (n_out + 63) // 64 * 64
    return n_in, k_in, hidden, n_out, n_op


def unpack(params, Z, hidden=256):
    n_in, k_in, H, n_out, n_op = shape(Z, hidden)
    o = 0
    out = {}
    for name, sh in (("W1", (k_in, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, n_op)), ("b3", (n_op,))):
        n = int(np.prod(sh))
        out[name] = params[o:o + n].reshape(sh)
        o += n
    assert o == params.size
    return out


def loglut(lut):
    v = np.log(np.maximum(np.asarray(lut, np.float32), np.float32(1e-30))).astype(np.float32)
    return (np.clip(v, -40.0, 40.0) * np.float32(0.25)).astype(np.float32)


def features(X, Y, Z, rem, lut, src, dst, hidden=256):
    """Feature rows for pairs (cell encodings src, dst) at remaining capacity rem."""
    n_in, k_in, _, _, _ = shape(Z, hidden)
    ll = loglut(lut)

    def dec(v):
        z = v // (X * Y)
        r = v - z * X * Y
        return (z, r // Y, r % Y) if z & 1 else (z, r % X, r // X)

    def enc(z, x, y):
        return z * X * Y + (x * Y + y if z & 1 else y * X + x)

    out = np.zeros((len(src), k_in), np.float32)
    for i, (s, d) in enumerate(zip(src, dst)):
        z1, x1, y1 = dec(int(s))
        z2, x2, y2 = dec(int(d))
        dx, dy = x2 - x1, y2 - y1
        f = [dx / 64.0, dy / 64.0, float(z1), float(z2)]
        xs = [x1 + int(math.trunc((dx * u + (2 if dx >= 0 else -2)) / 4)) for u in range(5)]
        ys = [y1 + int(math.trunc((dy * u + (2 if dy >= 0 else -2)) / 4)) for u in range(5)]
        for l in range(Z):
            for u in range(5):
                for v in range(5):
                    r = int(rem[enc(l, xs[u], ys[v])])
                    f.append(ll[min(max(r, -256), 255) + 256])
        out[i, :n_in] = np.asarray(f, np.float32)
    return bf16(out)


def forward(p, x, exact=False):
    """Logits (fp32) and bf16 activations of rows x.  exact=True: float64
    throughout with no bf16 rounding (for checking gradients against finite
    differences)."""
    r = (lambda v: np.asarray(v, np.float64)) if exact else bf16
    xb = r(x)
    W1, W2, W3 = r(p["W1"]), r(p["W2"]), r(p["W3"])
    h1 = r(np.maximum(xb @ W1 + p["b1"], 0))
    h2 = r(np.maximum(h1 @ W2 + p["b2"], 0))
    logits = h2 @ W3 + p["b3"]
    return (logits if exact else logits.astype(np.float32)), (xb, h1, h2)


def gradient(p, x, y, n_out, exact=False):
    """Mean cross-entropy over the first n_out logits and its gradient (flat,
    in parameter order W1, b1, W2, b2, W3, b3).  exact=True: float64, no bf16
    rounding."""
    r = (lambda v: np.asarray(v, np.float64)) if exact else bf16
    logits, (xb, h1, h2) = forward(p, x, exact)
    n = x.shape[0]
    l = logits[:, :n_out].astype(np.float64)
    l = l - l.max(axis=1, keepdims=True)
    e = np.exp(l)
    sm = e / e.sum(axis=1, keepdims=True)
    loss = float(-np.log(sm[np.arange(n), y]).mean())
    d = sm.copy()
    d[np.arange(n), y] -= 1.0
    dL = np.zeros_like(logits)
    dL[:, :n_out] = r(d / n) if exact else r((d / n).astype(np.float32))
    W2, W3 = r(p["W2"]), r(p["W3"])
    g = {}
    g["W3"] = h2.T @ dL
    g["b3"] = dL.sum(axis=0)
    dz2 = r((dL @ W3.T) * (h2 > 0))
    g["W2"] = h1.T @ dz2
    g["b2"] = dz2.sum(axis=0)
    dz1 = r((dz2 @ W2.T) * (h1 > 0))
    g["W1"] = xb.T @ dz1
    g["b1"] = dz1.sum(axis=0)
    ft = np.float64 if exact else np.float32
    flat = np.concatenate([g[k].astype(ft).ravel() for k in ("W1", "b1", "W2", "b2", "W3", "b3")])
    return flat, loss
