"""The learned proposal's CPU reference (oracle/mlp_ref.py), pinned against
float64 central differences of its own loss with bf16 rounding switched off.

The GPU gradient kernels are checked against mlp_ref.gradient; this pins
that gradient to the mathematics: a dropped 1/n, a wrong ReLU mask, a
transposed weight or a missing bias term fails here.
"""
import numpy as np
import pytest

from oracle import mlp_ref


def _params(rng, k, H, n_op):
    return {"W1": rng.normal(0, 0.6, (k, H)), "b1": rng.normal(0, 0.3, H), "W2": rng.normal(0, 0.5, (H, H)),
            "b2": rng.normal(0, 0.3, H), "W3": rng.normal(0, 0.5, (H, n_op)), "b3": rng.normal(0, 0.3, n_op)}


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_gradient_equals_central_differences(seed):
    rng = np.random.default_rng(seed)
    k, H, n_op, n_out, n = 7, 6, 9, 6, 5
    p = _params(rng, k, H, n_op)
    x = rng.normal(0, 1.0, (n, k))
    y = rng.integers(0, n_out, n)
    g, loss = mlp_ref.gradient(p, x, y, n_out, exact=True)
    names = ("W1", "b1", "W2", "b2", "W3", "b3")
    h = 1e-6
    fd = []
    for name in names:
        a = p[name]
        for idx in np.ndindex(a.shape):
            old = a[idx]
            a[idx] = old + h
            lp = mlp_ref.gradient(p, x, y, n_out, exact=True)[1]
            a[idx] = old - h
            lm = mlp_ref.gradient(p, x, y, n_out, exact=True)[1]
            a[idx] = old
            fd.append((lp - lm) / (2 * h))
    fd = np.array(fd)
    # pre-activations stay away from the ReLU kink by far more than h
    z1 = x @ p["W1"] + p["b1"]
    z2 = np.maximum(z1, 0) @ p["W2"] + p["b2"]
    assert np.abs(z1).min() > 1e-4 and np.abs(z2).min() > 1e-4
    assert np.abs(g - fd).max() <= 1e-7, np.abs(g - fd).max()
    # columns past n_out carry no loss: their W3 / b3 gradient is exactly 0
    G = dict(zip(names, np.split(g, np.cumsum([p[m].size for m in names])[:-1])))
    assert not G["W3"].reshape(H, n_op)[:, n_out:].any() and not G["b3"][n_out:].any()
    assert np.isfinite(loss) and loss > 0


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -6, 1.0, -2.5, np.float32(3.0e38)], np.float32)
    got = mlp_ref.bf16(v)
    assert got[0] == 1.0 and got[1] == 1.0  # tie at 1 + 2^-8 rounds to even (1.0)
    assert got[2] == np.float32(1.0 + 2 ** -6)  # tie at 1 + 3*2^-8 rounds to even (1 + 2^-6)
    assert got[3] == 1.0 and got[4] == -2.5
    assert abs(got[5] - want[5]) <= 2 ** 120
