"""Pins of the learned-light network oracle (oracle/mlp_oracle.py).

The network has no reference counterpart, so the oracle is pinned by what
the mathematics fixes: central finite differences of the FP64 loss for
every parameter block (a dropped SiLU' factor, a transposed weight or a
missing bias term fails), closed forms of SiLU, exact BF16 rounding cases,
and the quantized forward against a hand-unrolled scalar evaluation.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as N


def _params(seed=3, scale=1.0):
    rng = np.random.default_rng(seed)
    d = {
        "W1": rng.uniform(-1, 1, (9, 64)) * np.sqrt(3 / 9) * scale,
        "b1": rng.uniform(-0.2, 0.2, 64),
        "W2": rng.uniform(-1, 1, (64, 64)) * np.sqrt(3 / 64) * scale,
        "b2": rng.uniform(-0.2, 0.2, 64),
        "W3": rng.uniform(-1, 1, (64, 64)) * np.sqrt(3 / 64) * scale,
        "b3": rng.uniform(-0.2, 0.2, 64),
        "w4": rng.uniform(-1, 1, 64) * np.sqrt(3 / 64),
        "b4": 0.3,
    }
    return N.pack(d)


def _data(n=48, seed=5):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, 1, (n, 9))
    x[:, 5:8] /= np.linalg.norm(x[:, 5:8], axis=1, keepdims=True)
    y = -rng.uniform(0, 3, n)
    return x, y


def test_silu_closed_forms():
    assert N.silu(0.0) == 0.0
    assert abs(N.silu(1.0) - 1.0 / (1.0 + np.exp(-1.0))) < 1e-15
    # silu(z) - silu(-z) = z
    z = np.linspace(-6, 6, 25)
    assert np.allclose(N.silu(z) - N.silu(-z), z, atol=1e-12)
    # derivative by central differences
    h = 1e-6
    assert np.allclose(N.silu_grad(z), (N.silu(z + h) - N.silu(z - h)) / (2 * h), atol=1e-8)


def test_bf16_rounding_cases():
    # exactly representable values are unchanged
    for v in (0.0, 1.0, -2.5, 0.15625, 2.0 ** 100, -(2.0 ** -100)):
        assert N.bf16(v) == np.float32(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins
    assert N.bf16(1.0 + 2.0 ** -8) == 1.0
    # 1 + 3 * 2^-8 ties between 1 + 2^-7 and 1 + 2^-6: 1 + 2^-6 is even
    assert N.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    # just above the tie rounds up
    assert N.bf16(np.float32(1.0 + 2.0 ** -8 + 2.0 ** -20)) == 1.0 + 2.0 ** -7
    # relative error bound 2^-9
    x = np.random.default_rng(0).uniform(-100, 100, 1000)
    assert np.all(np.abs(N.bf16(x) - x) <= np.abs(x) * 2.0 ** -8)


def test_forward_matches_scalar_unrolled_evaluation():
    p = _params()
    x, _ = _data(4)
    d = N.unpack(p)
    got = N.forward(p, x, quantized=False)
    for r in range(4):
        a = list(x[r])
        for W, b in ((d["W1"], d["b1"]), (d["W2"], d["b2"]), (d["W3"], d["b3"])):
            z = [sum(a[k] * W[k][n] for k in range(len(a))) + b[n] for n in range(64)]
            a = [zi / (1.0 + np.exp(-zi)) for zi in z]
        y = sum(a[k] * d["w4"][k] for k in range(64)) + d["b4"]
        assert abs(got[r] - y) < 1e-12 * max(1.0, abs(y))


def test_pack_unpack_round_trip_and_offsets():
    p = np.arange(N.N_PARAMS, dtype=np.float64)
    d = N.unpack(p)
    assert d["W1"][0, 1] == 1 and d["W1"][1, 0] == 64  # W[k][n] row-major by input
    assert d["b1"][0] == 576 and d["W2"][0, 0] == 640 and d["b4"] == 9024
    assert np.array_equal(N.pack(d), p)


@pytest.mark.parametrize("block", ["W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4"])
def test_fp64_gradient_matches_central_differences(block):
    p = _params()
    x, y = _data()
    _, g = N.loss_and_grad(p, x, y, quantized=False)
    o, shape = N.OFF[block]
    n = int(np.prod(shape)) if shape else 1
    rng = np.random.default_rng(11)
    idx = o + rng.choice(n, size=min(n, 12), replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (N.mse(p + e, x, y) - N.mse(p - e, x, y)) / (2 * h)
        assert abs(g[i] - fd) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (block, i, g[i], fd)


def test_quantized_forward_stays_close_to_fp64():
    p = _params()
    x, _ = _data(256)
    a = N.forward(p, x, quantized=True)
    b = N.forward(p, x, quantized=False)
    # BF16 (8-bit mantissa) on inputs, weights and activations
    assert np.max(np.abs(a - b)) < 0.05 * max(1.0, np.max(np.abs(b)))
    assert np.max(np.abs(a - b)) > 0.0
