"""The calibration emulator's kernels (covid-sim_amd/csrc/cs_mlp.hip).

There is no reference counterpart (the reference calibrates by re-running
the simulator).  The emulator is pinned in two steps:
  * the CPU restatement (oracle/mlp_oracle.py) -- its analytic gradient
    against central finite differences in float64, its BF16 rounding against
    known bit patterns;
  * the GPU against the restatement: each layer's product from the GPU's own
    inputs of that layer (FP32 accumulation vs float64: 1e-5 of the scale of
    the products), the BF16 activations (at most one BF16 step apart, from
    the FP32 vs float64 GELU), the gradient of every parameter block from the
    GPU's forward values (relative norm error 1e-3: BF16 rounding of the
    output gradients may differ by one step where the FP32 and float64
    products straddle a rounding point), and training that lowers the loss.
The shape is the one the calibration uses: 10 -> 512 -> 512 -> 730.
"""
import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mlp_oracle as mo  # noqa: E402

SHAPE = (10, 512, 730)


def test_bf16_rounding_bits():
    x = np.array([1.0, -2.5, 1.00390625, 1.01171875, 3.14159265, 1e-3, 65504.0], np.float32)
    got = mo.bf16(x).view(np.uint32) >> 16
    # 1.00390625 = 1 + 2^-8 is a tie between 0x3f80 and 0x3f81: to even 0x3f80;
    # 1.01171875 = 1 + 3*2^-8 ties between 0x3f81 and 0x3f82: to even 0x3f82
    assert list(got) == [0x3F80, 0xC020, 0x3F80, 0x3F82, 0x4049, 0x3A83, 0x4780]


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    n_in, h, n_out, b = 10, 16, 12, 7
    p = rng.normal(0, 0.4, mo.n_params(n_in, h, n_out))
    x = rng.normal(size=(b, n_in))
    t = rng.normal(size=(b, n_out))
    _, g, _ = mo.exact(p, x, t, n_in, h, n_out)
    eps = 1e-6
    idx = rng.choice(p.size, 80, replace=False)
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp = mo.exact(q, x, t, n_in, h, n_out)[0]
        q[i] -= 2 * eps
        lm = mo.exact(q, x, t, n_in, h, n_out)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_rounded_oracle_tracks_exact():
    """Rounding to BF16 moves the loss gradient only by BF16-sized amounts."""
    rng = np.random.default_rng(5)
    n_in, h, n_out, b = 10, 64, 40, 33
    p = rng.normal(0, 0.2, mo.n_params(n_in, h, n_out)).astype(np.float32)
    x = rng.normal(size=(b, n_in)).astype(np.float32)
    t = rng.normal(size=(b, n_out)).astype(np.float32)
    le, ge, _ = mo.exact(p.astype(np.float64), x, t, n_in, h, n_out)
    z1, h1, z2, h2, y = mo.rounded_forward(p, x, n_in, h, n_out)
    lr, gr = mo.rounded_backward(p, x, t, z1, h1, z2, h2, y, n_in, h, n_out)
    assert abs(lr - le) < 0.02 * le
    assert np.linalg.norm(gr - ge) < 0.05 * np.linalg.norm(ge)


def emulator(seed=11):
    return importlib.import_module("covid-sim_amd.emulator").Emulator(*SHAPE[:2], SHAPE[2], seed=seed)


def rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.gpu
def test_gpu_forward_layer_by_layer():
    n_in, h, n_out = SHAPE
    em = emulator()
    p = em.params()
    W1, b1, W2, b2, W3, b3 = mo.split(p, n_in, h, n_out)
    rng = np.random.default_rng(1)
    x = rng.normal(size=(300, n_in)).astype(np.float32)  # 300: not a multiple of the 64-row tile
    z1, h1b, z2, h2b, y = em.forward_trace(x)
    h1, h2 = mo.bf16_bits_to_f32(h1b), mo.bf16_bits_to_f32(h2b)
    # each product from the GPU's own inputs of that layer
    for got, inp, W, b, K in [(z1, mo.bf16(x), W1, b1, n_in), (z2, h1, W2, b2, h), (y, h2, W3, b3, h)]:
        want = mo.layer(inp, W, b)
        scale = np.abs(np.asarray(inp, np.float64)) @ np.abs(mo.bf16(W).astype(np.float64)).T + np.abs(b)
        assert np.all(np.abs(got - want) <= 1e-5 * scale + 1e-6), K
    # activations: BF16 of GELU of the GPU's pre-activations, within one BF16 step
    for hb, z in [(h1, z1), (h2, z2)]:
        want = mo.bf16(mo.gelu(z.astype(np.float64)).astype(np.float32))
        step = np.abs(want) * 2.0 ** -7 + 1e-30
        assert np.all(np.abs(hb - want) <= step)
        assert np.mean(hb == want) > 0.99
    # end to end against the rounded restatement
    _, _, _, _, yo = mo.rounded_forward(p, x, n_in, h, n_out)
    assert rel(y, yo) < 1e-2


@pytest.mark.gpu
def test_gpu_gradient_matches_oracle():
    n_in, h, n_out = SHAPE
    em = emulator(seed=12)
    p = em.params()
    rng = np.random.default_rng(2)
    b = 200
    x = rng.normal(size=(b, n_in)).astype(np.float32)
    t = rng.normal(size=(b, n_out)).astype(np.float32)
    z1, h1b, z2, h2b, y = em.forward_trace(x)
    loss, g = em.loss_grad(x, t)
    lo, go = mo.rounded_backward(p, x, t, z1, mo.bf16_bits_to_f32(h1b), z2, mo.bf16_bits_to_f32(h2b), y, n_in, h, n_out)
    assert abs(loss - lo) <= 1e-6 * lo
    sizes = [h * n_in, h, h * h, h, n_out * h, n_out]
    o = 0
    for k, s in enumerate(sizes):
        assert rel(g[o:o + s], go[o:o + s]) < 1e-3, ("block", k, rel(g[o:o + s], go[o:o + s]))
        o += s


@pytest.mark.gpu
def test_gpu_training_lowers_loss():
    """A smooth synthetic map from 10 inputs to 730 outputs: Adam on the GPU
    lowers the held-out error well below the error of the mean predictor."""
    n_in, h, n_out = SHAPE
    rng = np.random.default_rng(4)
    n = 2048
    x = rng.uniform(-1, 1, size=(n, n_in)).astype(np.float32)
    days = np.arange(n_out // 2)
    peak = 150 + 60 * x[:, :1]
    width = 25 + 8 * x[:, 1:2]
    curve = np.exp(-0.5 * ((days - peak) / width) ** 2)
    t = np.concatenate([curve, 0.5 * np.roll(curve, 14, axis=1)], axis=1).astype(np.float32)
    em = emulator(seed=13)
    em.set_data(x[:1792], t[:1792])
    losses = em.train(epochs=60, batch=256, lr=2e-3)
    assert losses[-1] < 0.2 * losses[0]
    pred = em.forward(x[1792:])
    mse = np.mean((pred - t[1792:]) ** 2)
    base = np.mean((t[1792:] - t[:1792].mean(0)) ** 2)
    assert mse < 0.25 * base, (mse, base)
