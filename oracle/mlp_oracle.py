"""CPU restatement of the calibration emulator (covid-sim_amd/csrc/cs_mlp.hip).
TEST INFRASTRUCTURE: only tests/ use it, as the checker of the GPU kernels.

The network n_in -> H -> H -> n_out with GELU (erf form) between layers and
a linear output; loss = mean squared error over all outputs of a batch.
The parameter vector is W1 [H][n_in], b1, W2 [H][H], b2, W3 [n_out][H], b3.

Two forms:
  * exact(): float64 throughout, no rounding -- its analytic gradient is
    pinned by central finite differences;
  * rounded(): the GPU's numerics -- operands of every product rounded to
    BF16 (round to nearest even), products summed in float64 (the GPU sums
    in FP32), BF16 activations and output gradients at the points the GPU
    rounds them.
"""
import math

import numpy as np
from scipy.special import erf


def bf16(a):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def bf16_bits_to_f32(b):
    return (np.asarray(b, np.uint16).astype(np.uint32) << 16).view(np.float32)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def split(p, n_in, h, n_out):
    sizes = [h * n_in, h, h * h, h, n_out * h, n_out]
    out, o = [], 0
    for s in sizes:
        out.append(p[o:o + s])
        o += s
    W1, b1, W2, b2, W3, b3 = out
    return W1.reshape(h, n_in), b1, W2.reshape(h, h), b2, W3.reshape(n_out, h), b3


def n_params(n_in, h, n_out):
    return h * n_in + h + h * h + h + n_out * h + n_out


def exact(p, x, t, n_in, h, n_out):
    """Loss and gradient in float64 without rounding."""
    W1, b1, W2, b2, W3, b3 = [np.asarray(a, np.float64) for a in split(np.asarray(p, np.float64), n_in, h, n_out)]
    x = np.asarray(x, np.float64)
    z1 = x @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    y = h2 @ W3.T + b3
    e = y - t
    loss = np.mean(e * e)
    dy = 2.0 * e / e.size
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dz2 = (dy @ W3) * gelu_grad(z2)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * gelu_grad(z1)
    gW1 = dz1.T @ x
    gb1 = dz1.sum(0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grad, y


def rounded_forward(p, x, n_in, h, n_out):
    """The GPU forward: returns z1, h1, z2, h2, y (h1, h2 BF16-valued)."""
    W1, b1, W2, b2, W3, b3 = split(np.asarray(p, np.float32), n_in, h, n_out)
    xb = bf16(x).astype(np.float64)
    z1 = xb @ bf16(W1).astype(np.float64).T + b1
    h1 = bf16(gelu(z1).astype(np.float32))
    z2 = h1.astype(np.float64) @ bf16(W2).astype(np.float64).T + b2
    h2 = bf16(gelu(z2).astype(np.float32))
    y = h2.astype(np.float64) @ bf16(W3).astype(np.float64).T + b3
    return z1, h1, z2, h2, y


def layer(inp_bf16, W, b):
    """One product of the GPU forward from given BF16 inputs (float64 sum)."""
    return np.asarray(inp_bf16, np.float64) @ bf16(np.asarray(W, np.float32)).astype(np.float64).T + b


def rounded_backward(p, x, t, z1, h1, z2, h2, y, n_in, h, n_out):
    """The GPU backward from given forward values (z FP32, h BF16-valued)."""
    W1, b1, W2, b2, W3, b3 = split(np.asarray(p, np.float32), n_in, h, n_out)
    y = np.asarray(y, np.float32)
    e = y - np.asarray(t, np.float32)
    loss = np.mean(e.astype(np.float64) ** 2)
    dy = bf16(np.float32(2.0 / e.size) * e).astype(np.float64)
    gW3 = dy.T @ np.asarray(h2, np.float64)
    gb3 = dy.sum(0)
    dz2 = bf16(((dy @ bf16(W3).astype(np.float64)) * gelu_grad(np.asarray(z2, np.float64))).astype(np.float32)).astype(np.float64)
    gW2 = dz2.T @ np.asarray(h1, np.float64)
    gb2 = dz2.sum(0)
    dz1 = bf16(((dz2 @ bf16(W2).astype(np.float64)) * gelu_grad(np.asarray(z1, np.float64))).astype(np.float32)).astype(np.float64)
    gW1 = dz1.T @ bf16(x).astype(np.float64)
    gb1 = dz1.sum(0)
    grad = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])
    return loss, grad
