"""CPU oracle of the KI-cooling surrogate network (test infrastructure only).

Restates, in float64 numpy, the MLP 3 -> 64 -> 64 -> 1 (SiLU) of
code_tohoku_amd/csrc/mcf_mlp.h and the loss / gradient of its training
kernel (code_tohoku_amd/csrc/mcf_cool.hip, k_mlp_grad):

  z1 = W1b xb + b1b          xb, W1b, b1b rounded to BF16 (MFMA operands)
  h1 = bf16(silu(z1))        rounded: it is the next MFMA operand
  z2 = W2b h1 + b2           W2b rounded to BF16, b2 FP32
  y  = W3 . silu(z2) + b3
  L  = mean((y - t)^2)

The gradient treats the BF16 roundings as identity (straight-through) and
uses the rounded values the forward pass used.  With ``bf16=False`` every
rounding is skipped and the function is smooth, which the finite-difference
check of the tests uses to pin the analytic gradient.

The surrogate has no counterpart in the reference; the sample targets come
from IMC (3DMHD-CfCA/MCF3DSGincfca3.f90:2898-2921), restated in
oracle/mcf_oracle.c (mcfo_imc).
"""
from __future__ import annotations

import numpy as np

H = 64
OW1, OB1, OW2, OB2, OW3, OB3, NP = 0, 192, 256, 4352, 4416, 4480, 4481
MASK = 0xFFFFFFFFFFFFFFFF


def bf16(a):
    """Round to nearest even BF16 (returned as float64)."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    return (w[OW1:OB1].reshape(H, 3), w[OB1:OW2], w[OW2:OB2].reshape(H, H), w[OB2:OW3], w[OW3:OB3], w[OB3])


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(w, X, bf=True):
    W1, b1, W2, b2, W3, b3 = unpack(w)
    r = bf16 if bf else (lambda a: np.asarray(a, dtype=np.float64))
    xb = r(X)
    z1 = xb @ r(W1).T + r(b1)
    h1 = r(silu(z1))
    z2 = h1 @ r(W2).T + b2
    h2 = silu(z2)
    y = h2 @ W3 + b3
    return y, (xb, z1, h1, z2, h2)


def loss_grad(w, X, t, bf=True):
    """Mean-squared-error loss and its gradient (flat, NP values)."""
    W1, b1, W2, b2, W3, b3 = unpack(w)
    y, (xb, z1, h1, z2, h2) = forward(w, X, bf)
    n = len(t)
    e = y - np.asarray(t, dtype=np.float64)
    L = float(np.mean(e * e))
    dy = 2.0 * e / n
    g = np.zeros(NP)
    g[OB3] = dy.sum()
    g[OW3:OB3] = h2.T @ dy
    dz2 = (dy[:, None] * W3[None, :]) * dsilu(z2)
    g[OB2:OW3] = dz2.sum(0)
    g[OW2:OB2] = (dz2.T @ h1).ravel()
    W2r = bf16(W2) if bf else W2
    dz1 = (dz2 @ W2r) * dsilu(z1)
    g[OB1:OW2] = dz1.sum(0)
    g[OW1:OB1] = (dz1.T @ xb).ravel()
    return L, g


# ---------------------------------------------------------------------------
# training samples (mcf_cool.hip, k_cool_samples)
# ---------------------------------------------------------------------------
def _splitmix(seed: int, idx: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * (idx.astype(np.uint64) + np.uint64(1)))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def sample_inputs(n: int, seed: int):
    """(N, T[10^3 K], dt/t_cool) of samples 0..n-1 of k_cool_samples."""
    i = np.arange(n, dtype=np.uint64)
    lN = -1.0 + 5.0 * _splitmix(seed, 3 * i)
    lT = 1.0 + 3.0 * _splitmix(seed, 3 * i + 1)
    lr = -3.0 + 6.0 * _splitmix(seed, 3 * i + 2)
    return 10.0 ** lN, 10.0 ** lT * 1.0e-3, 10.0 ** lr
