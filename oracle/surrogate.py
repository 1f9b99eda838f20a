"""CPU oracle for the learned Born-radius surrogate (test infrastructure only).

The surrogate has no counterpart in origami; this file restates, in FP64
numpy, what the GPU computes, so the GPU's descriptor, forward and backward
kernels can be checked:

* ``descriptors``: 5 element classes (H, C, N, O, S/P by mass) x 32 radial
  bins of neighbour counts (same system, j != i, r < 12 Å, bin =
  floor(r / 0.375 Å) decided on r^2 against the exact edges 0.140625 k^2),
  scaled by 1/8, plus (rho - 1.5) * 4; rounded to bfloat16 like the GPU's
  stored inputs.
* ``forward`` / ``gradient``: MLP 161 -> 128 -> 128 -> 1 with SiLU, the
  parameter layout of og_surrogate_params, and (``emulate=True``) bfloat16
  rounding at the points where the GPU stores bfloat16 operands: inputs,
  both weight matrices, the layer-1 activations, and dZ2 / dZ1 before the
  GEMMs that consume them.  ``emulate=False`` is the exact FP64 network,
  which the tests pin against central finite differences.

Target radii come from origami's plain GB (gb.rs:203-342), whose radii equal
rho - 0.09 for every atom; the surrogate learns that mapping.
"""
from __future__ import annotations

import numpy as np

F = 161
K1 = 192
H = 128
OFF_W1 = 0
OFF_B1 = OFF_W1 + K1 * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
P = OFF_B3 + 1
R0 = 1.5
RS = 0.25


def bf16(a):
    """Round to bfloat16 (round to nearest even) via float32; returns float64."""
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def element_class(mass):
    m = np.asarray(mass, dtype=np.float64)
    return np.select([m < 2.0, m < 13.0, m < 15.0, m < 17.0], [0, 1, 2, 3], 4)


def _bins(r2):
    k = np.floor(np.sqrt(r2 / 0.140625)).astype(np.int64)
    k = np.minimum(k, 31)
    up = (k < 31) & (0.140625 * ((k + 1) ** 2).astype(np.float64) <= r2)
    k = k + up
    down = (k > 0) & (0.140625 * (k ** 2).astype(np.float64) > r2)
    return k - down


def descriptors(pos, mass, rho, offsets):
    """[N, 161] descriptors as the GPU stores them (bfloat16 values)."""
    pos = np.asarray(pos, dtype=np.float64)
    cls = element_class(mass)
    n = len(pos)
    X = np.zeros((n, F))
    for s in range(len(offsets) - 1):
        a, b = int(offsets[s]), int(offsets[s + 1])
        p = pos[a:b]
        dx = p[None, :, 0] - p[:, None, 0]
        dy = p[None, :, 1] - p[:, None, 1]
        dz = p[None, :, 2] - p[:, None, 2]
        r2 = dx * dx + dy * dy + dz * dz
        ok = r2 < 144.0
        np.fill_diagonal(ok, False)
        ii, jj = np.nonzero(ok)
        feat = cls[a + jj] * 32 + _bins(r2[ii, jj])
        np.add.at(X, (a + ii, feat), 1.0)
    X[:, :160] *= 0.125
    X[:, 160] = np.float32((np.asarray(rho) - 1.5) * 4.0)
    return bf16(X)


def descriptors_bruteforce(pos, mass, rho, offsets):
    """Plain-loop restatement (small inputs) that pins ``descriptors``."""
    cls = element_class(mass)
    n = len(pos)
    X = np.zeros((n, F))
    for s in range(len(offsets) - 1):
        for i in range(int(offsets[s]), int(offsets[s + 1])):
            for j in range(int(offsets[s]), int(offsets[s + 1])):
                if i == j:
                    continue
                r = float(np.linalg.norm(np.asarray(pos[j]) - np.asarray(pos[i])))
                if r < 12.0:
                    X[i, cls[j] * 32 + min(int(r / 0.375), 31)] += 1.0
    X[:, :160] /= 8.0
    X[:, 160] = np.float32((np.asarray(rho) - 1.5) * 4.0)
    return bf16(X)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return {"W1": p[OFF_W1:OFF_B1].reshape(K1, H), "b1": p[OFF_B1:OFF_W2],
            "W2": p[OFF_W2:OFF_B2].reshape(H, H), "b2": p[OFF_B2:OFF_W3],
            "w3": p[OFF_W3:OFF_B3], "b3": p[OFF_B3]}


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * _sig(z)


def dsilu(z):
    s = _sig(z)
    return s * (1.0 + z * (1.0 - s))


def forward(p, X, emulate=True):
    """y (normalised output) and the intermediates, for X of shape [n, 161]."""
    q = unpack(p)
    rnd = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    x = np.zeros((len(X), K1))
    x[:, :F] = rnd(X)
    W1 = rnd(q["W1"])
    W2 = rnd(q["W2"])
    z1 = x @ W1 + q["b1"]
    h1 = rnd(silu(z1))
    z2 = h1 @ W2 + q["b2"]
    h2 = silu(z2)
    y = h2 @ q["w3"] + q["b3"]
    return y, dict(x=x, W1=W1, W2=W2, z1=z1, h1=h1, z2=z2, h2=h2)


def radius(y):
    return R0 + RS * np.asarray(y)


def gradient(p, X, r_target, emulate=True):
    """Gradient of mean((y - t)^2), t = (R_target - 1.5) / 0.25, and the loss."""
    q = unpack(p)
    rnd = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    y, c = forward(p, X, emulate)
    t = (np.asarray(r_target, dtype=np.float64) - R0) / RS
    n = len(y)
    d = y - t
    dy = 2.0 * d / n
    g = np.zeros(P)
    g[OFF_W3:OFF_B3] = c["h2"].T @ dy
    g[OFF_B3] = dy.sum()
    dz2 = dy[:, None] * q["w3"][None, :] * dsilu(c["z2"])
    dz2r = rnd(dz2)
    g[OFF_W2:OFF_B2] = (c["h1"].T @ dz2r).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dh1 = dz2r @ c["W2"].T
    dz1 = dh1 * dsilu(c["z1"])
    dz1r = rnd(dz1)
    g[OFF_W1:OFF_B1] = (c["x"].T @ dz1r).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return g, float(np.mean(d * d))


def init_params(seed=0):
    """A random parameter vector with the GPU's layout (rows 161..191 of W1 zero)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(P)
    W1 = np.zeros((K1, H))
    W1[:F] = rng.normal(scale=1 / np.sqrt(F), size=(F, H))
    p[OFF_W1:OFF_B1] = W1.ravel()
    p[OFF_B1:OFF_W2] = rng.normal(scale=0.1, size=H)
    p[OFF_W2:OFF_B2] = rng.normal(scale=1 / np.sqrt(H), size=H * H)
    p[OFF_B2:OFF_W3] = rng.normal(scale=0.1, size=H)
    p[OFF_W3:OFF_B3] = rng.normal(scale=1 / np.sqrt(H), size=H)
    p[OFF_B3] = 0.05
    return p.astype(np.float32).astype(np.float64)
