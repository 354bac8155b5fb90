"""The learned-h network restated on the CPU (test infrastructure only).

NumPy float64 versions of what the HIP kernels of the learned starting h compute: the forward
pass (optionally rounding inputs, weights and the first hidden layer to bfloat16 as the MFMA
inference does), the mean squared error on the normalised target, and its gradient by
backpropagation. The parameter and normalisation layouts are those of include/nebula_hip.h
(nb_mlp_*). The gradient itself is checked against central finite differences of the loss.
"""

import numpy as np

K_IN, K_HID = 8, 64
OFF_W1 = 0
OFF_B1 = OFF_W1 + K_IN * K_HID
OFF_W2 = OFF_B1 + K_HID
OFF_B2 = OFF_W2 + K_HID * K_HID
OFF_W3 = OFF_B2 + K_HID
OFF_B3 = OFF_W3 + K_HID
N_PARAMS = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(K_IN, K_HID), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(K_HID, K_HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])  # fmt: skip


def bf16_round(a):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float64."""
    u = np.ascontiguousarray(np.asarray(a, np.float32)).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def normalise(norm, feat):
    norm = np.asarray(norm, np.float64)
    return (np.asarray(feat, np.float64) - norm[:K_IN]) * norm[K_IN : 2 * K_IN]


def forward(params, norm, feat, bf16=False):
    """delta = target mean + target scale * MLP(normalised features)."""
    w1, b1, w2, b2, w3, b3 = unpack(params)
    x = normalise(norm, feat)
    if bf16:
        x, w1, w2 = bf16_round(x), bf16_round(w1), bf16_round(w2)
    h1 = np.tanh(x @ w1 + b1)
    if bf16:
        h1 = bf16_round(h1)
    h2 = np.tanh(h1 @ w2 + b2)
    y = h2 @ w3 + b3
    norm = np.asarray(norm, np.float64)
    return norm[2 * K_IN] + norm[2 * K_IN + 1] * y


def loss_and_grad(params, norm, feat, target):
    """Mean over the batch of (y - t)^2, t the normalised target, and its gradient."""
    w1, b1, w2, b2, w3, b3 = unpack(params)
    norm = np.asarray(norm, np.float64)
    x = normalise(norm, feat)
    t = (np.asarray(target, np.float64) - norm[2 * K_IN]) / norm[2 * K_IN + 1]
    n = len(t)
    h1 = np.tanh(x @ w1 + b1)
    h2 = np.tanh(h1 @ w2 + b2)
    y = h2 @ w3 + b3
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    dz2 = np.outer(dy, w3) * (1.0 - h2 * h2)
    dz1 = (dz2 @ w2.T) * (1.0 - h1 * h1)
    g = np.empty(N_PARAMS)
    g[OFF_W1:OFF_B1] = (x.T @ dz1).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(axis=0)
    g[OFF_W2:OFF_B2] = (h1.T @ dz2).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(axis=0)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    return loss, g
