"""NumPy reference of the learned cost model's network (test infrastructure).

Mirrors the GPU kernels' arithmetic: BF16-rounded GEMM operands (inputs,
weights, hidden activations, output-layer gradients), FP32/FP64 accumulation,
MSE loss.  Parameter layout: W1[64][128] | b1 | W2[128][128] | b2 | w3 | b3."""
import numpy as np

IN, HID = 64, 128


ROUND = True


def bf16(x):
    if not ROUND:
        return np.asarray(x, np.float64)
    x = np.asarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    o = 0
    W1 = p[o:o + IN * HID].reshape(IN, HID); o += IN * HID
    b1 = p[o:o + HID]; o += HID
    W2 = p[o:o + HID * HID].reshape(HID, HID); o += HID * HID
    b2 = p[o:o + HID]; o += HID
    w3 = p[o:o + HID]; o += HID
    b3 = p[o]
    return W1, b1, W2, b2, w3, b3


def forward(p, X):
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, np.float64))
    Xb = bf16(X).astype(np.float64)
    z1 = Xb @ bf16(W1).astype(np.float64) + b1
    h1 = bf16(np.maximum(z1, 0)).astype(np.float64)
    z2 = h1 @ bf16(W2).astype(np.float64) + b2
    h2 = bf16(np.maximum(z2, 0)).astype(np.float64)
    y = h2 @ w3 + b3
    return y, (Xb, h1, h2)


def gradients(p, X, t):
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(p, np.float64))
    y, (Xb, h1, h2) = forward(p, X)
    N = len(y)
    dy = 2.0 * (y - t) / N
    loss = float(np.mean((y - t) ** 2))
    dw3 = h2.T @ dy
    db3 = dy.sum()
    dz2 = np.where(h2 > 0, dy[:, None] * w3[None, :], 0.0)
    db2 = dz2.sum(0)
    dz2b = bf16(dz2).astype(np.float64)
    dW2 = h1.T @ dz2b
    dh1 = dz2b @ bf16(W2).astype(np.float64).T
    dz1 = np.where(h1 > 0, dh1, 0.0)
    db1 = dz1.sum(0)
    dW1 = Xb.T @ bf16(dz1).astype(np.float64)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dw3, [db3]])
    return g, loss
