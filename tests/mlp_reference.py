"""CPU reference of the candidate scorer (test infrastructure only).

A 126-256-256-1 perceptron with ReLU hidden layers and a mean squared error
loss; parameter blob layout W1[128][256], b1, W2[256][256], b2, w3, b3 (the
C ABI's).  `bf16=True` rounds what the device keeps in BF16 (the features, W1,
W2, and the activations H1, H2, dZ1, dZ2 that feed the next matrix product) to
nearest-even, as the device kernels do; `bf16=False` is plain FP32, whose
analytic gradient the tests pin with central finite differences.
"""
import numpy as np

IN, IN_PAD, HID = 126, 128, 256
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN_PAD * HID
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + HID * HID
OFF_W3 = OFF_B2 + HID
OFF_B3 = OFF_W3 + HID
PARAMS = OFF_B3 + 1


def round_bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p, dtype=np.float32)
    return (p[OFF_W1:OFF_B1].reshape(IN_PAD, HID), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(HID, HID),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def init_params(seed, scale=1.0):
    """He-style initial parameters (W1 rows 126, 127 zero)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(PARAMS, dtype=np.float32)
    w1 = rng.normal(0, scale * np.sqrt(2.0 / IN), (IN_PAD, HID)).astype(np.float32)
    w1[IN:] = 0
    p[OFF_W1:OFF_B1] = w1.reshape(-1)
    p[OFF_W2:OFF_B2] = rng.normal(0, scale * np.sqrt(2.0 / HID), HID * HID).astype(np.float32)
    p[OFF_W3:OFF_B3] = rng.normal(0, scale * np.sqrt(1.0 / HID), HID).astype(np.float32)
    p[OFF_B1:OFF_W2] = 0.01
    p[OFF_B2:OFF_W3] = 0.01
    return p


def forward(p, X, bf16=True):
    r = round_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float32))
    W1, b1, W2, b2, w3, b3 = unpack(p)
    Xp = np.zeros((X.shape[0], IN_PAD), dtype=np.float32)
    Xp[:, :IN] = X
    Xp, W1, W2 = r(Xp), r(W1), r(W2)
    H1 = r(np.maximum(Xp @ W1 + b1, 0))
    H2 = r(np.maximum(H1 @ W2 + b2, 0))
    s = H2 @ w3 + b3
    return s.astype(np.float32), (Xp, W1, W2, H1, H2)


def loss_and_grads(p, X, y, bf16=True):
    r = round_bf16 if bf16 else (lambda a: np.asarray(a, dtype=np.float32))
    s, (Xp, W1, W2, H1, H2) = forward(p, X, bf16)
    n = X.shape[0]
    e = s - y
    loss = float(np.mean(e * e))
    ds = (2.0 * e / n).astype(np.float32)
    _, _, _, _, w3, _ = unpack(p)
    g = np.zeros(PARAMS, dtype=np.float32)
    g[OFF_W3:OFF_B3] = H2.T @ ds
    g[OFF_B3] = ds.sum()
    dZ2 = r(np.outer(ds, w3) * (H2 > 0))
    g[OFF_B2:OFF_W3] = dZ2.sum(0)
    g[OFF_W2:OFF_B2] = (H1.T @ dZ2).reshape(-1)
    dZ1 = r((dZ2 @ W2.T) * (H1 > 0))
    g[OFF_B1:OFF_W2] = dZ1.sum(0)
    g[OFF_W1:OFF_B1] = (Xp.T @ dZ1).reshape(-1)
    return loss, g
