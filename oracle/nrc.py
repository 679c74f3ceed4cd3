"""CPU oracle for the neural radiance cache (test infrastructure only).

The cache has no counterpart in the reference renderer; this numpy
restatement defines what the GPU kernels must compute and is itself pinned
by finite differences (tests/test_oracle_nrc.py): the query encoding, the
bias-free MLP (96 -> 5 x 64 ReLU -> 16), the relative-L2 loss with its
constant-denominator gradient, and the Adam update. `bf16=True` rounds
weights, inputs, hidden activations and back-propagated gradients to
bfloat16 (round to nearest even) at the points where the GPU stores them.
"""
import numpy as np

IN, HID, OUT = 96, 64, 16
SHAPES = [(HID, IN)] + [(HID, HID)] * 4 + [(OUT, HID)]
NPARAMS = sum(o * i for o, i in SHAPES)


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def split(w):
    out, off = [], 0
    for o, i in SHAPES:
        out.append(np.asarray(w[off: off + o * i], np.float64).reshape(o, i))
        off += o * i
    return out


def join(ws):
    return np.concatenate([np.asarray(w).reshape(-1) for w in ws]).astype(np.float32)


def init_weights(seed=0):
    """He-uniform weights with zero padding rows/columns (input features >= 83, outputs >= 3)."""
    rng = np.random.default_rng(seed)
    ws = []
    for k, (o, i) in enumerate(SHAPES):
        fan_in = 83 if k == 0 else HID
        b = np.sqrt(6.0 / fan_in) * (0.1 if k == 5 else 1.0)
        w = rng.uniform(-b, b, (o, i))
        if k == 0:
            w[:, 83:] = 0
        if k == 5:
            w[3:, :] = 0
        ws.append(w)
    return join(ws)


def encode(q, lo, hi):
    """Query records (n x 16) -> features (n x 96), float64."""
    q = np.asarray(q, np.float64)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    ext = hi - lo
    inv = np.where(ext > 0, 1.0 / np.where(ext > 0, ext, 1), 0.0)
    pn = np.clip((q[:, 0:3] - lo) * inv, 0.0, 1.0)
    n = len(q)
    x = np.zeros((n, IN))
    for d in range(3):
        for i in range(12):
            a = np.pi * pn[:, d] * (2.0 ** i)
            x[:, 24 * d + 2 * i] = np.sin(a)
            x[:, 24 * d + 2 * i + 1] = np.cos(a)
    x[:, 72:83] = q[:, 3:14]
    return x


def forward(w, x, use_bf16=True):
    ws = split(w)
    r = bf16 if use_bf16 else (lambda v: np.asarray(v, np.float64))
    hs = [r(x)]
    h = hs[0]
    for k in range(5):
        h = r(np.maximum(h @ r(ws[k]).T, 0.0))
        hs.append(h)
    y = h @ r(ws[5]).T
    return hs, y


def loss_grad(w, x, t, weight=None, use_bf16=True):
    """Loss (sum over samples of per-sample losses / (3n)) and dL/dw (flat)."""
    ws = split(w)
    r = bf16 if use_bf16 else (lambda v: np.asarray(v, np.float64))
    n = len(x)
    weight = np.ones(n) if weight is None else np.asarray(weight, np.float64)
    hs, y = forward(w, x, use_bf16)
    yr = y[:, :3]
    den = yr * yr + 0.01
    loss = (weight[:, None] * (yr - t[:, :3]) ** 2 / den).sum() / (3 * n)
    g = np.zeros((n, OUT))
    g[:, :3] = 2 * (yr - t[:, :3]) / den * weight[:, None] / (3 * n)
    g = r(g)
    grads = [None] * 6
    grads[5] = g.T @ hs[5]
    d = g @ r(ws[5])
    for k in range(4, -1, -1):
        gk = r(d * (hs[k + 1] > 0))
        grads[k] = gk.T @ hs[k]
        d = gk @ r(ws[k])
    return loss, join(grads).astype(np.float64)


def adam(w, g, m, v, step, lr=1e-3, b1=0.9, b2=0.99, eps=1e-8):
    w, g = np.asarray(w, np.float64), np.asarray(g, np.float64)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
