"""Pins the neural-radiance-cache oracle (oracle/nrc.py) with finite
differences in float64 and closed forms for the encoding and Adam."""
import numpy as np

from oracle import nrc


def _queries(n, rng):
    q = np.zeros((n, 16))
    q[:, 0:3] = rng.uniform(-1, 1, (n, 3))
    q[:, 3:7] = rng.uniform(0, 1, (n, 4))
    q[:, 7] = rng.uniform(0.1, 1, n)
    q[:, 8:14] = rng.uniform(0, 1, (n, 6))
    return q


def test_encoding_closed_form():
    q = np.zeros((1, 16))
    q[0, 0:3] = [0.0, 0.5, 1.0]            # normalised (0.5, 0.75, 1.0) with bounds [-1, 1]
    q[0, 3:14] = np.arange(11) / 10
    x = nrc.encode(q, [-1, -1, -1], [1, 1, 1])[0]
    assert np.isclose(x[0], np.sin(np.pi * 0.5)) and np.isclose(x[1], np.cos(np.pi * 0.5))
    assert np.isclose(x[24 + 2], np.sin(2 * np.pi * 0.75))      # d = 1, frequency 1
    assert np.isclose(x[48 + 23], np.cos(2048 * np.pi * 1.0))   # d = 2, frequency 11
    assert np.allclose(x[72:83], np.arange(11) / 10) and (x[83:] == 0).all()


def test_bf16_rounding():
    v = nrc.bf16(np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-3]))
    assert v[0] == 1.0 and v[1] == 1.0                 # tie to even
    assert v[2] == 1.0 + 2 ** -6                         # 1 + 3/256 rounds up to 1 + 4/256
    assert v[3] == -2.5
    assert abs(v[4] - 1e-3) < 1e-3 * 2 ** -8


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    q = _queries(6, rng)
    x = nrc.encode(q, [-1, -1, -1], [1, 1, 1])
    t = rng.uniform(0, 2, (6, 4))
    w = nrc.init_weights(3).astype(np.float64)
    loss, g = nrc.loss_grad(w, x, t, use_bf16=False)
    idx = rng.choice(np.nonzero(w)[0], 60, replace=False)
    for k in idx:
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (nrc.loss_grad(wp, x, t, use_bf16=False)[0] - nrc.loss_grad(wm, x, t, use_bf16=False)[0]) / (2 * h)
        # the gradient treats the loss denominator as constant; compare against
        # the finite difference of the same surrogate by freezing the denominator
        assert np.isfinite(fd)
    # surrogate with frozen denominator: L(w) = sum (y - t)^2 / c, c fixed at w0
    _, y0 = nrc.forward(w, x, use_bf16=False)
    c = y0[:, :3] ** 2 + 0.01

    def surrogate(wv):
        _, y = nrc.forward(wv, x, use_bf16=False)
        return ((y[:, :3] - t[:, :3]) ** 2 / c).sum() / (3 * len(x))

    for k in idx:
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[k] += h
        wm[k] -= h
        fd = (surrogate(wp) - surrogate(wm)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 + 1e-4 * abs(fd), (k, fd, g[k])


def test_adam_first_step_is_signed_lr():
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([3.0, -0.2, 1e-3])
    w1, m, v = nrc.adam(w, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    assert np.allclose(w1, w - 0.01 * np.sign(g), atol=1e-7)
