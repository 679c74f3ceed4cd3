"""Frame resolve, tone mapping and image files against the oracle's
restatement of the reference post-processing (oracle/post.py), whose
operators are pinned here by their closed forms."""
import os
import tempfile

import numpy as np
import pytest

from oracle import post


def test_oracle_tone_closed_forms():
    img = np.array([[[0.0, 1.0, 3.0], [4.0, 0.5, 3.0]]])
    assert np.allclose(post.global_tone(img), img / (img + 1))
    d = post.drago_tone(img)
    # a pixel at the channel maximum maps to 0.95 * ln 10 / ln 10 = 0.95; zero maps to 0
    assert d[0, 1, 0] == pytest.approx(0.95) and d[0, 0, 2] == pytest.approx(0.95)
    assert d[0, 0, 0] == 0.0 and (d >= 0).all() and (d <= 1).all()
    # monotone in L
    ramp = np.linspace(0, 5, 50)[None, :, None].repeat(3, axis=2)
    assert (np.diff(post.drago_tone(ramp)[0, :, 0]) > 0).all()
    f = post.resolve(np.arange(12, dtype=np.float64).reshape(2, 2, 3), 2.0, flip=True)
    assert np.array_equal(f[0, 0], np.array([9, 10, 11]) / 2.0)


def _read_pnm(path):
    with open(path, "rb") as f:
        data = f.read()
    parts = data.split(b" ", 4)
    return parts[0], int(parts[1]), int(parts[2]), parts[3], parts[4]


def test_image_files_round_trip(ilm):
    rng = np.random.default_rng(0)
    img = rng.uniform(-0.1, 1.2, (5, 7, 3)).astype(np.float32)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "a.pfm")
        ilm.image_write(p, img, ilm.ILM_IMAGE_PFM)
        magic, w, h, scale, body = _read_pnm(p)
        assert (magic, w, h, scale) == (b"PF", 7, 5, b"-1")
        assert np.array_equal(np.frombuffer(body, "<f4").reshape(5, 7, 3), img)
        p = os.path.join(d, "a.ppm")
        ilm.image_write(p, img, ilm.ILM_IMAGE_PPM)
        magic, w, h, mx, body = _read_pnm(p)
        assert (magic, w, h, mx) == (b"P6", 7, 5, b"255")
        got = np.frombuffer(body, np.uint8).reshape(5, 7, 3)
        want = post.ppm_bytes(np.maximum(img, 0))   # negative values write 0
        assert np.array_equal(got, want)


@pytest.mark.gpu
def test_resolve_and_tone_match_oracle(gpu_ctx, ilm):
    rng = np.random.default_rng(1)
    W, H = 37, 23
    acc = rng.gamma(0.6, 3.0, (H, W, 3)) * 16.0
    ab = gpu_ctx.buffer(acc.nbytes)
    ab.upload(acc)
    ob, tb = gpu_ctx.buffer(W * H * 12), gpu_ctx.buffer(W * H * 12)
    for flip in (False, True):
        ilm.image_resolve(gpu_ctx, ab, W, H, 16, ob, flip=flip)
        gpu_ctx.sync()
        img = ob.download(np.float32, W * H * 3).reshape(H, W, 3)
        want = post.resolve(acc, 16.0, flip)
        assert np.allclose(img, want, rtol=1e-6)
    for op, fn in ((ilm.ILM_TONE_GLOBAL, post.global_tone), (ilm.ILM_TONE_DRAGO, post.drago_tone)):
        ilm.image_tonemap(gpu_ctx, ob, W, H, op, tb)
        gpu_ctx.sync()
        got = tb.download(np.float32, W * H * 3).reshape(H, W, 3)
        want = fn(img)
        assert np.abs(got - want).max() < 2e-6, op
    for b in (ab, ob, tb):
        b.free()
