"""Batches of scenes (ilm_batch_*): several scenes rendered by one frame's
launch sequence give the same images as rendering each scene alone.

The first view of a batch uses the same per-pixel random streams as a
single-scene frame (its pixels come first), so its image must be identical.
Later views use other streams, so they are compared statistically: the
closed-form furnace mean, and relRMSE against a high-spp single-scene render
at the noise level of an equal-spp single-scene render."""
import numpy as np
import pytest

from tests.test_gpu_pt import _closed_box

pytestmark = pytest.mark.gpu


def _scene(ilm, ctx, arrays):
    pos, nrm, idx, tm, tl, mats, lights = arrays
    m = ilm.Mesh(ctx, pos, idx)
    m.build()
    return m, ilm.Scene(ctx, m, nrm, tm, tl, mats, lights)


def _single(ilm, ctx, sc, cam, W, H, spp, seed, max_depth=5):
    acc = ctx.buffer(W * H * 24).zero()
    sc.render(cam, W, H, spp, acc, seed=seed, max_depth=max_depth)
    img = acc.download(np.float64, W * H * 3).reshape(H, W, 3) / spp
    acc.free()
    return img


def test_batch_views_match_single_scenes(gpu_ctx, ilm):
    from illuminaprt_amd import workloads as w
    room = w.box_room()
    glossy = list(w.box_room())
    glossy[5] = np.concatenate([glossy[5], np.array([[0.4, 0.4, 0.4, 0.3, 0.3, 0.3, 60.0, 1]], np.float32)])
    glossy[3] = glossy[3].copy()
    glossy[3][10:34] = 4
    rho, le = 0.5, 1.0
    furnace = _closed_box(rho, le)
    objs = [_scene(ilm, gpu_ctx, a) for a in (room, glossy, furnace)]
    scenes = [sc for _, sc in objs]
    cam_room = w.room_camera()
    cam_box = ilm.Camera.make((0.5, 0.5, 0.1), (0.5, 0.5, 1.0), (0, 1, 0), 60.0, 1.0)
    # the glossy view's mean is dominated by its highlights' variance: 32 spp
    # keeps the mean comparison below well clear of the noise
    views = [(0, cam_room, 64, 64, 8), (1, cam_room, 48, 40, 32), (2, cam_box, 32, 32, 128), (0, cam_room, 40, 64, 8)]
    b = ilm.Batch(gpu_ctx, scenes)
    npix = sum(v[2] * v[3] for v in views)
    acc = gpu_ctx.buffer(npix * 24).zero()
    st = b.render(views, acc, seed=77, batch_paths=20000)   # several wavefront batches spanning views
    out = acc.download(np.float64, npix * 3)
    imgs, off = [], 0
    for (s, cam, W, H, spp) in views:
        imgs.append(out[off:off + W * H * 3].reshape(H, W, 3) / spp)
        off += W * H * 3
    assert st["camera_rays"] == sum(v[2] * v[3] * v[4] for v in views)
    # first view: the same streams as the scene alone
    alone = _single(ilm, gpu_ctx, scenes[0], cam_room, 64, 64, 8, seed=77)
    assert np.array_equal(imgs[0], alone)
    # furnace view: its closed form (only its own six lights are sampled).
    # The image mean's standard deviation at 32 x 32 x 128 paths is about
    # 0.3% (0.8% at 16 spp, measured on the oracle over 12 seeds): 1.2% is 4 sigma.
    want = le * sum(rho ** k for k in range(6))
    print(f"[batch furnace] {imgs[2].mean():.5f} vs {want:.5f}", flush=True)
    assert abs(imgs[2].mean() - want) < 0.012 * want
    # the other room views: the noise level of equal-spp single renders
    from illuminaprt_amd import workloads as ww
    for k, (s, cam, W, H, spp) in enumerate(views):
        if s == 2 or k == 0:
            continue
        ref = _single(ilm, gpu_ctx, scenes[s], cam, W, H, 256, seed=5)
        same = _single(ilm, gpu_ctx, scenes[s], cam, W, H, spp, seed=6)
        e_b, e_s = ww.rel_rmse(imgs[k], ref), ww.rel_rmse(same, ref)
        mb, ms = imgs[k].mean(), same.mean()
        print(f"[batch view {k}] relRMSE {e_b:.4f} vs single {e_s:.4f}; mean {mb:.4f} vs {ms:.4f}", flush=True)
        assert e_b < 1.3 * e_s and abs(mb - ms) < 0.03 * ms
    # graph replays of the batch frame equal the eager frame exactly (a
    # vertex's shadow contributions are summed by one thread in a fixed order)
    b.set_graph(True)
    for rep in range(3):
        acc.zero()
        b.render(views, acc, seed=77, batch_paths=20000)
        assert np.array_equal(acc.download(np.float64, npix * 3), out), rep
    b.set_graph(False)
    # the same frame over every tree: BVH2 and BVH8 regroup the BVH4's nodes
    # over the same leaves and quantised boxes only grow, so the images agree
    # up to paths sent another way at exact ties
    with pytest.raises(ilm.IlmError):
        b.configure(ilm.ILM_NODES_QUANTIZED, width=8)
    for f, wd in ((ilm.ILM_NODES_FLOAT, 2), (ilm.ILM_NODES_FLOAT, 8), (ilm.ILM_NODES_QUANTIZED, 4),
                  (ilm.ILM_NODES_FLOAT, 4)):
        b.configure(f, width=wd)
        acc.zero()
        b.render(views, acc, seed=77, batch_paths=20000)
        img = acc.download(np.float64, npix * 3)
        rel = np.abs(img - out).sum() / out.sum()
        print(f"[batch tree {f}/{wd}] mean |diff| / mean {rel:.2e}", flush=True)
        assert rel < 2e-3
    # the autotuner returns a table over trees (node format x width) and batch
    # sizes and keeps the fastest
    t = b.autotune(views)
    print(f"[batch autotune] {t}", flush=True)
    assert len(t["table"]) == 16 and t["best_ms"] == min(ms for *_, ms in t["table"])
    assert {bp for _, _, bp, _ in t["table"]} == {1 << 20, 1 << 22, 1 << 24, 1 << 26}
    assert {(f, wd) for f, wd, _, _ in t["table"]} == {(ilm.ILM_NODES_FLOAT, 4), (ilm.ILM_NODES_QUANTIZED, 4),
                                                       (ilm.ILM_NODES_FLOAT, 2), (ilm.ILM_NODES_FLOAT, 8)}
    acc.zero()
    b.render(views, acc, seed=77)
    tuned = acc.download(np.float64, npix * 3)
    # the same first view up to paths that a different traversal order sends
    # another way at exact ties (quantised boxes reorder children)
    a0, b0 = tuned[:64 * 64 * 3].reshape(64, 64, 3) / 8, out[:64 * 64 * 3].reshape(64, 64, 3) / 8
    assert ww.rel_rmse(a0, b0) < 0.02
    acc.free()
    b.close()
    for m, sc in objs:
        sc.close()
        m.close()
