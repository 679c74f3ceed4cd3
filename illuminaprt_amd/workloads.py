"""Benchmark workloads built with the library's own generators and kernels.

Intersection-only workload (1,334,720 triangles, three 4,194,304-ray sets):
  * mesh: icosphere level 8 (1,310,720 triangles) displaced radially by
    4-octave value noise of amplitude 0.1 (seed 7), plus 2000 boxes
    (12 triangles each, edges uniform in [0.01, 0.1], uniform random
    rotation, centres uniform in a radius-3 ball, seed 11);
  * primary rays: pinhole camera at (0, 0, -4.5) looking at the origin,
    fov 40 deg, 2048 x 2048. Pixels whose centre ray misses are replaced by
    jittered rays through pixels whose centre ray hits, so every primary ray
    hits geometry;
  * shadow rays: segments from the primary hit points to uniform points on a
    0.5 x 0.5 light at y = 2.5 (tmin 1e-4, tmax 1 - 1e-4, the visibility
    query's convention);
  * incoherent rays: cosine-distributed about the geometric normal at the
    primary hit points, origin offset by 1e-4 along that normal.
"""
from __future__ import annotations

import numpy as np

from . import (ILM_SPAWN_COSINE, ILM_SPAWN_SHADOW, Camera, Context, DeviceBuffer, Mesh, boxes, icosphere,
               RAY_BYTES, HIT_BYTES, HIT_DTYPE)

C2_LIGHT = np.array([-0.25, 2.5, -0.25, 0.5, 0.0, 0.0, 0.0, 0.0, 0.5], np.float32)  # corner, edge u, edge v


def intersection_scene(level=8, nboxes=2000, noise_seed=7, box_seed=11):
    sp, si = icosphere(level, 0.1, 4, noise_seed)
    bp, bi = boxes(nboxes, 0.01, 0.1, 3.0, box_seed, vertex_offset=len(sp))
    return np.concatenate([sp, bp]), np.concatenate([si, bi])


def intersection_camera():
    return Camera.make((0.0, 0.0, -4.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, 1.0)


class RaySets:
    def __init__(self, primary, shadow, incoherent, primary_hits, n):
        self.primary, self.shadow, self.incoherent, self.primary_hits, self.n = \
            primary, shadow, incoherent, primary_hits, n

    def free(self):
        for b in (self.primary, self.shadow, self.incoherent, self.primary_hits):
            b.free()


def intersection_rays(ctx: Context, mesh: Mesh, width=2048, height=2048, seed=1234) -> RaySets:
    n = width * height
    cam = intersection_camera()
    r0 = ctx.buffer(n * RAY_BYTES)
    h0 = ctx.buffer(n * HIT_BYTES)
    ctx.raygen_camera(cam, width, height, r0)
    mesh.trace_closest(r0, n, h0)
    idx = ctx.buffer(n * 4)
    nhit = ctx.compact_hits(h0, n, idx)
    if nhit == 0:
        raise RuntimeError("no primary ray hits the scene")
    prim = ctx.buffer(n * RAY_BYTES)
    ctx.raygen_camera_list(cam, width, height, idx, nhit, n, prim, seed=seed)
    hits = ctx.buffer(n * HIT_BYTES)
    mesh.trace_closest(prim, n, hits)
    h = hits.download(HIT_DTYPE, n)
    miss = np.nonzero(h["prim"] < 0)[0]
    if len(miss):
        # jittered rays that slipped past a silhouette: reuse the pixel-centre ray
        rays = prim.download(np.float32, 8 * n).reshape(n, 8)
        rays[miss] = rays[miss % nhit]
        h[miss] = h[miss % nhit]
        prim.upload(rays)
        hits.upload(h)
    shadow = ctx.buffer(n * RAY_BYTES)
    mesh.spawn(prim, hits, n, ILM_SPAWN_SHADOW, shadow, light=C2_LIGHT, seed=seed + 1)
    inco = ctx.buffer(n * RAY_BYTES)
    mesh.spawn(prim, hits, n, ILM_SPAWN_COSINE, inco, eps=1e-4, seed=seed + 2)
    ctx.sync()
    for b in (r0, h0, idx):
        b.free()
    rs = RaySets(prim, shadow, inco, hits, n)
    rs.replaced = int(len(miss))
    rs.centre_hits = nhit
    return rs
