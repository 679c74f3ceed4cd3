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


def large_scene(level=9, nboxes=50000, noise_seed=7, box_seed=11):
    """Two displaced level-9 icospheres (5,242,880 triangles each, centres at
    x = -1.2 and x = +1.2) and 50,000 boxes (600,000 triangles) in a radius-4
    ball: 11,085,760 triangles."""
    sp, si = icosphere(level, 0.1, 4, noise_seed)
    sp2, si2 = icosphere(level, 0.1, 4, noise_seed + 1)
    a = sp + np.float32([-1.2, 0.0, 0.0])
    b = sp2 + np.float32([1.2, 0.0, 0.0])
    bp, bi = boxes(nboxes, 0.01, 0.1, 4.0, box_seed, vertex_offset=2 * len(sp))
    return np.concatenate([a, b, bp]), np.concatenate([si, si2 + len(sp), bi])


def large_camera(aspect=16 / 9):
    return Camera.make((0.0, 0.3, -6.5), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 50.0, aspect)


def intersection_rays(ctx: Context, mesh: Mesh, width=2048, height=2048, seed=1234, cam=None,
                      light=None) -> RaySets:
    n = width * height
    cam = cam if cam is not None else intersection_camera()
    light = light if light is not None else C2_LIGHT
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
    mesh.spawn(prim, hits, n, ILM_SPAWN_SHADOW, shadow, light=light, seed=seed + 1)
    inco = ctx.buffer(n * RAY_BYTES)
    mesh.spawn(prim, hits, n, ILM_SPAWN_COSINE, inco, eps=1e-4, seed=seed + 2)
    ctx.sync()
    for b in (r0, h0, idx):
        b.free()
    rs = RaySets(prim, shadow, inco, hits, n)
    rs.replaced = int(len(miss))
    rs.centre_hits = nhit
    return rs


# ---------------------------------------------------------------- box room
# Plain path-tracing workload: procedural box room in the unit cube, open at
# z = 0: floor, back, left and right walls (2 triangles each), the ceiling as
# four quads around a 0.2 x 0.2 opening (8 triangles), two axis-aligned boxes
# (24 triangles), Lambertian; one emitter quad (2 triangles, Le = (17, 12, 4),
# facing down) filling the opening flush with the ceiling. The emitter is
# flush because the reference's mesh-light sampling turns the emitter normal
# toward the shaded point (Shape/TriangleMesh.cpp:150-161): an emitter hung
# below the ceiling would light the ceiling around it at grazing angles with
# an unbounded 1/d^4-like estimator. Geometric normals face the room interior
# / box exterior; vertex normals equal the face normals. Pinhole camera at
# (0.5, 0.5, -1.4) looking at (0.5, 0.5, 0.5), fov 40 deg.
ROOM_LE = (17.0, 12.0, 4.0)
ROOM_MATS = np.array([
    [0.73, 0.73, 0.73, 0, 0, 0, 1, 0],    # 0 white
    [0.63, 0.065, 0.05, 0, 0, 0, 1, 0],   # 1 red (left wall)
    [0.14, 0.45, 0.09, 0, 0, 0, 1, 0],    # 2 green (right wall)
    [0.0, 0.0, 0.0, 0, 0, 0, 1, 0],       # 3 emitter surface (black)
], np.float32)


def _quad(p0, p1, p2, p3, want_normal):
    p = np.array([p0, p1, p2, p3], np.float64)
    n = np.cross(p[1] - p[0], p[2] - p[0])
    n /= np.linalg.norm(n)
    if np.dot(n, want_normal) < 0:
        p = p[[0, 3, 2, 1]]
        n = -n
    return p, n


def _box_faces(lo, hi):
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    c = (lo + hi) / 2
    faces = []
    for axis in range(3):
        for side in (0, 1):
            v = hi[axis] if side else lo[axis]
            a1, a2 = (axis + 1) % 3, (axis + 2) % 3
            corners = []
            for (s1, s2) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                q = np.zeros(3)
                q[axis] = v
                q[a1] = hi[a1] if s1 else lo[a1]
                q[a2] = hi[a2] if s2 else lo[a2]
                corners.append(q)
            want = np.zeros(3)
            want[axis] = 1.0 if side else -1.0
            faces.append(_quad(*corners, want))
    return faces


def box_room():
    """Returns pos, nrm, idx, tri_mat, tri_light, mats, lights (42 triangles)."""
    quads = []  # (points, normal, material, light)
    quads.append((*_quad([0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1], [0, 1, 0]), 0, -1))    # floor
    for (x0, x1, z0, z1) in ((0, 1, 0, 0.4), (0, 1, 0.6, 1), (0, 0.4, 0.4, 0.6), (0.6, 1, 0.4, 0.6)):  # ceiling
        quads.append((*_quad([x0, 1, z0], [x1, 1, z0], [x1, 1, z1], [x0, 1, z1], [0, -1, 0]), 0, -1))
    quads.append((*_quad([0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1], [0, 0, -1]), 0, -1))   # back
    quads.append((*_quad([0, 0, 0], [0, 1, 0], [0, 1, 1], [0, 0, 1], [1, 0, 0]), 1, -1))    # left
    quads.append((*_quad([1, 0, 0], [1, 1, 0], [1, 1, 1], [1, 0, 1], [-1, 0, 0]), 2, -1))   # right
    for lo, hi in (([0.13, 0.0, 0.15], [0.43, 0.30, 0.45]), ([0.55, 0.0, 0.45], [0.85, 0.60, 0.75])):
        for p, n in _box_faces(lo, hi):
            quads.append((p, n, 0, -1))
    quads.append((*_quad([0.4, 1, 0.4], [0.6, 1, 0.4], [0.6, 1, 0.6], [0.4, 1, 0.6], [0, -1, 0]), 3, 0))
    pos, nrm, idx, tm, tl = [], [], [], [], []
    for p, n, m, l in quads:
        b = len(pos)
        pos += list(p)
        nrm += [n] * 4
        idx += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
        tm += [m, m]
        tl += [l, l]
    lights = np.array([[*ROOM_LE, 0.0]], np.float32)
    return (np.array(pos, np.float32), np.array(nrm, np.float32), np.array(idx, np.int32),
            np.array(tm, np.int32), np.array(tl, np.int32), ROOM_MATS.copy(), lights)


def box_room_reference(light_y=0.999):
    """The box room as the path-tracing workload specifies it, lit the
    reference's way: 5 walls (floor, ceiling, back, left, right; 10
    triangles) and the 2 boxes (24 triangles), 34 triangles, plus one ceiling
    quad light (2 triangles, Le = (17, 12, 4)) hung just below the ceiling
    and sampled as the reference's Quad (light mode 1): its triangles are
    (v1, v3, v2), (v1, v4, v3) (Shape/Quad.cpp:11-31) with the winding
    normal cross(v3 - v1, v2 - v1) facing down. Returns pos, nrm, idx,
    tri_mat, tri_light, mats, lights (36 triangles)."""
    quads = []
    quads.append((*_quad([0, 0, 0], [1, 0, 0], [1, 0, 1], [0, 0, 1], [0, 1, 0]), 0))    # floor
    quads.append((*_quad([0, 1, 0], [1, 1, 0], [1, 1, 1], [0, 1, 1], [0, -1, 0]), 0))   # ceiling
    quads.append((*_quad([0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1], [0, 0, -1]), 0))   # back
    quads.append((*_quad([0, 0, 0], [0, 1, 0], [0, 1, 1], [0, 0, 1], [1, 0, 0]), 1))    # left
    quads.append((*_quad([1, 0, 0], [1, 1, 0], [1, 1, 1], [1, 0, 1], [-1, 0, 0]), 2))   # right
    for lo, hi in (([0.13, 0.0, 0.15], [0.43, 0.30, 0.45]), ([0.55, 0.0, 0.45], [0.85, 0.60, 0.75])):
        for p, n in _box_faces(lo, hi):
            quads.append((p, n, 0))
    pos, nrm, idx, tm, tl = [], [], [], [], []
    for p, n, m in quads:
        b = len(pos)
        pos += list(p)
        nrm += [n] * 4
        idx += [[b, b + 1, b + 2], [b, b + 2, b + 3]]
        tm += [m, m]
        tl += [-1, -1]
    y = light_y
    v1, v2, v3, v4 = [0.4, y, 0.4], [0.4, y, 0.6], [0.6, y, 0.6], [0.6, y, 0.4]
    b = len(pos)
    pos += [v1, v2, v3, v4]
    nrm += [[0, -1, 0]] * 4
    idx += [[b, b + 2, b + 1], [b, b + 3, b + 2]]   # (v1, v3, v2), (v1, v4, v3)
    tm += [3, 3]
    tl += [0, 0]
    lights = np.array([[*ROOM_LE, 1.0]], np.float32)
    return (np.array(pos, np.float32), np.array(nrm, np.float32), np.array(idx, np.int32),
            np.array(tm, np.int32), np.array(tl, np.int32), ROOM_MATS.copy(), lights)


def room_camera():
    return Camera.make((0.5, 0.5, -1.4), (0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 40.0, 1.0)


def rel_rmse(img, ref):
    """sqrt(mean over pixels and channels of (I - R)^2 / (R^2 + 0.01))."""
    img, ref = np.asarray(img, np.float64), np.asarray(ref, np.float64)
    return float(np.sqrt(np.mean((img - ref) ** 2 / (ref ** 2 + 0.01))))


# ---------------------------------------------------------------- radiance-cache workload
# The displaced level-8 icosphere of the intersection workload (1,310,720
# triangles) inside a closed 12-triangle room (x, z in [-3, 3], y in
# [-1.2, 3]), lit by three emitter panels hanging from the ceiling: thin
# boxes (12 triangles) whose underside (2 triangles) emits. A panel's sides
# shadow the ceiling next to its emitting face, so no surface sees an
# emitter at grazing angles. Sphere triangles are glossy (Lambertian + Phong
# lobe, exponent 60) with probability 0.3 (seeded per triangle), else
# Lambertian. Camera inside the room, 16:9, fov 60 deg.
CACHE_MATS = np.array([
    [0.70, 0.70, 0.70, 0, 0, 0, 1, 0],        # 0 room white
    [0.60, 0.20, 0.15, 0, 0, 0, 1, 0],        # 1 left wall
    [0.15, 0.25, 0.60, 0, 0, 0, 1, 0],        # 2 right wall
    [0.60, 0.55, 0.50, 0, 0, 0, 1, 0],        # 3 sphere, diffuse
    [0.30, 0.28, 0.25, 0.45, 0.45, 0.45, 60, 1],  # 4 sphere, glossy
    [0.0, 0.0, 0.0, 0, 0, 0, 1, 0],           # 5 panel body / emitter surface
], np.float32)
CACHE_LIGHTS = np.array([[14, 12, 9, 0], [8, 10, 15, 0], [12, 12, 12, 0]], np.float32)
CACHE_PANELS = [((-1.6, 2.9, -1.0), (-0.8, 3.0, -0.2)), ((0.6, 2.9, 0.4), (1.4, 3.0, 1.2)),
                ((-0.3, 2.9, 1.8), (0.3, 3.0, 2.4))]


def cache_scene(level=8, glossy_fraction=0.3, seed=21):
    """Returns pos, nrm, idx, tri_mat, tri_light, mats, lights."""
    from . import vertex_normals
    sp, si = icosphere(level, 0.1, 4, 7)
    sn = vertex_normals(sp, si)
    rng = np.random.default_rng(seed)
    smat = np.where(rng.random(len(si)) < glossy_fraction, 4, 3).astype(np.int32)
    quads = []
    lo, hi = np.array([-3.0, -1.2, -3.0]), np.array([3.0, 3.0, 3.0])
    c = (lo + hi) / 2
    for p, n in _box_faces(lo, hi):
        n = -n                                     # room faces look inward
        p = p[[0, 3, 2, 1]]
        m = 1 if n[0] > 0.5 else (2 if n[0] < -0.5 else 0)
        quads.append((p, n, m, -1))
    for li, (plo, phi) in enumerate(CACHE_PANELS):
        for p, n in _box_faces(plo, phi):
            emit = n[1] < -0.5                     # the underside emits
            quads.append((p, n, 5, li if emit else -1))
    pos, nrm, idx, tm, tl = [sp], [sn], [si], [smat], [np.full(len(si), -1, np.int32)]
    base = len(sp)
    for p, n, m, l in quads:
        pos.append(np.array(p, np.float32))
        nrm.append(np.tile(np.array(n, np.float32), (4, 1)))
        idx.append(np.array([[base, base + 1, base + 2], [base, base + 2, base + 3]], np.int32))
        tm.append(np.array([m, m], np.int32))
        tl.append(np.array([l, l], np.int32))
        base += 4
    return (np.concatenate(pos).astype(np.float32), np.concatenate(nrm).astype(np.float32),
            np.concatenate(idx), np.concatenate(tm), np.concatenate(tl), CACHE_MATS.copy(), CACHE_LIGHTS.copy())


CACHE_BOUNDS = ((-3.0, -1.2, -3.0), (3.0, 3.0, 3.0))


def cache_camera(aspect=16 / 9):
    return Camera.make((0.0, 0.6, -2.8), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, aspect)


# ---------------------------------------------------------------- batched scenes
# Batched path-tracing workload: n box rooms (the plain path-tracing room),
# each holding a displaced icosphere (levels 3-8) and small boxes so that its
# triangle count is log-uniform in [1k, 2M]; object materials are 70%
# Lambertian and 30% Lambertian + Phong with exponent log-uniform in
# [10, 1000]; each room is seen by its own camera at a square resolution
# uniform in [128, 1024]; 32 spp, depth 8.
def batch_scene(rng):
    """One room of the batched workload: (arrays for Scene, camera, side, target triangles)."""
    from . import vertex_normals
    target = int(np.exp(rng.uniform(np.log(1e3), np.log(2e6))))
    pos, nrm, idx, tm, tl, mats, lights = box_room()
    nbase = len(ROOM_MATS)
    nobj = 8
    omats = []
    for _ in range(nobj):
        kd = rng.uniform(0.2, 0.8, 3)
        if rng.random() < 0.3:
            ks = rng.uniform(0.1, 0.4, 3)
            omats.append([*(kd * 0.6), *ks, float(np.exp(rng.uniform(np.log(10), np.log(1000)))), 1])
        else:
            omats.append([*kd, 0, 0, 0, 1, 0])
    mats = np.concatenate([mats, np.array(omats, np.float32)])
    room_tris = len(idx)
    level = int(np.clip(np.floor(np.log(max(target - room_tris, 20) / 20.0) / np.log(4.0)), 3, 8))
    sp, si = icosphere(level, 0.1, 4, int(rng.integers(1 << 30)))
    sp = (sp * 0.16 + np.float32([0.5, 0.25, 0.55])).astype(np.float32)
    sn = vertex_normals(sp, si)
    parts_p, parts_n, parts_i = [pos, sp], [nrm, sn], [idx, si + len(pos)]
    parts_m = [tm, np.full(len(si), nbase + int(rng.integers(nobj)), np.int32)]
    parts_l = [tl, np.full(len(si), -1, np.int32)]
    nb = max(0, (target - room_tris - len(si)) // 12)
    if nb:
        bp, bi = boxes(nb, 0.004, 0.03, 0.3, int(rng.integers(1 << 30)))
        bp = (bp + np.float32([0.5, 0.45, 0.5])).astype(np.float32)
        bn = vertex_normals(bp, bi)
        off = len(pos) + len(sp)
        parts_p.append(bp)
        parts_n.append(bn)
        parts_i.append(bi + off)
        parts_m.append((nbase + rng.integers(nobj, size=nb).repeat(12)).astype(np.int32))
        parts_l.append(np.full(len(bi), -1, np.int32))
    arrays = (np.concatenate(parts_p).astype(np.float32), np.concatenate(parts_n).astype(np.float32),
              np.concatenate(parts_i).astype(np.int32), np.concatenate(parts_m).astype(np.int32),
              np.concatenate(parts_l).astype(np.int32), mats, lights)
    side = int(rng.integers(128, 1025))
    return arrays, room_camera(), side, target


# ---------------------------------------------------------------- large rendered scene
# The 11M-triangle scene made renderable: the two displaced level-9 spheres and
# 50,000 boxes of large_scene() inside the radiance-cache room (scaled by 1.6)
# with its three panel lights; 30% of the sphere triangles and of the boxes are
# glossy. Eight cameras on a circle around the spheres look at the origin.
LARGE_BOUNDS = ((-4.8, -1.92, -4.8), (4.8, 4.8, 4.8))


def large_render_scene(glossy_fraction=0.3, seed=31):
    from . import vertex_normals
    rng = np.random.default_rng(seed)
    pos, idx = large_scene()
    nt_obj = len(idx)
    nrm = vertex_normals(pos, idx)
    tm = np.where(rng.random(nt_obj) < glossy_fraction, 4, 3).astype(np.int32)
    rp, rn, ri, rm, rl, mats, lights = cache_scene(level=3)
    nsph = 20 * 4 ** 3
    rp, rn, ri, rm, rl = rp[642:] * 1.6, rn[642:], ri[nsph:] - 642, rm[nsph:], rl[nsph:]   # the room and panels only
    off = len(pos)
    return (np.concatenate([pos, rp]).astype(np.float32), np.concatenate([nrm, rn]).astype(np.float32),
            np.concatenate([idx, ri + off]).astype(np.int32), np.concatenate([tm, rm]).astype(np.int32),
            np.concatenate([np.full(nt_obj, -1, np.int32), rl]).astype(np.int32), mats, lights)


def large_cameras(n=8, aspect=16 / 9):
    cams = []
    for k in range(n):
        a = 2 * np.pi * (k + 0.5) / n
        cams.append(Camera.make((4.0 * np.sin(a), 0.8, -4.0 * np.cos(a)), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0),
                                55.0, aspect))
    return cams


# ---------------------------------------------------------------- instanced scene
# The flat instance table's workload: 2,048 instances of three meshes (a
# displaced level-5 icosphere of 20,480 triangles, a cluster of 200 boxes of
# 2,400 triangles, a level-3 icosphere of 1,280 triangles) with uniformly
# random rotations, scalings in [0.2, 0.6] and positions in a 16 x 16 x 16
# cube; a camera outside the cube looks at its centre.
def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def instanced_scene(n_inst=2048, seed=13):
    """(meshes [(pos, idx)], placements [(mesh, rot9, scl3, trn3)])."""
    rng = np.random.default_rng(seed)
    meshes = [icosphere(5, 0.1, 4, 3), boxes(200, 0.05, 0.3, 1.0, 5), icosphere(3, 0.0, 1, 0)]
    placements = []
    for _ in range(n_inst):
        m = int(rng.integers(len(meshes)))
        placements.append((m, _rotation(rng).astype(np.float32).reshape(-1),
                           rng.uniform(0.2, 0.6, 3).astype(np.float32), rng.uniform(-8, 8, 3).astype(np.float32)))
    return meshes, placements


def flatten_placements(meshes, placements):
    """World-space positions and indices of every instance (the space's
    flattening: world_i = sum_j rot[3j+i] scl_j p_j + trn_i, in float64)."""
    ps, ix, off = [], [], 0
    for m, rot, scl, trn in placements:
        pos, idx = meshes[m]
        p = (pos.astype(np.float64) * scl.astype(np.float64)) @ rot.astype(np.float64).reshape(3, 3) + trn
        ps.append(p.astype(np.float32))
        ix.append(idx + off)
        off += len(pos)
    return np.concatenate(ps), np.concatenate(ix).astype(np.int32)


def instanced_camera():
    return Camera.make((0.0, 3.0, -24.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 45.0, 1.0)
