"""Pins the CPU oracle (the C++ restatement of the reference hot path).

Each test checks the oracle against something independent of its own code: a
closed form, a textbook or library routine, an invariant, or brute force.
The reference ships no tests, golden vectors or fixtures for this path, and it
cannot be compiled here (boost is absent), so these are the pins.
"""
import math
import random

import numpy as np
import pytest

from tests import meshgen

F32MAX = float(np.finfo(np.float32).max)


# ---------------------------------------------------------------- RNGs
@pytest.mark.parametrize("ctr,key", [
    ([0, 0, 0, 0], [0, 0]),
    ([5, 17, 3, 9], [7, 9]),
    ([2 ** 64 - 2, 2 ** 63, 1, 2 ** 40], [0xDEADBEEF, 0x12345678ABCDEF]),
])
def test_philox_matches_numpy(orc, ctr, key):
    # numpy.random.Philox (4x64-10) increments the counter before each block.
    bg = np.random.Philox(key=np.array(key, np.uint64), counter=np.array(ctr, np.uint64))
    want = bg.random_raw(4)
    c = list(ctr)
    c[0] = (c[0] + 1) % 2 ** 64
    if c[0] == 0:
        c[1] = (c[1] + 1) % 2 ** 64
    got = orc.philox4x64(np.array(c, np.uint64), np.array(key, np.uint64))
    assert [int(x) for x in got] == [int(x) for x in want]


@pytest.mark.parametrize("seed", [5489, 1234, 0xFFFFFFFF])
def test_mt19937_matches_python_random(orc, seed):
    # Random::Seed fills s[i] = 69069*s[i-1] (Random.cpp:36-43); Python's random
    # module is a textbook MT19937 and accepts that 624-word state directly.
    s = [seed & 0xFFFFFFFF]
    for _ in range(623):
        s.append((69069 * s[-1]) & 0xFFFFFFFF)
    r = random.Random()
    r.setstate((3, tuple(s) + (624,), None))
    # The reference's last tempering step is y ^= y >> 15 (Random.cpp:87), not
    # the textbook y ^= y >> 18: undo the textbook step, apply the reference's.
    want = []
    for _ in range(2000):
        y = r.getrandbits(32)
        z = y ^ (y >> 18)
        want.append(z ^ (z >> 15))
    got = orc.mt_words(seed, 2000)
    assert got.tolist() == want


def test_nextfloat_range(orc):
    f = orc.mt_floats(1, 200000)
    assert f.min() >= 0.0 and f.max() < 1.0
    assert abs(f.mean() - 0.5) < 3e-3


# ---------------------------------------------------------------- triangle
TRI = ([0, 0, 0], [1, 0, 0], [0, 1, 0])


@pytest.mark.parametrize("b,g,h", [(0.25, 0.25, 1.0), (0.125, 0.5, 2.0), (0.0, 0.5, 1.0), (0.5, 0.0, 1.0),
                                   (0.5, 0.5, 1.0), (0.0, 0.0, 0.5)])
def test_triangle_known_answers(orc, b, g, h):
    # Point a*p0 + b*p1 + g*p2 = (b, g, 0); ray straight down from height h.
    hit, t, beta, gamma = orc.tri_intersect(*TRI, [b, g, h], [0, 0, -1])
    assert hit
    assert t == h and beta == b and gamma == g  # exact in binary


def test_triangle_outside_and_back_face(orc):
    assert not orc.tri_intersect(*TRI, [0.6, 0.6, 1], [0, 0, -1])[0]      # beta+gamma > 1
    assert not orc.tri_intersect(*TRI, [-0.01, 0.5, 1], [0, 0, -1])[0]   # beta < 0
    assert orc.tri_intersect(*TRI, [0.25, 0.25, -1], [0, 0, 1])[0]       # two-sided


def test_triangle_epsilon_window(orc):
    f = np.float32
    # t >= 1e-5 (double) accepted; below rejected. Ray.Min is not consulted.
    assert not orc.tri_intersect(*TRI, [0.25, 0.25, f(0.9e-5)], [0, 0, -1])[0]
    assert orc.tri_intersect(*TRI, [0.25, 0.25, f(1.1e-5)], [0, 0, -1])[0]
    # t > Max - 1e-5 rejected
    assert not orc.tri_intersect(*TRI, [0.25, 0.25, 1.0], [0, 0, -1], ray_max=f(1.0 + 5e-6))[0]
    assert orc.tri_intersect(*TRI, [0.25, 0.25, 1.0], [0, 0, -1], ray_max=f(1.0 + 2e-5))[0]


def test_triangle_parallel_and_degenerate(orc):
    assert not orc.tri_intersect(*TRI, [0.25, 0.25, 1.0], [1, 0, 0])[0]
    line = ([0, 0, 0], [1, 0, 0], [2, 0, 0])
    assert not orc.tri_intersect(*line, [0.5, 0.3, 1.0], [0.2, -0.25, -1])[0]
    # Reference quirk: for a zero-area triangle n = 0, so rcp = inf and t = 0*inf
    # = NaN, which passes every comparison (IndexedTriangle.cpp:114-128). When the
    # ray direction is perpendicular to the triangle's line, beta and gamma are
    # NaN too and the test reports a hit with t = NaN. The GPU path rejects
    # det == 0 instead; the benchmark scenes hold no degenerate triangles.
    hit, t, _, _ = orc.tri_intersect(*line, [0.5, 0.3, 1.0], [0, -0.3, -1])
    assert hit and math.isnan(t)


def test_ray_min_is_ignored(orc):
    p, i = meshgen.quad([0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0])
    m = orc.Mesh(p, i)
    r = np.array([[0.5, 0.5, 1.0, 5.0, 0, 0, -1, F32MAX]], np.float32)  # tmin = 5 > t = 1
    prim, t = m.trace(r, accel=0)
    assert prim[0] in (0, 1) and t[0] == 1.0


# ---------------------------------------------------------------- AABB
def test_aabb_known_answers(orc):
    lo, hi = [0, 0, 0], [1, 1, 1]
    assert orc.aabb_intersect(lo, hi, [0.5, 0.5, -1], [0, 0, 1]) == (True, 1.0, 2.0)
    assert not orc.aabb_intersect(lo, hi, [0.5, 0.5, 2], [0, 0, 1])[0]          # behind
    assert not orc.aabb_intersect(lo, hi, [0.5, 0.5, -1], [0, 0, 1], 0, 0.5)[0]  # beyond Max
    hit, a, b = orc.aabb_intersect(lo, hi, [0.5, 0.5, 0.5], [0, 0, 1])          # inside
    assert hit and a == -0.5 and b == 0.5
    # origin on a slab plane with a zero direction component: (0-0)*inf = NaN.
    # Min(NaN, inf) returns inf by argument order, so the entry distance becomes
    # inf and the box is missed (BoundingBox.cpp:216-226, Maths.h:22-30).
    assert not orc.aabb_intersect(lo, hi, [0.0, 0.5, -1], [0, 0, 1])[0]
    hit, a, b = orc.aabb_intersect(lo, hi, [0.25, 0.5, -1], [0.5, 0, 1])
    assert hit and a == 1.0 and b == 1.5  # exits through x = 1 at t = 1.5


def zero_direction_fans():
    """Two fans of 16 thin triangles in the plane z = 1, one on each side of
    x = 0, sharing edges on that plane, and 16 rays along the plane
    (origin x = 0, direction (0, 0, 1)) that hit those edges at t = 2."""
    P, idx = [], []
    for s in (-1.0, 1.0):
        for k in range(16):
            y0, y1 = k / 64, (k + 1) / 64
            b = len(P)
            P += [[0.0, y0, 1.0], [s * 0.5, y0, 1.0], [0.0, y1, 1.0]]
            idx.append([b, b + 1, b + 2])
    rays = np.zeros((16, 8), np.float32)
    rays[:, 1] = (np.arange(16) + 0.3) / 64
    rays[:, 2] = -1.0
    rays[:, 6] = 1.0
    rays[:, 7] = F32MAX
    return np.array(P, np.float32), np.array(idx, np.int32), rays


def test_bvhmesh_loses_hits_along_a_box_face(orc):
    """The reference's BVHMesh misses what its BasicMesh hits when a ray with
    a zero direction component runs in the plane of a node's box face: the
    first split of the fans is the plane x = 0, both children's boxes have a
    face there, and (0 - 0) * inf = NaN makes the slab test reject both
    (BoundingBox.cpp:212-232, Maths.h:22-30). Moving the rays off the plane
    by 1e-3 restores the hits."""
    P, idx, rays = zero_direction_fans()
    m = orc.Mesh(P, idx)
    m.build(10, 35)
    pb, tb = m.trace(rays, 0, 0)
    assert (pb >= 0).all() and (pb < 16).all() and (tb == 2.0).all()   # BasicMesh: the first fan (index order)
    pv, _ = m.trace(rays, 0, 1)
    assert (pv == -1).all()                                             # BVHMesh: lost
    off = rays.copy()
    off[:, 0] = 1e-3
    po, to = m.trace(off, 0, 1)
    assert (po >= 16).all() and (to == 2.0).all()


# ---------------------------------------------------------------- warps
def test_cosine_hemisphere_moments(orc):
    rng = np.random.default_rng(0)
    uv = rng.random((20000, 2)).astype(np.float32)
    w = np.array([orc.cosine_hemisphere(u, v) for u, v in uv])
    assert np.allclose(np.linalg.norm(w, axis=1), 1, atol=1e-6)
    assert abs(w[:, 2].mean() - 2 / 3) < 0.01            # E[cos] = 2/3
    assert abs((w[:, 2] ** 2).mean() - 0.5) < 0.01       # E[cos^2] = 1/2
    assert abs(w[:, 0].mean()) < 0.01 and abs(w[:, 1].mean()) < 0.01


def test_uniform_triangle_is_uniform(orc):
    rng = np.random.default_rng(1)
    uv = rng.random((30000, 2)).astype(np.float32)
    b = np.array([orc.uniform_triangle(u, v) for u, v in uv])
    b3 = 1 - b[:, 0] - b[:, 1]
    assert (b >= 0).all() and (b3 >= -1e-6).all()
    assert abs(b[:, 0].mean() - 1 / 3) < 0.01 and abs(b[:, 1].mean() - 1 / 3) < 0.01
    # chi-square over the 4 sub-triangles of the midpoint subdivision (equal areas)
    cell = np.where(b[:, 0] > 0.5, 0, np.where(b[:, 1] > 0.5, 1, np.where(b3 > 0.5, 2, 3)))
    cnt = np.bincount(cell, minlength=4)
    e = len(cell) / 4
    assert ((cnt - e) ** 2 / e).sum() < 16.3  # p = 0.001, 3 dof


# ---------------------------------------------------------------- camera
def test_camera_centre_and_corner(orc):
    cam = [0.5, 0.5, -1.4, 0.5, 0.5, 0.5, 0, 1, 0, 40.0, 1.0]
    o, d = orc.camera_ray(cam, 256, 256, 128, 128, 0.5, 0.5)
    assert np.allclose(o, [0.5, 0.5, -1.4]) and np.allclose(d, [0, 0, 1], atol=1e-7)
    # pixel (0,0) with sample (0.5,0.5) goes through the frame corner u0,v0 = -tan(20deg)
    _, d = orc.camera_ray(cam, 256, 256, 0, 0, 0.5, 0.5)
    t = math.tan(math.radians(20))
    want = np.array([-t, -t, 1.0]) / math.sqrt(2 * t * t + 1)
    assert np.allclose(d, want, atol=1e-6)
    # the box filter shifts samples by -0.5: sample (0,0) of pixel (1,1) equals
    # sample (1,1) of pixel (0,0) up to rounding
    _, d1 = orc.camera_ray(cam, 256, 256, 1, 1, 0.0, 0.0)
    _, d0 = orc.camera_ray(cam, 256, 256, 0, 0, 1.0, 1.0)
    assert np.allclose(d0, d1, atol=1e-6)


# ---------------------------------------------------------------- BVHMesh
def _check_bvh_invariants(m, pos, idx, max_leaf, max_depth):
    box, child, lf, lc, refs = m.export()
    n = len(box)
    tri = pos[idx]
    tmin, tmax = tri.min(axis=1), tri.max(axis=1)
    seen = np.zeros(len(idx), np.int64)
    for i in range(n):
        if child[i, 0] < 0:
            r = refs[lf[i]: lf[i] + lc[i]]
            seen[r] += 1
            if lc[i]:
                assert (box[i, :3] <= tmin[r]).all() and (box[i, 3:] >= tmax[r]).all()
        else:
            for c in child[i]:
                if lc[c] > 0 or child[c, 0] >= 0:
                    assert (box[i, :3] <= box[c, :3]).all() and (box[i, 3:] >= box[c, 3:]).all()
    assert (seen == 1).all()  # DistributeUnique: every triangle in exactly one leaf
    st = m.stats()
    assert st["refs"] == len(idx)
    assert st["internal"] + st["leaves"] == n and st["leaves"] == st["internal"] + 1
    assert st["max_depth"] <= max_depth


@pytest.mark.parametrize("kind", ["sphere", "soup"])
def test_bvhmesh_invariants(orc, kind):
    if kind == "sphere":
        pos, idx = meshgen.icosphere(4)
    else:
        pos, idx = meshgen.random_triangles(3000, seed=3)
    m = orc.Mesh(pos, idx)
    m.build(10, 35)
    _check_bvh_invariants(m, pos, idx, 10, 35)
    st = m.stats()
    # leaves hold <= 10 triangles unless the radius stop fired first
    assert st["max_leaf"] <= 10 or kind == "soup"


@pytest.mark.parametrize("mode", [0, 1])
def test_bvhmesh_traversal_equals_brute_force(orc, mode):
    sp, si = meshgen.icosphere(4)                          # 5120 triangles
    tp, ti = meshgen.random_triangles(2000, seed=5, extent=1.5)
    pos, idx = meshgen.merge([(sp, si), (tp, ti)])
    m = orc.Mesh(pos, idx)
    m.build(10, 35)
    rays = np.concatenate([meshgen.rays_toward(6000, pos, seed=7), meshgen.random_rays(4000, seed=8)])
    pb, tb = m.trace(rays, mode=mode, accel=0)
    pv, tv = m.trace(rays, mode=mode, accel=1)
    if mode == 1:
        assert ((pb >= 0) == (pv >= 0)).all()
        return
    assert ((pb >= 0) == (pv >= 0)).all()
    diff = pb != pv
    # first-found-wins within the 1e-5 absolute slop is the only allowed difference
    assert (np.abs(tb[diff] - tv[diff]) <= 1e-5).all()
    assert diff.mean() < 1e-3
    assert (pb >= 0).mean() > 0.5


# ---------------------------------------------------------------- kd-trees
def _kd_scene():
    sp, si = meshgen.icosphere(4)
    tp, ti = meshgen.random_triangles(2000, seed=5, extent=1.5)
    return meshgen.merge([(sp, si), (tp, ti)])


@pytest.mark.parametrize("kind", [1, 2])
def test_kdtree_invariants(orc, kind):
    pos, idx = _kd_scene()
    m = orc.Mesh(pos, idx)
    m.kd_build(kind, 10, 35)
    box, child, split, lf, lc, refs = m.kd_export()
    n = len(child)
    tb_lo = pos[idx].min(axis=1)
    tb_hi = pos[idx].max(axis=1)
    seen = np.zeros(len(idx), np.int64)
    depth = np.zeros(n, np.int64)
    for i in range(n):
        if child[i, 0] < 0:
            ids = refs[lf[i]:lf[i] + lc[i]]
            seen[ids] += 1
            # every referenced triangle's box reaches the leaf cell
            assert (tb_lo[ids] <= box[i, 3:] + 1e-6).all() and (tb_hi[ids] >= box[i, :3] - 1e-6).all()
            continue
        a, plane = int(split[i, 0]), split[i, 1]
        l, r = child[i]
        depth[l] = depth[r] = depth[i] + 1
        # the two children are the parent cell cut at the plane
        lo, hi = box[i, :3].copy(), box[i, 3:].copy()
        hi_l = hi.copy(); hi_l[a] = plane
        lo_r = lo.copy(); lo_r[a] = plane
        assert np.array_equal(box[l], np.concatenate([lo, hi_l]))
        assert np.array_equal(box[r], np.concatenate([lo_r, hi]))
        assert lo[a] <= plane <= hi[a]
    assert (seen >= 1).all()          # Distribute keeps every triangle
    assert seen.sum() == m.kd_stats()["refs"] > len(idx)   # and duplicates straddlers
    # KDTreeMeshEx's depth limit is the leaf size (KDTreeMeshEx.cpp:36-44)
    assert depth.max() <= (35 if kind == 1 else 10)
    if kind == 2:
        # cell-centre planes
        for i in np.nonzero(child[:, 0] >= 0)[0][:200]:
            a = int(split[i, 0])
            c = box[i, a] + (box[i, 3 + a] - box[i, a]) * np.float32(0.5)
            assert split[i, 1] == np.float32(c)


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("mode", [0, 1])
def test_kdtree_traversal_equals_brute_force(orc, kind, mode):
    pos, idx = _kd_scene()
    m = orc.Mesh(pos, idx)
    m.kd_build(kind, 10, 35)
    rays = np.concatenate([meshgen.rays_toward(6000, pos, seed=7), meshgen.random_rays(4000, seed=8)])
    pb, tb = m.trace(rays, mode=mode, accel=0)
    pk, tk = m.trace(rays, mode=mode, accel=2)
    assert ((pb >= 0) == (pk >= 0)).all()
    if mode == 0:
        diff = pb != pk
        assert (np.abs(tb[diff] - tk[diff]) <= 1e-5).all()
        assert diff.mean() < 1e-3
    assert (pb >= 0).mean() > 0.5


def test_kdtree_segments(orc):
    # KDTreeMesh keeps the ray window; KDTreeMeshEx replaces ray.Max by the
    # root cell's exit (KDTreeMeshEx.cpp:258-259,371-372), so a segment that
    # ends before a triangle still reports it.
    pos, idx = meshgen.quad([-1, -1, 2], [1, -1, 2], [1, 1, 2], [-1, 1, 2])
    tp, ti = meshgen.random_triangles(400, seed=2, extent=0.5)
    pos, idx = meshgen.merge([(pos, idx), (tp + np.float32([0, 0, -3]), ti)])
    ray = np.array([[0.05, 0.07, 0.0, 0.0, 0.0, 0.0, 1.0, 1.0]], np.float32)   # ends at z = 1
    m = orc.Mesh(pos, idx)
    assert m.trace(ray, mode=1, accel=0)[0][0] == -1
    m.kd_build(1, 2, 35)
    assert m.trace(ray, mode=1, accel=2)[0][0] == -1
    assert m.trace(ray, mode=0, accel=2)[0][0] == -1
    m.kd_build(2, 2, 35)
    assert m.trace(ray, mode=1, accel=2)[0][0] >= 0
    p, t = m.trace(ray, mode=0, accel=2)
    assert p[0] >= 0 and abs(t[0] - 2.0) < 1e-5
    # full-length rays agree with brute force
    ray[0, 7] = F32MAX
    assert m.trace(ray, mode=0, accel=2)[0][0] == m.trace(ray, mode=0, accel=0)[0][0]


# ---------------------------------------------------------------- spaces
def test_space_matches_flattened_brute_force(orc):
    # BasicSpace transforms each ray into each primitive's object space
    # (Ray::ApplyInverse); brute force over the primitives transformed to
    # world space in float64 must find the same (primitive, triangle).
    inst = meshgen.instances(seed=1)
    space = orc.Space()
    meshes = []
    for pos, idx, rot, scl, trn in inst:
        m = orc.Mesh(pos, idx)
        m.build(10, 35)
        meshes.append(m)
        space.add(m, rot, scl, trn)
    wp, wi, owner, local = meshgen.flatten_instances(inst)
    rays = meshgen.rays_toward(20000, wp, seed=3, dist=6.0)
    p, tr, t = space.trace(rays, mode=0)
    pb, tb = orc.Mesh(wp, wi).trace(rays, mode=0, accel=0)
    assert ((p >= 0) == (pb >= 0)).all()
    both = (p >= 0) & (pb >= 0)
    same = both & (owner[np.maximum(pb, 0)] == p) & (local[np.maximum(pb, 0)] == tr)
    assert same.sum() / both.sum() > 0.9995
    assert np.abs(t[both] - tb[both]).max() < 1e-4 * max(1.0, float(tb[both].max()))
    pa, _, _ = space.trace(rays, mode=1)
    assert ((pa >= 0) == (pb >= 0)).all()
    assert both.mean() > 0.5


# ---------------------------------------------------------------- integrator
def _closed_box(albedo, le):
    """Unit cube seen from inside: 6 walls, normals inward, each wall a light."""
    c = [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]]
    c = np.array(c, np.float32)
    faces = [(0, 1, 2, 3), (4, 7, 6, 5), (0, 4, 5, 1), (3, 2, 6, 7), (0, 3, 7, 4), (1, 5, 6, 2)]
    pos, nrm, idx, tl = [], [], [], []
    for k, (a, b, cc, d) in enumerate(faces):
        p = c[[a, b, cc, d]]
        n = np.cross(p[1] - p[0], p[2] - p[0])
        n /= np.linalg.norm(n)
        if np.dot(n, np.array([0.5, 0.5, 0.5]) - p[0]) < 0:   # make it face inward
            p = p[[0, 3, 2, 1]]
            n = -n
        base = len(pos)
        pos += list(p)
        nrm += [n] * 4
        idx += [[base, base + 1, base + 2], [base, base + 2, base + 3]]
        tl += [k, k]
    mats = np.array([[albedo] * 3 + [0, 0, 0, 1, 0]], np.float32)
    lights = np.array([[le, le, le, 0]] * 6, np.float32)
    return np.array(pos), np.array(nrm), np.array(idx), np.zeros(12, np.int32), np.array(tl, np.int32), mats, lights


@pytest.mark.parametrize("rr_depth", [4, 1])
def test_furnace_closed_form(orc, rr_depth):
    # Emission is counted at depth 0 and NEE covers every other wall, whose
    # irradiance sums to pi*Le; each shaded vertex j adds rho^(j+1) Le, so
    # E[L] = Le * sum_{k=0..D} rho^k. Roulette must not move the mean.
    rho, le, depth = 0.5, 1.0, 5
    s = orc.Scene(*_closed_box(rho, le))
    cam = [0.5, 0.5, 0.1, 0.5, 0.5, 1.0, 0, 1, 0, 60.0, 1.0]
    img = s.render(cam, 32, 32, 64, max_depth=depth, rr_depth=rr_depth, seed=11)
    mean = img.mean() / 64
    want = le * sum(rho ** k for k in range(depth + 1))
    assert abs(mean - want) < 0.01 * want, (mean, want)


def _lambert_polygon_irradiance(p, n, poly, le):
    e = 0.0
    k = len(poly)
    for i in range(k):
        a, b = poly[i] - p, poly[(i + 1) % k] - p
        th = math.acos(np.clip(np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b), -1, 1))
        u = np.cross(a, b)
        e += th * np.dot(n, u / np.linalg.norm(u))
    return abs(e) * le / 2


def test_direct_lighting_matches_lambert_formula(orc):
    # A floor of albedo rho under a square emitter: with depth 1 the pixel is
    # rho/pi * E(P), E from Lambert's polygon formula.
    rho, le = 0.6, 3.0
    fp, fi = meshgen.quad([-4, 0, -4], [-4, 0, 4], [4, 0, 4], [4, 0, -4])     # normal +y
    lp, li = meshgen.quad([-0.3, 1, -0.3], [0.3, 1, -0.3], [0.3, 1, 0.3], [-0.3, 1, 0.3])  # normal -y
    pos, idx = meshgen.merge([(fp, fi), (lp, li)])
    nrm = np.array([[0, 1, 0]] * 4 + [[0, -1, 0]] * 4, np.float32)
    mats = np.array([[rho] * 3 + [0, 0, 0, 1, 0], [0, 0, 0, 0, 0, 0, 1, 0]], np.float32)
    s = orc.Scene(pos, nrm, idx, np.array([0, 0, 1, 1], np.int32), np.array([-1, -1, 0, 0], np.int32), mats,
                  np.array([[le, le, le, 0]], np.float32))
    cam = [0.0, 0.8, 0.0, 0.0, 0.0, 0.0, 0, 0, 1, 60.0, 1.0]  # below the emitter, looking down
    W = 16
    spp = 4000
    img = s.render(cam, W, W, spp, max_depth=1, seed=3, window=(4, 4, 12, 12))
    poly = np.array([[-0.3, 1, -0.3], [0.3, 1, -0.3], [0.3, 1, 0.3], [-0.3, 1, 0.3]])
    for y in range(4, 12, 3):
        for x in range(4, 12, 3):
            o, d = orc.camera_ray(cam, W, W, x, y, 0.5, 0.5)
            t = -o[1] / d[1]
            P = o + t * d
            want = rho / math.pi * _lambert_polygon_irradiance(P, np.array([0, 1, 0]), poly, le)
            got = img[y, x, 0] / spp
            # pixel footprint: the sub-pixel jitter moves P by < 1 pixel; allow 3%
            assert abs(got - want) < 0.03 * want, (x, y, got, want)


def test_quad_sample_map(orc):
    """Quad::SamplePoint (Shape/Quad.cpp:146-153) on the unit square v1..v4 =
    (0,0), (1,0), (1,1), (0,1): the corners of the parameter square go to v1,
    v3, v1 + v4 - v3 and v1 + v2 - v3, so the map's image is not the quad
    (its centre goes to v1, not to the quad's centre); the normal is the
    normalised cross(v3 - v1, v2 - v1)."""
    v = [np.array(q, np.float32) for q in ([0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0])]
    want = {(0, 0): v[0], (1, 0): v[2], (0, 1): v[0] + v[3] - v[2], (1, 1): v[0] + v[1] - v[2]}
    for (a, b), q in want.items():
        p, n = orc.quad_sample(*v, a, b)
        assert np.allclose(p, q, atol=1e-7) and np.allclose(n, [0, 0, -1])
    p, _ = orc.quad_sample(*v, 0.5, 0.5)
    assert np.allclose(p, v[0], atol=1e-7)


def test_quad_light_lights_the_image_of_its_map(orc):
    """A quad light (mode 1) whose map is affine (v2 = v3 + v4 - v1) samples
    the parallelogram v1, v3, v4, v1 + v4 - v3 uniformly with density
    1 / area (the map's Jacobian equals the two triangles' area), so with
    depth 1 a floor pixel is rho/pi * E(P), E from Lambert's formula over that
    parallelogram -- here the square [-0.3, 0.3]^2 at y = 1 -- not over the
    emitter's own triangles; sampling the same emitter as a mesh light (mode
    0) gives Lambert's formula over its two (equal-area) triangles instead."""
    rho, le = 0.6, 3.0
    fp, fi = meshgen.quad([-4, 0, -4], [-4, 0, 4], [4, 0, 4], [4, 0, -4])     # floor, normal +y
    v1, v3, v4 = np.array([-0.3, 1, -0.3]), np.array([0.3, 1, -0.3]), np.array([0.3, 1, 0.3])
    v2 = v3 + v4 - v1
    lp = np.array([v1, v2, v3, v4], np.float32)
    li = np.array([[0, 2, 1], [0, 3, 2]], np.int32)                           # (v1, v3, v2), (v1, v4, v3)
    assert np.cross(v3 - v1, v2 - v1)[1] < 0                                  # winding normal faces down
    pos, idx = meshgen.merge([(fp, fi), (lp, li)])
    nrm = np.array([[0, 1, 0]] * 4 + [[0, -1, 0]] * 4, np.float32)
    mats = np.array([[rho] * 3 + [0, 0, 0, 1, 0], [0, 0, 0, 0, 0, 0, 1, 0]], np.float32)
    cam = [0.0, 0.8, 0.0, 0.0, 0.0, 0.0, 0, 0, 1, 60.0, 1.0]
    W, spp = 16, 4000
    square = np.array([[-0.3, 1, -0.3], [0.3, 1, -0.3], [0.3, 1, 0.3], [-0.3, 1, 0.3]])
    imgs, apart = {}, 0.0
    for mode in (1, 0):
        s = orc.Scene(pos, nrm, idx, np.array([0, 0, 1, 1], np.int32), np.array([-1, -1, 0, 0], np.int32), mats,
                      np.array([[le, le, le, mode]], np.float32))
        imgs[mode] = s.render(cam, W, W, spp, max_depth=1, seed=3, window=(4, 4, 12, 12))
    for y in range(4, 12, 3):
        for x in range(4, 12, 3):
            o, d = orc.camera_ray(cam, W, W, x, y, 0.5, 0.5)
            P = o + (-o[1] / d[1]) * d
            up = np.array([0, 1, 0])
            want = rho / math.pi * _lambert_polygon_irradiance(P, up, square, le)
            got = imgs[1][y, x, 0] / spp
            assert abs(got - want) < 0.03 * want, (x, y, got, want)
            # mesh-light sampling of the same (equal-area) triangles: Lambert over each triangle
            want0 = rho / math.pi * (_lambert_polygon_irradiance(P, up, np.array([v1, v3, v2]), le) +
                                     _lambert_polygon_irradiance(P, up, np.array([v1, v4, v3]), le))
            got0 = imgs[0][y, x, 0] / spp
            assert abs(got0 - want0) < 0.03 * want0, (x, y, got0, want0)
            apart = max(apart, abs(want0 - want) / want)
    assert apart > 0.1   # the two closed forms are told apart
