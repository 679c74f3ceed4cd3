"""Small numpy mesh builders for CPU tests (not the product's scene generators)."""
import numpy as np


def icosphere(level, radius=1.0):
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(level):
        cache = {}
        nf = []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return (np.array(verts) * radius).astype(np.float32), np.array(faces, np.int32)


def random_triangles(n, seed=0, extent=1.0, size=0.1):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-extent, extent, (n, 1, 3))
    p = (c + rng.normal(0, size, (n, 3, 3))).astype(np.float32)
    return p.reshape(-1, 3), np.arange(3 * n, dtype=np.int32).reshape(n, 3)


def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2),(p0,p2,p3); normal = (p1-p0)x(p2-p0)."""
    return np.array([p0, p1, p2, p3], np.float32), np.array([[0, 1, 2], [0, 2, 3]], np.int32)


def merge(parts):
    pos, idx, off = [], [], 0
    for p, i in parts:
        pos.append(p)
        idx.append(i + off)
        off += len(p)
    return np.concatenate(pos), np.concatenate(idx)


def random_rays(n, seed=1, origin_extent=2.0, tmax=np.finfo(np.float32).max):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-origin_extent, origin_extent, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 3], r[:, 4:7], r[:, 7] = o, 0.0, d, tmax
    return r


def rays_toward(n, targets_pos, seed=2, dist=3.0):
    """Rays from random points on a sphere of radius `dist` toward random targets."""
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * dist
    tgt = targets_pos[rng.integers(0, len(targets_pos), n)] + rng.normal(0, 0.05, (n, 3))
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3], r[:, 3], r[:, 4:7], r[:, 7] = o, 0.0, d, np.finfo(np.float32).max
    return r


def random_rotation(rng):
    """A uniformly random rotation matrix (QR of a Gaussian matrix)."""
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def instances(seed=0, count=8):
    """Primitives for space tests: (pos, idx, rot9, scl, trn), alternating a
    sphere and a triangle soup, the first untransformed. rot9 holds the
    reference Matrix3x3 elements in storage order, applied as
    world_i = sum_j rot9[3j+i] * scl_j * p_j + trn_i."""
    rng = np.random.default_rng(seed)
    sp, si = icosphere(3)
    tp, ti = random_triangles(300, seed=seed + 2, extent=0.5)
    out = []
    for k in range(count):
        pos, idx = (sp, si) if k % 2 == 0 else (tp, ti)
        if k == 0:
            out.append((pos, idx, None, None, None))
            continue
        rot = random_rotation(rng).astype(np.float32).reshape(-1)
        scl = rng.uniform(0.3, 0.8, 3).astype(np.float32)
        trn = rng.uniform(-3, 3, 3).astype(np.float32)
        out.append((pos, idx, rot, scl, trn))
    return out


def flatten_instances(inst):
    """World-space positions (float64 transform), indices, and the owning
    primitive and local triangle of every flattened triangle."""
    wp, wi, owner, local = [], [], [], []
    off = 0
    for k, (pos, idx, rot, scl, trn) in enumerate(inst):
        p = pos.astype(np.float64)
        if scl is not None:
            p = p * scl.astype(np.float64)
        if rot is not None:
            p = p @ rot.astype(np.float64).reshape(3, 3)
        if trn is not None:
            p = p + trn.astype(np.float64)
        wp.append(p)
        wi.append(idx + off)
        off += len(pos)
        owner.append(np.full(len(idx), k, np.int32))
        local.append(np.arange(len(idx), dtype=np.int32))
    return (np.concatenate(wp).astype(np.float32), np.concatenate(wi).astype(np.int32),
            np.concatenate(owner), np.concatenate(local))
