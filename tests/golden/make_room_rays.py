"""Writes the box-room ray fixture: 4,096 closest-hit rays and 4,096 shadow
segments against the 42-triangle box room, with the hits of the oracle's
brute-force traversal (the BasicMesh restatement, Shape/BasicMesh.cpp:16-64,
with the reference's triangle test, Shape/IndexedTriangle.cpp:92-289).

- closest rays: origins uniform in [0.02, 0.98]^3 (some inside the two
  boxes), directions uniform on the sphere, tmin 1e-4, tmax FLT_MAX;
- shadow segments: origins as above, ends uniform on the ceiling light
  (y = 1, x and z in [0.4, 0.6]); direction = end - origin (unnormalised),
  tmin 1e-4, tmax 1 - 1e-4 (VisibilityQuery::SetSegment, Visibility.cpp:50-53).

Output: tests/golden/room_rays_4096.npz with pos, idx, closest_rays (n x 8:
ox oy oz tmin dx dy dz tmax), closest_prim, closest_t, shadow_rays,
shadow_occluded.

Run: python tests/golden/make_room_rays.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def make_rays(n=4096, seed=2024):
    rng = np.random.default_rng(seed)
    o = rng.uniform(0.02, 0.98, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    closest = np.zeros((n, 8), np.float32)
    closest[:, 0:3], closest[:, 3], closest[:, 4:7], closest[:, 7] = o, 1e-4, d, np.finfo(np.float32).max
    o2 = rng.uniform(0.02, 0.98, (n, 3)).astype(np.float32)
    q = np.stack([rng.uniform(0.4, 0.6, n), np.ones(n), rng.uniform(0.4, 0.6, n)], axis=1).astype(np.float32)
    shadow = np.zeros((n, 8), np.float32)
    shadow[:, 0:3], shadow[:, 3], shadow[:, 4:7], shadow[:, 7] = o2, 1e-4, q - o2, np.float32(1 - 1e-4)
    return closest, shadow


def main():
    import oracle
    from illuminaprt_amd import workloads as w
    oracle.build()
    pos, _, idx, *_ = w.box_room()
    closest, shadow = make_rays()
    om = oracle.Mesh(pos, idx)
    prim, t = om.trace(closest, mode=0, accel=0)
    occ, _ = om.trace(shadow, mode=1, accel=0)
    out = os.path.join(HERE, "room_rays_4096.npz")
    np.savez_compressed(out, pos=pos, idx=idx, closest_rays=closest, closest_prim=prim, closest_t=t,
                        shadow_rays=shadow, shadow_occluded=(occ >= 0).astype(np.uint8))
    print(f"wrote {out}: closest hit rate {np.mean(prim >= 0):.3f}, occluded {np.mean(occ >= 0):.3f}")


if __name__ == "__main__":
    main()
