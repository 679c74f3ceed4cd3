"""The box-room ray fixture (tests/golden/room_rays_4096.npz, written by
tests/golden/make_room_rays.py with the oracle's brute-force traversal).

The fixture is pinned by geometry independent of the triangle test: every
surface of the room is an axis-aligned rectangle made of two triangles
(quad q = triangles 2q and 2q + 1), so the closest hit and the occlusion of a
segment follow from ray/plane distances in float64. The oracle's BVHMesh and
KDTreeMesh traversals must then reproduce the fixture.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "room_rays_4096.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


def _rects(pos):
    """(axis, plane, lo[2], hi[2]) of each quad of 4 consecutive vertices."""
    out = []
    for q in range(len(pos) // 4):
        v = pos[4 * q:4 * q + 4].astype(np.float64)
        a = int(np.argmin(v.max(axis=0) - v.min(axis=0)))
        o = [k for k in range(3) if k != a]
        out.append((a, v[0, a], v[:, o].min(axis=0), v[:, o].max(axis=0), o))
    return out


def _analytic(rays, rects, tmin, tmax_of):
    """Closest quad and t per ray in float64 (t window open at both ends)."""
    o, d = rays[:, 0:3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    best_t = np.full(len(rays), np.inf)
    best_q = np.full(len(rays), -1)
    tmax = tmax_of(rays)
    for q, (a, c, lo, hi, oth) in enumerate(rects):
        with np.errstate(divide="ignore", invalid="ignore"):
            t = (c - o[:, a]) / d[:, a]
        p = o[:, oth] + t[:, None] * d[:, oth]
        inside = np.all((p >= lo - 1e-12) & (p <= hi + 1e-12), axis=1)
        ok = inside & (t > tmin) & (t < tmax) & np.isfinite(t) & (t < best_t)
        best_t = np.where(ok, t, best_t)
        best_q = np.where(ok, q, best_q)
    return best_q, best_t


def test_fixture_closest_matches_geometry(gold):
    rays = gold["closest_rays"]
    q, t = _analytic(rays, _rects(gold["pos"]), 1e-5, lambda r: r[:, 7].astype(np.float64) - 1e-5)
    prim, tf = gold["closest_prim"], gold["closest_t"]
    hit = prim >= 0
    assert np.array_equal(hit, q >= 0)
    # the same rectangle, or a tie at a shared edge (equal distances)
    same = (prim[hit] // 2) == q[hit]
    tie = np.abs(tf[hit] - t[hit]) <= 1e-6 * t[hit]
    assert (same | tie).all()
    assert np.abs(tf[hit] - t[hit]).max() <= 2e-6 * t[hit].max()
    assert 0.8 < hit.mean() < 0.95


def test_fixture_shadow_matches_geometry(gold):
    rays = gold["shadow_rays"]
    q, _ = _analytic(rays, _rects(gold["pos"]), 1e-5, lambda r: r[:, 7].astype(np.float64) - 1e-5)
    assert np.array_equal(gold["shadow_occluded"], (q >= 0).astype(np.uint8))
    assert 0.05 < gold["shadow_occluded"].mean() < 0.5


@pytest.mark.parametrize("accel", [1, 2])
def test_oracle_accelerators_reproduce_fixture(gold, orc, accel):
    om = orc.Mesh(gold["pos"], gold["idx"])
    if accel == 2:
        om.kd_build(1, 10, 35)
    else:
        om.build()
    prim, t = om.trace(gold["closest_rays"], mode=0, accel=accel)
    # the boxes' bottom faces lie on the floor: an exact tie may go either way
    gp, gt = gold["closest_prim"], gold["closest_t"]
    assert np.array_equal(prim >= 0, gp >= 0)
    tie = np.abs(t - gt) <= 1e-6 * gt
    assert ((prim == gp) | tie).all() and (prim == gp).mean() > 0.99
    occ, _ = om.trace(gold["shadow_rays"], mode=1, accel=accel)
    assert np.array_equal((occ >= 0).astype(np.uint8), gold["shadow_occluded"])
