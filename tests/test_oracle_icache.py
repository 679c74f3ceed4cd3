"""The irradiance-cache restatement (oracle/icache.py) against what its
definitions fix: a lookup finds exactly the records a brute-force scan over
all records accepts (an accepted record lies within RiClamp of the point, so
its sphere reaches the leaf holding the point, and Insert put it on that
leaf's root path); every record sits only in cells its sphere overlaps and
covers the part of its sphere inside the root; the strata of the cosine
sampler tile the unit square and average to the cosine moments."""
import numpy as np
import pytest

from oracle import icache as ic


def _records(n, seed):
    rng = np.random.default_rng(seed)
    P = rng.random((n, 3)).astype(np.float32) * 2 - 1
    N = rng.normal(size=(n, 3))
    N = (N / np.linalg.norm(N, axis=1, keepdims=True)).astype(np.float32)
    E = rng.random((n, 3)).astype(np.float32)
    rc = (0.02 + 0.3 * rng.random(n) ** 2).astype(np.float32)
    return P, N, E, rc


def _tree(P, N, E, rc, depth=8, variant="ward"):
    lo, hi, _ = ic.root_cell(np.array([[-1, -1, -1], [1, 1, 1]], np.float32))
    t = ic.Octree(lo, hi, depth, variant)
    for i in range(len(P)):
        t.insert(P[i], N[i], E[i], rc[i])
    return t


def test_root_cell():
    lo, hi, L = ic.root_cell(np.array([[0, 0, 0], [2, 1, 0.5]], np.float32))
    assert L == 2
    np.testing.assert_allclose(lo, [-1, -1.5, -1.75], rtol=1e-6)
    np.testing.assert_allclose(hi, [3, 2.5, 2.25], rtol=1e-6)


@pytest.mark.parametrize("variant", ["ward", "tabelion"])
def test_lookup_equals_brute_force(variant):
    P, N, E, rc = _records(400, 1)
    t = _tree(P, N, E, rc, variant=variant)
    rng = np.random.default_rng(2)
    for threshold in (1.0, 0.5, 0.2):
        inv = np.float32(1) / np.float32(threshold)
        for _ in range(300):
            p = (rng.random(3) * 2 - 1).astype(np.float32)
            k = int(rng.integers(len(P)))
            n = N[k] if rng.random() < 0.5 else (rng.normal(size=3)).astype(np.float32)
            n = (n / np.linalg.norm(n)).astype(np.float32)
            _, den, ids = t.lookup(p, n, threshold)
            brute = [r for r in range(len(P)) if t.weight(p, n, r, inv) > 0]
            assert sorted(ids) == brute


@pytest.mark.parametrize("variant", ["ward", "tabelion"])
def test_records_cover_their_spheres(variant):
    P, N, E, rc = _records(120, 3)
    t = _tree(P, N, E, rc, depth=6, variant=variant)
    holders = {r: [] for r in range(len(P))}
    for node, recs in enumerate(t.recs):
        for r in recs:
            holders[r].append(node)
            assert ic.sphere_box(t.lo[node], t.hi[node], P[r], rc[r])
    rng = np.random.default_rng(4)
    for r in range(len(P)):
        d = rng.normal(size=(50, 3))
        pts = P[r] + (d / np.linalg.norm(d, axis=1, keepdims=True)) * rng.random((50, 1)) ** (1 / 3) * rc[r] * 0.999
        for q in pts.astype(np.float32):
            if not t._contains(t.lo[0], t.hi[0], q):
                continue
            assert any(t._contains(t.lo[n], t.hi[n], q) for n in holders[r])


def test_stratified_cosine_directions():
    alt_n, az_n = 4, 12
    rng = np.random.default_rng(5)
    u, v = rng.random(200000), rng.random(200000)
    alt, az = rng.integers(alt_n, size=u.shape), rng.integers(az_n, size=u.shape)
    d = ic.cosine_stratified(u, v, alt, az, alt_n, az_n)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1, atol=1e-12)
    # cosine-weighted hemisphere: E[cos] = 2/3, E[cos^2] = 1/2, azimuth uniform
    assert abs(d[:, 2].mean() - 2 / 3) < 0.003 and abs((d[:, 2] ** 2).mean() - 0.5) < 0.003
    assert abs(d[:, 0].mean()) < 0.005 and abs(d[:, 1].mean()) < 0.005
    # each (altitude, azimuth) stratum maps to its own cell of the disk
    a = 1 - d[:, 2] ** 2
    assert np.array_equal(np.floor(a * alt_n + 1e-12).astype(int), alt)


def test_tabelion_weight_closed_forms():
    lo, hi, _ = ic.root_cell(np.array([[-1, -1, -1], [1, 1, 1]], np.float32))
    t = ic.Octree(lo, hi, 4, "tabelion")
    t.insert([0, 0, 0], [0, 0, 1], [1, 1, 1], 0.5)
    # (the last argument is 1 / threshold) at the record: w = 1; at half the
    # radius along the plane: 1 - 0.5 / threshold; tilted normals: the angle
    # term 9.0124 (1 - cos)
    assert t.weight(np.float32([0, 0, 0]), np.float32([0, 0, 1]), 0, 1.0) == 1.0
    assert abs(t.weight(np.float32([0.25, 0, 0]), np.float32([0, 0, 1]), 0, 1.0) - 0.5) < 1e-6
    c = np.cos(0.2)
    n = np.float32([np.sin(0.2), 0, c])
    assert abs(t.weight(np.float32([0, 0, 0]), n, 0, 0.5) - (1 - 0.5 * 9.0124 * (1 - c))) < 1e-5
    assert t.weight(np.float32([0, 0, 0]), np.float32([0, 0, -1]), 0, 1.0) < 0


def test_record_formula():
    E, ri, rc = ic.record([[1, 2, 3], [3, 2, 1]], [1.0, 3.0], 0.1, 1.0)
    assert np.allclose(E, [2, 2, 2]) and abs(ri - 1.5) < 1e-12 and rc == 1.0
    assert ic.record([[0, 0, 0]], [100.0], 0.1, 1000.0)[2] == 100.0
    assert ic.record([[0, 0, 0]], [0.001], 0.1, 1000.0)[2] == 0.1
    assert ic.record([[0, 0, 0], [1, 1, 1]], [2.0, 0.5], 0.1, 10.0, "tabelion")[1] == 0.5


def test_ic_integrator_matches_path_tracing(orc):
    """The ICIntegrator restatement (orc_ic_render) on the box room: with an
    error threshold so small that nearly every sample computes its own record
    (48 stratified gather paths), its image estimates what the path
    integrator converges to (indirect factor 1: the radiance a Lambertian
    surface reflects); with the factory's threshold 1, records are reused
    (far fewer records than samples) and the image mean stays close. (At one
    sample per pixel the emitter's edge pixels dominate a per-pixel error, so
    the means are compared.)"""
    from illuminaprt_amd import workloads as w
    sa = w.box_room()
    cam = w.room_camera().as_list()
    s = orc.Scene(*sa)
    W = H = 48
    ref = s.render(cam, W, H, 256, max_depth=6, seed=5) / 256
    fine = s.ic_render(cam, W, H, 1, seed=7, error_threshold=1e-3, ray_depth=5, indirect_factor=1.0,
                       miss_radiance=(0, 0, 0))
    cf = s.last_counts
    coarse = s.ic_render(cam, W, H, 1, seed=7, error_threshold=1.0, ray_depth=5, indirect_factor=1.0,
                         miss_radiance=(0, 0, 0))
    cc = s.last_counts
    print(f"[cpu ic] means pt {ref.mean():.4f} fine {fine.mean():.4f} coarse {coarse.mean():.4f}; "
          f"records {cf[3]} / {cc[3]} of {W * H}; relRMSE {w.rel_rmse(fine, ref):.3f} / {w.rel_rmse(coarse, ref):.3f}")
    assert cf[0] == cc[0] == W * H
    assert cf[3] >= 0.9 * W * H and cc[3] <= 0.05 * W * H
    assert abs(fine.mean() - ref.mean()) < 0.06 * ref.mean()
    assert abs(coarse.mean() - ref.mean()) < 0.08 * ref.mean()
    # the reference's defaults: indirect = irradiance x albedo / pi, misses (20, 50, 100)
    d = s.ic_render(cam, W, H, 1, seed=7, error_threshold=1.0, ray_depth=5)
    miss = (d == np.array([20.0, 50.0, 100.0])).all(axis=2)
    assert miss.any() and (coarse[miss] == 0).all()
    assert d[~miss].mean() < coarse[~miss].mean()
