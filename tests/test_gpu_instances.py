"""The flat instance table (two-level BVH over shared meshes) against the
oracle's BasicSpace restatement, which loops over every primitive and traces
the ray transformed into its object space (Space/BasicSpace.cpp:58-114,
Scene/GeometricPrimitive.cpp:68-200). 1,200 instances share three meshes;
after the transforms change (ISpace::Update) only the top-level tree is
rebuilt and the hits still agree. Device memory holds each mesh once."""
import numpy as np
import pytest

from tests import meshgen
from tests.test_gpu_trace import _compare_closest

pytestmark = pytest.mark.gpu


def _meshes():
    sp, si = meshgen.icosphere(4)
    tp, ti = meshgen.random_triangles(2000, seed=31, extent=0.5)
    bp, bi = meshgen.icosphere(1)
    return [(sp, si), (tp, ti), (bp, bi)]


def _placements(rng, n, nmeshes):
    out = []
    for k in range(n):
        m = int(rng.integers(nmeshes))
        if k % 100 == 0:
            out.append((m, None, None, rng.uniform(-4, 4, 3).astype(np.float32)))   # translation only
            continue
        rot = meshgen.random_rotation(rng).astype(np.float32).reshape(-1)
        scl = rng.uniform(0.08, 0.25, 3).astype(np.float32)
        trn = rng.uniform(-4, 4, 3).astype(np.float32)
        out.append((m, rot, scl, trn))
    return out


def _world_points(meshes, placements):
    pts = []
    for m, rot, scl, trn in placements:
        p = meshes[m][0][::7].astype(np.float64)
        if scl is not None:
            p = p * scl.astype(np.float64)
        if rot is not None:
            p = p @ rot.astype(np.float64).reshape(3, 3)
        p = p + trn.astype(np.float64)
        pts.append(p)
    return np.concatenate(pts).astype(np.float32)


def _oracle_space(orc, omeshes, placements):
    space = orc.Space()
    for m, rot, scl, trn in placements:
        space.add(omeshes[m], rot, scl, trn)
    return space


def test_instances_match_basicspace(gpu_ctx, ilm, orc):
    rng = np.random.default_rng(77)
    meshes = _meshes()
    gm = []
    for pos, idx in meshes:
        m = ilm.Mesh(gpu_ctx, pos, idx)
        m.build()
        gm.append(m)
    om = []
    for pos, idx in meshes:
        o = orc.Mesh(pos, idx)
        o.build(10, 35)
        om.append(o)
    table = ilm.Instances(gpu_ctx, gm)
    n_inst = 1200
    mesh_bytes = [m.info()["node_bytes"] + m.info()["tri_bytes"] for m in gm]
    for rnd in range(2):   # the second round moves every instance (ISpace::Update)
        placements = _placements(rng, n_inst, len(meshes))
        table.set(placements)
        info = table.info()
        flat = sum(mesh_bytes[m] for m, *_ in placements)
        print(f"[instances round {rnd}] {info}; the same instances flattened would take {flat} bytes of tree and "
              f"triangles", flush=True)
        assert info["instances"] == n_inst and info["table_bytes"] == 64 * n_inst
        assert info["mesh_bytes"] == sum(mesh_bytes)
        assert info["mesh_bytes"] + info["table_bytes"] + info["tlas_bytes"] < flat / 10
        wp = _world_points(meshes, placements)
        rays = np.concatenate([meshgen.rays_toward(30000, wp, seed=5 + rnd, dist=7.0),
                               meshgen.random_rays(10000, seed=9 + rnd)])
        n = len(rays)
        rb = ilm.rays_to_device(gpu_ctx, rays)
        hb, ib, ob = gpu_ctx.buffer(n * 16), gpu_ctx.buffer(n * 4), gpu_ctx.buffer(n)
        table.trace_closest(rb, n, hb, ib)
        table.trace_any(rb, n, ob)
        table.check()
        h = hb.download(ilm.HIT_DTYPE, n)
        gi, occ = ib.download(np.int32, n), ob.download(np.uint8, n)
        space = _oracle_space(orc, om, placements)
        p, tr, t = space.trace(rays, mode=0)
        key_g = np.where(h["prim"] >= 0, gi.astype(np.int64) * 100000 + h["prim"], -1)
        key_c = np.where(p >= 0, p.astype(np.int64) * 100000 + tr, -1)
        assert ((h["prim"] >= 0) == (gi >= 0)).all()
        frac = _compare_closest(key_g, h["t"], key_c, t, f"instances round {rnd}")
        assert frac >= 0.9999
        assert np.mean(h["prim"] >= 0) > 0.3
        pa, _, _ = space.trace(rays, mode=1)
        agree = (occ == (pa >= 0)).mean()
        print(f"[instances round {rnd}] any-hit agreement {agree:.6f}, occluded {occ.mean():.4f}", flush=True)
        assert agree >= 0.9999
        for b in (rb, hb, ib, ob):
            b.free()
    table.close()
    for m in gm:
        m.close()
