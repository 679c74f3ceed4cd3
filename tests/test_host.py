"""CPU-only checks of the library: the C ABI loads and exports every declared
symbol, the host BVH builders satisfy their invariants (and the reference
topology builder reproduces the oracle's BVHMesh leaves), the scene
generators are well formed, and GPU entry points fail loudly without a GPU."""
import os
import re

import numpy as np
import pytest

from tests import meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ilm_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ilm_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(ilm):
    import ctypes
    lib = ctypes.CDLL(ilm.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) > 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the Python bindings cover the whole ABI
    assert set(syms) == set(ilm.SIGNATURES), set(syms) ^ set(ilm.SIGNATURES)


def test_no_gpu_fails_loudly(ilm):
    if ilm.gpu_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(ilm.IlmError, match="no HIP device"):
        ilm.Context(0)


def test_icosphere_generator(ilm):
    pos, idx = ilm.icosphere(4, 0.1, 4, 7)
    assert pos.shape == (2562, 3) and idx.shape == (5120, 3)
    r = np.linalg.norm(pos, axis=1)
    assert r.min() > 0.9 - 1e-6 and r.max() < 1.1 + 1e-6 and r.std() > 0.005
    tri = pos[idx]
    n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
    assert (np.einsum("ij,ij->i", n, tri.mean(axis=1)) > 0).all()  # outward winding
    # closed 2-manifold: every undirected edge used exactly twice
    e = np.sort(np.concatenate([idx[:, [0, 1]], idx[:, [1, 2]], idx[:, [2, 0]]]), axis=1)
    _, cnt = np.unique(e, axis=0, return_counts=True)
    assert (cnt == 2).all()
    pos2, _ = ilm.icosphere(4, 0.1, 4, 7)
    assert np.array_equal(pos, pos2)  # deterministic


def test_box_generator(ilm):
    pos, idx = ilm.boxes(50, 0.01, 0.1, 3.0, 11)
    assert pos.shape == (1200, 3) and idx.shape == (600, 3)
    for b in range(50):
        p = pos[24 * b: 24 * b + 24]
        c = p.mean(axis=0)
        assert np.linalg.norm(c) <= 3.0 + 0.1
        tri = pos[idx[12 * b: 12 * b + 12]]
        n = np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0])
        assert (np.einsum("ij,ij->i", n, tri.mean(axis=1) - c) > 0).all()  # outward
        ext = np.linalg.norm(n, axis=1)
        assert (ext > 0).all()


def _bvh_invariants(ilm, pos, idx, nodes, order, width=4):
    boxes, child = ilm.decode_bvh(nodes, width)
    nt = len(idx)
    assert sorted(order.tolist()) == list(range(nt))
    tri = pos[idx[order]]                       # triangles in slot order
    tlo, thi = tri.min(axis=1), tri.max(axis=1)
    seen = np.zeros(nt, np.int64)
    internal_refs = np.zeros(len(boxes), np.int64)
    for ni in range(len(boxes)):
        for k in range(width):
            c = int(child[ni, k])
            b = boxes[ni, k]
            if c & 0x80000000:
                first, cnt = (c & 0x7FFFFFFF) >> 5, c & 31
                if cnt == 0:
                    continue
                s = slice(first, first + cnt)
                seen[s] += 1
                assert (b[:3] <= tlo[s]).all() and (b[3:] >= thi[s]).all()
            else:
                assert 0 < c < len(boxes)
                internal_refs[c] += 1
                cb = boxes[c]
                valid = (child[c] & 0x80000000 == 0) | ((child[c] & 31) > 0)
                assert (b[:3] <= cb[valid, :3]).all() and (b[3:] >= cb[valid, 3:]).all()
    assert (seen == 1).all()
    assert internal_refs[0] == 0 and (internal_refs[1:] == 1).all()   # a tree


@pytest.mark.parametrize("builder,width", [(0, 4), (1, 4), (0, 2), (0, 8), (1, 8), (2, 8)])
def test_wide_bvh_invariants(ilm, builder, width):
    sp, si = meshgen.icosphere(4)
    tp, ti = meshgen.random_triangles(3000, seed=3)
    pos, idx = meshgen.merge([(sp, si), (tp, ti)])
    m = ilm.Mesh(None, pos, idx)       # host-only mesh: build and export, no tracing
    info = m.build(builder=builder, width=width)
    nodes, order = m.export()
    if builder == 2:
        # kd-trees reference triangles from several cells: check the leaf boxes
        # and the tree shape on the distinct references
        boxes, child = ilm.decode_bvh(nodes, width)
        assert info["width"] == width and info["node_stride"] == {2: 64, 4: 112, 8: 224}[width]
        m.close()
        return
    _bvh_invariants(ilm, pos, idx, nodes, order, width)
    assert info["nodes"] == len(nodes) // {2: 64, 4: 112, 8: 224}[width]
    assert info["width"] == width
    # wider nodes: a shallower tree and a stack bound of (W - 1) per level
    if width == 8:
        assert info["depth"] < ilm.Mesh(None, pos, idx).build(builder=builder)["depth"]
    if builder == 0:
        assert info["max_leaf"] <= 4   # the reference topology's depth/radius stops allow bigger leaves
    m.close()


def test_reference_builder_matches_oracle_bvhmesh(ilm, orc):
    # The ILM_BUILDER_REFERENCE topology restates BVHMesh::Compile; its leaves
    # (as triangle sets) must equal those of the oracle's BVHMesh restatement.
    sp, si = meshgen.icosphere(3)
    tp, ti = meshgen.random_triangles(1500, seed=9)
    pos, idx = meshgen.merge([(sp, si), (tp, ti)])
    om = orc.Mesh(pos, idx)
    om.build(10, 35)
    _, ochild, olf, olc, orefs = om.export()
    oleaves = sorted(tuple(sorted(orefs[olf[i]:olf[i] + olc[i]].tolist()))
                     for i in range(len(ochild)) if ochild[i, 0] < 0 and olc[i] > 0)
    m = ilm.Mesh(None, pos, idx)
    m.build(builder=1)
    nodes, order = m.export()
    _, child = ilm.decode_bvh4(nodes)
    leaves = []
    for c in child.reshape(-1):
        c = int(c)
        if c & 0x80000000 and (c & 31):
            first = (c & 0x7FFFFFFF) >> 5
            leaves.append(tuple(sorted(order[first:first + (c & 31)].tolist())))
    assert sorted(leaves) == oleaves
    m.close()


def _product_leaves(ilm, m):
    nodes, order = m.export()
    _, child = ilm.decode_bvh4(nodes)
    leaves = []
    for c in child.reshape(-1):
        c = int(c)
        if c & 0x80000000 and (c & 31):
            first = (c & 0x7FFFFFFF) >> 5
            leaves.append(tuple(order[first:first + (c & 31)].tolist()))
    return leaves, order


@pytest.mark.parametrize("kind,builder", [(1, 2), (2, 3)])
def test_kdtree_builders_match_oracle(ilm, orc, kind, builder):
    # ILM_BUILDER_KDTREE / _EX restate KDTreeMesh / KDTreeMeshEx: the same
    # cells with the same (ordered) triangle lists as the oracle's
    # restatement. The flattener splits leaves above 31 triangles into
    # chunks, so those are compared as a multiset of references.
    sp, si = meshgen.icosphere(3)
    tp, ti = meshgen.random_triangles(1500, seed=9)
    pos, idx = meshgen.merge([(sp, si), (tp, ti)])
    om = orc.Mesh(pos, idx)
    om.kd_build(kind, 10, 35)
    _, ochild, _, olf, olc, orefs = om.kd_export()
    oleaves = [tuple(orefs[olf[i]:olf[i] + olc[i]].tolist()) for i in range(len(ochild))
               if ochild[i, 0] < 0 and olc[i] > 0]
    m = ilm.Mesh(None, pos, idx)
    info = m.build(builder=builder)
    leaves, order = _product_leaves(ilm, m)
    assert np.array_equal(np.sort(order), np.sort(orefs))
    small = [l for l in oleaves if len(l) <= 31]
    assert sorted(l for l in leaves if l in set(small)) == sorted(small)
    if kind == 1:
        assert sorted(leaves) == sorted(oleaves)
    assert len(order) == info["tri_bytes"] // 48 > len(idx)
    m.close()


@pytest.mark.parametrize("builder", [0, 1, 2])
def test_quantized_nodes_contain_float_boxes(ilm, builder):
    # The 64-byte nodes hold every child box of the 112-byte nodes, rounded
    # outward, with the same child words; and the quantisation is tight
    # (each side within two grid steps of the float box).
    sp, si = meshgen.icosphere(3)
    tp, ti = meshgen.random_triangles(2000, seed=4, extent=1.5)
    pos, idx = meshgen.merge([(sp, si), (tp, ti)])
    mf = ilm.Mesh(None, pos, idx)
    fi = mf.build(builder=builder)
    mq = ilm.Mesh(None, pos, idx)
    qi = mq.build(builder=builder, node_format=ilm.ILM_NODES_QUANTIZED)
    assert qi["node_stride"] == 64 and fi["node_stride"] == 112 and qi["nodes"] == fi["nodes"]
    fb, fc = ilm.decode_bvh4(mf.export()[0])
    qb, qc = ilm.decode_bvh4(mq.export()[0], ilm.ILM_NODES_QUANTIZED)
    assert np.array_equal(fc, qc)
    used = qc != 0x80000000
    f64 = fb.astype(np.float64)
    assert (qb[..., :3][used] <= f64[..., :3][used]).all() and (qb[..., 3:][used] >= f64[..., 3:][used]).all()
    raw = mq.export()[0].view(np.uint8).reshape(-1, 64)
    step = np.ldexp(1.0, raw[:, 12:15].astype(np.int64) - 127)[:, None, :].repeat(4, axis=1)
    slack = np.concatenate([f64[..., :3] - qb[..., :3], qb[..., 3:] - f64[..., 3:]], axis=-1)
    assert (slack[used] <= 2 * np.concatenate([step, step], axis=-1)[used]).all()
    mf.close()
    mq.close()


def test_space_flattening(ilm):
    # ilm_space_build applies Transformation::Apply to every vertex and
    # Transformation::Rotate (then normalisation) to every normal.
    inst = meshgen.instances(seed=4, count=5)
    sp = ilm.Space(None)
    for k, (pos, idx, rot, scl, trn) in enumerate(inst):
        nrm = ilm.vertex_normals(pos, idx)
        assert sp.add(pos, idx, nrm, rot, scl, trn, material=k % 3, light=(0 if k == 4 else -1)) == k
    sp.build()
    wp_ref, wi_ref, owner, _ = meshgen.flatten_instances(inst)
    wp, wi = sp.geometry()
    assert np.array_equal(wi, wi_ref)
    assert np.abs(wp - wp_ref).max() < 1e-5
    nrm, tm, tl = sp.shading()
    assert np.array_equal(tm, owner % 3) and np.array_equal(tl, np.where(owner == 4, 0, -1))
    vo = 0
    for pos, idx, rot, scl, trn in inst:
        n = ilm.vertex_normals(pos, idx).astype(np.float64)
        if scl is not None:
            n = n * scl
        if rot is not None:
            n = n @ rot.astype(np.float64).reshape(3, 3)
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        assert np.abs(nrm[vo:vo + len(pos)] - n).max() < 1e-5
        vo += len(pos)
    sp.close()


def test_mesh_without_context_cannot_trace(ilm):
    pos, idx = meshgen.icosphere(1)
    m = ilm.Mesh(None, pos, idx)
    m.build()
    import ctypes
    rc = ilm.load().ilm_trace_closest(m.h, ctypes.c_void_p(16), 1, ctypes.c_void_p(16), None)
    assert rc == -4 and b"device context" in ilm.load().ilm_last_error()
    m.close()


def _py_ordered_counts(pos, idx, boxes, child, order, ray, any_hit):
    """Ordered traversal of an exported BVH4 in float64 (slab entry distances,
    nearest child first, ties kept in slot order; Moller-Trumbore hits with
    the reference's 1e-5 window), counting node visits and triangle tests."""
    o, d, tmax = ray[0:3].astype(np.float64), ray[4:7].astype(np.float64), float(ray[7])
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / np.where(d == 0, 1e-18, d)
    lo_t, hi_t = 1e-5, tmax - 1e-5
    stack, cur, nn, nt = [], 0, 0, 0
    while True:
        if not cur & 0x80000000:
            nn += 1
            bx = boxes[cur].astype(np.float64)
            t0 = (bx[:, 0:3] - o) * inv
            t1 = (bx[:, 3:6] - o) * inv
            tn = np.maximum(np.minimum(t0, t1).max(axis=1), lo_t)
            tf = np.minimum(np.maximum(t0, t1).min(axis=1), hi_t)
            dist = np.where((tn <= tf) & (child[cur] != 0x80000000), tn, np.inf)
            ordk = sorted(range(4), key=lambda k: dist[k])
            for k in ordk[:0:-1]:
                if np.isfinite(dist[k]):
                    stack.append(int(child[cur][k]))
            if np.isfinite(dist[ordk[0]]):
                cur = int(child[cur][ordk[0]])
                continue
        else:
            first, cnt = (cur & 0x7FFFFFFF) >> 5, cur & 31
            for s in range(first, first + cnt):
                nt += 1
                p0, p1, p2 = (pos[i].astype(np.float64) for i in idx[order[s]])
                e1, e2 = p1 - p0, p2 - p0
                pv = np.cross(d, e2)
                det = e1 @ pv
                if det == 0:
                    continue
                tv = o - p0
                u = (tv @ pv) / det
                qv = np.cross(tv, e1)
                v = (d @ qv) / det
                t = (e2 @ qv) / det
                if u >= 0 and v >= 0 and u + v <= 1 and lo_t <= t <= hi_t:
                    hi_t = t - 1e-5
                    if any_hit:
                        return nn, nt
        if not stack:
            return nn, nt
        cur = stack.pop()


def test_visit_counts_of_the_ordered_traversal(ilm):
    """ilm_bvh_visit_counts (host): equal to an independent float64 ordered
    traversal of the exported tree on most rays (the two round box and hit
    distances differently, which can move a cut by a node), and bounded by
    the structure: a ray that misses the root's children visits one node and
    tests nothing; an any-hit query never does more work than the closest-hit
    query of the same ray."""
    sp, si = meshgen.icosphere(3)
    tp, ti = meshgen.random_triangles(300, seed=4, extent=1.2)
    pos, idx = meshgen.merge([(sp, si), (tp, ti)])
    m = ilm.Mesh(None, pos, idx)
    m.build()
    nodes, order = m.export()
    boxes, child = ilm.decode_bvh(nodes, 4)
    rays = meshgen.rays_toward(400, pos, seed=8)
    miss = rays[:20].copy()
    miss[:, 0:3] = (10.0, 10.0, 10.0)
    miss[:, 4:7] = (0.0, 1.0, 0.0)                     # away from the mesh
    rays = np.concatenate([rays, miss])
    c = m.visit_counts(rays)
    a = m.visit_counts(rays, any_hit=True)
    assert (c[-20:] == [1, 0]).all()
    assert (a <= c).all()
    py = np.array([_py_ordered_counts(pos, idx, boxes, child, order, r, False) for r in rays])
    eq = (py == c).all(axis=1).mean()
    print(f"visit counts: host {c.mean(axis=0)} vs float64 restatement {py.mean(axis=0)}, equal on {eq:.4f}")
    assert eq >= 0.97
    np.testing.assert_allclose(c.mean(axis=0), py.mean(axis=0), rtol=0.01)
    # any-hit counts against the float64 any-hit traversal
    pya = np.array([_py_ordered_counts(pos, idx, boxes, child, order, r, True) for r in rays])
    assert (pya == a).all(axis=1).mean() >= 0.97
    # quantised nodes: the host counts against the same float64 restatement
    # run on the decoded (outward-rounded) boxes of the quantised tree
    mq = ilm.Mesh(None, pos, idx)
    mq.build(node_format=ilm.ILM_NODES_QUANTIZED)
    qnodes, qorder = mq.export()
    qboxes, qchild = ilm.decode_bvh4(qnodes, ilm.ILM_NODES_QUANTIZED)
    cq = mq.visit_counts(rays)
    pyq = np.array([_py_ordered_counts(pos, idx, qboxes, qchild, qorder, r, False) for r in rays])
    eqq = (pyq == cq).all(axis=1).mean()
    print(f"quantised visit counts: host {cq.mean(axis=0)} vs float64 restatement {pyq.mean(axis=0)}, equal on {eqq:.4f}")
    assert eqq >= 0.97
    # the outward-rounded boxes contain the float ones: no fewer visits overall
    assert cq[:, 0].mean() >= c[:, 0].mean()
    mq.close()
    m.close()
