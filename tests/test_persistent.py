"""PersistentMesh import (ilm_persistent_*): the reference's on-disk mesh
and kd-tree (Shape/PersistentMesh.h:22-36, PersistentMesh.cpp:77-89,
171-245; written by KDTreeMesh's export, KDTreeMesh.cpp:143-257) loaded as
a mesh whose BVH is the file's own tree.

The files are written here from the oracle's KDTreeMesh restatement, with
the reference writer's layout: 48-byte vertex and triangle records and the
tree in breadth-first order with byte offsets to the children. The import
must reproduce the oracle tree's cells and leaf lists, and tracing the
imported mesh must agree with brute force."""
import os

import numpy as np
import pytest

from tests import meshgen


def write_persistent(trunk, pos, nrm, idx, group, box, child, split, lf, lc, refs):
    nv, nt = len(pos), len(idx)
    ver = np.zeros((nv, 12), np.float32)
    ver[:, 0:3] = pos
    ver[:, 4:7] = nrm
    ver.tofile(trunk + ".ver")
    tri = np.zeros((nt, 12), np.int32)
    tri[:, 0:3] = idx
    tri[:, 3] = group
    e = np.zeros((nt, 8), np.float32)
    P = pos[idx]
    e[:, 0:3] = P[:, 1] - P[:, 0]
    e[:, 4:7] = P[:, 2] - P[:, 0]
    tri[:, 4:12] = e.view(np.int32)
    tri.tofile(trunk + ".tri")
    # breadth-first, as the reference writer: header, then children offsets
    # (patched in a second pass) or triangle indices
    order, q = [], [0]
    while q:
        n = q.pop(0)
        order.append(n)
        if child[n, 0] >= 0:
            q += [int(child[n, 0]), int(child[n, 1])]
    offset, o = {}, 0
    for n in order:
        offset[n] = o
        o += 8 + (8 if child[n, 0] >= 0 else 4 * int(lc[n]))
    buf = bytearray(o)
    for n in order:
        b = offset[n]
        if child[n, 0] >= 0:
            axis, part, count = int(split[n, 0]), float(split[n, 1]), 0
            tail = np.array([offset[int(child[n, 0])], offset[int(child[n, 1])]], np.int32).tobytes()
        else:
            axis, part, count = 3, 666.666, int(lc[n])
            tail = refs[lf[n]:lf[n] + lc[n]].astype(np.int32).tobytes()
        head = np.array([part], np.float32).tobytes() + np.array([axis | (count << 2)], np.uint32).tobytes()
        buf[b:b + 8 + len(tail)] = head + tail
    with open(trunk + ".tre", "wb") as f:
        f.write(bytes(buf))
    return order, offset


def _scene():
    sp, si = meshgen.icosphere(3)
    tp, ti = meshgen.random_triangles(800, seed=4)
    pos, idx = meshgen.merge([(sp, si), (tp, ti)])
    from illuminaprt_amd import vertex_normals
    nrm = vertex_normals(pos, idx)
    group = (np.arange(len(idx)) % 3).astype(np.int32)
    return pos.astype(np.float32), nrm.astype(np.float32), idx.astype(np.int32), group


def _export(orc, pos, idx):
    om = orc.Mesh(pos, idx)
    om.kd_build(1, 10, 35)
    return om.kd_export()


def test_import_reproduces_the_tree(tmp_path, ilm, orc):
    pos, nrm, idx, group = _scene()
    box, child, split, lf, lc, refs = _export(orc, pos, idx)
    trunk = os.path.join(str(tmp_path), "object")
    write_persistent(trunk, pos, nrm, idx, group, box, child, split, lf, lc, refs)
    p = ilm.Persistent(trunk)
    ppos, pnrm, pidx, pgroup = p.arrays()
    assert np.array_equal(ppos, pos) and np.array_equal(pnrm, nrm)
    assert np.array_equal(pidx, idx) and np.array_equal(pgroup, group)
    assert p.tree_nodes == len(child) and p.tree_refs == len(refs)
    m = ilm.Mesh(None, ppos, pidx)
    p.build(m)
    # the same leaves (ordered triangle lists) as the oracle's KDTreeMesh
    from tests.test_host import _product_leaves
    leaves, order = _product_leaves(ilm, m)
    oleaves = [tuple(refs[lf[i]:lf[i] + lc[i]].tolist()) for i in range(len(child)) if child[i, 0] < 0 and lc[i] > 0]
    assert np.array_equal(np.sort(order), np.sort(refs))
    small = [l for l in oleaves if len(l) <= 31]
    assert sorted(l for l in leaves if l in set(small)) == sorted(small)
    m.close()
    p.close()


def _corrupt(trunk, edit):
    data = bytearray(open(trunk + ".tre", "rb").read())
    edit(data)
    with open(trunk + ".tre", "wb") as f:
        f.write(bytes(data))


def test_malformed_files_are_refused(tmp_path, ilm, orc):
    pos, nrm, idx, group = _scene()
    ex = _export(orc, pos, idx)
    trunk = os.path.join(str(tmp_path), "object")

    def fresh():
        return write_persistent(trunk, pos, nrm, idx, group, *ex)

    order, offset = fresh()
    root_children = np.frombuffer(open(trunk + ".tre", "rb").read()[8:16], np.int32)
    cases = {
        "outside": lambda d: d.__setitem__(slice(8, 12), np.array([len(d) + 64], np.int32).tobytes()),
        "twice": lambda d: d.__setitem__(slice(12, 16), np.array([root_children[0]], np.int32).tobytes()),
        "misaligned": lambda d: d.__setitem__(slice(8, 12), np.array([root_children[0] + 2], np.int32).tobytes()),
        "truncated": lambda d: d.__delitem__(slice(len(d) - 4, len(d))),
    }
    for name, edit in cases.items():
        fresh()
        _corrupt(trunk, edit)
        with pytest.raises(ilm.IlmError):
            ilm.Persistent(trunk)
    # a leaf's triangle index out of range
    fresh()
    leaf = next(n for n in order if ex[1][n, 0] < 0 and ex[4][n] > 0)
    _corrupt(trunk, lambda d: d.__setitem__(slice(offset[leaf] + 8, offset[leaf] + 12),
                                            np.array([len(idx)], np.int32).tobytes()))
    with pytest.raises(ilm.IlmError):
        ilm.Persistent(trunk)
    # records that are not whole
    fresh()
    with open(trunk + ".ver", "ab") as f:
        f.write(b"\0" * 4)
    with pytest.raises(ilm.IlmError):
        ilm.Persistent(trunk)
    with pytest.raises(ilm.IlmError):
        ilm.Persistent(os.path.join(str(tmp_path), "missing"))


@pytest.mark.gpu
def test_imported_tree_traces_like_brute_force(tmp_path, gpu_ctx, ilm, orc):
    pos, nrm, idx, group = _scene()
    trunk = os.path.join(str(tmp_path), "object")
    write_persistent(trunk, pos, nrm, idx, group, *_export(orc, pos, idx))
    p = ilm.Persistent(trunk)
    m = ilm.Mesh(gpu_ctx, pos, idx)
    p.build(m)
    rays = np.concatenate([meshgen.rays_toward(30000, pos, seed=12), meshgen.random_rays(20000, seed=13)])
    pc, tc = orc.Mesh(pos, idx).trace(rays, mode=0, accel=0)
    rb = ilm.rays_to_device(gpu_ctx, rays)
    hb = gpu_ctx.buffer(len(rays) * ilm.HIT_BYTES)
    m.trace_closest(rb, len(rays), hb)
    gpu_ctx.sync()
    h = hb.download(ilm.HIT_DTYPE, len(rays))
    from tests.test_gpu_trace import _compare_closest
    assert _compare_closest(h["prim"], h["t"], pc, tc, "persistent kd-tree") >= 0.9999
    for b in (rb, hb):
        b.free()
    m.close()
    p.close()


def test_scene_script_reads_persistent_mesh(tmp_path, ilm, orc):
    # a PersistentMesh shape in a script (ShapeFactories.h:337-370): the
    # geometry and group ids come from the files; the flattened space is
    # rebuilt as a KDTreeMesh
    from tests.test_environment import SPACE_ONE, _env, load
    pos, nrm, idx, group = _scene()
    write_persistent(os.path.join(str(tmp_path), "object"), pos, nrm, idx, group, *_export(orc, pos, idx))
    tri = "    Id = tri\n    Type = Triangle\n    P0 = {0, 0, 1}\n    P1 = {1, 0, 1}\n    P2 = {0, 1, 1}\n"
    text = _env(SPACE_ONE).replace(tri, "    Id = tri\n    Type = PersistentMesh\n    TrunkName = object\n")
    r, job = load(tmp_path, text)
    assert r.returncode == 0, r.stderr
    assert np.array_equal(job["pos"], pos) and np.array_equal(job["idx"], idx)
    np.testing.assert_allclose(job["nrm"], nrm / np.linalg.norm(nrm, axis=1, keepdims=True), atol=1e-6)
    assert job["builder"] == ilm.ILM_BUILDER_KDTREE and (job["tri_light"] == 0).all()
