"""Vectorised structural checks of exported BVH4 trees (numpy only), sized for
trees of millions of nodes: one or several trees packed in one node array,
each with its own root and its own range of triangles.

Node words follow the traversal kernels' layout: a child word is an inner
node index, or a leaf (bit 31 set, first triangle slot in bits 5-30, count in
bits 0-4); an empty slot is a leaf of 0 triangles with a far box.
"""
import numpy as np

LEAF = np.uint32(0x80000000)


def check_bvh4(boxes, child, order, tri_lo, tri_hi, roots, tri_ranges, max_leaf):
    """boxes (n,4,6) float (lo xyz, hi xyz per slot), child (n,4) uint32,
    order (slots,) original triangle index of each triangle slot,
    tri_lo / tri_hi (ntri,3) triangle bounds, roots (k,) root node per tree
    (increasing; tree j owns nodes [roots[j], roots[j+1])), tri_ranges (k,2)
    triangles [a, b) of tree j. Asserts:
    - every node but the roots has exactly one parent, a child index is larger
      than its parent's (so the trees are acyclic and every node is reached
      from a root), and a tree's inner children stay in its node range;
    - every child box contains the boxes of its node's slots;
    - every leaf holds at most max_leaf triangles, each inside the leaf's box;
    - every triangle of tree j sits in exactly one leaf of tree j.
    Returns a dict of counts."""
    n = len(child)
    roots = np.asarray(roots, np.int64)
    child = np.asarray(child, np.uint32)
    is_leaf = (child & LEAF) != 0
    inner = ~is_leaf
    # ---- topology
    parent_node = np.repeat(np.arange(n, dtype=np.int64), 4).reshape(n, 4)
    ci = child[inner].astype(np.int64)
    pi = parent_node[inner]
    assert (ci < n).all(), "inner child index out of range"
    assert (ci > pi).all(), "a child precedes its parent"
    nparents = np.bincount(ci, minlength=n)
    want = np.ones(n, np.int64)
    want[roots] = 0
    assert np.array_equal(nparents, want), "a node has no parent or several parents"
    tree_of = np.searchsorted(roots, np.arange(n), side="right") - 1
    assert (tree_of >= 0).all(), "nodes before the first root"
    assert np.array_equal(tree_of[ci], tree_of[pi]), "an inner child leaves its tree"
    # ---- nested boxes: the parent's slot box contains every used slot box of the child
    slot_box = boxes[pi.reshape(-1), np.nonzero(inner)[1]]            # (m,6) parent's box of each inner child
    cbox = boxes[ci]                                                  # (m,4,6)
    used = ~((child[ci] & LEAF != 0) & ((child[ci] & np.uint32(31)) == 0))   # empty slots aside
    lo_ok = (cbox[:, :, 0:3] >= slot_box[:, None, 0:3]).all(axis=2) | ~used
    hi_ok = (cbox[:, :, 3:6] <= slot_box[:, None, 3:6]).all(axis=2) | ~used
    assert lo_ok.all() and hi_ok.all(), "a child box exceeds its parent slot"
    # ---- leaves
    lw = child[is_leaf]
    cnt = (lw & np.uint32(31)).astype(np.int64)
    first = ((lw & np.uint32(0x7FFFFFFF)) >> np.uint32(5)).astype(np.int64)
    lnode = parent_node[is_leaf]
    lbox = boxes[lnode, np.nonzero(is_leaf)[1]]
    keep = cnt > 0
    cnt, first, lnode, lbox = cnt[keep], first[keep], lnode[keep], lbox[keep]
    assert cnt.max(initial=0) <= max_leaf, "leaf above max_leaf"
    # expand each leaf to its slots
    rep = np.repeat(np.arange(len(cnt)), cnt)
    off = np.arange(len(rep)) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    slots = first[rep] + off
    assert (slots < len(order)).all(), "leaf slot out of range"
    assert np.array_equal(np.bincount(slots, minlength=len(order)), np.ones(len(order), np.int64)), \
        "a triangle slot is in no leaf or in several"
    tri = np.asarray(order, np.int64)[slots]
    assert np.array_equal(np.bincount(tri, minlength=len(tri_lo)), np.ones(len(tri_lo), np.int64)), \
        "a triangle is in no leaf or in several"
    b = lbox[rep]
    assert (tri_lo[tri] >= b[:, 0:3]).all() and (tri_hi[tri] <= b[:, 3:6]).all(), "a triangle outside its leaf box"
    # each triangle in its own tree's leaves
    rng = np.asarray(tri_ranges, np.int64)
    t_tree = tree_of[lnode[rep]]
    assert ((tri >= rng[t_tree, 0]) & (tri < rng[t_tree, 1])).all(), "a triangle in another tree's leaf"
    return {"nodes": n, "leaves": int(len(cnt)), "triangles": int(len(tri)), "trees": int(len(roots))}
