// Host BVH construction: binned SAH, the reference BVHMesh topology and the
// reference kd-tree topologies, plus the collapse to wide nodes that the GPU
// traverses.
#include "bvh_build.hpp"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <thread>

#include "geom.hpp"

// The reference topologies depend on float comparisons of box radii and split
// planes; keep every product and sum rounded separately, as the reference's
// scalar code does.
#pragma clang fp contract(off)

namespace ilm {
namespace {

struct Prim {
    float lo[3], hi[3], c[3];
};

struct AABB {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
    float hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    void grow(const float* l, const float* h) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], l[k]); hi[k] = std::max(hi[k], h[k]); }
    }
    void grow(const AABB& b) { grow(b.lo, b.hi); }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

std::vector<Prim> make_prims(const float* pos, const int32_t* idx, int64_t nt) {
    std::vector<Prim> p(nt);
    for (int64_t i = 0; i < nt; ++i) {
        const float* a = pos + 3 * (int64_t)idx[3 * i];
        const float* b = pos + 3 * (int64_t)idx[3 * i + 1];
        const float* c = pos + 3 * (int64_t)idx[3 * i + 2];
        for (int k = 0; k < 3; ++k) {
            p[i].lo[k] = std::min(a[k], std::min(b[k], c[k]));
            p[i].hi[k] = std::max(a[k], std::max(b[k], c[k]));
            p[i].c[k] = 0.5f * (p[i].lo[k] + p[i].hi[k]);
        }
    }
    return p;
}

// ---------------------------------------------------------------- SAH builder
struct SahBuilder {
    const std::vector<Prim>& prims;
    std::vector<int32_t>& order;
    std::vector<Bvh2Node>& nodes;
    std::atomic<int32_t> next{1};
    int max_leaf;
    static constexpr int kBins = 32;
    static constexpr float kCi = 1.0f, kCt = 1.0f;

    SahBuilder(const std::vector<Prim>& p, std::vector<int32_t>& o, std::vector<Bvh2Node>& n, int ml)
        : prims(p), order(o), nodes(n), max_leaf(ml) {}

    void set_box(Bvh2Node& n, const AABB& b) {
        for (int k = 0; k < 3; ++k) { n.lo[k] = b.lo[k]; n.hi[k] = b.hi[k]; }
    }

    void build(int32_t ni, int32_t begin, int32_t end, int spawn_depth) {
        int32_t n = end - begin;
        AABB box, cbox;
        for (int32_t i = begin; i < end; ++i) {
            const Prim& p = prims[order[i]];
            box.grow(p.lo, p.hi);
            cbox.grow(p.c, p.c);
        }
        set_box(nodes[ni], box);
        if (n <= 1) { make_leaf(ni, begin, n); return; }
        // binned SAH over the three axes
        float best_cost = FLT_MAX;
        int best_axis = -1, best_split = -1;
        float parent_area = box.area();
        for (int axis = 0; axis < 3; ++axis) {
            float ext = cbox.hi[axis] - cbox.lo[axis];
            if (!(ext > 0.0f)) continue;
            float scale = (float)kBins / ext * (1.0f - 1e-6f);
            AABB bb[kBins];
            int cnt[kBins] = {0};
            for (int32_t i = begin; i < end; ++i) {
                const Prim& p = prims[order[i]];
                int b = std::min(kBins - 1, (int)((p.c[axis] - cbox.lo[axis]) * scale));
                cnt[b]++;
                bb[b].grow(p.lo, p.hi);
            }
            float right_area[kBins];
            int right_cnt[kBins];
            AABB acc;
            int c = 0;
            for (int b = kBins - 1; b > 0; --b) {
                acc.grow(bb[b]);
                c += cnt[b];
                right_area[b] = acc.area();
                right_cnt[b] = c;
            }
            AABB lacc;
            int lc = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                lacc.grow(bb[b]);
                lc += cnt[b];
                if (lc == 0 || right_cnt[b + 1] == 0) continue;
                float cost = kCi + kCt * (lacc.area() * lc + right_area[b + 1] * right_cnt[b + 1]) /
                                       std::max(parent_area, 1e-30f);
                if (cost < best_cost) { best_cost = cost; best_axis = axis; best_split = b; }
            }
        }
        float leaf_cost = kCt * n;
        if (n <= max_leaf && (best_axis < 0 || leaf_cost <= best_cost)) { make_leaf(ni, begin, n); return; }
        int32_t mid;
        if (best_axis < 0) {
            // all centroids coincide: split by index
            mid = begin + n / 2;
        } else {
            float ext = cbox.hi[best_axis] - cbox.lo[best_axis];
            float scale = (float)kBins / ext * (1.0f - 1e-6f);
            float lo = cbox.lo[best_axis];
            int ax = best_axis, sp = best_split;
            auto it = std::partition(order.begin() + begin, order.begin() + end, [&](int32_t pi) {
                int b = std::min(kBins - 1, (int)((prims[pi].c[ax] - lo) * scale));
                return b <= sp;
            });
            mid = (int32_t)(it - order.begin());
            if (mid == begin || mid == end) mid = begin + n / 2;
        }
        int32_t l = next.fetch_add(2);
        nodes[ni].left = l;
        nodes[ni].right = l + 1;
        nodes[ni].count = 0;
        if (spawn_depth > 0 && n > 65536) {
            std::thread th([=] { build(l, begin, mid, spawn_depth - 1); });
            build(l + 1, mid, end, spawn_depth - 1);
            th.join();
        } else {
            build(l, begin, mid, 0);
            build(l + 1, mid, end, 0);
        }
    }

    void make_leaf(int32_t ni, int32_t begin, int32_t n) {
        nodes[ni].left = nodes[ni].right = -1;
        nodes[ni].first = begin;
        nodes[ni].count = n;
    }
};

// ---------------------------------------------------------------- reference topology
struct RefBox {
    float mn[3], mx[3];
    float radius() const {
        float e0 = (mx[0] - mn[0]) * 0.5f, e1 = (mx[1] - mn[1]) * 0.5f, e2 = (mx[2] - mn[2]) * 0.5f;
        return (float)std::sqrt((double)(e0 * e0 + e1 * e1 + e2 * e2));
    }
};

inline int trunc_x86(float f) {
    if (!(f > -2147483649.0f && f < 2147483648.0f)) return INT_MIN;
    return (int)f;
}

// 128-bin count heuristic over the node box (TreeMesh.cpp:159-212), used by
// BVHMesh and KDTreeMesh. The reference's O(bins^2) loops sum the same
// integers as these prefix sums.
float count_plane(const std::vector<Prim>& prims, const int32_t* ids, int32_t n, const RefBox& box, int axis) {
    const int Bins = 128;
    int lo_bins[Bins] = {0}, hi_bins[Bins] = {0};
    float extent = box.mx[axis] - box.mn[axis], start = box.mn[axis];
    for (int32_t i = 0; i < n; ++i) {
        const Prim& p = prims[ids[i]];
        int l = trunc_x86(Bins * ((p.lo[axis] - start) / extent));
        int r = trunc_x86(Bins * ((p.hi[axis] - start) / extent));
        if (l >= 0 && l < Bins) lo_bins[l]++;
        if (r >= 0 && r < Bins) hi_bins[r]++;
    }
    int suffix[Bins + 1];
    suffix[Bins] = 0;
    for (int j = Bins - 1; j >= 0; --j) suffix[j] = suffix[j + 1] + hi_bins[j];
    int prefix = 0, best = 0;
    float best_cost = FLT_MAX;
    for (int j = 0; j < Bins; ++j) {
        prefix += lo_bins[j];
        float cost = (float)((1 + suffix[j]) * (Bins - j) + (1 + prefix) * j) / Bins;
        if (cost < best_cost) { best_cost = cost; best = j; }
    }
    return start + (best * extent) / Bins;
}

RefBox bounds_of(const std::vector<Prim>& prims, const int32_t* ids, int32_t n, float pad) {
    RefBox r;
    for (int k = 0; k < 3; ++k) { r.mn[k] = FLT_MAX; r.mx[k] = -FLT_MAX; }
    for (int32_t i = 0; i < n; ++i) {
        const Prim& p = prims[ids[i]];
        for (int k = 0; k < 3; ++k) {
            r.mn[k] = p.lo[k] < r.mn[k] ? p.lo[k] : r.mn[k];
            r.mx[k] = p.hi[k] > r.mx[k] ? p.hi[k] : r.mx[k];
        }
    }
    if (n > 0) for (int k = 0; k < 3; ++k) { r.mn[k] -= pad; r.mx[k] += pad; }
    return r;
}

struct RefBuilder {
    const std::vector<Prim>& prims;
    std::vector<int32_t>& order;
    std::vector<Bvh2Node>& nodes;
    int max_leaf, max_depth;
    float min_width = 0;

    RefBox bounds(int32_t b, int32_t e, float pad) const { return bounds_of(prims, order.data() + b, e - b, pad); }

    float partition_plane(int32_t b, int32_t e, const RefBox& box, int axis) const {
        return count_plane(prims, order.data() + b, e - b, box, axis);
    }

    void build(int32_t ni, int32_t b, int32_t e, const RefBox& box, int axis, int depth) {
        Bvh2Node& n = nodes[ni];
        for (int k = 0; k < 3; ++k) { n.lo[k] = box.mn[k]; n.hi[k] = box.mx[k]; }
        if (e - b <= max_leaf || depth == max_depth || box.radius() <= min_width) {
            n.left = n.right = -1;
            n.first = b;
            n.count = e - b;
            return;
        }
        float plane = partition_plane(b, e, box, axis);
        // midpoint rule, order-preserving (DistributeUnique, TreeMesh.cpp:66-85)
        auto it = std::stable_partition(order.begin() + b, order.begin() + e, [&](int32_t pi) {
            const Prim& p = prims[pi];
            float mid = p.lo[axis] + (p.hi[axis] - p.lo[axis]) * 0.5f;
            return plane >= mid;
        });
        int32_t m = (int32_t)(it - order.begin());
        int32_t l = (int32_t)nodes.size();
        nodes.emplace_back();
        nodes.emplace_back();
        nodes[ni].left = l;
        nodes[ni].right = l + 1;
        nodes[ni].count = 0;
        RefBox lb = bounds(b, m, 0.0f), rb = bounds(m, e, 0.0f);
        int na = (axis + 1) % 3;
        build(l, b, m, lb, na, depth + 1);
        build(l + 1, m, e, rb, na, depth + 1);
    }
};

// ---------------------------------------------------------------- kd-tree topologies
// KDTreeMesh::BuildHierarchy_S2 (KDTreeMesh.cpp:961-1005) and
// KDTreeMeshEx::BuildHierarchy_S2 (KDTreeMeshEx.cpp:68-112) as binary trees
// of cells: a node's box is its cell (not the tight box of its triangles),
// the split is the count-heuristic plane (KDTreeMesh) or the cell centre
// (KDTreeMeshEx), and a triangle goes to every half-cell its box reaches
// (Distribute, TreeMesh.cpp:91-106), so leaves may share triangles.
struct KdBuilder {
    const std::vector<Prim>& prims;
    Bvh2& out;
    int max_leaf, max_depth;
    bool median;
    float min_width = 0;

    void build(int32_t ni, std::vector<int32_t>& list, const RefBox& box, int axis, int depth) {
        for (int k = 0; k < 3; ++k) { out.nodes[ni].lo[k] = box.mn[k]; out.nodes[ni].hi[k] = box.mx[k]; }
        const int32_t n = (int32_t)list.size();
        if (n <= max_leaf || depth == max_depth || box.radius() <= min_width) {
            out.nodes[ni].left = out.nodes[ni].right = -1;
            out.nodes[ni].first = (int32_t)out.order.size();
            out.nodes[ni].count = n;
            if ((int64_t)out.order.size() + n > (int64_t)(1 << 26) - 1)
                throw std::invalid_argument("kd-tree references exceed 2^26 - 1 triangle slots");
            out.order.insert(out.order.end(), list.begin(), list.end());
            return;
        }
        // FindPartitionPlaneSpatialMedian returns the box centre, which the
        // reference keeps as min + (max - min) / 2 (BoundingBox.cpp:421-429).
        float plane = median ? box.mn[axis] + (box.mx[axis] - box.mn[axis]) * 0.5f
                             : count_plane(prims, list.data(), n, box, axis);
        std::vector<int32_t> left, right;
        for (int32_t id : list) {
            const Prim& p = prims[id];
            if (plane >= p.lo[axis]) left.push_back(id);
            if (plane <= p.hi[axis]) right.push_back(id);
        }
        std::vector<int32_t>().swap(list);
        int32_t l = (int32_t)out.nodes.size();
        out.nodes.emplace_back();
        out.nodes.emplace_back();
        out.nodes[ni].left = l;
        out.nodes[ni].right = l + 1;
        out.nodes[ni].count = 0;
        RefBox lb = box, rb = box;
        lb.mx[axis] = plane;
        rb.mn[axis] = plane;
        int na = (axis + 1) % 3;
        build(l, left, lb, na, depth + 1);
        build(l + 1, right, rb, na, depth + 1);
    }
};

}  // namespace

Bvh2 build_kdtree(const float* pos, const int32_t* idx, int64_t nt, int max_leaf, int max_depth, bool median) {
    if (nt <= 0) throw std::invalid_argument("empty mesh");
    if (nt > (1 << 26) - 1) throw std::invalid_argument("more than 2^26 - 1 triangles");
    Bvh2 out;
    std::vector<Prim> prims = make_prims(pos, idx, nt);
    std::vector<int32_t> all(nt);
    for (int64_t i = 0; i < nt; ++i) all[i] = (int32_t)i;
    out.nodes.emplace_back();
    KdBuilder b{prims, out, max_leaf, max_depth, median};
    RefBox tight = bounds_of(prims, all.data(), (int32_t)nt, 0.0f);
    b.min_width = tight.radius() / 1000.0f;                              // KDTreeMesh.cpp:129
    RefBox root = bounds_of(prims, all.data(), (int32_t)nt, 1e-4f);      // KDTreeMesh.cpp:951
    // Vector3::ArgMaxAbsComponent of the half-extent (Vector3.inl:140-149)
    float e[3];
    for (int k = 0; k < 3; ++k) e[k] = (root.mx[k] - root.mn[k]) * 0.5f;
    int axis = 0;
    float mx = std::fabs(e[0]);
    if (std::fabs(e[1]) > mx) { mx = std::fabs(e[1]); axis = 1; }
    if (std::fabs(e[2]) > mx) axis = 2;
    b.build(0, all, root, axis, 0);
    return out;
}

Bvh2 import_kdtree(const uint8_t* tre, size_t bytes, const float* pos, int64_t nv, int64_t nt) {
    if (nv <= 0 || nt <= 0) throw std::invalid_argument("empty mesh");
    if (nt > (1 << 26) - 1) throw std::invalid_argument("more than 2^26 - 1 triangles");
    RefBox root;
    for (int k = 0; k < 3; ++k) { root.mn[k] = FLT_MAX; root.mx[k] = -FLT_MAX; }
    for (int64_t v = 0; v < nv; ++v)
        for (int k = 0; k < 3; ++k) {
            root.mn[k] = std::min(root.mn[k], pos[3 * v + k]);
            root.mx[k] = std::max(root.mx[k], pos[3 * v + k]);
        }
    for (int k = 0; k < 3; ++k) { root.mn[k] -= 1e-4f; root.mx[k] += 1e-4f; }
    auto word = [&](size_t off) {
        uint32_t w;
        std::memcpy(&w, tre + off, 4);
        return w;
    };
    Bvh2 out;
    std::vector<uint8_t> seen((bytes + 3) / 4, 0);
    struct Item { size_t off; int32_t node; RefBox cell; int depth; };
    std::vector<Item> stack;
    out.nodes.emplace_back();
    stack.push_back({0, 0, root, 0});
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        if (it.off % 4 || it.off + 8 > bytes) throw std::invalid_argument("tree node outside the file");
        if (seen[it.off / 4]) throw std::invalid_argument("tree node reached twice");
        seen[it.off / 4] = 1;
        float part;
        std::memcpy(&part, tre + it.off, 4);
        const uint32_t bits = word(it.off + 4);
        const int axis = (int)(bits & 3u);
        const uint32_t count = bits >> 2;
        Bvh2Node& n = out.nodes[it.node];
        for (int k = 0; k < 3; ++k) { n.lo[k] = it.cell.mn[k]; n.hi[k] = it.cell.mx[k]; }
        if (axis == 3) {
            if (it.off + 8 + 4 * (size_t)count > bytes) throw std::invalid_argument("leaf runs past the file");
            n.left = n.right = -1;
            n.first = (int32_t)out.order.size();
            n.count = (int32_t)count;
            if ((int64_t)out.order.size() + count > (int64_t)(1 << 26) - 1)
                throw std::invalid_argument("kd-tree references exceed 2^26 - 1 triangle slots");
            for (uint32_t i = 0; i < count; ++i) {
                const int32_t t = (int32_t)word(it.off + 8 + 4 * (size_t)i);
                if (t < 0 || t >= nt) throw std::invalid_argument("leaf triangle index out of range");
                out.order.push_back(t);
            }
            continue;
        }
        if (it.off + 16 > bytes) throw std::invalid_argument("interior node runs past the file");
        const int32_t c0 = (int32_t)word(it.off + 8), c1 = (int32_t)word(it.off + 12);
        if (c0 < 0 || c1 < 0) throw std::invalid_argument("negative child offset");
        if (!std::isfinite(part)) throw std::invalid_argument("partition is not finite");
        const float cut = std::min(std::max(part, it.cell.mn[axis]), it.cell.mx[axis]);
        RefBox below = it.cell, above = it.cell;
        below.mx[axis] = cut;
        above.mn[axis] = cut;
        const int32_t l = (int32_t)out.nodes.size();
        out.nodes.emplace_back();
        out.nodes.emplace_back();
        out.nodes[it.node].left = l;
        out.nodes[it.node].right = l + 1;
        out.nodes[it.node].count = 0;
        // the reference traverses with a 64-entry stack (PersistentMesh.cpp:189-190)
        if (it.depth + 1 > 64) throw std::invalid_argument("tree deeper than 64 levels");
        stack.push_back({(size_t)c1, l + 1, above, it.depth + 1});
        stack.push_back({(size_t)c0, l, below, it.depth + 1});
    }
    return out;
}

Bvh2 build_sah(const float* pos, const int32_t* idx, int64_t nt, int max_leaf, int threads) {
    if (nt <= 0) throw std::invalid_argument("empty mesh");
    if (nt > (1 << 26) - 1) throw std::invalid_argument("more than 2^26 - 1 triangles");
    Bvh2 out;
    std::vector<Prim> prims = make_prims(pos, idx, nt);
    out.order.resize(nt);
    for (int64_t i = 0; i < nt; ++i) out.order[i] = (int32_t)i;
    out.nodes.resize(2 * nt);
    max_leaf = std::max(1, std::min(max_leaf, kMaxLeafTris));
    SahBuilder b(prims, out.order, out.nodes, max_leaf);
    int spawn = 0;
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    while ((1 << spawn) < threads) ++spawn;
    b.build(0, 0, (int32_t)nt, spawn + 1);
    out.nodes.resize(b.next.load());
    return out;
}

Bvh2 build_reference(const float* pos, const int32_t* idx, int64_t nt, int max_leaf, int max_depth) {
    if (nt <= 0) throw std::invalid_argument("empty mesh");
    Bvh2 out;
    std::vector<Prim> prims = make_prims(pos, idx, nt);
    out.order.resize(nt);
    for (int64_t i = 0; i < nt; ++i) out.order[i] = (int32_t)i;
    out.nodes.reserve(2 * nt / std::max(1, max_leaf) + 16);
    out.nodes.emplace_back();
    RefBuilder b{prims, out.order, out.nodes, max_leaf, max_depth};
    RefBox tight = b.bounds(0, (int32_t)nt, 0.0f);
    b.min_width = tight.radius() / 1000.0f;                 // BVHMesh.cpp:96
    RefBox root = b.bounds(0, (int32_t)nt, 0.0001f);         // BVHMesh.cpp:205
    float sx = (root.mx[0] - root.mn[0]) * 0.5f, sy = (root.mx[1] - root.mn[1]) * 0.5f,
          sz = (root.mx[2] - root.mn[2]) * 0.5f;
    int axis = (sx > sy) ? (sx > sz ? 0 : 2) : (sy > sz ? 1 : 2);
    b.build(0, 0, (int32_t)nt, root, axis, 0);
    return out;
}

// ---------------------------------------------------------------- collapse
namespace {

struct Collapser {
    const Bvh2& b;
    int W;
    std::vector<uint8_t>& out;
    int64_t count = 0;
    int max_depth = 0, max_leaf = 0, stack_need = 0;
    double cost_sum = 0;
    float root_area = 1;
    size_t node_bytes;

    Collapser(const Bvh2& bb, int w, std::vector<uint8_t>& o) : b(bb), W(w), out(o) {
        // the stride the kernels index by: the node struct's own size
        node_bytes = W == 2 ? sizeof(WideNode<2>) : W == 4 ? sizeof(WideNode<4>) : sizeof(WideNode<8>);
    }

    static float area(const float* lo, const float* hi) {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }

    int64_t alloc() {
        int64_t i = count++;
        out.resize((size_t)count * node_bytes);
        return i;
    }

    // float arrays of node i: lox, hix, loy, hiy, loz, hiz; then child words
    float* f(int64_t i) { return reinterpret_cast<float*>(out.data() + (size_t)i * node_bytes); }
    uint32_t* c(int64_t i) { return reinterpret_cast<uint32_t*>(out.data() + (size_t)i * node_bytes + (size_t)W * 24); }

    void set_slot(int64_t ni, int k, const float* lo, const float* hi, uint32_t child) {
        float* p = f(ni);
        p[0 * W + k] = lo[0]; p[1 * W + k] = hi[0];
        p[2 * W + k] = lo[1]; p[3 * W + k] = hi[1];
        p[4 * W + k] = lo[2]; p[5 * W + k] = hi[2];
        c(ni)[k] = child;
    }

    void set_empty(int64_t ni, int k) {
        const float far[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
        set_slot(ni, k, far, far, make_leaf(0, 0));
    }

    // leaf slot for a primitive range; ranges above kMaxLeafTris become a node
    uint32_t leaf_slot(int32_t first, int32_t n, const float* lo, const float* hi, int depth, int& need) {
        if (n <= kMaxLeafTris) {
            max_leaf = std::max(max_leaf, n);
            cost_sum += (double)area(lo, hi) / root_area * n;
            need = 0;
            return make_leaf((uint32_t)first, (uint32_t)n);
        }
        // split into W chunks (boxes are the parent's: conservative)
        int64_t ni = alloc();
        cost_sum += (double)area(lo, hi) / root_area;
        int per = (n + W - 1) / W;
        int used = 0, inner_need = 0;
        for (int k = 0; k < W; ++k) {
            int s = first + k * per, e = std::min(first + n, s + per);
            if (s >= e) { set_empty(ni, k); continue; }
            int cn = 0;
            uint32_t slot = leaf_slot(s, e - s, lo, hi, depth + 1, cn);
            set_slot(ni, k, lo, hi, slot);
            inner_need = std::max(inner_need, cn);
            ++used;
        }
        need = used - 1 + inner_need;
        max_depth = std::max(max_depth, depth + 1);
        return (uint32_t)ni;
    }

    // emits the wide node for binary node bi; returns its index
    int64_t emit(int32_t bi, int depth, int& need) {
        max_depth = std::max(max_depth, depth);
        int32_t kids[8];
        int nk = 0;
        const Bvh2Node& root = b.nodes[bi];
        kids[nk++] = root.left;
        kids[nk++] = root.right;
        while (nk < W) {
            int best = -1;
            float best_a = -1;
            for (int k = 0; k < nk; ++k) {
                const Bvh2Node& n = b.nodes[kids[k]];
                if (n.leaf()) continue;
                float a = area(n.lo, n.hi);
                if (a > best_a) { best_a = a; best = k; }
            }
            if (best < 0) break;
            int32_t open = kids[best];
            kids[best] = b.nodes[open].left;
            kids[nk++] = b.nodes[open].right;
        }
        int64_t ni = alloc();
        cost_sum += (double)area(root.lo, root.hi) / root_area;
        int child_need = 0;
        for (int k = 0; k < W; ++k) {
            if (k >= nk) { set_empty(ni, k); continue; }
            const Bvh2Node& n = b.nodes[kids[k]];
            if (n.leaf() && n.count == 0) { set_empty(ni, k); continue; }  // empty kd cell
            int cn = 0;
            uint32_t slot;
            if (n.leaf()) slot = leaf_slot(n.first, n.count, n.lo, n.hi, depth + 1, cn);
            else slot = (uint32_t)emit(kids[k], depth + 1, cn);
            set_slot(ni, k, n.lo, n.hi, slot);
            child_need = std::max(child_need, cn);
        }
        need = nk - 1 + child_need;
        return ni;
    }
};

}  // namespace

WideBvh collapse_from(const Bvh2& b, int32_t root, int width) {
    if (width != 2 && width != 4 && width != 8) throw std::invalid_argument("width must be 2, 4 or 8");
    WideBvh w;
    w.width = width;
    Collapser c(b, width, w.nodes);
    const Bvh2Node& r = b.nodes[root];
    c.root_area = std::max(Collapser::area(r.lo, r.hi), 1e-30f);
    int need = 0;
    if (r.leaf()) {
        int64_t ni = c.alloc();
        int cn = 0;
        uint32_t slot = c.leaf_slot(r.first, r.count, r.lo, r.hi, 1, cn);
        c.set_slot(ni, 0, r.lo, r.hi, slot);
        for (int k = 1; k < width; ++k) c.set_empty(ni, k);
        need = cn;
    } else {
        c.emit(root, 0, need);
    }
    w.nnodes = c.count;
    w.depth = c.max_depth + 1;
    w.max_leaf = c.max_leaf;
    w.stack_need = need + 1;
    w.sah_cost = c.cost_sum;
    return w;
}

namespace {

struct Item { float lo[3], hi[3]; uint32_t word; };

struct Unwider {
    const WideNode<4>* n;
    Bvh2& b;
    int32_t node_of(uint32_t ni, int depth) {
        if (depth > 128) throw std::invalid_argument("BVH4 deeper than 128 levels");
        Item it[4];
        int k = 0;
        for (int s = 0; s < 4; ++s) {
            const uint32_t w = n[ni].child[s];
            if (w == kEmptySlot) continue;
            it[k].lo[0] = n[ni].lox[s]; it[k].hi[0] = n[ni].hix[s];
            it[k].lo[1] = n[ni].loy[s]; it[k].hi[1] = n[ni].hiy[s];
            it[k].lo[2] = n[ni].loz[s]; it[k].hi[2] = n[ni].hiz[s];
            it[k].word = w;
            ++k;
        }
        if (k == 0) throw std::invalid_argument("BVH4 node without children");
        return items(it, k, depth);
    }
    int32_t items(const Item* it, int k, int depth) {
        if (k == 1) {
            if (is_leaf(it[0].word)) {
                int32_t i = (int32_t)b.nodes.size();
                b.nodes.emplace_back();
                Bvh2Node& nd = b.nodes[i];
                for (int a = 0; a < 3; ++a) { nd.lo[a] = it[0].lo[a]; nd.hi[a] = it[0].hi[a]; }
                nd.first = (int32_t)leaf_first(it[0].word);
                nd.count = (int32_t)leaf_count(it[0].word);
                return i;
            }
            return node_of(it[0].word, depth + 1);
        }
        const int h = k / 2;
        int32_t l = items(it, h, depth + 1), r = items(it + h, k - h, depth + 1);
        int32_t i = (int32_t)b.nodes.size();
        b.nodes.emplace_back();
        Bvh2Node& nd = b.nodes[i];
        for (int a = 0; a < 3; ++a) {
            nd.lo[a] = std::min(b.nodes[l].lo[a], b.nodes[r].lo[a]);
            nd.hi[a] = std::max(b.nodes[l].hi[a], b.nodes[r].hi[a]);
        }
        nd.left = l;
        nd.right = r;
        return i;
    }
};

}  // namespace

Bvh2 binary_from_bvh4(const uint8_t* nodes, const std::vector<uint32_t>& roots, std::vector<int32_t>& bin_roots) {
    Bvh2 b;
    Unwider u{reinterpret_cast<const WideNode<4>*>(nodes), b};
    bin_roots.clear();
    for (uint32_t r : roots) bin_roots.push_back(u.node_of(r, 0));
    return b;
}

WideBvh collapse(const Bvh2& b, int width) {
    if (width != 2 && width != 4 && width != 8) throw std::invalid_argument("width must be 2, 4 or 8");
    if (b.order.size() > (size_t)(1 << 26) - 1) throw std::invalid_argument("more than 2^26 - 1 triangle slots");
    WideBvh w;
    w.width = width;
    w.order = b.order;
    Collapser c(b, width, w.nodes);
    const Bvh2Node& r = b.nodes[0];
    c.root_area = std::max(Collapser::area(r.lo, r.hi), 1e-30f);
    int need = 0;
    if (r.leaf()) {
        int64_t ni = c.alloc();
        int cn = 0;
        uint32_t slot = c.leaf_slot(r.first, r.count, r.lo, r.hi, 1, cn);
        c.set_slot(ni, 0, r.lo, r.hi, slot);
        for (int k = 1; k < width; ++k) c.set_empty(ni, k);
        need = cn;
    } else {
        c.emit(0, 0, need);
    }
    w.nnodes = c.count;
    w.depth = c.max_depth + 1;
    w.max_leaf = c.max_leaf;
    w.stack_need = need + 1;
    w.sah_cost = c.cost_sum;
    return w;
}

void pack_triangles(const float* pos, const int32_t* idx, const std::vector<int32_t>& order, std::vector<float>& out) {
    out.resize(order.size() * 12);
    for (size_t k = 0; k < order.size(); ++k) {
        int32_t t = order[k];
        float* o = out.data() + 12 * k;
        for (int v = 0; v < 3; ++v) {
            const float* p = pos + 3 * (int64_t)idx[3 * (int64_t)t + v];
            o[4 * v + 0] = p[0];
            o[4 * v + 1] = p[1];
            o[4 * v + 2] = p[2];
            o[4 * v + 3] = 0.0f;
        }
        std::memcpy(&o[3], &t, 4);
    }
}

}  // namespace ilm
