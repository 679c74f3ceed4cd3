// CPU ORACLE — test infrastructure only (see oracle.h). Compiled with
// -O3 -march=native -ffp-contract=off, so every float expression is rounded
// exactly as written (the reference's release flags have no fast-math:
// Platform/OSX/build/Core/Makefile:37).
//
// Parity status of each part (how it is pinned, in tests/test_oracle.py):
//   triangle test, slab test ........ closed-form known answers + brute force
//   BVHMesh build ................... invariants (containment, partition,
//                                     leaf bounds, round-robin axes)
//   BVHMesh traversal ............... equality with the brute-force restatement
//   MT19937 ......................... std::mt19937 (same 624-word state)
//   Philox4x64-10 ................... numpy.random.Philox
//   warps / BSDF / light ............ moments, furnace and analytic irradiance
//   PathIntegrator .................. furnace closed form (truncated series)
//   camera .......................... closed-form corner/centre directions
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <random>
#include <shared_mutex>
#include <thread>
#include <vector>

namespace {

// ---------------------------------------------------------------- vectors
struct V3 {
    float x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V3 operator+(V3 a, V3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return V3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, float s) { return V3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(float s, V3 a) { return V3(s * a.x, s * a.y, s * a.z); }
inline V3 operator/(V3 a, float s) { return V3(a.x / s, a.y / s, a.z / s); }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return V3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Vector3::Length = Maths::Sqrt(LengthSquared) (double sqrt, Maths.h:71-74).
inline float length(V3 a) { return (float)std::sqrt((double)dot(a, a)); }
// Vector3::Normalize divides each component by the length (Vector3.inl:354-367).
inline V3 normalize(V3 a) { float l = length(a); return V3(a.x / l, a.y / l, a.z / l); }
// Maths::Min / Max (Maths.h:22-30): argument order decides NaN handling.
template <class T> inline T rmin(T a, T b) { return (a < b) ? a : b; }
template <class T> inline T rmax(T a, T b) { return (a > b) ? a : b; }
inline V3 vmin(V3 a, V3 b) { return V3(rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z)); }
inline V3 vmax(V3 a, V3 b) { return V3(rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z)); }

// Maths constants (Maths.cpp:20-33): derived from atan in double, rounded to float.
const float kPi = (float)(std::atan(1.0f) * 4.0f);
const float kPiTwo = (float)(std::atan(1.0f) * 8.0f);
const float kInvPi = 1.0f / (float)(std::atan(1.0f) * 4.0f);
const float kFltMax = FLT_MAX;

// ---------------------------------------------------------------- ray
struct Ray {
    V3 o, d, inv;  // inv = exact 1/d (Ray.cpp:27-34; SSE rcp path disabled)
    float tmin, tmax;
    Ray() : tmin(0), tmax(kFltMax) {}
    Ray(V3 oo, V3 dd, float mn, float mx) : o(oo), d(dd), tmin(mn), tmax(mx) {
        inv = V3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    }
    V3 at(float t) const { return o + t * d; }  // Ray.inl:32-34
};

// ---------------------------------------------------------------- AABB
// AxisAlignedBoundingBox (Geometry/BoundingBox.h:16-90, .cpp:61-107,287-306,421-429)
struct Box {
    V3 mn, mx, ext, ctr;
    float radius = 0;
    void invalidate() { mn = V3(kFltMax, kFltMax, kFltMax); mx = V3(-kFltMax, -kFltMax, -kFltMax); }
    void update() {
        ext = (mx - mn) * 0.5f;
        radius = length(ext);
        ctr = mn + ext;
    }
    void from_points(const V3* p, int n) {
        invalidate();
        for (int i = 0; i < n; ++i) { mn = vmin(mn, p[i]); mx = vmax(mx, p[i]); }
        update();
    }
    void unite(const Box& b) { mn = vmin(mn, b.mn); mx = vmax(mx, b.mx); update(); }
};

// BoundingBox.cpp:157-240 (slab test on the inverse cache).
inline bool box_hit(const Box& b, const Ray& r, float& tin, float& tout) {
    float rMin = (b.mn.x - r.o.x) * r.inv.x, rMax = (b.mx.x - r.o.x) * r.inv.x;
    float lmin = rmin(rMin, rMax), lmax = rmax(rMin, rMax);
    rMin = (b.mn.y - r.o.y) * r.inv.y; rMax = (b.mx.y - r.o.y) * r.inv.y;
    lmin = rmax(rmin(rMin, rMax), lmin);
    lmax = rmin(rmax(rMin, rMax), lmax);
    rMin = (b.mn.z - r.o.z) * r.inv.z; rMax = (b.mx.z - r.o.z) * r.inv.z;
    lmin = rmax(rmin(rMin, rMax), lmin);
    lmax = rmin(rmax(rMin, rMax), lmax);
    if (lmin > r.tmax || lmax < r.tmin) return false;
    tin = lmin; tout = lmax;
    return (lmax >= 0.0f) & (lmax >= lmin);
}

// ---------------------------------------------------------------- triangle
struct Tri {
    int v[3];
    V3 e1, e2;  // m_edge[0], m_edge[1] (IndexedTriangle.cpp:33-46)
    Box box;    // ComputeBoundingVolume (:75-84)
};

// IndexedTriangle::Intersects (IndexedTriangle.cpp:92-133). The t window is
// compared in double because 1e-5 is a double literal: (Max - 1e-5, 1e-5).
inline bool tri_hit(const V3& p0, const V3& te1, const V3& te2, const Ray& r,
                    float& t_out, float& beta_out, float& gamma_out) {
    V3 n = cross(te1, te2);
    const float rcp = 1.0f / dot(n, r.d);
    V3 edge2 = p0 - r.o;
    const float t = dot(n, edge2) * rcp;
    if ((double)t > (double)r.tmax - 1e-5 || (double)t < 1e-5) return false;
    V3 interm = cross(r.d, edge2);
    const float beta = dot(interm, te2) * -rcp;
    if (beta < 0.0f) return false;
    const float gamma = dot(interm, te1) * rcp;
    if (beta + gamma > 1.0f || gamma < 0.0f) return false;
    t_out = t; beta_out = beta; gamma_out = gamma;
    return true;
}

}  // namespace

// ---------------------------------------------------------------- mesh
struct orc_mesh {
    std::vector<V3> pos;
    std::vector<Tri> tris;
    // BVHMesh nodes (BVHMesh.h:29-51), stored in an array; pre-order.
    struct Node {
        Box box;
        int leaf = 0;
        int axis = 0;
        float partition = 0;
        int child[2] = {-1, -1};
        std::vector<int> list;  // triangle indices (leaves)
    };
    std::vector<Node> nodes;
    int max_leaf = 10, max_depth = 35;
    float min_node_width = 0;
    // KDTreeMesh / KDTreeMeshEx nodes (KDTreeMesh.h:25-100, KDTreeMeshEx.h:19-81):
    // cell boxes, split planes, duplicated triangle references.
    std::vector<Node> kd;
    int kd_kind = 0;      // 1: KDTreeMesh, 2: KDTreeMeshEx
    Box kd_bounds;        // KDTreeMeshEx::m_boundingBox (the padded root cell)
    int64_t kd_refs = 0, kd_leaves = 0, kd_maxdepth = 0;
    int64_t st_internal = 0, st_leaves = 0, st_maxdepth = 0, st_refs = 0, st_maxleaf = 0, st_empty = 0;
};

namespace {

// ITreeMesh::ComputeBounds (TreeMesh.cpp:44-60).
void compute_bounds(const orc_mesh& m, const std::vector<int>& list, Box& b, float emin, float emax) {
    b.invalidate();
    if (!list.empty()) {
        b = m.tris[list[0]].box;
        for (size_t i = 1; i < list.size(); ++i) b.unite(m.tris[list[i]].box);
        b.mn = b.mn - V3(emin, emin, emin); b.update();   // SetMinExtent (BoundingBox.cpp:376-380)
        b.mx = b.mx + V3(emax, emax, emax); b.update();   // SetMaxExtent (:386-390)
    }
}

// Float-to-int conversion as x86 cvttss2si does it: out-of-range and NaN give INT_MIN.
inline int f2i_x86(float f) {
    if (!(f > -2147483649.0f && f < 2147483648.0f)) return INT_MIN;
    return (int)f;
}

// ITreeMesh::FindPartitionPlaneSAH (TreeMesh.cpp:159-212): a 128-bin count heuristic.
float partition_sah(const orc_mesh& m, const std::vector<int>& list, const Box& b, int axis) {
    const int Bins = 128;
    int minBins[Bins], maxBins[Bins];
    for (int j = 0; j < Bins; ++j) minBins[j] = maxBins[j] = 0;
    float extent = b.mx[axis] - b.mn[axis], start = b.mn[axis];
    for (int idx : list) {
        const Box& tb = m.tris[idx].box;
        int left = f2i_x86(Bins * ((tb.mn[axis] - start) / extent));
        int right = f2i_x86(Bins * ((tb.mx[axis] - start) / extent));
        if (left >= 0 && left < Bins) minBins[left]++;
        if (right >= 0 && right < Bins) maxBins[right]++;
    }
    int bestSplit = 0;
    float bestCost = kFltMax;
    // prefix sums give the same integers as the reference's O(Bins^2) loops
    int prefix = 0, suffix[Bins + 1];
    suffix[Bins] = 0;
    for (int j = Bins - 1; j >= 0; --j) suffix[j] = suffix[j + 1] + maxBins[j];
    for (int j = 0; j < Bins; ++j) {
        prefix += minBins[j];
        int leftPrims = 1 + prefix, rightPrims = 1 + suffix[j];
        float cost = (float)(rightPrims * (Bins - j) + leftPrims * j) / Bins;
        if (cost < bestCost) { bestCost = cost; bestSplit = j; }
    }
    return start + (bestSplit * extent) / Bins;
}

// BVHMesh::BuildHierarchy_S2 (BVHMesh.cpp:216-269) with DistributeUnique
// (TreeMesh.cpp:66-85). Iterative over an explicit stack; pre-order ids.
void build_s2(orc_mesh& m, int root, std::vector<int>&& rootlist, int axis0) {
    struct Item { int node; std::vector<int> list; int axis; int depth; };
    std::vector<Item> stack;
    stack.push_back(Item{root, std::move(rootlist), axis0, 0});
    while (!stack.empty()) {
        Item it = std::move(stack.back());
        stack.pop_back();
        m.st_maxdepth = std::max<int64_t>(m.st_maxdepth, it.depth);
        orc_mesh::Node* node = &m.nodes[it.node];
        if ((int)it.list.size() <= m.max_leaf || it.depth == m.max_depth ||
            node->box.radius <= m.min_node_width) {
            node->leaf = 1;
            m.st_leaves++;
            m.st_refs += it.list.size();
            m.st_maxleaf = std::max<int64_t>(m.st_maxleaf, it.list.size());
            if (it.list.empty()) m.st_empty++;
            node->list = std::move(it.list);
            continue;
        }
        node->leaf = 0;
        node->axis = it.axis;
        node->partition = partition_sah(m, it.list, node->box, it.axis);
        std::vector<int> left, right;
        for (int idx : it.list) {
            const Box& tb = m.tris[idx].box;
            float mn = tb.mn[it.axis], mx = tb.mx[it.axis], mid = mn + (mx - mn) * 0.5f;
            if (node->partition >= mid) left.push_back(idx); else right.push_back(idx);
        }
        int c0 = (int)m.nodes.size();
        m.nodes.emplace_back();
        int c1 = (int)m.nodes.size();
        m.nodes.emplace_back();
        node = &m.nodes[it.node];
        node->child[0] = c0; node->child[1] = c1;
        compute_bounds(m, left, m.nodes[c0].box, 0, 0);
        compute_bounds(m, right, m.nodes[c1].box, 0, 0);
        m.st_internal++;
        int na = (it.axis + 1) % 3;
        // right pushed first so that the left subtree is built first
        stack.push_back(Item{c1, std::move(right), na, it.depth + 1});
        stack.push_back(Item{c0, std::move(left), na, it.depth + 1});
    }
}

// BVHMesh::Intersect closest (BVHMesh.cpp:169-198): every child is visited
// (bitwise |), child 0 first; each hit shrinks ray.Max.
bool bvh_closest(const orc_mesh& m, int ni, Ray& r, int& prim, float& t, float& b, float& g) {
    const orc_mesh::Node& n = m.nodes[ni];
    float tin, tout;
    if (!box_hit(n.box, r, tin, tout)) return false;
    if (!n.leaf) {
        bool a = bvh_closest(m, n.child[0], r, prim, t, b, g);
        bool c = bvh_closest(m, n.child[1], r, prim, t, b, g);
        return a | c;
    }
    bool hit = false;
    for (int idx : n.list) {
        const Tri& tr = m.tris[idx];
        float tt, bb, gg;
        if (tri_hit(m.pos[tr.v[0]], tr.e1, tr.e2, r, tt, bb, gg)) {
            hit = true; prim = idx; t = tt; b = bb; g = gg;
            r.tmax = rmin(r.tmax, tt);
        }
    }
    return hit;
}

// BVHMesh::Intersect any (BVHMesh.cpp:143-165): short-circuit, first hit wins.
bool bvh_any(const orc_mesh& m, int ni, const Ray& r, int& prim) {
    const orc_mesh::Node& n = m.nodes[ni];
    float tin, tout;
    if (!box_hit(n.box, r, tin, tout)) return false;
    if (!n.leaf) return bvh_any(m, n.child[0], r, prim) || bvh_any(m, n.child[1], r, prim);
    for (int idx : n.list) {
        const Tri& tr = m.tris[idx];
        float tt, bb, gg;
        if (tri_hit(m.pos[tr.v[0]], tr.e1, tr.e2, r, tt, bb, gg)) { prim = idx; return true; }
    }
    return false;
}

// ---------------------------------------------------------------- kd-trees
// KDTreeMesh::BuildHierarchy_S2 (KDTreeMesh.cpp:961-1005) and
// KDTreeMeshEx::BuildHierarchy_S2 (KDTreeMeshEx.cpp:68-112): leaf when
// n <= MaxLeafObjects, depth == MaxTreeDepth or the cell radius <= root
// radius / 1000; otherwise split the cell at the count heuristic's plane
// (KDTreeMesh) or at its centre (KDTreeMeshEx, FindPartitionPlaneSpatialMedian,
// TreeMesh.cpp:149-153), children are the two half-cells, and every triangle
// whose box reaches a half-cell goes there (Distribute, TreeMesh.cpp:91-106).
void kd_build(orc_mesh& m, int root, std::vector<int>&& rootlist, int axis0, bool median) {
    struct Item { int node; std::vector<int> list; int axis; int depth; };
    std::vector<Item> stack;
    stack.push_back(Item{root, std::move(rootlist), axis0, 0});
    while (!stack.empty()) {
        Item it = std::move(stack.back());
        stack.pop_back();
        m.kd_maxdepth = std::max<int64_t>(m.kd_maxdepth, it.depth);
        orc_mesh::Node* node = &m.kd[it.node];
        if ((int)it.list.size() <= m.max_leaf || it.depth == m.max_depth ||
            node->box.radius <= m.min_node_width) {
            node->leaf = 1;
            m.kd_leaves++;
            m.kd_refs += it.list.size();
            node->list = std::move(it.list);
            continue;
        }
        node->leaf = 0;
        node->axis = it.axis;
        node->partition = median ? node->box.ctr[it.axis] : partition_sah(m, it.list, node->box, it.axis);
        std::vector<int> left, right;
        for (int idx : it.list) {
            const Box& tb = m.tris[idx].box;
            if (node->partition >= tb.mn[it.axis]) left.push_back(idx);
            if (node->partition <= tb.mx[it.axis]) right.push_back(idx);
        }
        std::vector<int>().swap(it.list);
        int c0 = (int)m.kd.size();
        m.kd.emplace_back();
        int c1 = (int)m.kd.size();
        m.kd.emplace_back();
        node = &m.kd[it.node];
        node->child[0] = c0; node->child[1] = c1;
        Box lb = node->box, rb = node->box;
        lb.mx[it.axis] = node->partition; lb.update();  // SetMaxExtent(axis, partition)
        rb.mn[it.axis] = node->partition; rb.update();  // SetMinExtent(axis, partition)
        m.kd[c0].box = lb;
        m.kd[c1].box = rb;
        int na = (it.axis + 1) % 3;
        stack.push_back(Item{c1, std::move(right), na, it.depth + 1});
        stack.push_back(Item{c0, std::move(left), na, it.depth + 1});
    }
}

// KDTreeMesh::Intersect_New, closest (KDTreeMesh.cpp:467-577) and any
// (:313-388): ordered traversal with a todo stack; a node is left when the
// current hit is nearer than its entry distance.
bool kd_trace(const orc_mesh& m, const Ray& in, bool any, int& prim, float& t, float& b, float& g) {
    float tmin, tmax;
    if (!box_hit(m.kd[0].box, in, tmin, tmax)) return false;
    Ray ray(in);
    struct Todo { int node; float mn, mx; };
    Todo todo[64];
    int todoPos = 0;
    bool hit = false;
    int ni = 0;
    while (ni >= 0) {
        if (ray.tmax < tmin) break;
        const orc_mesh::Node& node = m.kd[ni];
        if (!node.leaf) {
            int axis = node.axis;
            float tplane = (node.partition - ray.o[axis]) * ray.inv[axis];
            bool belowFirst = (ray.o[axis] < node.partition) ||
                              (ray.o[axis] == node.partition && ray.d[axis] <= 0.0f);
            int first = belowFirst ? node.child[0] : node.child[1];
            int second = belowFirst ? node.child[1] : node.child[0];
            if (tplane > tmax || tplane <= 0) {
                ni = first;
            } else if (tplane < tmin) {
                ni = second;
            } else {
                if (todoPos == 64) std::abort();  // unreachable: depth <= 64 (orc_kd_build)
                todo[todoPos++] = Todo{second, tplane, tmax};
                ni = first;
                tmax = tplane;
            }
        } else {
            for (int idx : node.list) {
                const Tri& tr = m.tris[idx];
                float tt, bb, gg;
                if (tri_hit(m.pos[tr.v[0]], tr.e1, tr.e2, ray, tt, bb, gg)) {
                    prim = idx;
                    if (any) return true;
                    hit = true; t = tt; b = bb; g = gg;
                    ray.tmax = rmin(ray.tmax, tt);
                }
            }
            if (todoPos > 0) {
                --todoPos;
                ni = todo[todoPos].node;
                tmin = todo[todoPos].mn;
                tmax = todo[todoPos].mx;
            } else {
                break;
            }
        }
    }
    return hit;
}

// KDTreeMeshEx::Intersects_Stack, closest (KDTreeMeshEx.cpp:170-272) and any
// (:276-390). Each leaf resets the ray window to (entry, tTemp), where tTemp
// starts at the exit of the root cell: the caller's ray.Max is not kept, so a
// segment ray can report a hit beyond its end (a property of the reference
// traversal, restated as is).
bool kdex_trace(const orc_mesh& m, const Ray& in, bool any, int& prim, float& t, float& b, float& g) {
    float a, bb;
    if (!box_hit(m.kd_bounds, in, a, bb)) return false;
    Ray ray(in);
    float tTemp = bb;
    bool ret = false;
    struct Elem { int node; float a, b; };
    Elem st[30];
    unsigned sp = 0;
    st[sp++] = Elem{0, a, bb};
    while (sp) {
        const Elem& e = st[--sp];
        int cur = e.node;
        a = e.a;
        bb = e.b;
        while (!m.kd[cur].leaf) {
            const orc_mesh::Node& n = m.kd[cur];
            float diff = n.partition - ray.o[n.axis];
            if (diff == 0) {
                cur = ray.d[n.axis] > 0 ? n.child[1] : n.child[0];
                continue;
            }
            int nearc = diff > 0 ? n.child[0] : n.child[1];
            int farc = diff > 0 ? n.child[1] : n.child[0];
            if (ray.d[n.axis] != 0) {
                float tSplit = diff / ray.d[n.axis];
                if (tSplit > bb || tSplit < 0) {
                    cur = nearc;
                } else if (tSplit < a) {
                    cur = farc;
                } else {
                    if (sp + 1 >= 30) std::abort();  // unreachable: depth < 29 (orc_kd_build)
                    st[sp++] = Elem{farc, tSplit, bb};
                    cur = nearc;
                    bb = tSplit;
                }
            } else {
                cur = nearc;
            }
        }
        const orc_mesh::Node& leaf = m.kd[cur];
        if (!leaf.list.empty()) {
            ray.tmin = a;  // ignored by the triangle test
            ray.tmax = tTemp;
            for (int idx : leaf.list) {
                const Tri& tr = m.tris[idx];
                float tt, b2, g2;
                if (tri_hit(m.pos[tr.v[0]], tr.e1, tr.e2, ray, tt, b2, g2)) {
                    prim = idx;
                    if (any) return true;
                    ray.tmax = tTemp = rmin(tTemp, tt);
                    t = tt; b = b2; g = g2;
                    ret = true;
                }
            }
        }
        if (tTemp <= bb && tTemp >= a) return any ? false : ret;
    }
    return false;
}

// BasicMesh::Intersects (BasicMesh.cpp:16-30 closest, :47-56 any).
bool brute_closest(const orc_mesh& m, Ray r, int& prim, float& t, float& b, float& g) {
    bool hit = false;
    for (size_t i = 0; i < m.tris.size(); ++i) {
        const Tri& tr = m.tris[i];
        float tt, bb, gg;
        if (tri_hit(m.pos[tr.v[0]], tr.e1, tr.e2, r, tt, bb, gg)) {
            hit = true; prim = (int)i; t = tt; b = bb; g = gg;
            r.tmax = rmin(r.tmax, tt);
        }
    }
    return hit;
}
bool brute_any(const orc_mesh& m, const Ray& r, int& prim) {
    for (size_t i = 0; i < m.tris.size(); ++i) {
        const Tri& tr = m.tris[i];
        float tt, bb, gg;
        if (tri_hit(m.pos[tr.v[0]], tr.e1, tr.e2, r, tt, bb, gg)) { prim = (int)i; return true; }
    }
    return false;
}

bool mesh_closest(const orc_mesh& m, const Ray& ray, int accel, int& prim, float& t, float& b, float& g) {
    if (accel == 2 && m.kd_kind == 1) return kd_trace(m, ray, false, prim, t, b, g);
    if (accel == 2 && m.kd_kind == 2) return kdex_trace(m, ray, false, prim, t, b, g);
    if (accel == 0 || m.nodes.empty()) return brute_closest(m, ray, prim, t, b, g);
    Ray r(ray);  // BVHMesh::Intersects copies the ray (BVHMesh.cpp:114-119)
    return bvh_closest(m, 0, r, prim, t, b, g);
}
bool mesh_any(const orc_mesh& m, const Ray& ray, int accel, int& prim) {
    float t, b, g;
    if (accel == 2 && m.kd_kind == 1) return kd_trace(m, ray, true, prim, t, b, g);
    if (accel == 2 && m.kd_kind == 2) return kdex_trace(m, ray, true, prim, t, b, g);
    if (accel == 0 || m.nodes.empty()) return brute_any(m, ray, prim);
    return bvh_any(m, 0, ray, prim);
}

int resolve_threads(int n) {
    if (n > 0) return n;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)h : 1;
}

template <class F>
void parallel_chunks(int64_t n, int64_t chunk, int nthreads, F&& f) {
    std::atomic<int64_t> next(0);
    auto worker = [&]() {
        for (;;) {
            int64_t s = next.fetch_add(chunk);
            if (s >= n) break;
            f(s, std::min(n, s + chunk));
        }
    };
    if (nthreads <= 1) { worker(); return; }
    std::vector<std::thread> th;
    for (int i = 0; i < nthreads; ++i) th.emplace_back(worker);
    for (auto& x : th) x.join();
}

// ---------------------------------------------------------------- RNGs
// MT19937 (Core/Maths/Random.cpp:36-110, constants Random.h:13-22).
struct MT {
    uint32_t s[624];
    int i;
    explicit MT(uint32_t seed) {
        s[0] = seed;
        for (i = 1; i < 624; ++i) s[i] = 69069u * s[i - 1];
    }
    uint32_t next() {
        static const uint32_t mag[2] = {0u, 0x9908b0dfu};
        uint32_t y;
        if (i >= 624) {
            int k = 0;
            for (; k < 624 - 397; ++k) {
                y = (s[k] & 0x80000000u) | (s[k + 1] & 0x7fffffffu);
                s[k] = s[k + 397] ^ (y >> 1) ^ mag[y & 1u];
            }
            for (; k < 623; ++k) {
                y = (s[k] & 0x80000000u) | (s[k + 1] & 0x7fffffffu);
                s[k] = s[k + (397 - 624)] ^ (y >> 1) ^ mag[y & 1u];
            }
            y = (s[623] & 0x80000000u) | (s[0] & 0x7fffffffu);
            s[623] = s[396] ^ (y >> 1) ^ mag[y & 1u];
            i = 0;
        }
        y = s[i++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 15;
        return y;
    }
};

// Random::NextFloat (Random.cpp:117-120): u32 scaled by 1/(2^32-1) in double,
// rounded to float, times (1 - FLT_EPSILON).
inline float u32_to_float(uint32_t u) {
    return ((float)(u * 2.3283064370807974e-10)) * (1.0f - FLT_EPSILON);
}

inline void mulhilo64(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
    unsigned __int128 p = (unsigned __int128)a * b;
    hi = (uint64_t)(p >> 64); lo = (uint64_t)p;
}

void philox4x64(const uint64_t in[4], const uint64_t k_in[2], uint64_t out[4]) {
    uint64_t c0 = in[0], c1 = in[1], c2 = in[2], c3 = in[3];
    uint64_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B97F4A7C15ull; k1 += 0xBB67AE8584CAA73Bull; }
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ull, c0, hi0, lo0);
        mulhilo64(0xCA5A826395121157ull, c2, hi1, lo1);
        uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Counter-based sample block: 8 floats for (pixel, sample, bounce, stream).
struct Block8 { float v[8]; };
inline Block8 philox_block(uint32_t seed, uint64_t pixel, uint32_t sample, uint32_t bounce, uint32_t stream) {
    uint64_t ctr[4] = {pixel, ((uint64_t)bounce << 32) | sample, stream, 0};
    uint64_t key[2] = {seed, 0};
    uint64_t o[4];
    philox4x64(ctr, key, o);
    Block8 b;
    for (int i = 0; i < 4; ++i) {
        b.v[2 * i] = u32_to_float((uint32_t)o[i]);
        b.v[2 * i + 1] = u32_to_float((uint32_t)(o[i] >> 32));
    }
    return b;
}

// ---------------------------------------------------------------- warps
// Montecarlo::CosineSampleHemisphere (Montecarlo.h:30-37); Maths::Cos/Sin/Sqrt go via double.
inline V3 cosine_hemisphere(float u, float v) {
    float a = kPiTwo * v;
    float b = (float)std::sqrt((double)u);
    return V3(b * (float)std::cos((double)a), b * (float)std::sin((double)a), (float)std::sqrt((double)(1 - u)));
}
// Montecarlo::UniformSampleTriangle (Montecarlo.h:60-64)
inline void uniform_triangle(float u, float v, float& b1, float& b2) {
    float su1 = (float)std::sqrt((double)u);
    b1 = 1.f - su1;
    b2 = v * su1;
}

// OrthonormalBasis::InitFromW (Basis.cpp:85-99)
struct Basis { V3 U, V, W; };
inline Basis basis_from_w(V3 w) {
    Basis b;
    b.W = normalize(w);
    if (std::fabs(b.W.x) > std::fabs(b.W.y)) {
        float invLen = 1.f / (float)std::sqrt((double)(b.W.x * b.W.x + b.W.z * b.W.z));
        b.V = V3(-b.W.z * invLen, 0.f, b.W.x * invLen);
    } else {
        float invLen = 1.f / (float)std::sqrt((double)(b.W.y * b.W.y + b.W.z * b.W.z));
        b.V = V3(0.f, b.W.z * invLen, -b.W.y * invLen);
    }
    b.U = cross(b.W, b.V);
    return b;
}
inline V3 to_local(const Basis& b, V3 v) { return V3(dot(b.U, v), dot(b.V, v), dot(b.W, v)); }  // BSDF.cpp:37-52
inline V3 to_world(const Basis& b, V3 v) { return v.x * b.U + v.y * b.V + v.z * b.W; }          // Basis.cpp:205-208

// ---------------------------------------------------------------- camera
struct Camera { V3 centre, corner, across, up; };
// PerspectiveCamera ctor (PerspectiveCamera.cpp:15-25) + ICamera::SetFieldOfView
// (Camera.cpp:30-44) + OrthonormalBasis::FrowWV (Basis.cpp:28-44). Defaults of the
// factory (CameraFactories.h:46-49): distance 1, frame [-1,1]^2 replaced by the fov.
Camera make_camera(const float c[11]) {
    V3 pos(c[0], c[1], c[2]), tgt(c[3], c[4], c[5]), upv(c[6], c[7], c[8]);
    float fov = c[9], aspect = c[10];
    V3 dir = normalize(tgt - pos);
    Basis b;
    b.W = normalize(dir);
    b.V = normalize(upv);
    b.V = b.V - dot(b.V, b.W) * b.W;
    if (dot(b.V, b.V) <= FLT_EPSILON) {
        b.V = V3(b.W.z, b.W.x, b.W.y);
        b.V = normalize(cross(b.V, b.W));
        b.V = normalize(b.V - dot(b.V, b.W) * b.W);
    }
    b.U = cross(b.V, b.W);
    const float degToRad = (float)std::atan(1.0f) / 45.0f;  // Maths.cpp:38
    float half = fov * degToRad * 0.5f;
    float v1 = (float)std::tan((double)half), v0 = -v1;
    float u0 = v0 * aspect, u1 = v1 * aspect, d = 1.0f;
    Camera cam;
    cam.centre = pos;
    cam.corner = pos + u0 * b.U + v0 * b.V + d * b.W;
    cam.across = (u1 - u0) * b.U;
    cam.up = (v1 - v0) * b.V;
    return cam;
}
// PerspectiveCamera::GetRay (PerspectiveCamera.cpp:53-60), fed as in
// BaseRenderer::RenderRegion (BaseRenderer.cpp:79-81,106-107,117-118): the
// sample (u,v) in [0,1) is shifted by -0.5 (BoxFilter.cpp:20-28).
inline Ray camera_ray(const Camera& c, int W, int H, int px, int py, float u, float v) {
    float rcpW = 1.f / W, rcpH = 1.f / H;
    float su = u - 0.5f, sv = v - 0.5f;
    float nx = (float)px * rcpW, ny = (float)py * rcpH;
    float xi1 = su * rcpW, xi2 = sv * rcpH;
    V3 target = c.corner + c.across * (nx + xi1) + c.up * (ny + xi2);
    return Ray(c.centre, normalize(target - c.centre), 0.0f, kFltMax);
}

}  // namespace

// ---------------------------------------------------------------- scene
struct orc_scene {
    orc_mesh* mesh = nullptr;
    std::vector<V3> nrm;
    std::vector<int> tri_mat, tri_light;
    struct Mat { V3 kd, ks; float exponent; int type; };
    std::vector<Mat> mats;
    // quad: sampled as the reference's Quad (Shape/Quad.cpp:146-153) from its
    // two triangles (v1, v3, v2) and (v1, v4, v3); otherwise as ITriangleMesh
    struct Light { V3 Le; std::vector<int> tris; float area = 0; bool quad = false; };
    std::vector<Light> lights;
};

namespace {

inline float luminance(V3 s) { return 0.212671f * s.x + 0.715160f * s.y + 0.072169f * s.z; }  // RGBSpectrum.h:153-155
inline bool is_black(V3 s) { return s.x == 0 && s.y == 0 && s.z == 0; }
inline V3 mul(V3 a, V3 b) { return V3(a.x * b.x, a.y * b.y, a.z * b.z); }

// Phong lobe of the glossy material (no reference counterpart): normalised
// modified Phong, f = ks (n+2)/(2 pi) cos^n(alpha), alpha measured from the
// mirror direction of wOut about the local z axis; sampled by cos^n.
inline V3 phong_f(const orc_scene::Mat& m, V3 wo, V3 wi) {
    V3 r(-wo.x, -wo.y, wo.z);
    float c = dot(r, wi);
    if (c <= 0.f) return V3();
    float s = (m.exponent + 2.f) * (0.5f * kInvPi) * std::pow(c, m.exponent);
    return m.ks * s;
}
inline float phong_pdf(const orc_scene::Mat& m, V3 wo, V3 wi) {
    V3 r(-wo.x, -wo.y, wo.z);
    float c = dot(r, wi);
    if (c <= 0.f) return 0.f;
    return (m.exponent + 1.f) * (0.5f * kInvPi) * std::pow(c, m.exponent);
}

// BSDF::F (BSDF.cpp:215-237): reflection BxDFs only when wIn.z > 0.
inline V3 bsdf_f(const orc_scene::Mat& m, V3 wo, V3 wi) {
    if (!(wi.z > 0)) return V3();
    V3 f = m.kd * kInvPi;  // Lambertian::F (LambertianBxDF.cpp:41-44)
    if (m.type == 1) f = f + phong_f(m, wo, wi);
    return f;
}

// BSDF::SampleF (BSDF.cpp:141-211) over {Lambertian, [Phong]}.
inline V3 bsdf_sample(const orc_scene::Mat& m, V3 wo, V3& wi, float u, float v, float w, float& pdf) {
    int count = (m.type == 1) ? 2 : 1;
    int which = (int)(w * count) % count;
    if (which == 0) {
        // Lambertian::SampleF (LambertianBxDF.cpp:27-37)
        wi = cosine_hemisphere(u, v);
        if (wo.z < 0) wi.z = -wi.z;
        pdf = std::fabs(wi.z) * kInvPi;
    } else {
        V3 r(-wo.x, -wo.y, wo.z);
        float ct = std::pow(u, 1.f / (m.exponent + 1.f));
        float st = (float)std::sqrt((double)rmax(0.f, 1.f - ct * ct));
        float ph = kPiTwo * v;
        V3 l(st * (float)std::cos((double)ph), st * (float)std::sin((double)ph), ct);
        wi = to_world(basis_from_w(r), l);
        pdf = phong_pdf(m, wo, wi);
    }
    if (pdf == 0.0f) return V3();
    if (count > 1) {
        if (which == 0) pdf += phong_pdf(m, wo, wi);
        else pdf += std::fabs(wi.z) * kInvPi;
        pdf /= count;
    }
    return bsdf_f(m, wo, wi);
}

// Samples for one path; mode 0 = Philox (counter-based), mode 1 = MT stream.
struct Sampler {
    int mode;
    uint32_t seed;
    uint64_t pixel;
    uint32_t sample;
    MT* mt;
    MT* light_mt;
    Block8 blk;
    void begin(uint32_t bounce) { if (mode == 0) blk = philox_block(seed, pixel, sample, bounce, 0); }
    float cam(int i) { return mode == 0 ? blk.v[i] : u32_to_float(mt->next()); }
    // BSDF direction (2), BSDF select (1), roulette (1)
    float dir(int i) { return mode == 0 ? blk.v[2 + i] : u32_to_float(mt->next()); }
    float select() { return mode == 0 ? blk.v[4] : u32_to_float(mt->next()); }
    float rr() { return mode == 0 ? blk.v[5] : u32_to_float(mt->next()); }
    void light(uint32_t bounce, int l, float& u, float& v, float& pick) {
        if (mode == 0) {
            Block8 b = philox_block(seed, pixel, sample, bounce, 1u + (uint32_t)l);
            u = b.v[0]; v = b.v[1]; pick = b.v[2];
        } else {
            u = u32_to_float(mt->next()); v = u32_to_float(mt->next());
            pick = u32_to_float(light_mt->next());  // the mesh's own generator (TriangleMesh.cpp:145)
        }
    }
};

struct Hit {
    int prim = -1;
    float t = 0, b = 0, g = 0;
};

struct Surface {
    V3 P, Ng, Ns;
    int mat, light;
};

inline Surface make_surface(const orc_scene& s, const Ray& r, const Hit& h) {
    const orc_mesh& m = *s.mesh;
    const Tri& tr = m.tris[h.prim];
    Surface sf;
    sf.P = r.at(h.t);  // GeometricPrimitive.cpp:151
    float alpha = 1.f - h.b - h.g;
    const V3& n0 = s.nrm[tr.v[0]]; const V3& n1 = s.nrm[tr.v[1]]; const V3& n2 = s.nrm[tr.v[2]];
    // IndexedTriangle.cpp:225-243
    sf.Ns = normalize(V3(n0.x * alpha + n1.x * h.b + n2.x * h.g,
                         n0.y * alpha + n1.y * h.b + n2.y * h.g,
                         n0.z * alpha + n1.z * h.b + n2.z * h.g));
    sf.Ng = normalize(cross(tr.e1, tr.e2));
    sf.mat = s.tri_mat[h.prim];
    sf.light = s.tri_light[h.prim];
    return sf;
}

// DiffuseAreaLight::SampleRadiance (DiffuseAreaLight.cpp:52-69) through
// SamplePoint (:72-100) and ITriangleMesh::SamplePoint (TriangleMesh.cpp:143-161),
// IndexedTriangle::SamplePoint (IndexedTriangle.cpp:400-414).
// Quad::SamplePoint (Shape/Quad.cpp:146-153) over the tessellated triangles
// T0 = (v1, v3, v2), T1 = (v1, v4, v3): T0.V0 + (T0.V1 - T0.V0) u +
// (T1.V1 - T0.V1) v + (T0.V0 - T0.V1 + T0.V2 - T1.V1) u v, normal
// cross(T0.V1 - T0.V0, T0.V2 - T0.V0), normalised by DiffuseAreaLight::
// SamplePoint (DiffuseAreaLight.cpp:95-97) and not turned toward the viewer.
inline V3 quad_sample(V3 v1, V3 v2, V3 v3, V3 v4, float u, float v, V3& n) {
    n = normalize(cross(v3 - v1, v2 - v1));
    return v1 + (v3 - v1) * u + (v4 - v3) * v + (v1 - v3 + v2 - v4) * (u * v);
}

inline V3 light_sample(const orc_scene& s, int l, V3 P, float u, float v, float pick, V3& wIn, Ray& seg) {
    const orc_scene::Light& L = s.lights[l];
    V3 Q, nl;
    if (L.quad) {
        const Tri& t0 = s.mesh->tris[L.tris[0]];
        const Tri& t1 = s.mesh->tris[L.tris[1]];
        Q = quad_sample(s.mesh->pos[t0.v[0]], s.mesh->pos[t0.v[2]], s.mesh->pos[t0.v[1]], s.mesh->pos[t1.v[1]], u, v, nl);
    } else {
        int n = (int)L.tris.size();
        int k = (int)rmin((size_t)(n * pick), (size_t)(n - 1));
        const Tri& tr = s.mesh->tris[L.tris[k]];
        float b1, b2;
        uniform_triangle(u, v, b1, b2);
        const V3& p1 = s.mesh->pos[tr.v[0]]; const V3& p2 = s.mesh->pos[tr.v[1]]; const V3& p3 = s.mesh->pos[tr.v[2]];
        Q = b1 * p1 + b2 * p2 + (1.f - b1 - b2) * p3;
        nl = normalize(cross(p3 - p1, p2 - p1));
        if (dot(Q - P, nl) > 0) nl = -nl;
        nl = normalize(nl);
    }
    float pdf = 1.0f / L.area;
    seg = Ray(P, Q - P, 1e-4f, 1.0f - 1e-4f);  // Visibility.cpp:50-53, Ray::Epsilon
    wIn = P - Q;
    float d2 = dot(wIn, wIn);
    wIn = normalize(wIn);
    float c = rmax(0.f, dot(wIn, nl));
    V3 num = L.Le * c;
    float den = pdf * d2;
    return num / den;
}

struct PathStats { int64_t ext = 0, shadow = 0; };

// PathIntegrator::Radiance (PathIntegrator.cpp:44-201,357-367) with the helpers
// SampleAllLights / EstimateDirectLighting (Integrator.cpp:114-185).
V3 radiance(const orc_scene& s, Ray ray, Sampler& smp, const orc_render_params& p, PathStats& st) {
    const int accel = 1;
    V3 L;
    Hit h;
    st.ext++;
    if (!mesh_closest(*s.mesh, ray, accel, h.prim, h.t, h.b, h.g)) return L;  // area lights add nothing on a miss
    V3 beta(1.f, 1.f, 1.f);
    for (int depth = 0; depth < p.max_depth; ++depth) {
        Surface sf = make_surface(s, ray, h);
        if (sf.mat < 0) break;
        const orc_scene::Mat& mat = s.mats[sf.mat];
        if (depth > 0) smp.begin((uint32_t)depth);
        V3 wOut = -normalize(ray.d);
        // emission at depth 0 (no specular BxDFs here), DiffuseAreaLight::Radiance (:47-50)
        if (depth == 0 && sf.light >= 0) {
            if (dot(sf.Ns, wOut) > 0) L = L + mul(beta, s.lights[sf.light].Le);
        }
        // SampleAllLights, excluding the light that was hit (ShadowRays = 1)
        V3 Ls;
        Basis gb = basis_from_w(sf.Ng);
        for (int l = 0; l < (int)s.lights.size(); ++l) {
            if (l == sf.light) continue;  // the excluded light draws no samples (Integrator.cpp:160-163)
            float lu, lv, pick;
            smp.light((uint32_t)depth, l, lu, lv, pick);
            V3 wIn;
            Ray seg;
            V3 Le = light_sample(s, l, sf.P, lu, lv, pick, wIn, seg);
            if (is_black(Le)) continue;
            st.shadow++;
            int op;
            if (mesh_any(*s.mesh, seg, accel, op)) continue;
            wIn = -wIn;
            V3 f = bsdf_f(mat, to_local(gb, wOut), to_local(gb, wIn));
            V3 c = Le * rmax(0.f, dot(wIn, sf.Ns));
            Ls = Ls + mul(c, f);
        }
        L = L + mul(beta, Ls / 1.0f);
        // BSDF sampling in the shading frame
        float du = smp.dir(0), dv = smp.dir(1), dw = smp.select();
        Basis sb = basis_from_w(sf.Ns);
        V3 woL = to_local(sb, wOut), wiL;
        float pdf = 0;
        V3 f = bsdf_sample(mat, woL, wiL, du, dv, dw, pdf);
        if (is_black(f) || pdf == 0.0f) break;
        V3 wIn = to_world(sb, wiL);
        V3 tmp = f * std::fabs(dot(wIn, sf.Ns));
        tmp = tmp / pdf;
        beta = mul(beta, tmp);
        if (depth >= p.rr_depth) {
            float q = rmin(0.5f, luminance(beta));
            if (smp.rr() > q) break;
            beta = beta / q;
        }
        // the reference traces this ray also after the last shaded vertex; its
        // hit is never used, so the restatement stops here (same image)
        if (depth + 1 >= p.max_depth) break;
        ray = Ray(sf.P + sf.Ng * p.eps, wIn, p.eps, kFltMax);
        st.ext++;
        if (!mesh_closest(*s.mesh, ray, accel, h.prim, h.t, h.b, h.g)) break;
    }
    return L;
}

}  // namespace

// ====================================================================== C API
extern "C" {

int orc_tri_intersect(const float p0[3], const float p1[3], const float p2[3],
                      const float o[3], const float d[3], float ray_max,
                      float* t, float* beta, float* gamma) {
    V3 a(p0[0], p0[1], p0[2]), b(p1[0], p1[1], p1[2]), c(p2[0], p2[1], p2[2]);
    Ray r(V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), 0.0f, ray_max);
    float tt = 0, bb = 0, gg = 0;
    bool hit = tri_hit(a, b - a, c - a, r, tt, bb, gg);
    if (hit) { *t = tt; *beta = bb; *gamma = gg; }
    return hit ? 1 : 0;
}

int orc_aabb_intersect(const float bmin[3], const float bmax[3], const float o[3], const float d[3],
                       float ray_min, float ray_max, float* tin, float* tout) {
    Box b;
    b.mn = V3(bmin[0], bmin[1], bmin[2]);
    b.mx = V3(bmax[0], bmax[1], bmax[2]);
    b.update();
    Ray r(V3(o[0], o[1], o[2]), V3(d[0], d[1], d[2]), ray_min, ray_max);
    float a = 0, c = 0;
    bool hit = box_hit(b, r, a, c);
    *tin = a; *tout = c;
    return hit ? 1 : 0;
}

void orc_mt_words(uint32_t seed, uint32_t* out, int n) {
    MT mt(seed);
    for (int i = 0; i < n; ++i) out[i] = mt.next();
}
void orc_mt_floats(uint32_t seed, float* out, int n) {
    MT mt(seed);
    for (int i = 0; i < n; ++i) out[i] = u32_to_float(mt.next());
}
void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) { philox4x64(ctr, key, out); }
void orc_cosine_hemisphere(float u, float v, float out[3]) {
    V3 w = cosine_hemisphere(u, v);
    out[0] = w.x; out[1] = w.y; out[2] = w.z;
}
void orc_uniform_triangle(float u, float v, float* b1, float* b2) { uniform_triangle(u, v, *b1, *b2); }
void orc_quad_sample(const float quad[12], float u, float v, float point[3], float normal[3]) {
    V3 n;
    V3 p = quad_sample(V3(quad[0], quad[1], quad[2]), V3(quad[3], quad[4], quad[5]), V3(quad[6], quad[7], quad[8]),
                       V3(quad[9], quad[10], quad[11]), u, v, n);
    point[0] = p.x; point[1] = p.y; point[2] = p.z;
    normal[0] = n.x; normal[1] = n.y; normal[2] = n.z;
}

void orc_camera_ray(const float cam[11], int width, int height, int px, int py, float u, float v,
                    float o[3], float d[3]) {
    Camera c = make_camera(cam);
    Ray r = camera_ray(c, width, height, px, py, u, v);
    o[0] = r.o.x; o[1] = r.o.y; o[2] = r.o.z;
    d[0] = r.d.x; d[1] = r.d.y; d[2] = r.d.z;
}

orc_mesh* orc_mesh_create(const float* pos, int64_t nv, const int32_t* idx, int64_t nt) {
    orc_mesh* m = new orc_mesh();
    m->pos.resize(nv);
    for (int64_t i = 0; i < nv; ++i) m->pos[i] = V3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    m->tris.resize(nt);
    for (int64_t i = 0; i < nt; ++i) {
        Tri& t = m->tris[i];
        for (int k = 0; k < 3; ++k) t.v[k] = idx[3 * i + k];
        t.e1 = m->pos[t.v[1]] - m->pos[t.v[0]];
        t.e2 = m->pos[t.v[2]] - m->pos[t.v[0]];
        V3 p[3] = {m->pos[t.v[0]], m->pos[t.v[1]], m->pos[t.v[2]]};
        t.box.from_points(p, 3);
    }
    return m;
}

void orc_mesh_destroy(orc_mesh* m) { delete m; }

int64_t orc_bvh_build(orc_mesh* m, int max_leaf, int max_depth) {
    m->nodes.clear();
    m->max_leaf = max_leaf;
    m->max_depth = max_depth;
    m->st_internal = m->st_leaves = m->st_maxdepth = m->st_refs = m->st_maxleaf = m->st_empty = 0;
    std::vector<int> list(m->tris.size());
    for (size_t i = 0; i < list.size(); ++i) list[i] = (int)i;
    m->nodes.emplace_back();
    Box rootb;
    compute_bounds(*m, list, rootb, 0, 0);               // BVHMesh.cpp:92
    m->min_node_width = rootb.radius / 1000.0f;           // :96
    compute_bounds(*m, list, m->nodes[0].box, 0.0001f, 0.0001f);  // BuildHierarchy :205
    V3 size = m->nodes[0].box.ext;
    int axis;
    if (size.x > size.y) axis = size.x > size.z ? 0 : 2;
    else axis = size.y > size.z ? 1 : 2;
    build_s2(*m, 0, std::move(list), axis);
    return (int64_t)m->nodes.size();
}

int64_t orc_kd_build(orc_mesh* m, int kind, int max_leaf, int max_depth) {
    if (kind != 1 && kind != 2) return -1;
    // Traversal stacks: 64 entries (KDTreeMesh) and 30 (KDTreeMeshEx, which
    // throws when it fills); deeper trees are refused here.
    if ((kind == 1 && max_depth > 64) || (kind == 2 && max_leaf >= 29)) return -1;
    m->kd.clear();
    m->kd_kind = kind;
    m->kd_refs = m->kd_leaves = m->kd_maxdepth = 0;
    m->max_leaf = max_leaf;
    // KDTreeMeshEx's constructors set the depth limit from the leaf size
    // (KDTreeMeshEx.cpp:36-44); KDTreeMesh keeps MaximumTreeDepth.
    m->max_depth = kind == 2 ? max_leaf : max_depth;
    std::vector<int> list(m->tris.size());
    for (size_t i = 0; i < list.size(); ++i) list[i] = (int)i;
    Box tight;
    compute_bounds(*m, list, tight, 0, 0);                 // Compile (KDTreeMesh.cpp:125, KDTreeMeshEx.cpp:138)
    m->min_node_width = tight.radius / 1000.0f;
    m->kd.emplace_back();
    compute_bounds(*m, list, m->kd[0].box, 1e-4f, 1e-4f);  // BuildHierarchy (KDTreeMesh.cpp:951, KDTreeMeshEx.cpp:61)
    m->kd_bounds = m->kd[0].box;
    // Vector3::ArgMaxAbsComponent of the half-extent (Vector3.inl:140-149)
    V3 e = m->kd[0].box.ext;
    int axis = 0;
    float mx = std::fabs(e.x);
    if (std::fabs(e.y) > mx) { mx = std::fabs(e.y); axis = 1; }
    if (std::fabs(e.z) > mx) axis = 2;
    kd_build(*m, 0, std::move(list), axis, kind == 2);
    return (int64_t)m->kd.size();
}

void orc_kd_stats(const orc_mesh* m, int64_t out[4]) {
    out[0] = (int64_t)m->kd.size(); out[1] = m->kd_leaves; out[2] = m->kd_maxdepth; out[3] = m->kd_refs;
}

void orc_kd_export(const orc_mesh* m, float* box, int32_t* child, float* split, int32_t* leaf_first,
                   int32_t* leaf_count, int32_t* refs) {
    int32_t r = 0;
    for (size_t i = 0; i < m->kd.size(); ++i) {
        const orc_mesh::Node& n = m->kd[i];
        box[6 * i + 0] = n.box.mn.x; box[6 * i + 1] = n.box.mn.y; box[6 * i + 2] = n.box.mn.z;
        box[6 * i + 3] = n.box.mx.x; box[6 * i + 4] = n.box.mx.y; box[6 * i + 5] = n.box.mx.z;
        child[2 * i] = n.leaf ? -1 : n.child[0];
        child[2 * i + 1] = n.leaf ? -1 : n.child[1];
        split[2 * i] = n.leaf ? 0.0f : (float)n.axis;
        split[2 * i + 1] = n.leaf ? 0.0f : n.partition;
        leaf_first[i] = r;
        leaf_count[i] = n.leaf ? (int32_t)n.list.size() : 0;
        if (n.leaf) for (int x : n.list) refs[r++] = x;
    }
}

void orc_bvh_stats(const orc_mesh* m, int64_t out[6]) {
    out[0] = m->st_internal; out[1] = m->st_leaves; out[2] = m->st_maxdepth;
    out[3] = m->st_refs; out[4] = m->st_maxleaf; out[5] = m->st_empty;
}

void orc_bvh_export(const orc_mesh* m, float* box, int32_t* child, int32_t* leaf_first,
                    int32_t* leaf_count, int32_t* refs) {
    int32_t r = 0;
    for (size_t i = 0; i < m->nodes.size(); ++i) {
        const orc_mesh::Node& n = m->nodes[i];
        box[6 * i + 0] = n.box.mn.x; box[6 * i + 1] = n.box.mn.y; box[6 * i + 2] = n.box.mn.z;
        box[6 * i + 3] = n.box.mx.x; box[6 * i + 4] = n.box.mx.y; box[6 * i + 5] = n.box.mx.z;
        child[2 * i] = n.leaf ? -1 : n.child[0];
        child[2 * i + 1] = n.leaf ? -1 : n.child[1];
        leaf_first[i] = r;
        leaf_count[i] = n.leaf ? (int32_t)n.list.size() : 0;
        if (n.leaf) for (int x : n.list) refs[r++] = x;
    }
}

int orc_trace(const orc_mesh* m, const float* rays, int64_t n, int mode, int accel,
              int32_t* prim, float* t, float* bary, int nthreads) {
    int nt = resolve_threads(nthreads);
    parallel_chunks(n, 1024, nt, [&](int64_t s, int64_t e) {
        for (int64_t i = s; i < e; ++i) {
            const float* q = rays + 8 * i;
            Ray r(V3(q[0], q[1], q[2]), V3(q[4], q[5], q[6]), q[3], q[7]);
            int p = -1;
            float tt = 0, bb = 0, gg = 0;
            if (mode == 0) {
                if (!mesh_closest(*m, r, accel, p, tt, bb, gg)) p = -1;
            } else {
                if (!mesh_any(*m, r, accel, p)) p = -1;
            }
            prim[i] = p;
            if (t) t[i] = tt;
            if (bary) { bary[2 * i] = bb; bary[2 * i + 1] = gg; }
        }
    });
    return nt;
}

// ---------------------------------------------------------------- spaces
// BasicSpace over GeometricPrimitives (Space/BasicSpace.cpp:58-114,
// Scene/GeometricPrimitive.cpp:68-200): each primitive is a mesh placed by a
// Transformation (rotation elements in the reference's Matrix3x3 storage
// order, scaling, translation; Geometry/Transform.cpp:133-199).
struct orc_space {
    struct Prim {
        const orc_mesh* mesh;
        float rot[9], scl[3], trn[3];
        bool has_rot, has_scl, has_trn;
    };
    std::vector<Prim> prims;
};

namespace {

// Ray::ApplyInverse (Geometry/Ray.cpp:50-60) with Transformation::RotateInverse,
// ScaleInverse and ApplyInverse (Transform.cpp:147-176,196-199); the rotation
// inverse is the transpose (Transform.cpp:42), so Product(inverse, v) is the
// row-major product of the stored elements with v.
Ray inverse_ray(const orc_space::Prim& p, const Ray& r) {
    auto rot_inv = [&](V3 v) {
        return V3(p.rot[0] * v.x + p.rot[1] * v.y + p.rot[2] * v.z,
                  p.rot[3] * v.x + p.rot[4] * v.y + p.rot[5] * v.z,
                  p.rot[6] * v.x + p.rot[7] * v.y + p.rot[8] * v.z);
    };
    V3 d = rot_inv(r.d);
    if (p.has_scl) d = V3(d.x / p.scl[0], d.y / p.scl[1], d.z / p.scl[2]);
    V3 o = rot_inv(r.o - V3(p.trn[0], p.trn[1], p.trn[2]));
    float yz = p.scl[1] * p.scl[2], xz = p.scl[0] * p.scl[2], xy = p.scl[0] * p.scl[1];
    float ixyz = 1 / (xy * p.scl[2]);
    o.x *= (ixyz * yz);
    o.y *= (ixyz * xz);
    o.z *= (ixyz * xy);
    return Ray(o, d, r.tmin, r.tmax);
}

bool prim_closest(const orc_space::Prim& p, const Ray& r, int& tri, float& t) {
    float b, g;
    int accel = p.mesh->nodes.empty() ? 0 : 1;
    if (!(p.has_rot || p.has_scl || p.has_trn)) return mesh_closest(*p.mesh, r, accel, tri, t, b, g);
    return mesh_closest(*p.mesh, inverse_ray(p, r), accel, tri, t, b, g);
}

bool prim_any(const orc_space::Prim& p, const Ray& r, int& tri) {
    int accel = p.mesh->nodes.empty() ? 0 : 1;
    if (!(p.has_rot || p.has_scl || p.has_trn)) return mesh_any(*p.mesh, r, accel, tri);
    return mesh_any(*p.mesh, inverse_ray(p, r), accel, tri);
}

}  // namespace

orc_space* orc_space_create(void) { return new orc_space(); }
void orc_space_destroy(orc_space* s) { delete s; }

int orc_space_add(orc_space* s, const orc_mesh* m, const float* rot, const float* scl, const float* trn) {
    orc_space::Prim p;
    p.mesh = m;
    const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, one[3] = {1, 1, 1}, zero[3] = {0, 0, 0};
    std::memcpy(p.rot, rot ? rot : I, sizeof p.rot);
    std::memcpy(p.scl, scl ? scl : one, sizeof p.scl);
    std::memcpy(p.trn, trn ? trn : zero, sizeof p.trn);
    // Transformation::SetRotation / SetScaling / SetTranslation flags (Transform.cpp:48-50)
    p.has_rot = std::memcmp(p.rot, I, sizeof I) != 0;
    p.has_scl = !(p.scl[0] == 1 && p.scl[1] == 1 && p.scl[2] == 1);
    p.has_trn = !(p.trn[0] == 0 && p.trn[1] == 0 && p.trn[2] == 0);
    s->prims.push_back(p);
    return (int)s->prims.size() - 1;
}

int orc_space_trace(const orc_space* s, const float* rays, int64_t n, int mode, int32_t* prim, int32_t* tri,
                    float* t, int nthreads) {
    int nt = resolve_threads(nthreads);
    parallel_chunks(n, 1024, nt, [&](int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            const float* q = rays + 8 * i;
            Ray ray(V3(q[0], q[1], q[2]), V3(q[4], q[5], q[6]), q[3], q[7]);
            int hp = -1, ht = -1;
            float tt = 0;
            for (int k = 0; k < (int)s->prims.size(); ++k) {
                int tr = -1;
                float tk = 0;
                if (mode == 0) {
                    if (prim_closest(s->prims[k], ray, tr, tk)) {   // BasicSpace.cpp:58-73
                        hp = k; ht = tr; tt = tk;
                        ray.tmax = rmin(ray.tmax, tk);
                    }
                } else if (prim_any(s->prims[k], ray, tr)) {         // BasicSpace.cpp:92-101
                    hp = k; ht = tr;
                    break;
                }
            }
            prim[i] = hp;
            tri[i] = ht;
            if (t) t[i] = tt;
        }
    });
    return nt;
}

orc_scene* orc_scene_create(const float* pos, const float* nrm, int64_t nv, const int32_t* idx,
                            const int32_t* tri_mat, const int32_t* tri_light, int64_t nt,
                            const float* mats, int nmat, const float* lights, int nlight) {
    orc_scene* s = new orc_scene();
    s->mesh = orc_mesh_create(pos, nv, idx, nt);
    orc_bvh_build(s->mesh, 10, 35);  // BVHMesh factory defaults (ShapeFactories.h:311-315)
    s->nrm.resize(nv);
    for (int64_t i = 0; i < nv; ++i) s->nrm[i] = V3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]);
    s->tri_mat.assign(tri_mat, tri_mat + nt);
    s->tri_light.assign(tri_light, tri_light + nt);
    for (int i = 0; i < nmat; ++i) {
        const float* q = mats + 8 * i;
        s->mats.push_back(orc_scene::Mat{V3(q[0], q[1], q[2]), V3(q[3], q[4], q[5]), q[6], (int)q[7]});
    }
    s->lights.resize(nlight);
    for (int l = 0; l < nlight; ++l) {
        s->lights[l].Le = V3(lights[4 * l], lights[4 * l + 1], lights[4 * l + 2]);
        s->lights[l].quad = lights[4 * l + 3] == 1.0f;
    }
    for (int64_t i = 0; i < nt; ++i) {
        int l = tri_light[i];
        if (l >= 0 && l < nlight) {
            s->lights[l].tris.push_back((int)i);
            const Tri& t = s->mesh->tris[i];
            // IndexedTriangle::GetArea (IndexedTriangle.cpp:386-393), summed (TriangleMesh.cpp:163-169)
            s->lights[l].area += length(cross(s->mesh->pos[t.v[1]] - s->mesh->pos[t.v[0]],
                                              s->mesh->pos[t.v[2]] - s->mesh->pos[t.v[0]])) * 0.5f;
        }
    }
    return s;
}

void orc_scene_destroy(orc_scene* s) {
    if (!s) return;
    orc_mesh_destroy(s->mesh);
    delete s;
}

int orc_render(const orc_scene* s, const orc_render_params* p, double* accum, int64_t ray_counts[2],
               int nthreads) {
    const int W = p->width, H = p->height;
    int x0 = p->x0, y0 = p->y0, x1 = p->x1 ? p->x1 : W, y1 = p->y1 ? p->y1 : H;
    Camera cam = make_camera(p->cam);
    const int tile = 32;  // default tile size (Sandbox/main.cpp:248)
    int tx = (x1 - x0 + tile - 1) / tile, ty = (y1 - y0 + tile - 1) / tile;
    int nt = resolve_threads(nthreads);
    std::atomic<int64_t> ext(0), shd(0);
    parallel_chunks((int64_t)tx * ty, 1, nt, [&](int64_t a, int64_t b) {
        for (int64_t ti = a; ti < b; ++ti) {
            int bx = x0 + (int)(ti % tx) * tile, by = y0 + (int)(ti / tx) * tile;
            MT mt(p->seed ^ (uint32_t)(ti * 2654435761u));
            MT lmt(p->seed + 5489u + (uint32_t)ti);
            PathStats st;
            for (int y = by; y < std::min(by + tile, y1); ++y) {
                for (int x = bx; x < std::min(bx + tile, x1); ++x) {
                    double acc[3] = {0, 0, 0};
                    for (int si = 0; si < p->spp; ++si) {
                        Sampler smp{};
                        smp.mode = p->rng_mode;
                        smp.seed = p->seed;
                        smp.pixel = (uint64_t)p->pix0 + (uint64_t)y * W + x;
                        smp.sample = (uint32_t)(p->spp_offset + si);
                        smp.mt = &mt;
                        smp.light_mt = &lmt;
                        smp.begin(0);
                        float u = smp.cam(0), v = smp.cam(1);
                        Ray r = camera_ray(cam, W, H, x, y, u, v);
                        V3 L = radiance(*s, r, smp, *p, st);
                        acc[0] += L.x; acc[1] += L.y; acc[2] += L.z;
                    }
                    double* o = accum + 3 * ((int64_t)y * W + x);
                    o[0] += acc[0]; o[1] += acc[1]; o[2] += acc[2];
                }
            }
            ext += st.ext;
            shd += st.shadow;
        }
    });
    if (ray_counts) { ray_counts[0] = ext.load(); ray_counts[1] = shd.load(); }
    return nt;
}


// ---------------------------------------------------------------- irradiance cache integrator
// ICIntegrator (Integrator/ICIntegrator.cpp) on the host: the CPU comparator
// of the neural radiance cache. Radiance (:348-424) shades a camera hit with
// SampleAllLights plus GetIrradiance * albedo * indirect_factor; GetIrradiance
// (:438-484) averages the records FindRecords (:142-183) accepts with W_Ward
// (:185-196), or computes a record (ComputeRecord, :486-543: AltitudeStrata x
// AzimuthStrata stratified cosine gather rays from P + w eps, each followed by
// PathLi, :545-642) and inserts it (Insert, :107-140). The octree's root cell
// and RMin/RMax follow Initialise (:306-330). Random numbers come from the
// counter-based streams (a record's gather path k of a camera sample draws
// from bounce counter 64 + k * 32 + path length), not the shared sampler.
namespace {

struct IcRec { V3 P, N, E; float Ri, RiClamp; };
struct IcNode { V3 lo, hi; int child = -1; std::vector<int> recs; };

struct IcCache {
    std::vector<IcNode> nodes;
    std::vector<IcRec> recs;
    int depth = 16;
    float inv_threshold = 1.f;
    std::shared_mutex mu;

    static bool contains(const IcNode& n, V3 p) {
        return !(n.lo.x > p.x || n.lo.y > p.y || n.lo.z > p.z || n.hi.x < p.x || n.hi.y < p.y || n.hi.z < p.z);
    }
    // IrradianceCache::SetBounds (:54-86): child i takes the upper half in x
    // when bit 0 is set, in z for bit 1, in y for bit 2
    static void child_bounds(const IcNode& n, int i, V3& lo, V3& hi) {
        V3 c = n.lo + (n.hi - n.lo) * 0.5f;
        lo = V3((i & 1) ? c.x : n.lo.x, (i & 4) ? c.y : n.lo.y, (i & 2) ? c.z : n.lo.z);
        hi = V3((i & 1) ? n.hi.x : c.x, (i & 4) ? n.hi.y : c.y, (i & 2) ? n.hi.z : c.z);
    }
    static bool sphere_box(V3 lo, V3 hi, V3 c, float r) {  // SphereBoxOverlap (:88-101)
        float d = 0;
        const float cc[3] = {c.x, c.y, c.z}, ll[3] = {lo.x, lo.y, lo.z}, hh[3] = {hi.x, hi.y, hi.z};
        for (int i = 0; i < 3; ++i) {
            if (cc[i] < ll[i]) d += (cc[i] - ll[i]) * (cc[i] - ll[i]);
            else if (cc[i] > hh[i]) d += (cc[i] - hh[i]) * (cc[i] - hh[i]);
        }
        return d <= r * r;
    }
    void insert(int n, int r, int d) {
        const IcRec& rec = recs[r];
        V3 ext = (nodes[n].hi - nodes[n].lo) * 0.5f;   // AxisAlignedBoundingBox::GetExtent
        float half = rmax(std::fabs(ext.x), rmax(std::fabs(ext.y), std::fabs(ext.z)));
        if (d <= 0 || rec.RiClamp > half) {
            nodes[n].recs.push_back(r);
            return;
        }
        if (nodes[n].child < 0) {
            int first = (int)nodes.size();
            nodes[n].child = first;
            for (int i = 0; i < 8; ++i) {
                IcNode c;
                child_bounds(nodes[n], i, c.lo, c.hi);
                nodes.push_back(c);
            }
        }
        const int first = nodes[n].child;
        for (int i = 0; i < 8; ++i)
            if (sphere_box(nodes[first + i].lo, nodes[first + i].hi, rec.P, rec.RiClamp)) insert(first + i, r, d - 1);
    }
    float weight(V3 p, V3 nrm, const IcRec& r) const {  // W_Ward
        float dist = length(p - r.P) / r.RiClamp, c = dot(nrm, r.N);
        if (c <= 0) return -1.f;
        return 1.f / (dist + 1.f - c) - inv_threshold;
    }
    bool lookup(V3 p, V3 nrm, V3& E) const {
        int n = 0;
        if (!contains(nodes[0], p)) return false;
        double num[3] = {0, 0, 0}, den = 0;
        for (;;) {
            for (int r : nodes[n].recs) {
                float w = weight(p, nrm, recs[r]);
                if (w > 0.f) {
                    num[0] += w * recs[r].E.x; num[1] += w * recs[r].E.y; num[2] += w * recs[r].E.z;
                    den += w;
                }
            }
            const int c = nodes[n].child;
            if (c < 0) break;
            int k = 0;
            while (k < 8 && !contains(nodes[c + k], p)) ++k;
            if (k == 8) break;
            n = c + k;
        }
        if (den <= 0) return false;
        E = V3((float)(num[0] / den), (float)(num[1] / den), (float)(num[2] / den));
        return true;
    }
};

struct IcStats { int64_t camera = 0, gather = 0, shadow = 0, created = 0; };

// SampleAllLights at a surface, excluding the light that was hit (Integrator.cpp:114-185)
V3 direct_light(const orc_scene& s, const Surface& sf, const orc_scene::Mat& mat, V3 wOut, uint32_t seed,
                uint64_t pixel, uint32_t sample, uint32_t bounce, int64_t& shadow) {
    V3 Ls;
    Basis gb = basis_from_w(sf.Ng);
    for (int l = 0; l < (int)s.lights.size(); ++l) {
        if (l == sf.light) continue;
        Block8 b = philox_block(seed, pixel, sample, bounce, 1u + (uint32_t)l);
        V3 wIn;
        Ray seg;
        V3 Le = light_sample(s, l, sf.P, b.v[0], b.v[1], b.v[2], wIn, seg);
        if (is_black(Le)) continue;
        ++shadow;
        int op;
        if (mesh_any(*s.mesh, seg, 1, op)) continue;
        wIn = -wIn;
        V3 f = bsdf_f(mat, to_local(gb, wOut), to_local(gb, wIn));
        Ls = Ls + mul(Le * rmax(0.f, dot(wIn, sf.Ns)), f);
    }
    return Ls;
}

// ICIntegrator::PathLi (:545-642): NEE at every vertex, no emission (no
// specular bounces), the next ray from P with tmin eps, roulette after the
// third vertex with q = min(0.5, 0.33 (r + g + b)). first_t = the first hit's
// distance (1 when the gather ray finds no material, as PathLi leaves it).
V3 path_li(const orc_scene& s, Ray ray, int ray_depth, float eps, uint32_t seed, uint64_t pixel, uint32_t sample,
           uint32_t ctr0, float& first_t, IcStats& st) {
    V3 L, beta(1.f, 1.f, 1.f);
    first_t = 1.f;
    for (int len = 0;; ++len) {
        Hit h;
        ++st.gather;
        if (!mesh_closest(*s.mesh, ray, 1, h.prim, h.t, h.b, h.g)) break;
        Surface sf = make_surface(s, ray, h);
        if (sf.mat < 0) break;
        if (len == 0) first_t = h.t;
        const orc_scene::Mat& mat = s.mats[sf.mat];
        V3 wOut = -ray.d;
        L = L + mul(beta, direct_light(s, sf, mat, wOut, seed, pixel, sample, ctr0 + (uint32_t)len, st.shadow));
        if (len + 1 == ray_depth) break;
        Block8 b = philox_block(seed, pixel, sample, ctr0 + (uint32_t)len, 0);
        Basis sb = basis_from_w(sf.Ns);
        V3 woL = to_local(sb, wOut), wiL;
        float pdf = 0;
        V3 f = bsdf_sample(mat, woL, wiL, b.v[2], b.v[3], b.v[4], pdf);
        if (is_black(f) || pdf == 0.0f) break;
        V3 wIn = to_world(sb, wiL);
        ray = Ray(sf.P, wIn, eps, kFltMax);
        beta = mul(beta, f * std::fabs(dot(wIn, sf.Ns)) / pdf);
        if (len > 2) {
            float q = rmin(0.5f, 0.33f * (beta.x + beta.y + beta.z));
            if (b.v[5] > q) break;
            beta = beta / q;
        }
    }
    return L;
}

}  // namespace

extern "C" int orc_ic_render(const orc_scene* s, const orc_render_params* p, const orc_ic_params* icp,
                             double* accum, int64_t counts[4], int nthreads) {
    const int W = p->width, H = p->height;
    int x0 = p->x0, y0 = p->y0, x1 = p->x1 ? p->x1 : W, y1 = p->y1 ? p->y1 : H;
    Camera cam = make_camera(p->cam);
    const orc_mesh& m = *s->mesh;
    // Initialise (:306-330): root cell from the scene's bounds
    V3 lo(kFltMax, kFltMax, kFltMax), hi(-kFltMax, -kFltMax, -kFltMax);
    for (const V3& v : m.pos) {
        lo = V3(rmin(lo.x, v.x), rmin(lo.y, v.y), rmin(lo.z, v.z));
        hi = V3(rmax(hi.x, v.x), rmax(hi.y, v.y), rmax(hi.z, v.z));
    }
    V3 size = hi - lo, centre = lo + size * 0.5f;
    float edge = rmax(std::fabs(size.x), rmax(std::fabs(size.y), std::fabs(size.z)));
    V3 a = (centre - V3(edge, edge, edge)) * (1.0f + FLT_EPSILON), b = (centre + V3(edge, edge, edge)) * (1.0f + FLT_EPSILON);
    IcCache cache;
    IcNode root;
    root.lo = V3(rmin(a.x, b.x), rmin(a.y, b.y), rmin(a.z, b.z));
    root.hi = V3(rmax(a.x, b.x), rmax(a.y, b.y), rmax(a.z, b.z));
    cache.nodes.push_back(root);
    cache.depth = icp->cache_depth;
    cache.inv_threshold = 1.f / icp->error_threshold;
    const float rmin_ = edge * icp->ambient_resolution, rmax_ = edge * icp->ambient_multiplier;
    const int M = icp->altitude_strata, N = icp->azimuth_strata, mn = M * N;
    const int tile = 32;
    int tx = (x1 - x0 + tile - 1) / tile, ty = (y1 - y0 + tile - 1) / tile;
    int nt = resolve_threads(nthreads);
    std::atomic<int64_t> c_cam(0), c_gather(0), c_shadow(0), c_created(0);
    parallel_chunks((int64_t)tx * ty, 1, nt, [&](int64_t t0, int64_t t1) {
        IcStats st;
        for (int64_t ti = t0; ti < t1; ++ti) {
            int bx = x0 + (int)(ti % tx) * tile, by = y0 + (int)(ti / tx) * tile;
            for (int y = by; y < std::min(by + tile, y1); ++y) {
                for (int x = bx; x < std::min(bx + tile, x1); ++x) {
                    const uint64_t pixel = (uint64_t)p->pix0 + (uint64_t)y * W + x;
                    double acc[3] = {0, 0, 0};
                    for (int si = 0; si < p->spp; ++si) {
                        const uint32_t sample = (uint32_t)(p->spp_offset + si);
                        Block8 cb = philox_block(p->seed, pixel, sample, 0, 0);
                        Ray r = camera_ray(cam, W, H, x, y, cb.v[0], cb.v[1]);
                        ++st.camera;
                        Hit h;
                        V3 L;
                        if (!mesh_closest(m, r, 1, h.prim, h.t, h.b, h.g)) {
                            L = V3(icp->miss_radiance[0], icp->miss_radiance[1], icp->miss_radiance[2]);
                        } else {
                            Surface sf = make_surface(*s, r, h);
                            if (sf.mat >= 0) {
                                const orc_scene::Mat& mat = s->mats[sf.mat];
                                V3 wOut = -normalize(r.d);
                                if (sf.light >= 0) {
                                    if (dot(sf.Ns, wOut) > 0) L = s->lights[sf.light].Le;
                                } else {
                                    L = direct_light(*s, sf, mat, wOut, p->seed, pixel, sample, 0, st.shadow);
                                    V3 rho = mat.type == 1 ? mat.kd + mat.ks : mat.kd;
                                    V3 E;
                                    bool found;
                                    {
                                        std::shared_lock<std::shared_mutex> g(cache.mu);
                                        found = cache.lookup(sf.P, sf.Ns, E);
                                    }
                                    if (!found) {
                                        // ComputeRecord (:486-543)
                                        V3 sum;
                                        float inv_len = 0;
                                        Basis sb = basis_from_w(sf.Ns);
                                        for (int ai = 0; ai < M; ++ai) {
                                            for (int zi = 0; zi < N; ++zi) {
                                                const int k = ai * N + zi;
                                                Block8 gb = philox_block(p->seed, pixel, sample, 63, 1u + (uint32_t)k);
                                                // Montecarlo::CosineSampleHemisphere with strata (Montecarlo.h:20-29)
                                                float su = (gb.v[0] + (float)ai) / (float)M;
                                                float sv = (gb.v[1] + (float)zi) / (float)N;
                                                float ph = kPiTwo * sv, rr = std::sqrt(su);
                                                V3 vh(rr * std::cos(ph), rr * std::sin(ph), std::sqrt(rmax(0.f, 1.f - su)));
                                                V3 w = to_world(sb, vh);
                                                float ft;
                                                sum = sum + path_li(*s, Ray(sf.P + w * icp->eps, w, icp->eps, kFltMax),
                                                                    icp->ray_depth, icp->eps, p->seed, pixel, sample,
                                                                    64u + 32u * (uint32_t)k, ft, st);
                                                inv_len += 1.f / ft;
                                            }
                                        }
                                        IcRec rec;
                                        rec.P = sf.P;
                                        rec.N = sf.Ns;
                                        rec.E = sum / (float)mn;
                                        rec.Ri = (float)mn / inv_len;
                                        rec.RiClamp = rmax(rmin_, rmin(rmax_, rec.Ri));
                                        E = rec.E;
                                        ++st.created;
                                        std::unique_lock<std::shared_mutex> g(cache.mu);
                                        cache.recs.push_back(rec);
                                        cache.insert(0, (int)cache.recs.size() - 1, cache.depth);
                                    }
                                    L = L + mul(E, rho) * icp->indirect_factor;
                                }
                            }
                        }
                        acc[0] += L.x; acc[1] += L.y; acc[2] += L.z;
                    }
                    double* o = accum + 3 * ((int64_t)y * W + x);
                    o[0] += acc[0]; o[1] += acc[1]; o[2] += acc[2];
                }
            }
        }
        c_cam += st.camera;
        c_gather += st.gather;
        c_shadow += st.shadow;
        c_created += st.created;
    });
    if (counts) {
        counts[0] = c_cam.load();
        counts[1] = c_gather.load();
        counts[2] = c_shadow.load();
        counts[3] = c_created.load();
    }
    return nt;
}

// Local-frame BSDF of a material record (kd rgb, ks rgb, exponent, type):
// f, and the pdf BSDF::SampleF reports for a direction it sampled
// (BSDF.cpp:141-211: the chosen lobe's pdf plus the other lobe's Pdf,
// divided by the lobe count; LambertianBxDF.cpp:52-55).
void orc_bsdf_eval(const float mat[8], int64_t n, const float* wo, const float* wi, float* f, float* pdf) {
    orc_scene::Mat m{V3(mat[0], mat[1], mat[2]), V3(mat[3], mat[4], mat[5]), mat[6], (int)mat[7]};
    for (int64_t k = 0; k < n; ++k) {
        V3 o(wo[3 * k], wo[3 * k + 1], wo[3 * k + 2]), i(wi[3 * k], wi[3 * k + 1], wi[3 * k + 2]);
        V3 v = bsdf_f(m, o, i);
        f[3 * k] = v.x; f[3 * k + 1] = v.y; f[3 * k + 2] = v.z;
        float pl = std::fabs(i.z) * kInvPi;
        pdf[k] = m.type == 1 ? 0.5f * (pl + phong_pdf(m, o, i)) : pl;
    }
}

// BSDF::SampleF over n (wo, u, v, w): wi, f and pdf as the integrator uses them.
void orc_bsdf_sample(const float mat[8], int64_t n, const float* wo, const float* uvw, float* wi, float* f,
                     float* pdf) {
    orc_scene::Mat m{V3(mat[0], mat[1], mat[2]), V3(mat[3], mat[4], mat[5]), mat[6], (int)mat[7]};
    for (int64_t k = 0; k < n; ++k) {
        V3 o(wo[3 * k], wo[3 * k + 1], wo[3 * k + 2]), i;
        float pd = 0.f;
        V3 v = bsdf_sample(m, o, i, uvw[3 * k], uvw[3 * k + 1], uvw[3 * k + 2], pd);
        wi[3 * k] = i.x; wi[3 * k + 1] = i.y; wi[3 * k + 2] = i.z;
        f[3 * k] = v.x; f[3 * k + 1] = v.y; f[3 * k + 2] = v.z;
        pdf[k] = pd;
    }
}

}  // extern "C"
