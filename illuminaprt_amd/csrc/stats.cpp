// Traversal statistics on the host: the node visits and triangle tests of a
// deterministic ordered traversal of a mesh's flattened tree (the same float
// BVH2/4/8 or quantised BVH4 nodes and leaf-order triangles the traversal
// kernels read). Used to price a traversal launch by algorithmic bytes per
// ray (32 + 16 + N_node * S_node + N_tri * S_tri); it produces no hits.
//
// The traversal is the kernels' without speculation: at an inner node the
// children's slab entry distances are computed, the nearest child is visited
// next and the others are pushed farthest first; a leaf tests its triangles
// with the watertight test, a closest hit shrinks the t window, an any hit
// ends the ray.
#include <stdint.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "geom.hpp"
#include "internal.hpp"

namespace ilm {
namespace {

struct TriSoA {
    std::vector<f3> v0, v1, v2;
};

// Child entry distances of a float node (slab test; +inf: missed or empty).
template <int W>
void child_dists(const WideNode<W>& n, const RayState& r, float (&d)[W], uint32_t (&c)[W]) {
    for (int k = 0; k < W; ++k) {
        d[k] = n.child[k] == kEmptySlot ? INFINITY : slab<W>(n, k, r);
        c[k] = n.child[k];
    }
}

// The same for a quantised BVH4 node, by the traversal kernel's own formula:
// the plane at byte q of an axis is at t = q * (2^e / d) + (p - o) / d,
// evaluated with the same fma and products, so the host and the device
// decide each box alike.
void child_dists(const QNode4& n, const RayState& r, float (&d)[4], uint32_t (&c)[4]) {
    const float Sx = pow2_from_biased(n.exps & 0xFFu) * r.inv.x, Bx = fmaf(n.px, r.inv.x, -r.oinv.x);
    const float Sy = pow2_from_biased((n.exps >> 8) & 0xFFu) * r.inv.y, By = fmaf(n.py, r.inv.y, -r.oinv.y);
    const float Sz = pow2_from_biased((n.exps >> 16) & 0xFFu) * r.inv.z, Bz = fmaf(n.pz, r.inv.z, -r.oinv.z);
    for (int k = 0; k < 4; ++k) {
        const int sh = 8 * k;
        auto t = [&](uint32_t w, float S, float B) { return fmaf((float)((w >> sh) & 0xFFu), S, B); };
        const float tx0 = t(n.qlox, Sx, Bx), tx1 = t(n.qhix, Sx, Bx);
        const float ty0 = t(n.qloy, Sy, By), ty1 = t(n.qhiy, Sy, By);
        const float tz0 = t(n.qloz, Sz, Bz), tz1 = t(n.qhiz, Sz, Bz);
        const float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), r.tlo));
        const float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), r.thi));
        d[k] = (tn <= tf * 1.00000048f && n.child[k] != kEmptySlot) ? tn : INFINITY;
        c[k] = n.child[k];
    }
}

template <int W, class Node>
void count_rays(const Node* nodes, const TriSoA& tris, const ilm_ray* rays, int64_t b, int64_t e, bool any,
                int32_t* out) {
    std::vector<uint32_t> stack;
    stack.reserve(256);
    for (int64_t i = b; i < e; ++i) {
        const ilm_ray& q = rays[i];
        RayState r;
        ray_setup(r, mk(q.ox, q.oy, q.oz), mk(q.dx, q.dy, q.dz), q.tmax);
        int32_t nn = 0, nt = 0;
        stack.clear();
        uint32_t cur = 0;
        bool done = false;
        while (!done) {
            if (!is_leaf(cur)) {
                ++nn;
                float d[W];
                uint32_t c[W], key[W];
                child_dists(nodes[cur], r, d, c);
                for (int k = 0; k < W; ++k) key[k] = child_key<W>(d[k], k);
                // nearest first, by the kernels' own keys and exchange network
                // (so that near-ties in entry distance order the children alike)
                sort_keys<W>(key);
                for (int k = W - 1; k >= 1; --k)
                    if (key[k] < kMissKey) stack.push_back(c[key_slot<W>(key[k])]);
                if (key[0] < kMissKey) {
                    cur = c[key_slot<W>(key[0])];
                    continue;
                }
            } else {
                const uint32_t first = leaf_first(cur), cnt = leaf_count(cur);
                for (uint32_t k = 0; k < cnt; ++k) {
                    ++nt;
                    float t, u, v;
                    const uint32_t s = first + k;
                    if (tri_watertight(r, tris.v0[s], tris.v1[s], tris.v2[s], t, u, v)) {
                        r.thi = tmax_window(t);
                        if (any) { done = true; break; }
                    }
                }
                if (done) break;
            }
            if (stack.empty()) break;
            cur = stack.back();
            stack.pop_back();
        }
        out[2 * i] = nn;
        out[2 * i + 1] = nt;
    }
}

}  // namespace
}  // namespace ilm

using namespace ilm;

extern "C" int ilm_bvh_visit_counts(ilm_mesh* m, const ilm_ray* rays, int64_t n, int any_hit, int32_t* visits,
                                    int threads) {
    if (!m) return fail(ILM_ERR_INVALID, "null mesh");
    if (!m->built) return fail(ILM_ERR_STATE, "BVH not built");
    if (n < 0 || (n > 0 && (!rays || !visits))) return fail(ILM_ERR_INVALID, "bad ray or output buffer");
    if (n == 0) return ILM_OK;
    TriSoA tris;
    const size_t ns = m->h_order.size();
    tris.v0.resize(ns);
    tris.v1.resize(ns);
    tris.v2.resize(ns);
    for (size_t s = 0; s < ns; ++s) {
        const int32_t* t = &m->idx[3 * (size_t)m->h_order[s]];
        const float* p0 = &m->pos[3 * (size_t)t[0]];
        const float* p1 = &m->pos[3 * (size_t)t[1]];
        const float* p2 = &m->pos[3 * (size_t)t[2]];
        tris.v0[s] = mk(p0[0], p0[1], p0[2]);
        tris.v1[s] = mk(p1[0], p1[1], p1[2]);
        tris.v2[s] = mk(p2[0], p2[1], p2[2]);
    }
    int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    nt = (int)std::min<int64_t>(nt, (n + 1023) / 1024);
    nt = std::max(nt, 1);
    std::vector<std::thread> pool;
    const int64_t chunk = (n + nt - 1) / nt;
    for (int k = 0; k < nt; ++k) {
        const int64_t b = k * chunk, e = std::min<int64_t>(n, b + chunk);
        if (b >= e) break;
        pool.emplace_back([&, b, e]() {
            const bool any = any_hit != 0;
            if (m->node_format == ILM_NODES_QUANTIZED)
                count_rays<4>(reinterpret_cast<const QNode4*>(m->h_nodes.data()), tris, rays, b, e, any, visits);
            else if (m->width == 2)
                count_rays<2>(reinterpret_cast<const WideNode<2>*>(m->h_nodes.data()), tris, rays, b, e, any, visits);
            else if (m->width == 8)
                count_rays<8>(reinterpret_cast<const WideNode<8>*>(m->h_nodes.data()), tris, rays, b, e, any, visits);
            else
                count_rays<4>(reinterpret_cast<const WideNode<4>*>(m->h_nodes.data()), tris, rays, b, e, any, visits);
        });
    }
    for (auto& t : pool) t.join();
    return ILM_OK;
}
