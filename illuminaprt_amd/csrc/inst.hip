// The flat instance table: a two-level BVH over shared meshes. Each unique
// mesh's float BVH4 (the bottom level) and its triangles are stored once; an
// instance is a record (world-to-object transform, the mesh's root node, the
// instance's index) and a top-level BVH4 is built over the instances' world
// boxes. Replaces the reference's BasicSpace linear loop over primitives,
// each transforming the ray into its object space (Space/BasicSpace.cpp:58-114,
// Scene/GeometricPrimitive.cpp:68-200, Geometry/Ray.cpp:50-60): a ray walks
// the top-level tree, and at an instance continues in the mesh's tree with
// the ray transformed into object space. The direction is transformed
// without normalising, so a hit's t is the same in both spaces, as in the
// reference's object-space test. Updating the transforms (Space::Update, as
// the reference's frame loop does: Utility/Sandbox/Multithreaded.h:399)
// rebuilds only the top-level tree and the instance records. The traversal
// is the flat kernel's (k_trace4 in instance mode, trace.hip): persistent
// waves with XCD-aware ray fetch, the LDS stack ring, the branch-free node
// step and speculation inside the meshes' trees.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "geom.hpp"
#include "internal.hpp"

struct ilm_instances {
    ilm_ctx* ctx = nullptr;
    // bottom level: the meshes' trees packed into one node array (child words
    // rebased), their triangles into one array in leaf order
    std::vector<uint32_t> mesh_root;    // first node of each mesh's tree
    std::vector<float> mesh_box;        // root box of each mesh (lo xyz, hi xyz)
    void* d_blas = nullptr;
    float* d_tris = nullptr;
    int64_t blas_bytes = 0, tri_bytes = 0;
    int blas_need = 0;
    // top level and the instance records (in the top-level tree's leaf order)
    void* d_tlas = nullptr;
    float4* d_inst = nullptr;
    int64_t tlas_bytes = 0, table_bytes = 0, ninst = 0, tlas_cap = 0, inst_cap = 0;
    int tlas_need = 0;
};

namespace ilm {
namespace {

void rebase(uint32_t* child, uint32_t node_base, uint32_t slot_base) {
    for (int k = 0; k < 4; ++k) {
        const uint32_t c = child[k];
        if (c == kEmptySlot) continue;
        child[k] = is_leaf(c) ? make_leaf(leaf_first(c) + slot_base, leaf_count(c)) : c + node_base;
    }
}

// Transformation (Geometry/Transform.cpp:133-145, as ilm_space_add): world =
// R (S p) + T with R applied as out_i = sum_j rot[3j+i] v_j. The inverse
// (Ray::ApplyInverse) is S^-1 R^T (w - T): row i of the linear part is
// rot[3i..3i+2] / scl[i].
void world_to_object(const ilm_instance& in, double m[3][4]) {
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) m[i][j] = (double)in.rot[3 * i + j] / (double)in.scl[i];
        m[i][3] = -(m[i][0] * in.trn[0] + m[i][1] * in.trn[1] + m[i][2] * in.trn[2]);
    }
}

}  // namespace
}  // namespace ilm

using namespace ilm;

extern "C" {

int ilm_instances_destroy(ilm_instances* t) {
    if (!t) return ILM_OK;
    for (void* p : {t->d_blas, (void*)t->d_tris, t->d_tlas, (void*)t->d_inst})
        if (p) ilm_free(t->ctx, p);
    delete t;
    return ILM_OK;
}

int ilm_instances_create(ilm_ctx* ctx, ilm_mesh* const* meshes, int nmeshes, ilm_instances** out) {
    if (!ctx || !meshes || nmeshes <= 0 || !out) return fail(ILM_ERR_INVALID, "bad arguments");
    *out = nullptr;
    int64_t nodes = 0, slots = 0;
    int need = 0;
    for (int i = 0; i < nmeshes; ++i) {
        const ilm_mesh* m = meshes[i];
        if (!m || !m->built) return fail(ILM_ERR_STATE, "instanced meshes must be built");
        if (m->node_format != ILM_NODES_FLOAT || m->width != 4)
            return fail(ILM_ERR_INVALID, "instanced meshes need float BVH4 nodes");
        nodes += m->info.nodes;
        slots += (int64_t)m->h_order.size();
        need = std::max(need, m->info.stack_need);
    }
    if (slots >= (int64_t(1) << 26)) return fail(ILM_ERR_INVALID, "more than 2^26 triangle slots");
    std::vector<WideNode<4>> blas;
    std::vector<float> tris;
    blas.reserve(nodes);
    ilm_instances* t = new ilm_instances();
    t->ctx = ctx;
    t->blas_need = need;
    uint32_t nbase = 0, sbase = 0;
    for (int i = 0; i < nmeshes; ++i) {
        const ilm_mesh* m = meshes[i];
        t->mesh_root.push_back(nbase);
        const auto* src = reinterpret_cast<const WideNode<4>*>(m->h_nodes.data());
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int k = 0; k < 4; ++k) {
            if (src[0].child[k] == kEmptySlot) continue;
            lo[0] = std::min(lo[0], src[0].lox[k]); hi[0] = std::max(hi[0], src[0].hix[k]);
            lo[1] = std::min(lo[1], src[0].loy[k]); hi[1] = std::max(hi[1], src[0].hiy[k]);
            lo[2] = std::min(lo[2], src[0].loz[k]); hi[2] = std::max(hi[2], src[0].hiz[k]);
        }
        t->mesh_box.insert(t->mesh_box.end(), {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]});
        for (int64_t k = 0; k < m->info.nodes; ++k) {
            WideNode<4> w = src[k];
            rebase(w.child, nbase, sbase);
            blas.push_back(w);
        }
        std::vector<float> mt;
        pack_triangles(m->pos.data(), m->idx.data(), m->h_order, mt);   // original index = within the mesh
        tris.insert(tris.end(), mt.begin(), mt.end());
        nbase += (uint32_t)m->info.nodes;
        sbase += (uint32_t)m->h_order.size();
    }
    t->blas_bytes = (int64_t)(blas.size() * sizeof(WideNode<4>));
    t->tri_bytes = (int64_t)(tris.size() * sizeof(float));
    int rc = ilm_alloc(ctx, (size_t)t->blas_bytes, &t->d_blas);
    if (!rc) rc = ilm_alloc(ctx, (size_t)t->tri_bytes, (void**)&t->d_tris);
    if (!rc) rc = ilm_upload(ctx, t->d_blas, blas.data(), (size_t)t->blas_bytes);
    if (!rc) rc = ilm_upload(ctx, t->d_tris, tris.data(), (size_t)t->tri_bytes);
    if (rc) {
        ilm_instances_destroy(t);
        return rc;
    }
    *out = t;
    return ILM_OK;
}

int ilm_instances_set(ilm_instances* t, const ilm_instance* inst, int64_t n) {
    if (!t || n < 0 || (n > 0 && !inst)) return fail(ILM_ERR_INVALID, "bad arguments");
    if (n >= (int64_t(1) << 26)) return fail(ILM_ERR_INVALID, "more than 2^26 - 1 instances");
    const int nmeshes = (int)t->mesh_root.size();
    // world boxes of the instances as the bounds of the transformed corners of
    // their meshes' root boxes, widened by a relative 2^-20 so that rounding
    // of the object-space ray never leaves the box
    std::vector<float> pos((size_t)std::max<int64_t>(n, 1) * 9);
    std::vector<int32_t> idx((size_t)std::max<int64_t>(n, 1) * 3);
    for (int64_t i = 0; i < n; ++i) {
        const ilm_instance& in = inst[i];
        if (in.mesh < 0 || in.mesh >= nmeshes) return fail(ILM_ERR_INVALID, "instance mesh index out of range");
        if (in.scl[0] == 0.f || in.scl[1] == 0.f || in.scl[2] == 0.f) return fail(ILM_ERR_INVALID, "zero scaling");
        const float* b = &t->mesh_box[6 * (size_t)in.mesh];
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int c = 0; c < 8; ++c) {
            const double p[3] = {(c & 1) ? b[3] : b[0], (c & 2) ? b[4] : b[1], (c & 4) ? b[5] : b[2]};
            const double s[3] = {p[0] * in.scl[0], p[1] * in.scl[1], p[2] * in.scl[2]};
            for (int k = 0; k < 3; ++k) {
                const double w = in.rot[k] * s[0] + in.rot[3 + k] * s[1] + in.rot[6 + k] * s[2] + in.trn[k];
                lo[k] = std::min(lo[k], w);
                hi[k] = std::max(hi[k], w);
            }
        }
        for (int k = 0; k < 3; ++k) {
            const double pad = (std::fabs(lo[k]) + std::fabs(hi[k]) + (hi[k] - lo[k])) * 0x1p-20 + 1e-30;
            lo[k] -= pad;
            hi[k] += pad;
        }
        // a triangle whose bounds are the box: (lo, hi, (lo.x, hi.y, lo.z))
        float* v = &pos[9 * (size_t)i];
        v[0] = (float)lo[0]; v[1] = (float)lo[1]; v[2] = (float)lo[2];
        v[3] = (float)hi[0]; v[4] = (float)hi[1]; v[5] = (float)hi[2];
        v[6] = (float)lo[0]; v[7] = (float)hi[1]; v[8] = (float)lo[2];
        for (int k = 0; k < 3; ++k) {
            if ((double)v[k] > lo[k]) v[k] = std::nextafter(v[k], -INFINITY);
            if ((double)v[3 + k] < hi[k]) v[3 + k] = std::nextafter(v[3 + k], INFINITY);
        }
        v[6] = v[0]; v[7] = v[4]; v[8] = v[2];
        idx[3 * i] = (int32_t)(3 * i);
        idx[3 * i + 1] = (int32_t)(3 * i + 1);
        idx[3 * i + 2] = (int32_t)(3 * i + 2);
    }
    std::vector<uint8_t> tlas;
    std::vector<int32_t> order;
    int tneed = 1;
    if (n > 0) {
        WideBvh w;
        try {
            Bvh2 b = build_sah(pos.data(), idx.data(), n, 1, 0);
            w = collapse(b, 4);
        } catch (const std::exception& e) {
            return fail(ILM_ERR_INVALID, std::string("top-level build failed: ") + e.what());
        }
        tlas.swap(w.nodes);
        order.swap(w.order);
        tneed = w.stack_need;
    } else {
        // an empty table: one node of empty slots
        WideNode<4> e;
        for (int k = 0; k < 4; ++k) {
            e.lox[k] = e.loy[k] = e.loz[k] = INFINITY;
            e.hix[k] = e.hiy[k] = e.hiz[k] = -INFINITY;
            e.child[k] = kEmptySlot;
        }
        tlas.assign(reinterpret_cast<const uint8_t*>(&e), reinterpret_cast<const uint8_t*>(&e) + sizeof e);
    }
    // instance records in the top-level tree's leaf order
    std::vector<float4> rec((size_t)std::max<size_t>(order.size(), 1) * 4);
    for (size_t s = 0; s < order.size(); ++s) {
        const ilm_instance& in = inst[order[s]];
        double m[3][4];
        world_to_object(in, m);
        for (int r = 0; r < 3; ++r)
            rec[4 * s + r] = make_float4((float)m[r][0], (float)m[r][1], (float)m[r][2], (float)m[r][3]);
        int4 meta = make_int4((int)t->mesh_root[(size_t)in.mesh], order[s], in.mesh, 0);
        memcpy(&rec[4 * s + 3], &meta, sizeof meta);
    }
    const int64_t tb = (int64_t)tlas.size(), rb = (int64_t)(rec.size() * sizeof(float4));
    ilm_ctx* ctx = t->ctx;
    if (tb > t->tlas_cap) {
        if (t->d_tlas) ilm_free(ctx, t->d_tlas);
        t->d_tlas = nullptr;
        t->tlas_cap = 0;
        int rc = ilm_alloc(ctx, (size_t)tb, &t->d_tlas);
        if (rc) return rc;
        t->tlas_cap = tb;
    }
    if (rb > t->inst_cap) {
        if (t->d_inst) ilm_free(ctx, t->d_inst);
        t->d_inst = nullptr;
        t->inst_cap = 0;
        int rc = ilm_alloc(ctx, (size_t)rb, (void**)&t->d_inst);
        if (rc) return rc;
        t->inst_cap = rb;
    }
    int rc = ilm_upload(ctx, t->d_tlas, tlas.data(), (size_t)tb);
    if (!rc) rc = ilm_upload(ctx, t->d_inst, rec.data(), (size_t)rb);
    if (rc) return rc;
    t->tlas_bytes = tb;
    t->table_bytes = rb;
    t->ninst = n;
    t->tlas_need = tneed;
    return ILM_OK;
}

}  // extern "C"

namespace {

template <bool ANY>
int launch_inst(ilm_instances* t, const ilm_ray* rays, int64_t n, ilm_hit* hits, int32_t* inst, uint8_t* occ,
                void* stream) {
    if (!t) return fail(ILM_ERR_INVALID, "null instance table");
    if (!t->d_tlas) return fail(ILM_ERR_STATE, "instance table not set");
    // a lane's stack holds the top-level entries, the remainder of a
    // top-level leaf, kReturn and the mesh tree's entries
    return launch_trace_inst(t->ctx, t->d_tlas, t->d_blas, t->d_tris, t->d_inst, t->tlas_need + 2 + t->blas_need,
                             rays, n, hits, inst, occ, ANY, stream);
}

}  // namespace

extern "C" {

int ilm_instances_trace_closest(ilm_instances* t, const ilm_ray* rays, int64_t n, ilm_hit* hits, int32_t* inst,
                                void* stream) {
    return launch_inst<false>(t, rays, n, hits, inst, nullptr, stream);
}

int ilm_instances_trace_any(ilm_instances* t, const ilm_ray* rays, int64_t n, uint8_t* occluded, void* stream) {
    return launch_inst<true>(t, rays, n, nullptr, nullptr, occluded, stream);
}

int ilm_instances_info(ilm_instances* t, int64_t* ninst, int64_t* blas_bytes, int64_t* table_bytes,
                       int64_t* tlas_bytes, int32_t* stack_need) {
    if (!t) return fail(ILM_ERR_INVALID, "null instance table");
    if (ninst) *ninst = t->ninst;
    if (blas_bytes) *blas_bytes = t->blas_bytes + t->tri_bytes;
    if (table_bytes) *table_bytes = t->table_bytes;
    if (tlas_bytes) *tlas_bytes = t->tlas_bytes;
    if (stack_need) *stack_need = t->tlas_need + 1 + t->blas_need;
    return ILM_OK;
}

int ilm_instances_check(ilm_instances* t) {
    if (!t) return fail(ILM_ERR_INVALID, "null instance table");
    // the traversal flags a stack bound exceeded in its context's counters
    return ilm_sync(t->ctx);
}

}  // extern "C"
