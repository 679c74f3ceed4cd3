// PersistentMesh files as an import path for prebuilt kd-trees
// (Shape/PersistentMesh.h:22-36, PersistentMesh.cpp:77-89; the files are the
// ones KDTreeMesh's export writes, KDTreeMesh.cpp:143-257):
//   <trunk>.ver  Vertex records, 48 bytes: position (16-byte Vector3:
//                x, y, z, pad), normal (same), uv (2 floats), 8 bytes of
//                padding to the Vector3 alignment;
//   <trunk>.tri  PersistentIndexedTriangle records, 48 bytes: three vertex
//                ids, the group id, two edge vectors (16 bytes each);
//   <trunk>.tre  the kd-tree (ilm::import_kdtree).
// The record sizes are those of the reference's structs on x86-64, where
// Vector3 is 16-byte aligned (Geometry/Vector3.h:34-38).
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <exception>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "bvh_build.hpp"
#include "internal.hpp"

struct ilm_persistent {
    std::vector<float> pos, nrm;
    std::vector<int32_t> idx, group;
    std::vector<uint8_t> tree;
    int64_t tree_nodes = 0, tree_refs = 0;
};

namespace {

constexpr size_t kVertexBytes = 48, kTriangleBytes = 48;

bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    out.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
    return true;
}

}  // namespace

using ilm::fail;

extern "C" {

int ilm_persistent_load(const char* trunk, ilm_persistent** out) {
    if (!trunk || !out) return fail(ILM_ERR_INVALID, "null argument");
    *out = nullptr;
    std::vector<uint8_t> ver, tri;
    ilm_persistent* p = new ilm_persistent();
    const std::string t(trunk);
    if (!read_file(t + ".ver", ver) || !read_file(t + ".tri", tri) || !read_file(t + ".tre", p->tree)) {
        delete p;
        return fail(ILM_ERR_INVALID, "cannot open " + t + ".ver / .tri / .tre");
    }
    if (ver.size() % kVertexBytes || tri.size() % kTriangleBytes || ver.empty() || tri.empty()) {
        delete p;
        return fail(ILM_ERR_INVALID, "PersistentMesh files are not whole 48-byte records");
    }
    const int64_t nv = (int64_t)(ver.size() / kVertexBytes), nt = (int64_t)(tri.size() / kTriangleBytes);
    if (nv > INT32_MAX || nt > (1 << 26) - 1) {
        delete p;
        return fail(ILM_ERR_INVALID, "PersistentMesh too large");
    }
    p->pos.resize(3 * nv);
    p->nrm.resize(3 * nv);
    for (int64_t v = 0; v < nv; ++v) {
        memcpy(&p->pos[3 * v], ver.data() + kVertexBytes * v, 12);
        memcpy(&p->nrm[3 * v], ver.data() + kVertexBytes * v + 16, 12);
    }
    p->idx.resize(3 * nt);
    p->group.resize(nt);
    for (int64_t i = 0; i < nt; ++i) {
        memcpy(&p->idx[3 * i], tri.data() + kTriangleBytes * i, 12);
        memcpy(&p->group[i], tri.data() + kTriangleBytes * i + 12, 4);
        for (int k = 0; k < 3; ++k)
            if (p->idx[3 * i + k] < 0 || p->idx[3 * i + k] >= nv) {
                delete p;
                return fail(ILM_ERR_INVALID, "triangle vertex id out of range");
            }
    }
    try {
        ilm::Bvh2 b = ilm::import_kdtree(p->tree.data(), p->tree.size(), p->pos.data(), nv, nt);
        p->tree_nodes = (int64_t)b.nodes.size();
        p->tree_refs = (int64_t)b.order.size();
    } catch (const std::exception& e) {
        delete p;
        return fail(ILM_ERR_INVALID, std::string("PersistentMesh tree: ") + e.what());
    }
    *out = p;
    return ILM_OK;
}

int ilm_persistent_counts(ilm_persistent* p, int64_t* nv, int64_t* nt, int64_t* tree_nodes, int64_t* tree_refs) {
    if (!p) return fail(ILM_ERR_INVALID, "null object");
    if (nv) *nv = (int64_t)p->pos.size() / 3;
    if (nt) *nt = (int64_t)p->idx.size() / 3;
    if (tree_nodes) *tree_nodes = p->tree_nodes;
    if (tree_refs) *tree_refs = p->tree_refs;
    return ILM_OK;
}

int ilm_persistent_arrays(ilm_persistent* p, float* pos, float* nrm, int32_t* idx, int32_t* group) {
    if (!p) return fail(ILM_ERR_INVALID, "null object");
    if (pos) memcpy(pos, p->pos.data(), p->pos.size() * sizeof(float));
    if (nrm) memcpy(nrm, p->nrm.data(), p->nrm.size() * sizeof(float));
    if (idx) memcpy(idx, p->idx.data(), p->idx.size() * sizeof(int32_t));
    if (group) memcpy(group, p->group.data(), p->group.size() * sizeof(int32_t));
    return ILM_OK;
}

int ilm_persistent_build(ilm_persistent* p, ilm_mesh* m, int node_format) {
    if (!p || !m) return fail(ILM_ERR_INVALID, "null argument");
    if (node_format != ILM_NODES_FLOAT && node_format != ILM_NODES_QUANTIZED)
        return fail(ILM_ERR_INVALID, "unknown node format");
    if (m->nv != (int64_t)p->pos.size() / 3 || m->nt != (int64_t)p->idx.size() / 3 || m->idx != p->idx ||
        m->pos != p->pos)
        return fail(ILM_ERR_INVALID, "the mesh was not created from this PersistentMesh's arrays");
    auto t0 = std::chrono::steady_clock::now();
    ilm::Bvh2 b;
    try {
        b = ilm::import_kdtree(p->tree.data(), p->tree.size(), p->pos.data(), m->nv, m->nt);
    } catch (const std::exception& e) {
        return fail(ILM_ERR_INVALID, std::string("PersistentMesh tree: ") + e.what());
    }
    return ilm::install_bvh(m, b, node_format, t0);
}

int ilm_persistent_destroy(ilm_persistent* p) {
    delete p;
    return ILM_OK;
}

}  // extern "C"
