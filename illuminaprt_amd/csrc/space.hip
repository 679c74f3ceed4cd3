// Spaces: the primitives of a scene (a mesh placed by a transformation, with
// a material and an optional light) flattened into one world-space mesh, so
// that one BVH and one traversal launch serve the whole scene. Replaces the
// reference's BasicSpace (Space/BasicSpace.cpp:58-114), whose linear loop
// transforms every ray into every primitive's object space
// (Scene/GeometricPrimitive.cpp:68-200). A hit's global triangle index is
// resolved back to (primitive, triangle) by ilm_space_resolve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ilm_gpu.h"
#include "internal.hpp"

struct ilm_space {
    ilm_ctx* ctx = nullptr;
    struct Prim {
        std::vector<float> pos, nrm;  // object space (nrm empty: area-weighted normals)
        std::vector<int32_t> idx;
        float rot[9], scl[3], trn[3];
        bool has_rot, has_scl, has_trn;
        int32_t material, light;
    };
    std::vector<Prim> prims;
    ilm_mesh* mesh = nullptr;
    std::vector<int32_t> tri_off;  // first global triangle of each primitive, plus the total
    int32_t* d_off = nullptr;
    int64_t nv = 0, nt = 0;
    std::vector<float> wpos, wnrm;
    std::vector<int32_t> widx, tmat, tlight;
};

namespace ilm {
namespace {

// Transformation::Apply (Geometry/Transform.cpp:133-145): scale, then the
// Matrix3x3 product (Matrix3x3.h:586-597: out_i = sum_j m[3j+i] v_j), then
// translate. Normals follow Transformation::Rotate (:178-194): scale and
// rotate, normalised afterwards as Basis::InitFromW does.
inline void apply_linear(const ilm_space::Prim& p, const float* v, float* o) {
    float x = v[0], y = v[1], z = v[2];
    if (p.has_scl) { x *= p.scl[0]; y *= p.scl[1]; z *= p.scl[2]; }
    if (p.has_rot) {
        const float* m = p.rot;
        float a = m[0] * x + m[3] * y + m[6] * z;
        float b = m[1] * x + m[4] * y + m[7] * z;
        float c = m[2] * x + m[5] * y + m[8] * z;
        x = a; y = b; z = c;
    }
    o[0] = x; o[1] = y; o[2] = z;
}

__global__ void k_space_resolve(const ilm_hit* __restrict__ hits, int64_t n, const int32_t* __restrict__ off,
                                int32_t nprims, int32_t* __restrict__ prim, int32_t* __restrict__ local) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t g = hits[i].prim;
    if (g < 0) {
        prim[i] = -1;
        local[i] = -1;
        return;
    }
    // last primitive whose first triangle is <= g
    int32_t lo = 0, hi = nprims - 1;
    while (lo < hi) {
        int32_t mid = (lo + hi + 1) >> 1;
        if (off[mid] <= g) lo = mid; else hi = mid - 1;
    }
    prim[i] = lo;
    local[i] = g - off[lo];
}

}  // namespace
}  // namespace ilm

using ilm::fail;

extern "C" {

int ilm_space_create(ilm_ctx* ctx, ilm_space** out) {
    if (!out) return fail(ILM_ERR_INVALID, "null argument");
    *out = new ilm_space();
    (*out)->ctx = ctx;
    return ILM_OK;
}

int ilm_space_destroy(ilm_space* s) {
    if (!s) return ILM_OK;
    if (s->mesh) ilm_mesh_destroy(s->mesh);
    if (s->d_off) ilm_free(s->ctx, s->d_off);
    delete s;
    return ILM_OK;
}

int ilm_space_add(ilm_space* s, const float* pos, int64_t nv, const int32_t* idx, int64_t nt, const float* nrm,
                  const float* rotation, const float* scaling, const float* translation, int32_t material,
                  int32_t light, int32_t* prim_id) {
    if (!s || !pos || !idx || nv <= 0 || nt <= 0) return fail(ILM_ERR_INVALID, "bad primitive arrays");
    for (int64_t i = 0; i < 3 * nt; ++i)
        if (idx[i] < 0 || idx[i] >= nv) return fail(ILM_ERR_INVALID, "triangle index out of range");
    if (s->nt + nt > (int64_t)INT32_MAX / 2) return fail(ILM_ERR_INVALID, "space too large");
    ilm_space::Prim p;
    p.pos.assign(pos, pos + 3 * nv);
    p.idx.assign(idx, idx + 3 * nt);
    if (nrm) p.nrm.assign(nrm, nrm + 3 * nv);
    const float I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, one[3] = {1, 1, 1}, zero[3] = {0, 0, 0};
    std::memcpy(p.rot, rotation ? rotation : I, sizeof p.rot);
    std::memcpy(p.scl, scaling ? scaling : one, sizeof p.scl);
    std::memcpy(p.trn, translation ? translation : zero, sizeof p.trn);
    p.has_rot = std::memcmp(p.rot, I, sizeof I) != 0;
    p.has_scl = !(p.scl[0] == 1 && p.scl[1] == 1 && p.scl[2] == 1);
    p.has_trn = !(p.trn[0] == 0 && p.trn[1] == 0 && p.trn[2] == 0);
    if (p.has_scl && (p.scl[0] == 0 || p.scl[1] == 0 || p.scl[2] == 0))
        return fail(ILM_ERR_INVALID, "zero scaling");
    p.material = material;
    p.light = light;
    s->prims.push_back(std::move(p));
    s->nv += nv;
    s->nt += nt;
    if (prim_id) *prim_id = (int32_t)s->prims.size() - 1;
    return ILM_OK;
}

int ilm_space_build(ilm_space* s, const ilm_bvh_params* params) {
    if (!s) return fail(ILM_ERR_INVALID, "null space");
    if (s->prims.empty()) return fail(ILM_ERR_STATE, "space has no primitives");
    std::vector<float>& wpos = s->wpos;
    std::vector<int32_t>& widx = s->widx;
    wpos.assign((size_t)s->nv * 3, 0.0f);
    widx.assign((size_t)s->nt * 3, 0);
    s->wnrm.assign((size_t)s->nv * 3, 0.0f);
    s->tmat.resize(s->nt);
    s->tlight.resize(s->nt);
    s->tri_off.clear();
    int64_t vo = 0, to = 0;
    for (const ilm_space::Prim& p : s->prims) {
        const int64_t nv = (int64_t)p.pos.size() / 3, nt = (int64_t)p.idx.size() / 3;
        s->tri_off.push_back((int32_t)to);
        std::vector<float> on;
        if (p.nrm.empty()) {
            on.resize(p.pos.size());
            int rc = ilm_scenegen_vertex_normals(p.pos.data(), nv, p.idx.data(), nt, on.data());
            if (rc) return rc;
        }
        const float* objn = p.nrm.empty() ? on.data() : p.nrm.data();
        for (int64_t v = 0; v < nv; ++v) {
            float w[3], n[3];
            ilm::apply_linear(p, &p.pos[3 * v], w);
            if (p.has_trn) { w[0] += p.trn[0]; w[1] += p.trn[1]; w[2] += p.trn[2]; }
            std::memcpy(&wpos[3 * (vo + v)], w, sizeof w);
            ilm::apply_linear(p, objn + 3 * v, n);
            float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
            if (l > 0) { n[0] /= l; n[1] /= l; n[2] /= l; }
            std::memcpy(&s->wnrm[3 * (vo + v)], n, sizeof n);
        }
        for (int64_t t = 0; t < nt; ++t) {
            for (int k = 0; k < 3; ++k) widx[3 * (to + t) + k] = (int32_t)(p.idx[3 * t + k] + vo);
            s->tmat[to + t] = p.material;
            s->tlight[to + t] = p.light;
        }
        vo += nv;
        to += nt;
    }
    s->tri_off.push_back((int32_t)to);
    if (s->mesh) { ilm_mesh_destroy(s->mesh); s->mesh = nullptr; }
    int rc = ilm_mesh_create(s->ctx, wpos.data(), s->nv, widx.data(), s->nt, &s->mesh);
    if (rc) return rc;
    rc = ilm_bvh_build(s->mesh, params);
    if (rc) return rc;
    if (!s->ctx) return ILM_OK;  // host-only space: flattened and built, not traceable
    if (s->d_off) { ilm_free(s->ctx, s->d_off); s->d_off = nullptr; }
    rc = ilm_alloc(s->ctx, s->tri_off.size() * sizeof(int32_t), (void**)&s->d_off);
    if (rc) return rc;
    return ilm_upload(s->ctx, s->d_off, s->tri_off.data(), s->tri_off.size() * sizeof(int32_t));
}

int ilm_space_mesh(ilm_space* s, ilm_mesh** mesh) {
    if (!s || !mesh) return fail(ILM_ERR_INVALID, "null argument");
    if (!s->mesh) return fail(ILM_ERR_STATE, "space not built");
    *mesh = s->mesh;
    return ILM_OK;
}

int ilm_space_size(ilm_space* s, int64_t* nprims, int64_t* nv, int64_t* nt) {
    if (!s) return fail(ILM_ERR_INVALID, "null space");
    if (nprims) *nprims = (int64_t)s->prims.size();
    if (nv) *nv = s->nv;
    if (nt) *nt = s->nt;
    return ILM_OK;
}

int ilm_space_shading(ilm_space* s, float* nrm, int32_t* tri_mat, int32_t* tri_light) {
    if (!s) return fail(ILM_ERR_INVALID, "null space");
    if (!s->mesh) return fail(ILM_ERR_STATE, "space not built");
    if (nrm) std::memcpy(nrm, s->wnrm.data(), s->wnrm.size() * sizeof(float));
    if (tri_mat) std::memcpy(tri_mat, s->tmat.data(), s->tmat.size() * sizeof(int32_t));
    if (tri_light) std::memcpy(tri_light, s->tlight.data(), s->tlight.size() * sizeof(int32_t));
    return ILM_OK;
}

int ilm_space_geometry(ilm_space* s, float* pos, int32_t* idx) {
    if (!s) return fail(ILM_ERR_INVALID, "null space");
    if (!s->mesh) return fail(ILM_ERR_STATE, "space not built");
    if (pos) std::memcpy(pos, s->wpos.data(), s->wpos.size() * sizeof(float));
    if (idx) std::memcpy(idx, s->widx.data(), s->widx.size() * sizeof(int32_t));
    return ILM_OK;
}

int ilm_space_resolve(ilm_space* s, const ilm_hit* hits, int64_t n, int32_t* prim, int32_t* local, void* stream) {
    if (!s || (n > 0 && (!hits || !prim || !local)) || n < 0) return fail(ILM_ERR_INVALID, "bad arguments");
    if (!s->mesh || !s->d_off) return fail(ILM_ERR_STATE, "space not built on a device context");
    if (n == 0) return ILM_OK;
    hipStream_t st = ilm::pick_stream(s->ctx, stream);
    const int block = 256;
    hipLaunchKernelGGL(ilm::k_space_resolve, dim3((unsigned)((n + block - 1) / block)), dim3(block), 0, st, hits, n,
                       s->d_off, (int32_t)s->prims.size(), prim, local);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? ILM_OK : ilm::fail_hip(e, "space resolve launch");
}

}  // extern "C"
