// Wavefront OBJ / MTL loading into the arrays the library's meshes and
// scenes take. Follows WavefrontSceneLoader (Scene/WavefrontSceneLoader.cpp):
//   - statements v, vn, vt, f, o, usemtl, mtllib (:445-561); newmtl, illum,
//     Ka, Kd, Ks, Ke, Ns and the map_* names (:203-300);
//   - faces of 3 or 4 corners only (larger polygons are skipped, :489-490);
//     a quad becomes (0,1,2), (0,2,3) (:528-531);
//   - a corner is "p", "p/t/n" or "p//n" (empty fields are dropped, so a
//     two-field corner is read as position/normal, :496-507); a corner with
//     one field keeps the texture and normal of the previous corner of the
//     same face (one vertex record is reused across corners, :492-507);
//   - corners with equal (position, texture, normal) share one vertex (:510-523);
//   - indices are 1-based; usemtl selects the material of later faces.
// Vertices without a normal get the area-weighted normal of their triangles
// (ITriangleMesh::UpdateNormals, Shape/TriangleMesh.cpp:96-133) instead of
// the zero vector the reference keeps. Materials become the library's 8-float
// records (kd, ks, Ns, type 1 when ks is not black, else 0); a material with a
// non-black Ke also becomes a light (Le = Ke) on its triangles.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <tuple>
#include <vector>

#include "ilm_gpu.h"
#include "internal.hpp"

struct ilm_obj {
    std::vector<float> pos, nrm;
    std::vector<int32_t> idx, tri_mat, tri_light;
    std::vector<float> mats, lights;
    std::vector<int32_t> illum;  // per material, 2 when absent
    std::string mtllib;          // file name of the first material library
};

namespace {

struct Mtl {
    std::string name;
    float kd[3] = {0, 0, 0}, ks[3] = {0, 0, 0}, ke[3] = {0, 0, 0};
    float ns = 0;
    int illum = 2;
};

std::vector<std::string> split(const std::string& s, const char* seps) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : s) {
        if (std::strchr(seps, c)) {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

bool to_float(const std::string& s, float& v) {
    char* end = nullptr;
    v = std::strtof(s.c_str(), &end);
    return end && *end == 0;
}
bool to_int(const std::string& s, long& v) {
    char* end = nullptr;
    v = std::strtol(s.c_str(), &end, 10);
    return end && *end == 0;
}

std::string dir_of(const std::string& path) {
    size_t p = path.find_last_of('/');
    return p == std::string::npos ? std::string() : path.substr(0, p + 1);
}

bool load_mtl(const std::string& path, std::vector<Mtl>& out) {
    std::ifstream f(path);
    if (!f) return false;
    std::string line;
    while (std::getline(f, line)) {
        std::vector<std::string> t = split(line, " \t\r\n");
        if (t.empty()) continue;
        if (t[0] == "newmtl" && t.size() == 2) {
            out.emplace_back();
            out.back().name = t[1];
            continue;
        }
        if (out.empty()) continue;
        Mtl& m = out.back();
        float* dst = t[0] == "Kd" ? m.kd : t[0] == "Ks" ? m.ks : t[0] == "Ke" ? m.ke : nullptr;
        if (dst && t.size() == 4) {
            for (int k = 0; k < 3; ++k) to_float(t[1 + k], dst[k]);
        } else if (t[0] == "Ns" && t.size() == 2) {
            to_float(t[1], m.ns);
        } else if (t[0] == "illum" && t.size() == 2) {
            long v = 2;
            if (to_int(t[1], v)) m.illum = (int)v;
        }
    }
    return true;
}

}  // namespace

using ilm::fail;

extern "C" {

int ilm_obj_load(const char* path, ilm_obj** out) {
    if (!path || !out) return fail(ILM_ERR_INVALID, "null argument");
    std::ifstream f(path);
    if (!f) return fail(ILM_ERR_INVALID, std::string("cannot open ") + path);
    std::vector<float> P, N, T;  // positions, normals, texture coordinates (counted only)
    std::vector<Mtl> mtls;
    std::map<std::string, int> mtl_index;
    std::map<std::tuple<long, long, long>, int32_t> vmap;
    ilm_obj* o = new ilm_obj();
    std::vector<char> has_n;
    int32_t cur_mat = 0;
    std::string line;
    long lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        std::vector<std::string> t = split(line, " \t\r\n");
        if (t.empty()) continue;
        if (t[0] == "v" && t.size() == 4) {
            for (int k = 0; k < 3; ++k) {
                float v;
                if (!to_float(t[1 + k], v)) { delete o; return fail(ILM_ERR_INVALID, "bad vertex at line " + std::to_string(lineno)); }
                P.push_back(v);
            }
        } else if (t[0] == "vn" && t.size() == 4) {
            for (int k = 0; k < 3; ++k) {
                float v;
                if (!to_float(t[1 + k], v)) { delete o; return fail(ILM_ERR_INVALID, "bad normal at line " + std::to_string(lineno)); }
                N.push_back(v);
            }
        } else if (t[0] == "vt" && t.size() >= 3) {
            T.push_back(0.0f);
        } else if (t[0] == "f" && (t.size() == 4 || t.size() == 5)) {
            int32_t corner[4];
            long tex = 0, nor = 0;  // kept across corners of one face
            for (size_t c = 1; c < t.size(); ++c) {
                std::vector<std::string> ft = split(t[c], "/");
                long p = 0;
                if (ft.empty() || !to_int(ft[0], p)) { delete o; return fail(ILM_ERR_INVALID, "bad face at line " + std::to_string(lineno)); }
                if (ft.size() == 2) {
                    if (!to_int(ft[1], nor)) { delete o; return fail(ILM_ERR_INVALID, "bad face at line " + std::to_string(lineno)); }
                    tex = 0;
                } else if (ft.size() == 3) {
                    if (!to_int(ft[1], tex) || !to_int(ft[2], nor)) { delete o; return fail(ILM_ERR_INVALID, "bad face at line " + std::to_string(lineno)); }
                }
                if (p < 1 || p > (long)P.size() / 3 || nor < 0 || nor > (long)N.size() / 3 || tex < 0 ||
                    tex > (long)T.size()) {
                    delete o;
                    return fail(ILM_ERR_INVALID, "face index out of range at line " + std::to_string(lineno));
                }
                auto key = std::make_tuple(p, tex, nor);
                auto it = vmap.find(key);
                if (it == vmap.end()) {
                    int32_t vi = (int32_t)(o->pos.size() / 3);
                    for (int k = 0; k < 3; ++k) o->pos.push_back(P[3 * (p - 1) + k]);
                    for (int k = 0; k < 3; ++k) o->nrm.push_back(nor ? N[3 * (nor - 1) + k] : 0.0f);
                    has_n.push_back(nor ? 1 : 0);
                    vmap[key] = vi;
                    corner[c - 1] = vi;
                } else {
                    corner[c - 1] = it->second;
                }
            }
            const int32_t tris[2][3] = {{corner[0], corner[1], corner[2]}, {corner[0], corner[2], corner[3]}};
            for (int k = 0; k < (t.size() == 5 ? 2 : 1); ++k) {
                o->idx.insert(o->idx.end(), tris[k], tris[k] + 3);
                o->tri_mat.push_back(cur_mat);
            }
        } else if (t[0] == "mtllib" && t.size() == 2) {
            if (o->mtllib.empty()) o->mtllib = t[1];
            size_t before = mtls.size();
            load_mtl(dir_of(path) + t[1], mtls);
            for (size_t k = before; k < mtls.size(); ++k)
                if (!mtl_index.count(mtls[k].name)) mtl_index[mtls[k].name] = (int)k;
        } else if (t[0] == "usemtl" && t.size() == 2) {
            auto it = mtl_index.find(t[1]);
            cur_mat = it == mtl_index.end() ? 0 : it->second;
        }
    }
    if (o->idx.empty()) {
        delete o;
        return fail(ILM_ERR_INVALID, std::string("no triangles in ") + path);
    }
    // vertices without a normal: area-weighted normals of their triangles
    bool need = false;
    for (char h : has_n) need |= !h;
    if (need) {
        std::vector<float> wn(o->pos.size());
        int rc = ilm_scenegen_vertex_normals(o->pos.data(), (int64_t)o->pos.size() / 3, o->idx.data(),
                                             (int64_t)o->idx.size() / 3, wn.data());
        if (rc) { delete o; return rc; }
        for (size_t v = 0; v < has_n.size(); ++v)
            if (!has_n[v])
                for (int k = 0; k < 3; ++k) o->nrm[3 * v + k] = wn[3 * v + k];
    }
    if (mtls.empty()) {
        Mtl grey;
        grey.name = "default";
        grey.kd[0] = grey.kd[1] = grey.kd[2] = 0.5f;
        mtls.push_back(grey);
    }
    std::vector<int32_t> light_of(mtls.size(), -1);
    for (size_t k = 0; k < mtls.size(); ++k) {
        const Mtl& m = mtls[k];
        bool spec = m.ks[0] > 0 || m.ks[1] > 0 || m.ks[2] > 0;
        const float rec[8] = {m.kd[0], m.kd[1], m.kd[2], m.ks[0], m.ks[1], m.ks[2], m.ns, spec ? 1.0f : 0.0f};
        o->mats.insert(o->mats.end(), rec, rec + 8);
        o->illum.push_back(m.illum);
        if (m.ke[0] > 0 || m.ke[1] > 0 || m.ke[2] > 0) {
            light_of[k] = (int32_t)(o->lights.size() / 4);
            const float le[4] = {m.ke[0], m.ke[1], m.ke[2], 0.0f};
            o->lights.insert(o->lights.end(), le, le + 4);
        }
    }
    for (int32_t m : o->tri_mat) o->tri_light.push_back(light_of[m]);
    *out = o;
    return ILM_OK;
}

int ilm_obj_counts(ilm_obj* o, int64_t* nv, int64_t* nt, int32_t* nmat, int32_t* nlight) {
    if (!o) return fail(ILM_ERR_INVALID, "null object");
    if (nv) *nv = (int64_t)o->pos.size() / 3;
    if (nt) *nt = (int64_t)o->idx.size() / 3;
    if (nmat) *nmat = (int32_t)(o->mats.size() / 8);
    if (nlight) *nlight = (int32_t)(o->lights.size() / 4);
    return ILM_OK;
}

int ilm_obj_arrays(ilm_obj* o, float* pos, float* nrm, int32_t* idx, int32_t* tri_mat, int32_t* tri_light,
                   float* mats, float* lights) {
    if (!o) return fail(ILM_ERR_INVALID, "null object");
    auto cp = [](void* dst, const void* src, size_t bytes) { if (dst && bytes) std::memcpy(dst, src, bytes); };
    cp(pos, o->pos.data(), o->pos.size() * sizeof(float));
    cp(nrm, o->nrm.data(), o->nrm.size() * sizeof(float));
    cp(idx, o->idx.data(), o->idx.size() * sizeof(int32_t));
    cp(tri_mat, o->tri_mat.data(), o->tri_mat.size() * sizeof(int32_t));
    cp(tri_light, o->tri_light.data(), o->tri_light.size() * sizeof(int32_t));
    cp(mats, o->mats.data(), o->mats.size() * sizeof(float));
    cp(lights, o->lights.data(), o->lights.size() * sizeof(float));
    return ILM_OK;
}

int ilm_obj_material_info(ilm_obj* o, int32_t* illum, char* mtllib, int mtllib_len) {
    if (!o) return fail(ILM_ERR_INVALID, "null object");
    if (illum && !o->illum.empty()) std::memcpy(illum, o->illum.data(), o->illum.size() * sizeof(int32_t));
    if (mtllib && mtllib_len > 0) {
        std::strncpy(mtllib, o->mtllib.c_str(), (size_t)mtllib_len - 1);
        mtllib[mtllib_len - 1] = 0;
    }
    return ILM_OK;
}

int ilm_obj_destroy(ilm_obj* o) {
    delete o;
    return ILM_OK;
}

}  // extern "C"
