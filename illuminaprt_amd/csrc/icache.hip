// Irradiance cache on the GPU: the comparator of the neural radiance cache,
// restating ICIntegrator (Integrator/ICIntegrator.cpp:29-668) on the
// wavefront path integrator.
//
// Kept from the reference:
//   - the octree: root = scene centre +- the scene's longest edge (each corner
//     scaled by 1 + FLT_EPSILON, :306-323); children ordered as SetBounds
//     (:54-86); a record goes into every child its sphere (radius RiClamp)
//     overlaps (SphereBoxOverlap, :88-101) until the depth is spent or
//     RiClamp exceeds the node's half-size (Insert, :107-140);
//   - lookup: Ward's weight w = 1 / (d / RiClamp + 1 - cos) - 1 / error
//     threshold, records with w > 0 on the root-to-leaf path of the point
//     averaged by weight (FindRecords, W_Ward, GetIrradiance, :142-200, 455-496);
//   - a record: AltitudeStrata x AzimuthStrata stratified cosine-distributed
//     gather paths (Montecarlo::CosineSampleHemisphere with strata) from
//     P + w * eps, each traced as PathLi (next-event estimation at every
//     vertex, no emission, roulette after the third vertex with
//     q = min(0.5, 0.33 (r + g + b)), RayDepth vertices); E = their mean,
//     Ri = mn / sum(1 / first-hit distance), RiClamp = clamp(Ri, RMin, RMax)
//     with RMin/RMax = longest edge * AmbientResolution / AmbientMultiplier
//     (ComputeRecord, :498-543, PathLi :545-642);
//   - shading: direct lighting (next-event estimation) plus
//     irradiance * albedo * 1/pi at the first hit, Le at emitters, a
//     constant (20, 50, 100) for camera rays that hit nothing (Radiance, :348-441).
// The MLICIntegrator variant as the reference compiles it (Tabelion weights,
// Integrator/MLICIntegrator.cpp:28-34) differs in the weight
// (1 - max(d / RiClamp, 9.0124 (1 - cos)) / threshold), Ri (the nearest gather
// hit), the insertion test (half-size along x) and the gather origin (P).
// Different on purpose: the reference fills the cache lazily, one pixel after
// another, so a record serves the pixels traced after it. Here records are
// placed in passes over the image (every 16th pixel, then every 8th, ... then
// every sample), each pass computing records for all the samples the cache
// does not yet cover, in parallel; every sample is then looked up in the
// final cache of the frame. The cache persists across frames.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include <hipcub/device/device_select.hpp>

#include "geom.hpp"
#include "internal.hpp"

namespace ilm {

namespace {

constexpr float kPiF = 3.14159265358979323846f;

// 3 float4 per record: (P, RiClamp), (N, Ri), (E, 0)
// 3 int4/float4 per node: (lo, child), (hi, first ref), (count, 0, 0, 0)
struct IcArgs {
    const float4* nodes;
    const int32_t* refs;
    const float4* recs;
    float inv_threshold;
    int variant;  // ILM_IC_WARD or ILM_IC_TABELION
};

__device__ __forceinline__ bool contains(float4 lo, float4 hi, f3 p) {
    // AxisAlignedBoundingBox::Contains (Geometry/BoundingBox.cpp:254-266)
    return !(lo.x > p.x || hi.x < p.x || lo.y > p.y || hi.y < p.y || lo.z > p.z || hi.z < p.z);
}

// IrradianceCache::FindRecords + GetIrradiance's weighted mean; returns the
// weight sum (0: no record covers the point).
__device__ float ic_lookup(const IcArgs& c, f3 p, f3 n, f3& e) {
    e = mk(0, 0, 0);
    float den = 0.f;
    int node = 0;
    if (!contains(c.nodes[0], c.nodes[1], p)) return 0.f;
    for (int level = 0; level < 64; ++level) {
        const float4 lo = c.nodes[3 * node], hi = c.nodes[3 * node + 1], cnt = c.nodes[3 * node + 2];
        const int first = __float_as_int(hi.w), count = __float_as_int(cnt.x);
        for (int k = 0; k < count; ++k) {
            const int r = c.refs[first + k];
            const float4 pr = c.recs[3 * r], nr = c.recs[3 * r + 1], er = c.recs[3 * r + 2];
            const float dx = p.x - pr.x, dy = p.y - pr.y, dz = p.z - pr.z;
            const float dist = sqrtf(dx * dx + dy * dy + dz * dz) / pr.w;
            const float cosv = n.x * nr.x + n.y * nr.y + n.z * nr.z;
            float w;
            if (c.variant == ILM_IC_TABELION) {
                // MLIrradianceCache::W_Tabelion (MLICIntegrator.cpp:285-318)
                w = 1.f - c.inv_threshold * fmaxf(dist, (1.f - cosv) * 9.0124f);
            } else {
                // IrradianceCache::W_Ward (ICIntegrator.cpp:185-196)
                if (cosv <= 0.f) continue;
                w = 1.f / (dist + 1.f - cosv) - c.inv_threshold;
            }
            if (w > 0.f) {
                e = e + w * mk(er.x, er.y, er.z);
                den += w;
            }
        }
        const int child = __float_as_int(lo.w);
        if (child < 0) break;
        int next = -1;
        for (int i = 0; i < 8; ++i) {
            const int ch = child + i;
            if (contains(c.nodes[3 * ch], c.nodes[3 * ch + 1], p)) {
                next = ch;
                break;
            }
        }
        if (next < 0) break;
        node = next;
    }
    if (den > 0.f) e = e * (1.f / den);
    return den;
}

// Lookups of a placement pass's candidates: with stride > 0 the first sample
// of every stride-th pixel of every stride-th row (row-major), with stride 0
// every sample. Candidate i's sample (an index into the G-buffer) goes to
// list[i], and miss[i] = 1 for a reflector sample that no record covers.
__global__ void k_ic_probe_pass(IcArgs c, const float4* gbuf, int stride, int width, int spp, int64_t nx, int64_t nc,
                                int32_t* list, uint8_t* miss) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nc) return;
    const int64_t g = stride > 0 ? (((i / nx) * stride) * (int64_t)width + (i % nx) * stride) * spp : i;
    list[i] = (int32_t)g;
    const float4 a = gbuf[3 * g], b = gbuf[3 * g + 1];
    uint8_t m = 0;
    if (a.w == 1.f) {
        f3 e;
        m = ic_lookup(c, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), e) > 0.f ? 0 : 1;
    }
    miss[i] = m;
}

// Gather rays of the new records: record r (sample list[r]), stratum
// j = altitude * azimuth_strata + azimuth, in ComputeRecord's loop order.
__global__ void k_ic_gather_rays(const float4* gbuf, const int32_t* list, int64_t nrec, int az_n, int alt_n,
                                 float eps, uint32_t seed, uint32_t pass, ilm_ray* out) {
    // ICIntegrator starts gather rays at P + w eps; MLICIntegrator at P with
    // tmin 2 eps (MLICIntegrator.cpp:697) -- pass eps = 0 for that
    const int mn = az_n * alt_n;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrec * mn) return;
    const int64_t r = i / mn;
    const int j = (int)(i % mn);
    const int alt = j / az_n, az = j % az_n;
    const int64_t g = list[r];
    const float4 a = gbuf[3 * g], b = gbuf[3 * g + 1];
    Block8 u = sample_block(seed ^ 0x1c0ffeeu, (uint64_t)g, pass, (uint32_t)j, 9u);
    // Montecarlo::CosineSampleHemisphere(u, v, altitude, azimuth, m, n) (Maths/Montecarlo.h:20-29)
    const float sa = (u.v[0] + (float)alt) / (float)alt_n, sb = (u.v[1] + (float)az) / (float)az_n;
    const float c = 2.f * kPiF * sb, d = sqrtf(sa);
    const f3 local = mk(d * cosf(c), d * sinf(c), sqrtf(fmaxf(0.f, 1.f - sa)));
    const f3 w = to_world(basis_from_w(mk(b.x, b.y, b.z)), local);
    ilm_ray ray;
    ray.ox = a.x + w.x * eps; ray.oy = a.y + w.y * eps; ray.oz = a.z + w.z * eps; ray.tmin = 0.f;
    ray.dx = w.x; ray.dy = w.y; ray.dz = w.z; ray.tmax = FLT_MAX;
    out[i] = ray;
}

// New records from their gather paths' radiance sums and first-hit distances.
__global__ void k_ic_reduce(const float4* gbuf, const int32_t* list, int64_t nrec, int mn, const double* L,
                            const float* first_t, float rmin, float rmax, int variant, float4* recs_out) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrec) return;
    double e0 = 0, e1 = 0, e2 = 0;
    float inv = 0.f, tmin = FLT_MAX;
    for (int j = 0; j < mn; ++j) {
        const int64_t k = r * mn + j;
        e0 += L[3 * k];
        e1 += L[3 * k + 1];
        e2 += L[3 * k + 2];
        inv += 1.f / first_t[k];
        tmin = fminf(tmin, first_t[k]);
    }
    // Ward: harmonic mean distance; Tabelion: the nearest hit (MLICIntegrator.cpp:717-725)
    const float ri = variant == ILM_IC_TABELION ? tmin : (float)mn / inv;
    const float rc = fmaxf(rmin, fminf(rmax, ri));
    const int64_t g = list[r];
    const float4 a = gbuf[3 * g], b = gbuf[3 * g + 1];
    recs_out[3 * r] = make_float4(a.x, a.y, a.z, rc);
    recs_out[3 * r + 1] = make_float4(b.x, b.y, b.z, ri);
    recs_out[3 * r + 2] = make_float4((float)(e0 / mn), (float)(e1 / mn), (float)(e2 / mn), 0.f);
}

// Indirect light of every sample (its own record when it made one, else the
// cache's weighted mean) and the miss constant, summed per pixel into accum.
__global__ void k_ic_shade(IcArgs c, const float4* gbuf, const int32_t* own, int64_t npix, int spp, float factor,
                           float3 miss, double* accum, unsigned long long* lookups) {
    int64_t px = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= npix) return;
    double s0 = 0, s1 = 0, s2 = 0;
    int looked = 0;
    for (int s = 0; s < spp; ++s) {
        const int64_t g = px * spp + s;
        const float4 a = gbuf[3 * g], b = gbuf[3 * g + 1], al = gbuf[3 * g + 2];
        if (a.w == 3.f) {
            s0 += miss.x;
            s1 += miss.y;
            s2 += miss.z;
            continue;
        }
        if (a.w != 1.f) continue;
        f3 e;
        const int r = own[g];
        if (r >= 0) {
            const float4 er = c.recs[3 * r + 2];
            e = mk(er.x, er.y, er.z);
        } else {
            ic_lookup(c, mk(a.x, a.y, a.z), mk(b.x, b.y, b.z), e);
            ++looked;
        }
        s0 += (double)(e.x * al.x * factor);
        s1 += (double)(e.y * al.y * factor);
        s2 += (double)(e.z * al.z * factor);
    }
    accum[3 * px] += s0;
    accum[3 * px + 1] += s1;
    accum[3 * px + 2] += s2;
    if (looked) atomicAdd(lookups, (unsigned long long)looked);
}

__global__ void k_ic_lookup_points(IcArgs c, const float* pts, int64_t n, float* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 e;
    const float* q = pts + 6 * i;
    const float w = ic_lookup(c, mk(q[0], q[1], q[2]), mk(q[3], q[4], q[5]), e);
    out[4 * i] = e.x;
    out[4 * i + 1] = e.y;
    out[4 * i + 2] = e.z;
    out[4 * i + 3] = w;
}

__global__ void k_ic_own(const int32_t* list, int64_t nrec, int32_t base, int32_t* own) {
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrec) own[list[r]] = base + (int32_t)r;
}

__global__ void k_ic_fill_i32(int32_t* p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

unsigned grid_of(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

}  // namespace

}  // namespace ilm

// ---------------------------------------------------------------- host side
// The octree's host arithmetic rounds every operation, as the reference's
// float code does (no fused multiply-adds).
#pragma clang fp contract(off)

struct ilm_ic {
    struct Rec { float P[3], N[3], E[3], ri, rc; };
    struct Node { float lo[3], hi[3]; int32_t child = -1; std::vector<int32_t> recs; };
    ilm_scene* scene = nullptr;
    ilm_ctx* ctx = nullptr;
    ilm_ic_params prm{};
    float rmin = 0, rmax = 0;
    std::vector<Rec> recs;
    std::vector<Node> nodes;
    // device copies (rebuilt after each insertion pass)
    float4* d_nodes = nullptr;
    int32_t* d_refs = nullptr;
    float4* d_recs = nullptr;
    size_t cap_nodes = 0, cap_refs = 0, cap_recs = 0;
    uint32_t passes = 0;
    unsigned long long* d_count = nullptr;
    // frame buffers, grown on demand and kept across frames
    struct Buf {
        void* p = nullptr;
        size_t bytes = 0;
    };
    Buf gbuf, own, list, miss, sel, nsel, tmp, rays, L, ft, newrecs;
};

namespace {

using ilm::fail;

float half_max(const ilm_ic::Node& n) {
    float e = 0.f;
    for (int k = 0; k < 3; ++k) e = std::max(e, std::fabs((n.hi[k] - n.lo[k]) * 0.5f));
    return e;
}

// IrradianceCache::SetBounds (ICIntegrator.cpp:54-86)
void child_bounds(const ilm_ic::Node& p, int i, float lo[3], float hi[3]) {
    float c[3];
    for (int k = 0; k < 3; ++k) c[k] = p.lo[k] + (p.hi[k] - p.lo[k]) * 0.5f;
    const int bx = i & 1, bz = (i >> 1) & 1, by = (i >> 2) & 1;
    const int b[3] = {bx, by, bz};
    for (int k = 0; k < 3; ++k) {
        lo[k] = b[k] ? c[k] : p.lo[k];
        hi[k] = b[k] ? p.hi[k] : c[k];
    }
}

// IrradianceCache::SphereBoxOverlap (ICIntegrator.cpp:88-101)
bool sphere_box(const float lo[3], const float hi[3], const float c[3], float r) {
    float d = 0.f;
    for (int i = 0; i < 3; ++i) {
        if (c[i] < lo[i]) d += (c[i] - lo[i]) * (c[i] - lo[i]);
        else if (c[i] > hi[i]) d += (c[i] - hi[i]) * (c[i] - hi[i]);
    }
    return d <= r * r;
}

// IrradianceCache::Insert (ICIntegrator.cpp:107-140); MLIrradianceCache::Insert
// compares RiClamp with the half-size along x only (MLICIntegrator.cpp:172-176)
void insert(ilm_ic* ic, int32_t node, int32_t rec, int depth) {
    const ilm_ic::Rec& r = ic->recs[rec];
    const ilm_ic::Node& nd = ic->nodes[node];
    const float ext = ic->prm.variant == ILM_IC_TABELION ? (nd.hi[0] - nd.lo[0]) * 0.5f : half_max(nd);
    if (depth <= 0 || r.rc > ext) {
        ic->nodes[node].recs.push_back(rec);
        return;
    }
    if (ic->nodes[node].child < 0) {
        const int32_t first = (int32_t)ic->nodes.size();
        ic->nodes.resize(ic->nodes.size() + 8);
        ilm_ic::Node& parent = ic->nodes[node];
        parent.child = first;
        for (int i = 0; i < 8; ++i) child_bounds(parent, i, ic->nodes[first + i].lo, ic->nodes[first + i].hi);
    }
    const int32_t first = ic->nodes[node].child;
    for (int i = 0; i < 8; ++i) {
        const ilm_ic::Node& ch = ic->nodes[first + i];
        if (sphere_box(ch.lo, ch.hi, r.P, r.rc)) insert(ic, first + i, rec, depth - 1);
    }
}

int grow(ilm_ctx* ctx, void** p, size_t& cap, size_t bytes) {
    if (bytes <= cap && *p) return ILM_OK;
    if (*p) ilm_free(ctx, *p);
    *p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes + bytes / 2, 256);
    int rc = ilm_alloc(ctx, want, p);
    if (!rc) cap = want;
    return rc;
}

// Flattened octree and records to the device.
int upload(ilm_ic* ic) {
    std::vector<float> nodes(12 * ic->nodes.size());
    std::vector<int32_t> refs;
    for (size_t i = 0; i < ic->nodes.size(); ++i) {
        const ilm_ic::Node& n = ic->nodes[i];
        float* o = &nodes[12 * i];
        int32_t first = (int32_t)refs.size(), count = (int32_t)n.recs.size();
        refs.insert(refs.end(), n.recs.begin(), n.recs.end());
        o[0] = n.lo[0]; o[1] = n.lo[1]; o[2] = n.lo[2]; std::memcpy(&o[3], &n.child, 4);
        o[4] = n.hi[0]; o[5] = n.hi[1]; o[6] = n.hi[2]; std::memcpy(&o[7], &first, 4);
        std::memcpy(&o[8], &count, 4); o[9] = o[10] = o[11] = 0.f;
    }
    std::vector<float> recs(12 * ic->recs.size());
    for (size_t i = 0; i < ic->recs.size(); ++i) {
        const ilm_ic::Rec& r = ic->recs[i];
        float* o = &recs[12 * i];
        o[0] = r.P[0]; o[1] = r.P[1]; o[2] = r.P[2]; o[3] = r.rc;
        o[4] = r.N[0]; o[5] = r.N[1]; o[6] = r.N[2]; o[7] = r.ri;
        o[8] = r.E[0]; o[9] = r.E[1]; o[10] = r.E[2]; o[11] = 0.f;
    }
    if (refs.empty()) refs.push_back(0);
    if (recs.empty()) recs.assign(12, 0.f);
    int rc = grow(ic->ctx, (void**)&ic->d_nodes, ic->cap_nodes, nodes.size() * 4);
    if (!rc) rc = grow(ic->ctx, (void**)&ic->d_refs, ic->cap_refs, refs.size() * 4);
    if (!rc) rc = grow(ic->ctx, (void**)&ic->d_recs, ic->cap_recs, recs.size() * 4);
    if (!rc) rc = ilm_upload(ic->ctx, ic->d_nodes, nodes.data(), nodes.size() * 4);
    if (!rc) rc = ilm_upload(ic->ctx, ic->d_refs, refs.data(), refs.size() * 4);
    if (!rc) rc = ilm_upload(ic->ctx, ic->d_recs, recs.data(), recs.size() * 4);
    return rc;
}

ilm::IcArgs args_of(const ilm_ic* ic) {
    ilm::IcArgs a;
    a.nodes = ic->d_nodes;
    a.refs = ic->d_refs;
    a.recs = ic->d_recs;
    a.inv_threshold = 1.0f / ic->prm.error_threshold;
    a.variant = ic->prm.variant;
    return a;
}

int add_records(ilm_ic* ic, const std::vector<ilm_ic::Rec>& add) {
    for (const ilm_ic::Rec& r : add) {
        ic->recs.push_back(r);
        insert(ic, 0, (int32_t)ic->recs.size() - 1, ic->prm.cache_depth);
    }
    return upload(ic);
}

}  // namespace

extern "C" {

int ilm_ic_create(ilm_scene* scene, const ilm_ic_params* params, ilm_ic** out) {
    if (!scene || !out) return fail(ILM_ERR_INVALID, "null argument");
    *out = nullptr;
    ilm_ic_params p{};
    if (params) p = *params;
    // ICIntegratorFactory defaults (Integrator/IntegratorFactories.h:159-183)
    if (p.cache_depth <= 0) p.cache_depth = 16;
    if (!(p.error_threshold > 0)) p.error_threshold = 1.0f;
    if (!(p.ambient_resolution > 0)) p.ambient_resolution = 1e-3f;
    if (!(p.ambient_multiplier > 0)) p.ambient_multiplier = 1e3f;
    if (p.azimuth_strata <= 0) p.azimuth_strata = 12;
    if (p.altitude_strata <= 0) p.altitude_strata = 4;
    if (p.ray_depth <= 0) p.ray_depth = 6;
    if (!(p.eps > 0)) p.eps = 1e-4f;
    if (!(p.indirect_factor > 0)) p.indirect_factor = 1.0f / ilm::kPiF;
    if (p.placement_stride <= 0) p.placement_stride = 16;
    if (p.variant != ILM_IC_WARD && p.variant != ILM_IC_TABELION) return fail(ILM_ERR_INVALID, "unknown cache variant");
    if (p.cache_depth > 40) return fail(ILM_ERR_INVALID, "cache depth above 40");
    ilm_mesh* m = ilm::scene_mesh(scene);
    if (!m || m->pos.empty()) return fail(ILM_ERR_STATE, "scene has no geometry");
    ilm_ic* ic = new ilm_ic();
    ic->scene = scene;
    ic->ctx = ilm::scene_ctx(scene);
    ic->prm = p;
    // ICIntegrator::Initialise (:306-323): the space's bounds
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (size_t v = 0; v < m->pos.size() / 3; ++v)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], m->pos[3 * v + k]);
            hi[k] = std::max(hi[k], m->pos[3 * v + k]);
        }
    float ext[3], c[3], L = 0.f;
    for (int k = 0; k < 3; ++k) {
        ext[k] = (hi[k] - lo[k]) * 0.5f;
        c[k] = lo[k] + ext[k];
        L = std::max(L, std::fabs(hi[k] - lo[k]));
    }
    ilm_ic::Node root;
    for (int k = 0; k < 3; ++k) {
        float a = (c[k] - L) * (1.0f + FLT_EPSILON), b = (c[k] + L) * (1.0f + FLT_EPSILON);
        root.lo[k] = std::min(a, b);
        root.hi[k] = std::max(a, b);
    }
    ic->nodes.push_back(root);
    ic->rmin = L * p.ambient_resolution;
    ic->rmax = L * p.ambient_multiplier;
    int rc = ilm_alloc(ic->ctx, 64, (void**)&ic->d_count);
    if (!rc) rc = upload(ic);
    if (rc) {
        ilm_ic_destroy(ic);
        return rc;
    }
    *out = ic;
    return ILM_OK;
}

int ilm_ic_destroy(ilm_ic* ic) {
    if (!ic) return ILM_OK;
    for (void* p : {(void*)ic->d_nodes, (void*)ic->d_refs, (void*)ic->d_recs, (void*)ic->d_count})
        if (p) ilm_free(ic->ctx, p);
    for (ilm_ic::Buf* b : {&ic->gbuf, &ic->own, &ic->list, &ic->miss, &ic->sel, &ic->nsel, &ic->tmp, &ic->rays, &ic->L,
                           &ic->ft, &ic->newrecs})
        if (b->p) ilm_free(ic->ctx, b->p);
    delete ic;
    return ILM_OK;
}

int ilm_ic_reset(ilm_ic* ic) {
    if (!ic) return fail(ILM_ERR_INVALID, "null cache");
    ilm_ic::Node root = ic->nodes[0];
    root.child = -1;
    root.recs.clear();
    ic->nodes.assign(1, root);
    ic->recs.clear();
    return upload(ic);
}

int ilm_ic_info(ilm_ic* ic, int64_t* records, int64_t* nodes, int64_t* refs, float* rmin, float* rmax,
                float* root_lo_hi) {
    if (!ic) return fail(ILM_ERR_INVALID, "null cache");
    int64_t nref = 0;
    for (const ilm_ic::Node& n : ic->nodes) nref += (int64_t)n.recs.size();
    if (records) *records = (int64_t)ic->recs.size();
    if (nodes) *nodes = (int64_t)ic->nodes.size();
    if (refs) *refs = nref;
    if (rmin) *rmin = ic->rmin;
    if (rmax) *rmax = ic->rmax;
    if (root_lo_hi)
        for (int k = 0; k < 3; ++k) {
            root_lo_hi[k] = ic->nodes[0].lo[k];
            root_lo_hi[3 + k] = ic->nodes[0].hi[k];
        }
    return ILM_OK;
}

int ilm_ic_insert(ilm_ic* ic, const float* records, int64_t n) {
    if (!ic || (n > 0 && !records) || n < 0) return fail(ILM_ERR_INVALID, "bad arguments");
    std::vector<ilm_ic::Rec> add((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        const float* q = records + 10 * i;
        ilm_ic::Rec& r = add[i];
        for (int k = 0; k < 3; ++k) {
            r.P[k] = q[k];
            r.N[k] = q[3 + k];
            r.E[k] = q[6 + k];
        }
        r.ri = q[9];
        if (!(r.ri > 0)) return fail(ILM_ERR_INVALID, "record radius must be positive");
        r.rc = std::max(ic->rmin, std::min(ic->rmax, r.ri));
    }
    return add_records(ic, add);
}

int ilm_ic_records(ilm_ic* ic, float* out) {
    if (!ic || !out) return fail(ILM_ERR_INVALID, "null argument");
    for (size_t i = 0; i < ic->recs.size(); ++i) {
        const ilm_ic::Rec& r = ic->recs[i];
        float* o = out + 11 * i;
        for (int k = 0; k < 3; ++k) {
            o[k] = r.P[k];
            o[3 + k] = r.N[k];
            o[6 + k] = r.E[k];
        }
        o[9] = r.ri;
        o[10] = r.rc;
    }
    return ILM_OK;
}

int ilm_ic_node_records(ilm_ic* ic, int64_t node, int32_t* count, int32_t* child, int32_t* recs, int32_t cap) {
    if (!ic || node < 0 || node >= (int64_t)ic->nodes.size()) return fail(ILM_ERR_INVALID, "bad node");
    const ilm_ic::Node& n = ic->nodes[node];
    if (count) *count = (int32_t)n.recs.size();
    if (child) *child = n.child;
    if (recs)
        for (int32_t i = 0; i < std::min<int32_t>(cap, (int32_t)n.recs.size()); ++i) recs[i] = n.recs[i];
    return ILM_OK;
}

int ilm_ic_lookup(ilm_ic* ic, const float* points, int64_t n, float* out, void* stream) {
    if (!ic || n < 0 || (n > 0 && (!points || !out))) return fail(ILM_ERR_INVALID, "bad arguments");
    if (n == 0) return ILM_OK;
    hipStream_t st = stream ? (hipStream_t)stream : ic->ctx->stream;
    hipLaunchKernelGGL(ilm::k_ic_lookup_points, dim3(ilm::grid_of(n)), dim3(256), 0, st, args_of(ic), points, n, out);
    ILM_HIP(hipGetLastError());
    return ILM_OK;
}

int ilm_ic_render(ilm_ic* ic, const ilm_render_params* prm, double* accum, ilm_ic_stats* stats) {
    if (!ic || !prm || !accum) return fail(ILM_ERR_INVALID, "null argument");
    if (prm->width <= 0 || prm->height <= 0 || prm->spp <= 0)
        return fail(ILM_ERR_INVALID, "width, height and spp must be positive");
    ilm_ctx* ctx = ic->ctx;
    ILM_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    auto t0 = std::chrono::steady_clock::now();
    const int64_t npix = (int64_t)prm->width * prm->height, nsamp = npix * prm->spp;
    if (nsamp > INT32_MAX) return fail(ILM_ERR_INVALID, "more than 2^31 - 1 samples in one frame");
    const int mn = ic->prm.azimuth_strata * ic->prm.altitude_strata;
    auto grow = [&](ilm_ic::Buf& b, size_t bytes) -> int {
        if (b.bytes >= bytes && b.p) return ILM_OK;
        if (b.p) ilm_free(ctx, b.p);
        b.p = nullptr;
        b.bytes = 0;
        const size_t want = std::max<size_t>(bytes, 256);
        int rc = ilm_alloc(ctx, want, &b.p);
        if (!rc) b.bytes = want;
        return rc;
    };
    int rc = grow(ic->gbuf, (size_t)nsamp * 48);
    if (!rc) rc = grow(ic->own, (size_t)nsamp * 4);
    if (!rc) rc = grow(ic->list, (size_t)nsamp * 4);
    if (!rc) rc = grow(ic->miss, (size_t)nsamp);
    if (!rc) rc = grow(ic->sel, (size_t)nsamp * 4);
    if (!rc) rc = grow(ic->nsel, 8);
    if (rc) return rc;
    float* gbuf = (float*)ic->gbuf.p;
    int32_t* own = (int32_t*)ic->own.p;
    int32_t* list = (int32_t*)ic->list.p;
    int32_t* slist = (int32_t*)ic->sel.p;
    uint8_t* miss = (uint8_t*)ic->miss.p;
    int32_t* nsel = (int32_t*)ic->nsel.p;
    // 1. direct light (PathIntegrator with one vertex: emission at the first
    //    hit and next-event estimation there) and the first-hit records
    ilm_render_params p1 = *prm;
    p1.max_depth = 1;
    ilm::PathMode gm;
    gm.gbuf = gbuf;
    ilm_render_stats s1{};
    rc = ilm::render_paths(ic->scene, p1, gm, accum, &s1);
    if (rc) return rc;
    hipLaunchKernelGGL(ilm::k_ic_fill_i32, dim3(ilm::grid_of(nsamp)), dim3(256), 0, st, own, nsamp, -1);
    // 2. record placement passes: every stride-th pixel (first sample), the
    //    stride halving down to 1, then every sample. The samples that no
    //    record covers are selected on the device in candidate order, which
    //    is the order their records are inserted in.
    int64_t created = 0, gather = 0;
    std::vector<int> strides;
    for (int s = ic->prm.placement_stride; s >= 1; s /= 2) strides.push_back(s);
    strides.push_back(0);
    const bool has_stride1 = std::find(strides.begin(), strides.end(), 1) != strides.end();
    for (int stride : strides) {
        // with one sample per pixel the every-sample pass probes the stride-1
        // pass's candidates again, and each of those that was uncovered has
        // since made a record that covers it (weight at its own record: w > 0),
        // so the pass cannot select anything
        if (stride == 0 && prm->spp == 1 && has_stride1) continue;
        const int64_t nx = stride > 0 ? (prm->width + stride - 1) / stride : 0;
        const int64_t ny = stride > 0 ? (prm->height + stride - 1) / stride : 0;
        const int64_t nc = stride > 0 ? nx * ny : nsamp;
        hipLaunchKernelGGL(ilm::k_ic_probe_pass, dim3(ilm::grid_of(nc)), dim3(256), 0, st, args_of(ic),
                           reinterpret_cast<const float4*>(gbuf), stride, prm->width, prm->spp, nx, nc, list, miss);
        ILM_HIP(hipGetLastError());
        size_t tmp_bytes = 0;
        ILM_HIP(hipcub::DeviceSelect::Flagged(nullptr, tmp_bytes, list, miss, slist, nsel, (int)nc, st));
        rc = grow(ic->tmp, tmp_bytes);
        if (rc) return rc;
        ILM_HIP(hipcub::DeviceSelect::Flagged(ic->tmp.p, tmp_bytes, list, miss, slist, nsel, (int)nc, st));
        int32_t hsel = 0;
        rc = ilm_download(ctx, &hsel, nsel, sizeof(hsel));
        if (rc) return rc;
        if (hsel == 0) continue;
        const int64_t nrec = hsel, nray = nrec * mn;
        if (nray > INT32_MAX) return fail(ILM_ERR_INVALID, "too many gather rays in one pass");
        rc = grow(ic->rays, (size_t)nray * sizeof(ilm_ray));
        if (!rc) rc = grow(ic->L, (size_t)nray * 24);
        if (!rc) rc = grow(ic->ft, (size_t)nray * 4);
        if (!rc) rc = grow(ic->newrecs, (size_t)nrec * 48);
        if (rc) return rc;
        void *rays = ic->rays.p, *L = ic->L.p, *ft = ic->ft.p, *recs = ic->newrecs.p;
        rc = ilm_memset(ctx, L, 0, (size_t)nray * 24);
        if (rc) return rc;
        hipLaunchKernelGGL(ilm::k_ic_gather_rays, dim3(ilm::grid_of(nray)), dim3(256), 0, st,
                           reinterpret_cast<const float4*>(gbuf), (const int32_t*)slist, nrec,
                           ic->prm.azimuth_strata, ic->prm.altitude_strata,
                           ic->prm.variant == ILM_IC_TABELION ? 0.f : ic->prm.eps, prm->seed, ic->passes,
                           (ilm_ray*)rays);
        // PathLi for every gather ray: one row of nray paths
        ilm_render_params pg = *prm;
        pg.width = (int32_t)nray;
        pg.height = 1;
        pg.spp = 1;
        pg.spp_offset = (int32_t)ic->passes;
        pg.max_depth = ic->prm.ray_depth;
        pg.rr_depth = 3;
        pg.eps = ic->prm.eps;
        pg.seed = prm->seed ^ 0x9e3779b9u;
        ilm::PathMode mode;
        mode.gather = (const ilm_ray*)rays;
        mode.first_t = (float*)ft;
        mode.rr_rule = 1;
        mode.emit0 = false;
        rc = ilm::render_paths(ic->scene, pg, mode, (double*)L, nullptr);
        if (rc) return rc;
        hipLaunchKernelGGL(ilm::k_ic_reduce, dim3(ilm::grid_of(nrec)), dim3(256), 0, st,
                           reinterpret_cast<const float4*>(gbuf), (const int32_t*)slist, nrec, mn,
                           (const double*)L, (const float*)ft, ic->rmin, ic->rmax, ic->prm.variant, (float4*)recs);
        // the sample that made a record takes that record's irradiance
        // (GetIrradiance returns it, ICIntegrator.cpp:487-493)
        hipLaunchKernelGGL(ilm::k_ic_own, dim3(ilm::grid_of(nrec)), dim3(256), 0, st, (const int32_t*)slist, nrec,
                           (int32_t)ic->recs.size(), own);
        std::vector<float> hrec((size_t)nrec * 12);
        rc = ilm_download(ctx, hrec.data(), recs, hrec.size() * 4);
        if (rc) return rc;
        std::vector<ilm_ic::Rec> add((size_t)nrec);
        for (int64_t r = 0; r < nrec; ++r) {
            const float* q = &hrec[12 * r];
            ilm_ic::Rec& R = add[r];
            for (int k = 0; k < 3; ++k) {
                R.P[k] = q[k];
                R.N[k] = q[4 + k];
                R.E[k] = q[8 + k];
            }
            R.rc = q[3];
            R.ri = q[7];
        }
        rc = add_records(ic, add);
        if (rc) return rc;
        created += nrec;
        gather += nray;
        ++ic->passes;
    }
    // 3. indirect light of every sample from the final cache
    float3 missc = make_float3(ic->prm.miss_radiance[0], ic->prm.miss_radiance[1], ic->prm.miss_radiance[2]);
    ILM_HIP(hipMemsetAsync(ic->d_count, 0, 8, st));
    hipLaunchKernelGGL(ilm::k_ic_shade, dim3(ilm::grid_of(npix)), dim3(256), 0, st, args_of(ic),
                       reinterpret_cast<const float4*>(gbuf), own, npix, prm->spp, ic->prm.indirect_factor, missc,
                       accum, ic->d_count);
    ILM_HIP(hipGetLastError());
    unsigned long long looked = 0;
    rc = ilm_download(ctx, &looked, ic->d_count, 8);
    if (rc) return rc;
    auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        stats->records_created = created;
        stats->records_total = (int64_t)ic->recs.size();
        stats->gather_rays = gather;
        stats->lookups = (int64_t)looked;
        stats->ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    return ilm_sync(ctx);
}

}  // extern "C"
