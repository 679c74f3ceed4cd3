// Wavefront path integrator for CDNA4 (gfx950).
//
// Restates PathIntegrator::Radiance (Integrator/PathIntegrator.cpp:44-201)
// driven by BaseRenderer::RenderRegion (Renderer/BaseRenderer.cpp:62-137) as
// a wavefront: all paths of a batch advance one bounce at a time through
//   generate -> extend (closest hit) -> shade -> shadow (any hit) -> accumulate
// Queue lengths live in device memory, so a bounce is a fixed sequence of
// launches with no host round trip. Queues are appended with one atomic per
// wave (ballot + popcount + lane prefix). Random numbers are counter-based
// (Philox4x64-10 keyed by seed; counter = pixel, sample, bounce, stream), so a
// path's samples do not depend on scheduling and renders are reproducible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "bvh_build.hpp"
#include "geom.hpp"
#include "internal.hpp"

namespace ilm {

constexpr int kPtBlock = 256;
// Rounds of resident workgroups per shading pass: one (each thread walks its
// share of the queue) measured fastest; 2 and 4 rounds, and 8 workgroups per
// CU whatever fits, were slower.
constexpr int kShadeRounds = 1;
constexpr int kCounts = 16;
constexpr int kAccumSlots = 4;  // shadow slots per vertex summed from one batch of loads  // the integrator's device counters (PtArgs::counts), cleared per frame

struct Material {
    f3 kd, ks;
    float exponent;
    int type;  // 0 Lambertian, 1 Lambertian + Phong lobe
};

__device__ __forceinline__ Material load_material(const float* m, int i) {
    const float* q = m + 8 * i;
    Material r;
    r.kd = mk(q[0], q[1], q[2]);
    r.ks = mk(q[3], q[4], q[5]);
    r.exponent = q[6];
    r.type = (int)q[7];
    return r;
}

// Phong lobe (no reference counterpart): normalised modified Phong about the
// mirror direction of wOut in the local frame.
__device__ __forceinline__ f3 phong_f(const Material& m, f3 wo, f3 wi) {
    f3 r = mk(-wo.x, -wo.y, wo.z);
    float c = dot(r, wi);
    if (c <= 0.f) return mk(0, 0, 0);
    return m.ks * ((m.exponent + 2.f) * (0.5f * kInvPi) * powf(c, m.exponent));
}
__device__ __forceinline__ float phong_pdf(const Material& m, f3 wo, f3 wi) {
    f3 r = mk(-wo.x, -wo.y, wo.z);
    float c = dot(r, wi);
    if (c <= 0.f) return 0.f;
    return (m.exponent + 1.f) * (0.5f * kInvPi) * powf(c, m.exponent);
}
// BSDF::F (Material/BSDF.cpp:215-237): reflection only when wIn.z > 0.
__device__ __forceinline__ f3 bsdf_f(const Material& m, f3 wo, f3 wi) {
    if (!(wi.z > 0.f)) return mk(0, 0, 0);
    f3 f = m.kd * kInvPi;  // Lambertian::F (Material/LambertianBxDF.cpp:41-44)
    if (m.type == 1) f = f + phong_f(m, wo, wi);
    return f;
}
// BSDF::SampleF (Material/BSDF.cpp:141-211) over {Lambertian, [Phong]}.
__device__ __forceinline__ f3 bsdf_sample(const Material& m, f3 wo, f3& wi, float u, float v, float w, float& pdf) {
    int count = (m.type == 1) ? 2 : 1;
    int which = (int)(w * count) % count;
    if (which == 0) {
        // Lambertian::SampleF (Material/LambertianBxDF.cpp:27-37)
        wi = cosine_hemisphere(u, v);
        if (wo.z < 0.f) wi.z = -wi.z;
        pdf = fabsf(wi.z) * kInvPi;
    } else {
        f3 r = mk(-wo.x, -wo.y, wo.z);
        float ct = powf(u, 1.f / (m.exponent + 1.f));
        float st = sqrtf(fmaxf(0.f, 1.f - ct * ct));
        float ph = 2.0f * kPi * v;
        wi = to_world(basis_from_w(r), mk(st * cosf(ph), st * sinf(ph), ct));
        pdf = phong_pdf(m, wo, wi);
    }
    if (pdf == 0.0f) return mk(0, 0, 0);
    if (count > 1) {
        if (which == 0) pdf += phong_pdf(m, wo, wi);
        else pdf += fabsf(wi.z) * kInvPi;
        pdf /= count;
    }
    return bsdf_f(m, wo, wi);
}

__device__ __forceinline__ float luminance(f3 s) { return 0.212671f * s.x + 0.715160f * s.y + 0.072169f * s.z; }
__device__ __forceinline__ bool is_black(f3 s) { return s.x == 0.f && s.y == 0.f && s.z == 0.f; }


// Appends the block's wanting threads to a queue with one atomic and returns
// each one's slot (-1 for the others). Every thread of the block calls it
// together; cnt (one entry per wave) and base are the call site's own LDS.
__device__ __forceinline__ int64_t block_append(unsigned long long* counter, bool want, uint32_t* cnt,
                                                unsigned long long* base) {
    const unsigned long long m = __ballot(want);
    const int wv = threadIdx.x >> 6;
    if (__lane_id() == 0) cnt[wv] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long tot = 0;
        for (int w = 0; w < kPtBlock / 64; ++w) tot += cnt[w];
        *base = tot ? atomicAdd(counter, tot) : 0ull;
    }
    __syncthreads();
    unsigned long long off = *base;
    for (int w = 0; w < wv; ++w) off += cnt[w];
    return want ? (int64_t)(off + (unsigned long long)__popcll(m & __lanemask_lt())) : -1;
}

struct CamK {
    f3 centre, corner, across, up;
};

// One image of a frame: its camera, size and samples, the lights and BVH root
// of its scene, and where its paths and pixels start in the frame's global
// path and pixel numbering (views are stored in that order). A single-scene
// frame is one view with root 0 and offsets 0.
struct View {
    CamK cam;
    int W, H, spp;
    int l0, l1;       // lights [l0, l1) of the view's scene
    uint32_t root;    // BVH root node of the view's scene
    int64_t path0;    // first global path
    int64_t pix0;     // first global pixel (accumulator offset / 3)
};

struct PtArgs {
    // scene
    const float4* shade;     // 4 float4 per triangle (64 bytes): (Ng, mat), (n0, light), (n1, 0), (n2, 0)
    const float* mats;
    const float4* ltris;     // emitter triangles, 3 float4 each
    const float4* lights;    // (Le rgb, area)
    const int2* lrange;      // (first emitter triangle, count)
    int nlights;             // lights of all views' scenes
    int max_vl;              // shadow-queue slots per path vertex: the most lights of any view
    // views, and the batch of global paths / pixels this launch covers
    const View* views;
    int nviews;
    int64_t bpath0, bpix0;
    // per-frame constants in device memory, so that a captured frame can be
    // replayed with a new seed and sample offset: [0] seed, [1] spp_offset
    const int32_t* frame;
    int max_depth, rr_depth;
    float eps;
    // path state
    int64_t npaths;
    float4* beta;            // per path
    float* L;                // 3 floats per path
    // queues (ping-pong) and shadow queue
    ilm_ray* rays[2];
    int32_t* qpath[2];
    ilm_hit* hits;
    ilm_ray* srays;
    float4* scontrib;        // rgb contribution, path index in w
    float4* scontrib2;       // contribution to the cache target of a training path (rgb)
    uint8_t* socc;
    // [0], [1] ray queues, [2] shadow queue, [3] extension total, [4] shadow total,
    // [5] cache queries of this batch, [6] training samples of this frame,
    // [7] shadow segments written this bounce (the shadow queue's used slots),
    // [8] training paths at the cache's query vertex in this batch (the query
    // bounce's shadow queue holds their slots only)
    unsigned long long* counts;
    double* accum;
    // radiance cache (nrc_prefix < 0: plain path tracing)
    int nrc_prefix;          // depth of the vertex whose outgoing radiance the cache supplies
    int max_depth_train;     // depth limit of the training paths (prefix + suffix)
    float train_frac;        // fraction of paths extended as training paths
    int64_t train_cap;       // training record slots (train_count rounded up to whole training batches)
    int64_t total_paths;     // paths of the frame (all batches)
    int64_t train_count;     // training paths of the frame: floor(total_paths * train_frac)
    float4* tbeta;           // throughput of a training path relative to its query vertex
    float* T;                // 3 floats per path: radiance leaving the query vertex (training paths)
    int32_t* tslot;          // training-sample slot of a path, or -1
    float* queries;          // 16 floats per cache query (non-training paths)
    float4* qbeta;           // (throughput rgb, path index) per cache query
    float4* qout;            // cache predictions per query
    float* tq;               // training queries (train_cap x 16)
    float4* tt;              // training targets (train_cap x (rgb, weight))
    // irradiance-cache passes
    const ilm_ray* gather;   // non-null: path g starts from gather[g] instead of a camera ray
    float* first_t;          // non-null: first-hit distance of path g (1 when its first ray finds no material)
    float4* gbuf;            // non-null: first-hit record of path g (3 float4: P + kind, Ns, albedo)
    int rr_rule;             // roulette: 0 q = min(0.5, luminance), 1 q = min(0.5, 0.33 (r + g + b))
    int emit0;               // count emission at depth 0
};

// last view whose first path (or pixel) is <= g
__device__ __forceinline__ int find_view(const View* v, int n, int64_t g, bool by_pixel) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if ((by_pixel ? v[mid].pix0 : v[mid].path0) <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Path p of the batch: its view, global pixel (the RNG's pixel counter and
// the accumulator index) and sample index.
__device__ __forceinline__ void path_pixel(const PtArgs& a, int64_t p, int& vi, int64_t& pixel, uint32_t& sample) {
    const int64_t g = a.bpath0 + p;
    vi = find_view(a.views, a.nviews, g, false);
    const int64_t local = g - a.views[vi].path0;
    const int spp = a.views[vi].spp;
    pixel = a.views[vi].pix0 + local / spp;
    sample = (uint32_t)(a.frame[1] + (int)(local % spp));
}

// Training paths of the radiance cache: exactly train_count of the frame's
// total_paths paths, spread evenly over the global path order with a
// per-frame offset (from the frame seed): path g trains when
// floor(((g + 1) K + off) / T) > floor((g K + off) / T), and its record goes
// to slot floor((g K + off) / T), so every training path owns one slot and
// the slots do not depend on scheduling.
__device__ __forceinline__ int64_t train_offset(const PtArgs& a) {
    const Block8 b = sample_block((uint32_t)a.frame[0], ~0ull, 0u, 0u, 15u);
    return min((int64_t)((double)b.v[0] * (double)a.total_paths), a.total_paths - 1);
}
__device__ __forceinline__ bool is_training_path(const PtArgs& a, int64_t g, int64_t off) {
    if (a.train_count <= 0) return false;
    return ((g + 1) * a.train_count + off) / a.total_paths > (g * a.train_count + off) / a.total_paths;
}
__device__ __forceinline__ int64_t training_slot(const PtArgs& a, int64_t g, int64_t off) {
    return (g * a.train_count + off) / a.total_paths;
}

__global__ void k_generate(PtArgs a) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) {
        a.counts[0] = (unsigned long long)a.npaths;
        a.counts[1] = 0ull;
        a.counts[2] = 0ull;
        a.counts[8] = 0ull;
    }
    if (p >= a.npaths) return;
    int vi;
    int64_t pixel;
    uint32_t sample;
    path_pixel(a, p, vi, pixel, sample);
    const View& v = a.views[vi];
    Block8 b = sample_block((uint32_t)a.frame[0], (uint64_t)pixel, sample, 0, 0);
    if (a.gather) {
        // ICIntegrator::ComputeRecord's gather rays (ICIntegrator.cpp:512-519)
        ilm_ray r = a.gather[a.bpath0 + p];
        r.tmin = __uint_as_float(v.root);
        a.rays[0][p] = r;
        a.qpath[0][p] = (int32_t)p;
        a.beta[p] = make_float4(1.f, 1.f, 1.f, 0.f);
        a.L[3 * p] = 0.f;
        a.L[3 * p + 1] = 0.f;
        a.L[3 * p + 2] = 0.f;
        return;
    }
    const int64_t lp = pixel - v.pix0;
    int x = (int)(lp % v.W), y = (int)(lp / v.W);
    // PerspectiveCamera::GetRay (Camera/PerspectiveCamera.cpp:53-60) as fed by
    // RenderRegion with the box filter's -0.5 shift (BaseRenderer.cpp:106-118)
    float rcpW = 1.f / (float)v.W, rcpH = 1.f / (float)v.H;
    float nx = (float)x * rcpW, ny = (float)y * rcpH;
    float xi1 = (b.v[0] - 0.5f) * rcpW, xi2 = (b.v[1] - 0.5f) * rcpH;
    f3 t = v.cam.corner + v.cam.across * (nx + xi1) + v.cam.up * (ny + xi2);
    f3 d = normalize(t - v.cam.centre);
    // tmin carries the BVH root of the view's scene (the traversal's t window
    // starts at 1e-5 regardless of tmin, as the reference's triangle test does)
    ilm_ray r;
    r.ox = v.cam.centre.x; r.oy = v.cam.centre.y; r.oz = v.cam.centre.z; r.tmin = __uint_as_float(v.root);
    r.dx = d.x; r.dy = d.y; r.dz = d.z; r.tmax = FLT_MAX;
    a.rays[0][p] = r;
    a.qpath[0][p] = (int32_t)p;
    // beta.w flags a training path of the radiance cache
    float train = (a.nrc_prefix >= 0 && is_training_path(a, a.bpath0 + p, train_offset(a))) ? 1.f : 0.f;
    a.beta[p] = make_float4(1.f, 1.f, 1.f, train);
    if (a.nrc_prefix >= 0) a.tslot[p] = -1;
    a.L[3 * p] = 0.f;
    a.L[3 * p + 1] = 0.f;
    a.L[3 * p + 2] = 0.f;
}

// One bounce of PathIntegrator::Radiance for every path in queue `cur`.
// CACHE: the frame renders with the radiance cache (query vertices,
// training paths); plain frames use the variant without that state.
template <bool CACHE>
__global__ void __launch_bounds__(kPtBlock) k_shade(PtArgs a, int bounce, int cur) {
    const int nxt = cur ^ 1;
    const int64_t count = (int64_t)a.counts[cur];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // the cache's query bounce: only training paths go on, so the shadow
    // slots go to them alone, in append order (counter 8); its queue length
    // is set after this pass (k_query_shadow_count)
    const bool qbounce = CACHE && a.nrc_prefix >= 0 && bounce == a.nrc_prefix;
    // the shadow queue has max_vl fixed slots per queue entry (slot j: the
    // view's light j); unused slots hold an empty segment, so the queue needs
    // no append atomics. Its length is known up front.
    if (!qbounce && blockIdx.x == 0 && threadIdx.x == 0)
        a.counts[2] = (unsigned long long)count * (unsigned long long)a.max_vl;
    uint32_t nshadow = 0;  // shadow segments this lane wrote (for the frame's ray count)
    // LDS of the block appends (0 continuing paths, 1 cache queries, 2
    // training records, 3 the query bounce's shadow entries), double-buffered
    // by iteration parity: a wave may start the next iteration while another
    // still reads this one's
    __shared__ uint32_t s_cnt[4][2][kPtBlock / 64];
    __shared__ unsigned long long s_base[4][2];
    // iterate in whole-block steps so that every thread of the block reaches
    // the queue appends (and their barriers) together
    int par = 0;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < count; base += stride, par ^= 1) {
        const int64_t k = base + (int64_t)threadIdx.x;
        bool live = k < count;
        int32_t p = 0;
        ilm_hit h;
        ilm_ray ray;
        if (live) {
            p = a.qpath[cur][k];
            h = a.hits[k];
            ray = a.rays[cur][k];
            live = h.prim >= 0;  // a miss adds nothing for area lights (PathIntegrator.cpp:192-199)
        }
        int mat_id = -1, light_id = -1;
        f3 P = mk(0, 0, 0), Ng = mk(0, 0, 1), Ns = mk(0, 0, 1), wOut = mk(0, 0, 1);
        f3 beta = mk(0, 0, 0), tbeta = mk(0, 0, 0);
        bool train = false;
        int64_t pixel = 0;
        uint32_t sample = 0;
        int vi = 0;
        if (live) {
            const float4* sr = a.shade + 4 * (int64_t)h.prim;
            float4 G = sr[0], N0 = sr[1], N1 = sr[2], N2 = sr[3];
            mat_id = __float_as_int(G.w);
            light_id = __float_as_int(N0.w);
            f3 rd = mk(ray.dx, ray.dy, ray.dz);
            P = mk(ray.ox, ray.oy, ray.oz) + h.t * rd;  // GeometricPrimitive.cpp:151
            float alpha = 1.f - h.u - h.v;
            Ns = normalize(mk(N0.x * alpha + N1.x * h.u + N2.x * h.v, N0.y * alpha + N1.y * h.u + N2.y * h.v,
                              N0.z * alpha + N1.z * h.u + N2.z * h.v));
            Ng = mk(G.x, G.y, G.z);
            wOut = -normalize(rd);
            float4 b4 = a.beta[p];
            beta = mk(b4.x, b4.y, b4.z);
            train = CACHE && b4.w != 0.f;
            if (train && bounce > a.nrc_prefix) {
                float4 t4 = a.tbeta[p];
                tbeta = mk(t4.x, t4.y, t4.z);
            }
            path_pixel(a, p, vi, pixel, sample);
            if (mat_id < 0) live = false;  // no material: the path ends (PathIntegrator.cpp:110-111)
        }
        Material mat;
        if (live) mat = load_material(a.mats, mat_id);
        if (bounce == 0 && k < count) {
            const int64_t g = a.bpath0 + p;
            // ICIntegrator::PathLi keeps 1 as the first segment's length unless
            // its first hit has a material (ICIntegrator.cpp:566-573)
            if (a.first_t) a.first_t[g] = live ? h.t : 1.0f;
            if (a.gbuf) {
                // first-hit record for ICIntegrator::Radiance (ICIntegrator.cpp:380-407):
                // kind 0 a hit without material, 1 reflector, 2 emitter, 3 miss;
                // albedo = Rho (kd; kd + ks for the Phong lobe, which has no
                // reference counterpart)
                const float kind = h.prim < 0 ? 3.f : !live ? 0.f : (light_id >= 0 ? 2.f : 1.f);
                f3 rho = live ? (mat.type == 1 ? mat.kd + mat.ks : mat.kd) : mk(0, 0, 0);
                a.gbuf[3 * g] = make_float4(P.x, P.y, P.z, kind);
                a.gbuf[3 * g + 1] = make_float4(Ns.x, Ns.y, Ns.z, 0.f);
                a.gbuf[3 * g + 2] = make_float4(rho.x, rho.y, rho.z, 0.f);
            }
        }
        // radiance cache: at the query vertex a non-training path hands its
        // remaining radiance to the cache; a training path records the query and
        // keeps going, measuring that radiance from here on with tbeta
        bool to_cache = false, new_train = false;
        if (CACHE && live && a.nrc_prefix >= 0 && bounce == a.nrc_prefix) {
            to_cache = !train;
            new_train = train;
            if (train) tbeta = mk(1.f, 1.f, 1.f);
        }
        float qrec[16];
        if (CACHE && (to_cache || new_train)) {
            float ct = fminf(fmaxf(wOut.z, -1.f), 1.f), cn = fminf(fmaxf(Ns.z, -1.f), 1.f);
            qrec[0] = P.x; qrec[1] = P.y; qrec[2] = P.z;
            qrec[3] = acosf(ct) * kInvPi;
            qrec[4] = (atan2f(wOut.y, wOut.x) + kPi) * (0.5f * kInvPi);
            qrec[5] = acosf(cn) * kInvPi;
            qrec[6] = (atan2f(Ns.y, Ns.x) + kPi) * (0.5f * kInvPi);
            qrec[7] = mat.type == 1 ? sqrtf(2.f / (mat.exponent + 2.f)) : 1.f;
            qrec[8] = mat.kd.x; qrec[9] = mat.kd.y; qrec[10] = mat.kd.z;
            qrec[11] = mat.type == 1 ? mat.ks.x : 0.f;
            qrec[12] = mat.type == 1 ? mat.ks.y : 0.f;
            qrec[13] = mat.type == 1 ? mat.ks.z : 0.f;
            qrec[14] = 0.f; qrec[15] = 0.f;
        }
        int64_t sentry = k;  // the vertex's shadow-queue entry (slots sentry * max_vl + j)
        if (qbounce) {  // uniform: the appends hold barriers
            int64_t qs = block_append(&a.counts[5], to_cache, s_cnt[1][par], &s_base[1][par]);
            if (to_cache) {
                float4* d = reinterpret_cast<float4*>(a.queries + 16 * qs);
                for (int k = 0; k < 4; ++k) d[k] = make_float4(qrec[4 * k], qrec[4 * k + 1], qrec[4 * k + 2], qrec[4 * k + 3]);
                a.qbeta[qs] = make_float4(beta.x, beta.y, beta.z, __int_as_float(p));
                live = false;
            }
            // counts the frame's training records; the record's slot is the path's own
            (void)block_append(&a.counts[6], new_train, s_cnt[2][par], &s_base[2][par]);
            sentry = block_append(&a.counts[8], new_train, s_cnt[3][par], &s_base[3][par]);
            if (new_train) {
                const int64_t ts = training_slot(a, a.bpath0 + p, train_offset(a));
                if (ts < a.train_cap) {
                    float4* d = reinterpret_cast<float4*>(a.tq + 16 * ts);
                    for (int k = 0; k < 4; ++k) d[k] = make_float4(qrec[4 * k], qrec[4 * k + 1], qrec[4 * k + 2], qrec[4 * k + 3]);
                    a.tslot[p] = (int32_t)ts;
                }
                a.T[3 * p] = 0.f;
                a.T[3 * p + 1] = 0.f;
                a.T[3 * p + 2] = 0.f;
            }
        }
        const bool measure = CACHE && train && bounce >= a.nrc_prefix && a.nrc_prefix >= 0;
        const bool last = bounce + 1 >= (train ? a.max_depth_train : a.max_depth);
        // emission at depth 0 (no specular BxDFs), DiffuseAreaLight::Radiance (DiffuseAreaLight.cpp:47-50)
        if (live && a.emit0 && bounce == 0 && light_id >= 0 && dot(Ns, wOut) > 0.f) {
            float4 le = a.lights[light_id];
            atomicAdd(&a.L[3 * p], beta.x * le.x);
            atomicAdd(&a.L[3 * p + 1], beta.y * le.y);
            atomicAdd(&a.L[3 * p + 2], beta.z * le.z);
        }
        // next-event estimation over every light except the one hit
        // (SampleAllLights / EstimateDirectLighting, Integrator.cpp:114-185)
        Basis gb = basis_from_w(Ng);
        // the view's scene's lights: light j of the view fills shadow slot j
        const int vl0 = a.views[vi].l0, vl1 = a.views[vi].l1;
        const uint32_t root = a.views[vi].root;
        for (int j = 0; j < a.max_vl; ++j) {
            const int l = vl0 + j;
            bool want = false;
            ilm_ray sray;
            f3 contrib = mk(0, 0, 0), contrib2 = mk(0, 0, 0);
            if (live && l != light_id && l < vl1) {
                Block8 lb = sample_block((uint32_t)a.frame[0], (uint64_t)pixel, sample, (uint32_t)bounce,
                                         1u + (uint32_t)(l - vl0));
                int2 rg = a.lrange[l];
                float4 le = a.lights[l];
                f3 Q, nl;
                if (le.w < 0.f) {
                    // Quad::SamplePoint (Shape/Quad.cpp:146-153): the bilinear map
                    // v1 + (v3 - v1) u + (v4 - v3) v + (v1 - v3 + v2 - v4) u v and the
                    // winding normal cross(v3 - v1, v2 - v1), normalised by
                    // DiffuseAreaLight::SamplePoint (DiffuseAreaLight.cpp:95-97) and
                    // not turned toward the shaded point
                    const float4* lt = a.ltris + 3 * (int64_t)rg.x;
                    const f3 v1 = mk(lt[0].x, lt[0].y, lt[0].z), v3 = mk(lt[1].x, lt[1].y, lt[1].z);
                    const f3 v2 = mk(lt[2].x, lt[2].y, lt[2].z), v4 = mk(lt[4].x, lt[4].y, lt[4].z);
                    const float u = lb.v[0], v = lb.v[1];
                    Q = v1 + (v3 - v1) * u + (v4 - v3) * v + (v1 - v3 + v2 - v4) * (u * v);
                    nl = normalize(cross(v3 - v1, v2 - v1));
                } else {
                    int kk = min((int)((float)rg.y * lb.v[2]), rg.y - 1);
                    const float4* lt = a.ltris + 3 * (int64_t)(rg.x + kk);
                    float4 p1 = lt[0], p2 = lt[1], p3 = lt[2];
                    float su = sqrtf(lb.v[0]);
                    float b1 = 1.f - su, b2 = lb.v[1] * su;  // Montecarlo::UniformSampleTriangle
                    float b3 = 1.f - b1 - b2;
                    Q = mk(b1 * p1.x + b2 * p2.x + b3 * p3.x, b1 * p1.y + b2 * p2.y + b3 * p3.y,
                           b1 * p1.z + b2 * p2.z + b3 * p3.z);
                    nl = normalize(cross(mk(p3.x - p1.x, p3.y - p1.y, p3.z - p1.z),
                                         mk(p2.x - p1.x, p2.y - p1.y, p2.z - p1.z)));
                    if (dot(Q - P, nl) > 0.f) nl = -nl;  // ITriangleMesh::SamplePoint (TriangleMesh.cpp:150-161)
                }
                f3 wIn = P - Q;
                float d2 = dot(wIn, wIn);
                wIn = normalize(wIn);
                float c = fmaxf(0.f, dot(wIn, nl));
                float pdf = 1.0f / fabsf(le.w);
                f3 Ls = mk(le.x, le.y, le.z) * c / (pdf * d2);
                if (!is_black(Ls)) {
                    wIn = -wIn;
                    f3 f = bsdf_f(mat, to_local(gb, wOut), to_local(gb, wIn));
                    f3 local = mul(Ls * fmaxf(0.f, dot(wIn, Ns)), f);
                    contrib = mul(beta, local);
                    if (measure) contrib2 = mul(tbeta, local);
                    if (!is_black(contrib) || !is_black(contrib2)) {
                        want = true;
                        f3 sd = Q - P;  // VisibilityQuery::SetSegment (Visibility.cpp:50-53)
                        // tmin (1e-4 in the reference; the t window is fixed) carries the root
                        sray.ox = P.x; sray.oy = P.y; sray.oz = P.z; sray.tmin = __uint_as_float(root);
                        sray.dx = sd.x; sray.dy = sd.y; sray.dz = sd.z; sray.tmax = 1.0f - 1e-4f;
                    }
                }
            }
            if (sentry >= 0 && k < count) {
                const int64_t slot = sentry * a.max_vl + j;
                if (want) {
                    a.srays[slot] = sray;
                    a.scontrib[slot] = make_float4(contrib.x, contrib.y, contrib.z, __int_as_float(p));
                    if (CACHE && a.nrc_prefix >= 0)
                        a.scontrib2[slot] = make_float4(contrib2.x, contrib2.y, contrib2.z, 0.f);
                    ++nshadow;
                } else {
                    // an empty segment (t window below 0: no node is entered)
                    // and no path to credit
                    ilm_ray e;
                    e.ox = 0.f; e.oy = 0.f; e.oz = 0.f; e.tmin = __uint_as_float(0u);
                    e.dx = 0.f; e.dy = 0.f; e.dz = 1.f; e.tmax = -1.0f;
                    a.srays[slot] = e;
                    a.scontrib[slot] = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
                }
            }
        }
        // BSDF sampling in the shading frame (PathIntegrator.cpp:138-177)
        bool cont = false;
        ilm_ray nray;
        if (live) {
            Block8 b = sample_block((uint32_t)a.frame[0], (uint64_t)pixel, sample, (uint32_t)bounce, 0);
            Basis sb = basis_from_w(Ns);
            f3 wiL;
            float pdf = 0.f;
            f3 f = bsdf_sample(mat, to_local(sb, wOut), wiL, b.v[2], b.v[3], b.v[4], pdf);
            if (!is_black(f) && pdf != 0.0f) {
                f3 wIn = to_world(sb, wiL);
                f3 wgt = (f * fabsf(dot(wIn, Ns))) / pdf;
                beta = mul(beta, wgt);
                if (measure) tbeta = mul(tbeta, wgt);
                cont = true;
                if (bounce >= a.rr_depth) {
                    // PathIntegrator: luminance; ICIntegrator::PathLi: 0.33 (r + g + b)
                    // (PathIntegrator.cpp:164-171, ICIntegrator.cpp:626-632)
                    float q = fminf(0.5f, a.rr_rule ? 0.33f * (beta.x + beta.y + beta.z) : luminance(beta));
                    if (b.v[5] > q) cont = false;
                    else {
                        beta = beta / q;
                        if (measure) tbeta = tbeta / q;
                    }
                }
                if (last) cont = false;  // the last vertex's continuation ray is never used
                if (cont) {
                    f3 org = P + Ng * a.eps;  // offset along the unflipped geometric normal
                    nray.ox = org.x; nray.oy = org.y; nray.oz = org.z; nray.tmin = __uint_as_float(root);
                    nray.dx = wIn.x; nray.dy = wIn.y; nray.dz = wIn.z; nray.tmax = FLT_MAX;
                }
            }
        }
        // one atomic per block for the continuing paths
        const int64_t slot = block_append(&a.counts[nxt], cont, s_cnt[0][par], &s_base[0][par]);
        if (cont) {
            a.rays[nxt][slot] = nray;
            a.qpath[nxt][slot] = p;
            a.beta[p] = make_float4(beta.x, beta.y, beta.z, train ? 1.f : 0.f);
            if (measure) a.tbeta[p] = make_float4(tbeta.x, tbeta.y, tbeta.z, 0.f);
        }
    }
    // one atomic per wave for the frame's shadow-ray count
    for (int off = 32; off > 0; off >>= 1) nshadow += __shfl_xor(nshadow, off);
    if (__lane_id() == 0 && nshadow) atomicAdd(&a.counts[7], (unsigned long long)nshadow);
}

// Cache predictions at the query vertices: L += throughput * max(prediction, 0).
__global__ void k_nrc_resolve(PtArgs a) {
    const int64_t count = (int64_t)a.counts[5];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
        float4 b = a.qbeta[k], y = a.qout[k];
        int64_t p = __float_as_int(b.w);
        a.L[3 * p] += b.x * fmaxf(y.x, 0.f);
        a.L[3 * p + 1] += b.y * fmaxf(y.y, 0.f);
        a.L[3 * p + 2] += b.z * fmaxf(y.z, 0.f);
    }
}

// Training targets of the batch's training paths (radiance leaving the query vertex).
__global__ void k_nrc_collect(PtArgs a) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.npaths) return;
    int32_t s = a.tslot[p];
    if (s >= 0) a.tt[s] = make_float4(a.T[3 * p], a.T[3 * p + 1], a.T[3 * p + 2], 1.f);
}

// The query bounce's shadow queue: max_vl slots per training path there.
__global__ void k_query_shadow_count(unsigned long long* counts, int max_vl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counts[2] = counts[8] * (unsigned long long)max_vl;
}

__global__ void k_nrc_reset_queries(unsigned long long* counts) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counts[5] = 0ull;
}

// Start of a frame: no training samples yet (weights 0 mark unused slots).
// The queries of the slots are cleared too: a slot that no training path
// fills this frame still enters the Adam step with weight 0, and a
// non-finite query there (uninitialised memory) would make its zero-weight
// terms NaN instead of 0.
__global__ void k_nrc_clear(PtArgs a) {
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) a.counts[6] = 0ull;
    if (k < a.train_cap) {
        a.tt[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        float4* q = reinterpret_cast<float4*>(a.tq + 16 * k);
        for (int j = 0; j < 4; ++j) q[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

__global__ void k_shadow_accum(PtArgs a, int cur) {
    // The bounce's bookkeeping (k_end_bounce's, which scenes without lights
    // launch instead): this kernel reads none of these counters, and the
    // next shading pass, which sets the shadow queue's length, starts after it.
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.counts[3] += a.counts[cur];
        a.counts[4] += a.counts[7];
        a.counts[cur] = 0ull;
        a.counts[7] = 0ull;
    }
    // one thread per path vertex: its max_vl shadow slots are adjacent, and a
    // path has one vertex per bounce, so its radiance needs no atomics
    const int64_t count = (int64_t)(a.counts[2] / (unsigned long long)a.max_vl);
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count; k += stride) {
        int64_t p = -1;
        float lx = 0.f, ly = 0.f, lz = 0.f, tx = 0.f, ty = 0.f, tz = 0.f;
        if (a.max_vl <= kAccumSlots) {
            // the vertex's slots are loaded together, before any is tested, so
            // their shared cache lines are fetched once; then summed in slot
            // order as in the loop below
            uint8_t occ[kAccumSlots];
            float4 c[kAccumSlots], t[kAccumSlots] = {};
#pragma unroll
            for (int j = 0; j < kAccumSlots; ++j) {
                if (j < a.max_vl) {
                    const int64_t slot = k * a.max_vl + j;
                    occ[j] = a.socc[slot];
                    c[j] = a.scontrib[slot];
                    if (a.nrc_prefix >= 0) t[j] = a.scontrib2[slot];
                }
            }
#pragma unroll
            for (int j = 0; j < kAccumSlots; ++j) {
                if (j >= a.max_vl || occ[j]) continue;
                const int32_t pj = __float_as_int(c[j].w);
                if (pj < 0) continue;  // an unused slot
                p = pj;
                lx += c[j].x;
                ly += c[j].y;
                lz += c[j].z;
                if (a.nrc_prefix >= 0) {
                    tx += t[j].x;
                    ty += t[j].y;
                    tz += t[j].z;
                }
            }
        }
        for (int j = a.max_vl <= kAccumSlots ? a.max_vl : 0; j < a.max_vl; ++j) {
            const int64_t slot = k * a.max_vl + j;
            if (a.socc[slot]) continue;
            const float4 c = a.scontrib[slot];
            const int32_t pj = __float_as_int(c.w);
            if (pj < 0) continue;  // an unused slot
            p = pj;
            lx += c.x;
            ly += c.y;
            lz += c.z;
            if (a.nrc_prefix >= 0) {
                const float4 t = a.scontrib2[slot];
                tx += t.x;
                ty += t.y;
                tz += t.z;
            }
        }
        if (p < 0) continue;
        a.L[3 * p] += lx;
        a.L[3 * p + 1] += ly;
        a.L[3 * p + 2] += lz;
        if (a.nrc_prefix >= 0 && (tx != 0.f || ty != 0.f || tz != 0.f)) {
            a.T[3 * p] += tx;
            a.T[3 * p + 1] += ty;
            a.T[3 * p + 2] += tz;
        }
    }
}

// Bookkeeping between bounces (a one-thread kernel, so the bounce loop needs
// no host synchronisation and no memset nodes).
__global__ void k_clear_counts(unsigned long long* counts) {
    if (threadIdx.x < kCounts) counts[threadIdx.x] = 0ull;
}

__global__ void k_end_bounce(unsigned long long* counts, int cur) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        counts[3] += counts[cur];
        counts[4] += counts[7];
        counts[cur] = 0ull;
        counts[2] = 0ull;
        counts[7] = 0ull;
    }
}

// Sum of the spp samples of each pixel into the FP64 accumulator.
__global__ void k_accumulate(PtArgs a, int64_t npix) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const int64_t gp = a.bpix0 + i;
    const View& v = a.views[find_view(a.views, a.nviews, gp, true)];
    const int64_t p0 = v.path0 + (gp - v.pix0) * v.spp - a.bpath0;  // the pixel's first path in the batch
    double s0 = 0, s1 = 0, s2 = 0;
    for (int s = 0; s < v.spp; ++s) {
        int64_t p = p0 + s;
        s0 += (double)a.L[3 * p];
        s1 += (double)a.L[3 * p + 1];
        s2 += (double)a.L[3 * p + 2];
    }
    double* o = a.accum + 3 * gp;
    o[0] += s0;
    o[1] += s1;
    o[2] += s2;
}

}  // namespace ilm

// ------------------------------------------------------------------ scene
struct ilm_scene {
    ilm_ctx* ctx = nullptr;
    ilm_mesh* mesh = nullptr;
    float4* d_shade = nullptr;
    float* d_mats = nullptr;
    float4* d_ltris = nullptr;
    float4* d_lights = nullptr;
    int2* d_lrange = nullptr;
    int nlights = 0, nmat = 0;
    // path-state buffers (grown on demand)
    int64_t cap = 0;
    int state_nl = 0;
    bool cache_ready = false;  // radiance-cache buffers allocated for `cap` paths
    int64_t train_cap = 0;
    uint64_t train_round = 0;  // Adam steps taken by cache frames (picks the next training batch)
    std::vector<void*> bufs;
    ilm::PtArgs a{};
    // frame constants (device) and their host copy
    int32_t* d_frame = nullptr;
    int32_t h_frame[2] = {0, 0};
    // the frame's views (device; reallocated when more are needed) and host copy
    ilm::View* d_views = nullptr;
    int views_cap = 0;
    std::vector<ilm::View> h_views;
    // captured frame: replayed while the key (everything baked into the
    // launches) is unchanged
    bool use_graph = false;
    hipGraphExec_t graph = nullptr;
    // cache frames: the cache inference at the query vertices runs on a side
    // stream (default priority) while the training suffix traces on the
    // context's stream; created on first use
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    std::vector<int64_t> graph_key;
    // host copies of the shading arrays (merged by batches)
    std::vector<float> h_shade, h_mats, h_lt, h_lrec;
    std::vector<int32_t> h_lrange;
    bool owns_mesh = false;  // a batch's merged scene owns its packed mesh
};

namespace {

void drop_graph(ilm_scene* s) {
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    s->graph = nullptr;
    s->graph_key.clear();
}

void scene_free_state(ilm_scene* s) {
    drop_graph(s);
    for (void* p : s->bufs) ilm_free(s->ctx, p);
    s->bufs.clear();
    s->cap = 0;
    s->state_nl = 0;
    s->cache_ready = false;
    s->train_cap = 0;
}

int scene_alloc(ilm_scene* s, size_t bytes, void** out) {
    int rc = ilm_alloc(s->ctx, bytes, out);
    if (rc == ILM_OK) s->bufs.push_back(*out);
    return rc;
}

// Path state for `cap` paths, each sending at most `nl` shadow rays per bounce.
// Uploads the scene's shading arrays (its host copies) to the device.
int scene_upload(ilm_scene* s) {
    ilm_ctx* ctx = s->ctx;
    int rc = 0;
    rc = rc ? rc : scene_alloc(s, s->h_shade.size() * 4, (void**)&s->d_shade);
    rc = rc ? rc : scene_alloc(s, s->h_mats.size() * 4, (void**)&s->d_mats);
    rc = rc ? rc : scene_alloc(s, s->h_lt.size() * 4, (void**)&s->d_ltris);
    rc = rc ? rc : scene_alloc(s, s->h_lrec.size() * 4, (void**)&s->d_lights);
    rc = rc ? rc : scene_alloc(s, s->h_lrange.size() * 4, (void**)&s->d_lrange);
    rc = rc ? rc : scene_alloc(s, 4 * sizeof(int32_t), (void**)&s->d_frame);
    rc = rc ? rc : ilm_upload(ctx, s->d_shade, s->h_shade.data(), s->h_shade.size() * 4);
    rc = rc ? rc : ilm_upload(ctx, s->d_mats, s->h_mats.data(), s->h_mats.size() * 4);
    rc = rc ? rc : ilm_upload(ctx, s->d_ltris, s->h_lt.data(), s->h_lt.size() * 4);
    rc = rc ? rc : ilm_upload(ctx, s->d_lights, s->h_lrec.data(), s->h_lrec.size() * 4);
    rc = rc ? rc : ilm_upload(ctx, s->d_lrange, s->h_lrange.data(), s->h_lrange.size() * 4);
    if (rc) {
        std::string msg = ilm_last_error();
        for (void* p : s->bufs) ilm_free(ctx, p);
        s->bufs.clear();
        return ilm::fail(rc, msg);
    }
    // scene buffers are not path state: keep them out of the state list
    s->bufs.clear();
    return ILM_OK;
}

int ensure_state(ilm_scene* s, int64_t cap, int nl, bool cache = false, int64_t train_cap = 0) {
    nl = std::max(1, nl);
    if (s->cap >= cap && s->state_nl >= nl && (!cache || (s->cache_ready && s->train_cap == train_cap)))
        return ILM_OK;
    cap = std::max(cap, s->cap);
    nl = std::max(nl, s->state_nl);
    scene_free_state(s);
    ilm::PtArgs& a = s->a;
    int rc = 0;
    if (cache) {
        rc |= scene_alloc(s, cap * sizeof(float4), (void**)&a.tbeta);
        rc |= scene_alloc(s, cap * 3 * sizeof(float), (void**)&a.T);
        rc |= scene_alloc(s, cap * sizeof(int32_t), (void**)&a.tslot);
        rc |= scene_alloc(s, cap * 16 * sizeof(float), (void**)&a.queries);
        rc |= scene_alloc(s, cap * sizeof(float4), (void**)&a.qbeta);
        rc |= scene_alloc(s, cap * sizeof(float4), (void**)&a.qout);
        rc |= scene_alloc(s, cap * nl * sizeof(float4), (void**)&a.scontrib2);
        rc |= scene_alloc(s, std::max<int64_t>(train_cap, 1) * 16 * sizeof(float), (void**)&a.tq);
        rc |= scene_alloc(s, std::max<int64_t>(train_cap, 1) * sizeof(float4), (void**)&a.tt);
        s->cache_ready = true;
        s->train_cap = train_cap;
    }
    rc |= scene_alloc(s, cap * sizeof(float4), (void**)&a.beta);
    rc |= scene_alloc(s, cap * 3 * sizeof(float), (void**)&a.L);
    rc |= scene_alloc(s, cap * sizeof(ilm_ray), (void**)&a.rays[0]);
    rc |= scene_alloc(s, cap * sizeof(ilm_ray), (void**)&a.rays[1]);
    rc |= scene_alloc(s, cap * sizeof(int32_t), (void**)&a.qpath[0]);
    rc |= scene_alloc(s, cap * sizeof(int32_t), (void**)&a.qpath[1]);
    rc |= scene_alloc(s, cap * sizeof(ilm_hit), (void**)&a.hits);
    rc |= scene_alloc(s, cap * nl * sizeof(ilm_ray), (void**)&a.srays);
    rc |= scene_alloc(s, cap * nl * sizeof(float4), (void**)&a.scontrib);
    rc |= scene_alloc(s, cap * nl, (void**)&a.socc);
    rc |= scene_alloc(s, ilm::kCounts * sizeof(unsigned long long), (void**)&a.counts);
    if (rc) {
        scene_free_state(s);
        return ILM_ERR_NOMEM;
    }
    s->cap = cap;
    s->state_nl = nl;
    return ILM_OK;
}

}  // namespace

using ilm::fail;

extern "C" {

int ilm_scene_create(ilm_ctx* ctx, ilm_mesh* mesh, const float* nrm, const int32_t* tri_mat,
                     const int32_t* tri_light, const float* mats, int nmat, const float* lights, int nlight,
                     ilm_scene** out) {
    if (!out) return fail(ILM_ERR_INVALID, "null out");
    *out = nullptr;
    if (!ctx || !mesh || !nrm || !tri_mat || !tri_light || (nmat > 0 && !mats) || (nlight > 0 && !lights))
        return fail(ILM_ERR_INVALID, "null argument");
    if (mesh->ctx != ctx) return fail(ILM_ERR_INVALID, "mesh belongs to another context");
    if (!mesh->built) return fail(ILM_ERR_STATE, "build the mesh's BVH first");
    if (nmat <= 0) return fail(ILM_ERR_INVALID, "a scene needs at least one material");
    const int64_t nt = mesh->nt;
    // Shading record of a triangle, 64 bytes (half a cache line, one gather
    // per hit): the geometric normal normalize((v1 - v0) x (v2 - v0))
    // (IndexedTriangle.cpp:225-243), computed here once in float with
    // separate products and sums, then the vertex normals; the material
    // and light ids ride in the w words.
    std::vector<float> shade(16 * nt);
    std::vector<std::vector<int64_t>> ltri(nlight);
    for (int64_t i = 0; i < nt; ++i) {
        int m = tri_mat[i], l = tri_light[i];
        if (m >= nmat || l >= nlight) return fail(ILM_ERR_INVALID, "material or light index out of range");
        float* o = shade.data() + 16 * i;
        ilm::f3 p[3];
        for (int v = 0; v < 3; ++v) {
            int64_t vi = mesh->idx[3 * i + v];
            p[v] = ilm::mk(mesh->pos[3 * vi], mesh->pos[3 * vi + 1], mesh->pos[3 * vi + 2]);
            for (int k = 0; k < 3; ++k) o[4 + 4 * v + k] = nrm[3 * vi + k];
            o[4 + 4 * v + 3] = 0.f;
        }
        const ilm::f3 ng = ilm::normalize(ilm::cross(p[1] - p[0], p[2] - p[0]));
        o[0] = ng.x;
        o[1] = ng.y;
        o[2] = ng.z;
        int32_t mm = m < 0 ? -1 : m, ll = l < 0 ? -1 : l;
        memcpy(&o[3], &mm, 4);
        memcpy(&o[7], &ll, 4);
        if (l >= 0) ltri[l].push_back(i);
    }
    std::vector<float> lt, lrec(4 * std::max(1, nlight));
    std::vector<int32_t> lrange(2 * std::max(1, nlight));
    for (int l = 0; l < nlight; ++l) {
        if (ltri[l].empty()) return fail(ILM_ERR_INVALID, "light " + std::to_string(l) + " has no triangles");
        lrange[2 * l] = (int32_t)(lt.size() / 12);
        lrange[2 * l + 1] = (int32_t)ltri[l].size();
        double area = 0;
        for (int64_t t : ltri[l]) {
            float p[3][3];
            for (int v = 0; v < 3; ++v) {
                int64_t vi = mesh->idx[3 * t + v];
                for (int k = 0; k < 3; ++k) p[v][k] = mesh->pos[3 * vi + k];
                lt.insert(lt.end(), {p[v][0], p[v][1], p[v][2], 0.f});
            }
            ilm::f3 e1 = ilm::mk(p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]);
            ilm::f3 e2 = ilm::mk(p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]);
            area += (double)(ilm::length(ilm::cross(e1, e2)) * 0.5f);  // IndexedTriangle::GetArea
        }
        lrec[4 * l] = lights[4 * l];
        lrec[4 * l + 1] = lights[4 * l + 1];
        lrec[4 * l + 2] = lights[4 * l + 2];
        lrec[4 * l + 3] = (float)area;
        if (lights[4 * l + 3] == 1.0f) {
            // the reference's Quad (Shape/Quad.cpp:11-31): triangles (v1, v3, v2)
            // and (v1, v4, v3); sampled by its SamplePoint map, marked by a
            // negative area
            if (ltri[l].size() != 2) return fail(ILM_ERR_INVALID, "a quad light needs exactly two triangles");
            const int32_t* a0 = &mesh->idx[3 * ltri[l][0]];
            const int32_t* a1 = &mesh->idx[3 * ltri[l][1]];
            if (a0[0] != a1[0] || a0[1] != a1[2])
                return fail(ILM_ERR_INVALID, "a quad light's triangles must be (v1, v3, v2) and (v1, v4, v3)");
            lrec[4 * l + 3] = -(float)area;
        } else if (lights[4 * l + 3] != 0.0f) {
            return fail(ILM_ERR_INVALID, "light sampling mode must be 0 (mesh) or 1 (quad)");
        }
    }
    if (lt.empty()) lt.assign(12, 0.f);
    ilm_scene* s = new ilm_scene();
    s->ctx = ctx;
    s->mesh = mesh;
    s->nlights = nlight;
    s->nmat = nmat;
    s->h_shade.swap(shade);
    s->h_mats.assign(mats, mats + (size_t)nmat * 8);
    s->h_lt.swap(lt);
    s->h_lrec.swap(lrec);
    s->h_lrange.swap(lrange);
    int rc = scene_upload(s);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return ILM_OK;
}

int ilm_scene_destroy(ilm_scene* s) {
    if (!s) return ILM_OK;
    // a frame that failed after its fork may have left cache work on the side
    // stream, which reads the buffers freed below
    if (s->side) (void)hipStreamSynchronize(s->side);
    scene_free_state(s);
    for (void* p : {(void*)s->d_shade, (void*)s->d_mats, (void*)s->d_ltris, (void*)s->d_lights, (void*)s->d_lrange,
                    (void*)s->d_frame, (void*)s->d_views})
        if (p) ilm_free(s->ctx, p);
    if (s->owns_mesh) ilm_mesh_destroy(s->mesh);
    if (s->fork_ev) (void)hipEventDestroy(s->fork_ev);
    if (s->join_ev) (void)hipEventDestroy(s->join_ev);
    if (s->side) (void)hipStreamDestroy(s->side);
    delete s;
    return ILM_OK;
}

}  // extern "C"

namespace {

// One frame over `views` (all of the scene `s`, whose lights and BVH roots the
// views reference). Paths are numbered view by view, pixel-major, spp per
// pixel; batches of at most batch_paths paths end at pixel boundaries.
int render_views(ilm_scene* s, ilm_nrc* nrc, const std::vector<ilm::View>& views_in, const ilm_render_params& p,
                 const ilm_nrc_render_params* np, double* accum, ilm_render_stats* stats,
                 const ilm::PathMode* mode = nullptr) {
    if (!accum && !(np && !np->accumulate)) return fail(ILM_ERR_INVALID, "null accumulator");
    if (views_in.empty() || p.max_depth <= 0) return fail(ILM_ERR_INVALID, "no views or max_depth not positive");
    const bool cache = nrc != nullptr;
    int prefix = -1, max_depth_train = p.max_depth, train_steps = 0;
    int64_t train_cap = 0, train_batch = 0, train_count = 0;
    float train_frac = 0.f;
    if (cache) {
        if (!np) return fail(ILM_ERR_INVALID, "cache rendering needs ilm_nrc_render_params");
        prefix = np->prefix_depth > 0 ? np->prefix_depth : 2;
        max_depth_train = prefix + (np->suffix_depth > 0 ? np->suffix_depth : 8);
        train_frac = np->train_fraction > 0 ? np->train_fraction : 1.f / 16.f;
        train_batch = np->train_batch > 0 ? np->train_batch : 65536;
        if (train_frac > 1.f) return fail(ILM_ERR_INVALID, "train_fraction above 1");
        train_steps = np->train_steps >= 0 ? np->train_steps : 1;
    }
    ilm_ctx* ctx = s->ctx;
    ILM_HIP(hipSetDevice(ctx->device));
    // global path / pixel offsets of the views
    std::vector<ilm::View>& views = s->h_views;
    views = views_in;
    int64_t total_paths = 0, total_pix = 0;
    int max_spp = 1, max_vl = 1;
    for (ilm::View& v : views) {
        if (v.W <= 0 || v.H <= 0 || v.spp <= 0) return fail(ILM_ERR_INVALID, "width, height and spp must be positive");
        v.path0 = total_paths;
        v.pix0 = total_pix;
        total_pix += (int64_t)v.W * v.H;
        total_paths += (int64_t)v.W * v.H * v.spp;
        max_spp = std::max(max_spp, v.spp);
        max_vl = std::max(max_vl, v.l1 - v.l0);
    }
    // exactly floor(total * fraction) training paths, each with its own record
    // slot; the slots are grouped into training batches of train_batch records
    if (cache) {
        train_count = (int64_t)std::floor((double)total_paths * (double)train_frac);
        const int64_t nb = std::max<int64_t>(1, (train_count + train_batch - 1) / train_batch);
        train_cap = nb * train_batch;
        if (train_cap > (int64_t(1) << 31) - 1) return fail(ILM_ERR_INVALID, "too many training records");
    }
    const int64_t want = std::max<int64_t>(p.batch_paths > 0 ? p.batch_paths : (int64_t(1) << 22), max_spp);
    struct Batch { int64_t pix0, npix, path0, npaths; };
    std::vector<Batch> batches;
    {
        Batch b{0, 0, 0, 0};
        for (const ilm::View& v : views) {
            int64_t left = (int64_t)v.W * v.H, gp = v.pix0;
            while (left > 0) {
                int64_t k = std::min(left, (want - b.npaths) / v.spp);
                if (k == 0) {
                    batches.push_back(b);
                    b = Batch{gp, 0, v.path0 + (gp - v.pix0) * v.spp, 0};
                    continue;
                }
                if (b.npix == 0) b = Batch{gp, 0, v.path0 + (gp - v.pix0) * v.spp, 0};
                b.npix += k;
                b.npaths += k * v.spp;
                gp += k;
                left -= k;
            }
        }
        if (b.npix > 0) batches.push_back(b);
    }
    int64_t cap = 0;
    for (const Batch& b : batches) cap = std::max(cap, b.npaths);
    if (cap > (int64_t(1) << 31) - 1) return fail(ILM_ERR_INVALID, "batch too large");
    int rc = ensure_state(s, cap, max_vl, cache, train_cap);
    if (rc) return fail(rc, "path state allocation");
    if (s->views_cap < (int)views.size()) {
        drop_graph(s);
        if (s->d_views) ilm_free(ctx, s->d_views);
        s->d_views = nullptr;
        s->views_cap = 0;
        rc = ilm_alloc(ctx, views.size() * sizeof(ilm::View), (void**)&s->d_views);
        if (rc) return rc;
        s->views_cap = (int)views.size();
    }
    ilm::PtArgs a = s->a;
    a.nrc_prefix = prefix;
    a.max_depth_train = max_depth_train;
    a.train_frac = train_frac;
    a.train_cap = train_cap;
    a.total_paths = total_paths;
    a.train_count = train_count;
    const bool do_accum = !cache || np->accumulate;
    const int max_bounces = cache ? max_depth_train : p.max_depth;
    a.shade = s->d_shade;
    a.mats = s->d_mats;
    a.ltris = s->d_ltris;
    a.lights = s->d_lights;
    a.lrange = s->d_lrange;
    a.nlights = s->nlights;
    a.max_vl = max_vl;
    a.views = s->d_views;
    a.nviews = (int)views.size();
    a.max_depth = p.max_depth;
    a.rr_depth = p.rr_depth;
    a.eps = p.eps;
    a.accum = accum;
    a.frame = s->d_frame;
    a.gather = mode ? mode->gather : nullptr;
    a.first_t = mode ? mode->first_t : nullptr;
    a.gbuf = mode ? reinterpret_cast<float4*>(mode->gbuf) : nullptr;
    a.rr_rule = mode ? mode->rr_rule : 0;
    a.emit0 = mode ? (mode->emit0 ? 1 : 0) : 1;
    hipStream_t st = ctx->stream;
    int ev0 = 0, ev1 = 0;
    if (cache && !s->side) {
        // the side stream of cache frames, created before the frame's timing starts
        ILM_HIP(hipSetDevice(ctx->device));
        ILM_HIP(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
        ILM_HIP(hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming));
        ILM_HIP(hipEventCreateWithFlags(&s->join_ev, hipEventDisableTiming));
    }
    rc = ilm_timer_mark(ctx, &ev0);
    if (rc) return rc;
    // per-frame inputs: the only ones that may change between replays
    s->h_frame[0] = (int32_t)p.seed;
    s->h_frame[1] = p.spp_offset;
    ILM_HIP(hipMemcpyAsync(s->d_frame, s->h_frame, sizeof(s->h_frame), hipMemcpyHostToDevice, st));
    ILM_HIP(hipMemcpyAsync(s->d_views, views.data(), views.size() * sizeof(ilm::View), hipMemcpyHostToDevice, st));
    const int grid = ctx->cu_count * 8;
    // The shading pass's grid is whole rounds of resident workgroups (its
    // register use allows two or three per CU): a partial last round would
    // leave most of the chip idle at the end of each pass.
    int shade_per_cu = 0;
    ILM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &shade_per_cu, cache ? (const void*)ilm::k_shade<true> : (const void*)ilm::k_shade<false>, ilm::kPtBlock, 0));
    const int shade_grid = ctx->cu_count * std::max(shade_per_cu, 1) * ilm::kShadeRounds;
    // The frame's launch sequence (no host synchronisation inside, counters
    // cleared by a kernel rather than a memset node, so it can be captured).
    auto launch_frame = [&]() -> int {
        hipLaunchKernelGGL(ilm::k_clear_counts, dim3(1), dim3(64), 0, st, a.counts);
        if (cache)
            hipLaunchKernelGGL(ilm::k_nrc_clear, dim3((unsigned)((train_cap + 255) / 256)), dim3(256), 0, st, a);
        for (const Batch& bt : batches) {
            const int64_t npix = bt.npix;
            a.bpix0 = bt.pix0;
            a.bpath0 = bt.path0;
            a.npaths = bt.npaths;
            hipLaunchKernelGGL(ilm::k_generate, dim3((unsigned)((a.npaths + 255) / 256)), dim3(256), 0, st, a);
            int cur = 0;
            // camera rays are coherent: their waves refill only when 56 lanes
            // are idle (measured fastest for primary rays); later launches use
            // the context's threshold
            int r2 = ilm::trace_queue(s->mesh, a.rays[0], a.counts + 0, cap, a.hits, nullptr, st, true, 56);
            if (r2) return r2;
            bool forked = false;  // the cache inference of this batch runs on the side stream
            for (int b = 0; b < max_bounces; ++b) {
                if (cache)
                    hipLaunchKernelGGL(ilm::k_shade<true>, dim3(shade_grid), dim3(ilm::kPtBlock), 0, st, a, b, cur);
                else
                    hipLaunchKernelGGL(ilm::k_shade<false>, dim3(shade_grid), dim3(ilm::kPtBlock), 0, st, a, b, cur);
                if (cache && b == a.nrc_prefix) {
                    hipLaunchKernelGGL(ilm::k_query_shadow_count, dim3(1), dim3(1), 0, st, a.counts, max_vl);
                    // the batch's queries are complete: infer and resolve them
                    // beside the training suffix, whose launches carry few rays.
                    // Query paths end at the query vertex, so the suffix writes
                    // no radiance of theirs; the batch's end waits for the side
                    // stream before the training targets and the accumulation.
                    ILM_HIP(hipEventRecord(s->fork_ev, st));
                    ILM_HIP(hipStreamWaitEvent(s->side, s->fork_ev, 0));
                    r2 = ilm::nrc_infer_queue(nrc, a.queries, a.counts + 5, cap, reinterpret_cast<float*>(a.qout),
                                              s->side);
                    if (r2) return r2;
                    hipLaunchKernelGGL(ilm::k_nrc_resolve, dim3(grid), dim3(256), 0, s->side, a);
                    ILM_HIP(hipEventRecord(s->join_ev, s->side));
                    forked = true;
                }
                // this bounce's shadow segments and the next bounce's rays come
                // from the same shading pass: one launch traces both
                const bool next = b + 1 < max_bounces;
                if (s->nlights > 0 && next)
                    r2 = ilm::trace_queue_dual(s->mesh, a.rays[cur ^ 1], a.counts + (cur ^ 1), cap, a.hits, a.srays,
                                               a.counts + 2, cap * max_vl, a.socc, st, true);
                else if (s->nlights > 0)
                    r2 = ilm::trace_queue(s->mesh, a.srays, a.counts + 2, cap * max_vl, nullptr, a.socc, st, true);
                else if (next)
                    r2 = ilm::trace_queue(s->mesh, a.rays[cur ^ 1], a.counts + (cur ^ 1), cap, a.hits, nullptr, st,
                                          true);
                if (r2) return r2;
                if (s->nlights > 0)
                    hipLaunchKernelGGL(ilm::k_shadow_accum, dim3(grid), dim3(256), 0, st, a, cur);
                else
                    hipLaunchKernelGGL(ilm::k_end_bounce, dim3(1), dim3(1), 0, st, a.counts, cur);
                cur ^= 1;
            }
            if (cache) {
                // batched cache inference at the query vertices (unless the side
                // stream did it), then training targets
                if (forked) {
                    ILM_HIP(hipStreamWaitEvent(st, s->join_ev, 0));
                } else {
                    int r3 = ilm::nrc_infer_queue(nrc, a.queries, a.counts + 5, cap, reinterpret_cast<float*>(a.qout),
                                                  st);
                    if (r3) return r3;
                    hipLaunchKernelGGL(ilm::k_nrc_resolve, dim3(grid), dim3(256), 0, st, a);
                }
                hipLaunchKernelGGL(ilm::k_nrc_collect, dim3((unsigned)((a.npaths + 255) / 256)), dim3(256), 0, st, a);
                hipLaunchKernelGGL(ilm::k_nrc_reset_queries, dim3(1), dim3(1), 0, st, a.counts);
            }
            if (do_accum)
                hipLaunchKernelGGL(ilm::k_accumulate, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, st, a, npix);
            ILM_HIP(hipGetLastError());
        }
        return ILM_OK;
    };
    if (!s->use_graph || mode) {
        rc = launch_frame();
        if (rc) return rc;
    } else {
        // everything the captured launches bake in, including the context's
        // traversal spill area (a later launch that needs more reallocates it)
        // and the mesh's device arrays and stack bound
        auto make_key = [&]() {
            std::vector<int64_t> key = {
                (int64_t)(uintptr_t)nrc, (int64_t)(uintptr_t)accum, p.max_depth, p.rr_depth,
                (int64_t)ilm::float_bits(p.eps), cap, do_accum, prefix, max_depth_train,
                (int64_t)ilm::float_bits(train_frac), train_cap, (int64_t)(uintptr_t)s->mesh->d_nodes,
                (int64_t)(uintptr_t)s->mesh->d_tris, s->mesh->node_format, s->mesh->width, s->mesh->info.stack_need,
                (int64_t)(uintptr_t)ctx->spill, (int64_t)ctx->spill_words, (int64_t)ctx->spill_generation,
                ctx->trace_refill, want, (int64_t)views.size()};
            for (const ilm::View& v : views) {
                key.push_back(v.W);
                key.push_back(v.H);
                key.push_back(v.spp);
            }
            return key;
        };
        if (s->graph && s->graph_key == make_key()) {
            ILM_HIP(hipGraphLaunch(s->graph, st));
        } else {
            // first frame with these settings: run it, then capture the same
            // sequence for the next frames
            drop_graph(s);
            rc = launch_frame();
            if (rc) return rc;
            ILM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
            int crc = launch_frame();
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(st, &g);
            if (crc) {
                if (g) (void)hipGraphDestroy(g);
                return crc;
            }
            if (e != hipSuccess) return ilm::fail_hip(e, "frame capture");
            e = hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                s->graph = nullptr;
                return ilm::fail_hip(e, "frame graph instantiation");
            }
            s->graph_key = make_key();
        }
    }
    // Adam steps of train_batch records each; successive steps (across
    // frames) take the frame's batches in turn
    if (cache) {
        const int64_t nb = train_cap / train_batch;
        for (int k = 0; k < train_steps; ++k) {
            const int64_t bi = (int64_t)(s->train_round++ % (uint64_t)nb);
            rc = ilm_nrc_train_step(nrc, a.tq + 16 * bi * train_batch,
                                    reinterpret_cast<const float*>(a.tt + bi * train_batch), train_batch, nullptr, st);
            if (rc) return rc;
        }
    }
    rc = ilm_timer_mark(ctx, &ev1);
    if (rc) return rc;
    float ms = 0.f;
    rc = ilm_timer_elapsed(ctx, ev0, ev1, &ms);
    if (rc) return rc;
    unsigned long long cnt[8];
    ILM_HIP(hipMemcpy(cnt, a.counts, sizeof(cnt), hipMemcpyDeviceToHost));
    if (stats) {
        stats->camera_rays = total_paths;
        stats->extension_rays = (int64_t)cnt[3] - stats->camera_rays;
        stats->shadow_rays = (int64_t)cnt[4];
        stats->ms = ms;
        stats->train_records = cache ? (int64_t)cnt[6] : 0;
    }
    return ilm_sync(ctx);
}

ilm::View view_of(const ilm_render_params& p, int nlights) {
    ilm::View v{};
    float fr[12];
    ilm::camera_frame(p.cam, fr);  // same frame as ilm_raygen_camera
    v.cam.centre = ilm::mk(fr[0], fr[1], fr[2]);
    v.cam.corner = ilm::mk(fr[3], fr[4], fr[5]);
    v.cam.across = ilm::mk(fr[6], fr[7], fr[8]);
    v.cam.up = ilm::mk(fr[9], fr[10], fr[11]);
    v.W = p.width;
    v.H = p.height;
    v.spp = p.spp;
    v.l0 = 0;
    v.l1 = nlights;
    v.root = 0;
    return v;
}

int render_impl(ilm_scene* s, ilm_nrc* nrc, const ilm_render_params* prm, const ilm_nrc_render_params* np,
                double* accum, ilm_render_stats* stats) {
    if (!s || !prm) return fail(ILM_ERR_INVALID, "null argument");
    if (prm->width <= 0 || prm->height <= 0 || prm->spp <= 0 || prm->max_depth <= 0)
        return fail(ILM_ERR_INVALID, "width, height, spp and max_depth must be positive");
    std::vector<ilm::View> views(1, view_of(*prm, s->nlights));
    return render_views(s, nrc, views, *prm, np, accum, stats);
}

}  // namespace

namespace ilm {

ilm_ctx* scene_ctx(ilm_scene* s) { return s->ctx; }
ilm_mesh* scene_mesh(ilm_scene* s) { return s->mesh; }

int render_paths(ilm_scene* s, const ilm_render_params& p, const PathMode& mode, double* accum,
                 ilm_render_stats* stats) {
    if (!s) return fail(ILM_ERR_INVALID, "null scene");
    std::vector<View> views(1, view_of(p, s->nlights));
    return render_views(s, nullptr, views, p, nullptr, accum, stats, &mode);
}

}  // namespace ilm

extern "C" {

int ilm_scene_set_graph(ilm_scene* s, int enable) {
    if (!s) return fail(ILM_ERR_INVALID, "null scene");
    s->use_graph = enable != 0;
    if (!s->use_graph) drop_graph(s);
    return ILM_OK;
}

int ilm_render(ilm_scene* s, const ilm_render_params* prm, double* accum, ilm_render_stats* stats) {
    return render_impl(s, nullptr, prm, nullptr, accum, stats);
}

int ilm_render_nrc(ilm_scene* s, ilm_nrc* nrc, const ilm_render_params* prm, const ilm_nrc_render_params* np,
                   double* accum, ilm_render_stats* stats) {
    if (!nrc) return fail(ILM_ERR_INVALID, "null cache");
    return render_impl(s, nrc, prm, np, accum, stats);
}

}  // extern "C"

// ---------------------------------------------------------------- batches
// Several scenes rendered by one frame's launch sequence: their BVHs are
// packed into one node array (each scene's nodes keep their shape; child
// words are rebased), their triangles into one array with global indices,
// and their shading, materials and lights are concatenated with rebased
// indices. A view names its scene; the traversal starts each ray at its
// scene's root (carried in the ray's tmin bits) and a path samples only its
// scene's lights.
struct ilm_batch {
    ilm_ctx* ctx = nullptr;
    ilm_scene* merged = nullptr;
    std::vector<uint32_t> root;    // the scenes' roots in the node array in use
    std::vector<int32_t> lbase, nl;
    void* d_nodes_float = nullptr;
    void* d_nodes_quant = nullptr;
    int64_t nodes = 0;
    int node_format = ILM_NODES_FLOAT;
    int width = 4;
    int64_t batch_paths = 0;  // 0: the library default
    // BVH4 roots and stack bound, and the same trees at widths 2 and 8 over
    // the same leaf slots (built on first use)
    std::vector<uint32_t> root4;
    int need4 = 0, depth4 = 0;
    struct Wide { void* d_nodes = nullptr; int64_t nodes = 0; int need = 0, depth = 0; std::vector<uint32_t> root; };
    Wide w2, w8;
};

namespace {

// Rebases the child words of packed nodes: internal children by the scene's
// first node, leaves by its first triangle slot.
void rebase_children(uint32_t* child, uint32_t node_base, uint32_t slot_base, int width = 4) {
    for (int k = 0; k < width; ++k) {
        uint32_t c = child[k];
        if (c == ilm::kEmptySlot) continue;
        if (ilm::is_leaf(c)) child[k] = ilm::make_leaf(ilm::leaf_first(c) + slot_base, ilm::leaf_count(c));
        else child[k] = c + node_base;
    }
}

// The batch's trees at width 2 or 8: its BVH4 trees unwidened to binary
// trees over the same leaf slots and collapsed again, so the packed
// triangles serve every width.
int batch_build_width(ilm_batch* b, int width) {
    ilm_batch::Wide& w = width == 2 ? b->w2 : b->w8;
    if (w.d_nodes) return ILM_OK;
    const size_t stride = width == 2 ? sizeof(ilm::WideNode<2>) : sizeof(ilm::WideNode<8>);
    std::vector<uint8_t> all;
    w.root.clear();
    w.need = w.depth = 0;
    try {
        std::vector<int32_t> broots;
        ilm::Bvh2 bin = ilm::binary_from_bvh4(b->merged->mesh->h_nodes.data(), b->root4, broots);
        int64_t base = 0;
        for (int32_t r : broots) {
            ilm::WideBvh t = ilm::collapse_from(bin, r, width);
            // internal child words are local to the scene's tree: rebase them
            for (int64_t i = 0; i < t.nnodes; ++i) {
                uint32_t* ch = reinterpret_cast<uint32_t*>(t.nodes.data() + (size_t)i * stride + (size_t)width * 24);
                rebase_children(ch, (uint32_t)base, 0u, width);
            }
            w.root.push_back((uint32_t)base);
            w.need = std::max(w.need, t.stack_need);
            w.depth = std::max(w.depth, t.depth);
            all.insert(all.end(), t.nodes.begin(), t.nodes.end());
            base += t.nnodes;
        }
        w.nodes = base;
    } catch (const std::exception& e) {
        return fail(ILM_ERR_INVALID, std::string("batch tree of width ") + std::to_string(width) + ": " + e.what());
    }
    int rc = ilm_alloc(b->ctx, all.size(), &w.d_nodes);
    if (!rc) rc = ilm_upload(b->ctx, w.d_nodes, all.data(), all.size());
    return rc;
}

int batch_use_format(ilm_batch* b, int fmt, int width = 4) {
    if (fmt == ILM_NODES_QUANTIZED && width != 4) return fail(ILM_ERR_INVALID, "quantised nodes are BVH4");
    if (width != 2 && width != 4 && width != 8) return fail(ILM_ERR_INVALID, "BVH width must be 2, 4 or 8");
    if (width != 4) {
        int rc = batch_build_width(b, width);
        if (rc) return rc;
    }
    ilm_mesh* m = b->merged->mesh;
    const ilm_batch::Wide* w = width == 2 ? &b->w2 : width == 8 ? &b->w8 : nullptr;
    m->d_nodes = w ? w->d_nodes : fmt == ILM_NODES_QUANTIZED ? b->d_nodes_quant : b->d_nodes_float;
    m->node_format = fmt;
    m->width = width;
    m->info.width = width;
    m->info.node_format = fmt;
    m->info.node_stride = fmt == ILM_NODES_QUANTIZED ? (int32_t)sizeof(ilm::QNode4)
                          : width == 2                ? (int32_t)sizeof(ilm::WideNode<2>)
                          : width == 8                ? (int32_t)sizeof(ilm::WideNode<8>)
                                                      : (int32_t)sizeof(ilm::WideNode<4>);
    m->info.nodes = w ? w->nodes : b->nodes;
    m->info.node_bytes = m->info.nodes * m->info.node_stride;
    m->info.stack_need = w ? w->need : b->need4;
    m->info.depth = w ? w->depth : b->depth4;
    b->root = w ? w->root : b->root4;
    b->node_format = fmt;
    b->width = width;
    drop_graph(b->merged);
    return ILM_OK;
}

}  // namespace

extern "C" {

int ilm_batch_destroy(ilm_batch* b) {
    if (!b) return ILM_OK;
    if (b->merged) {
        ilm_mesh* m = b->merged->mesh;
        // the mesh frees whichever node array it points at; free the others here
        for (void* p : {b->d_nodes_float, b->d_nodes_quant, b->w2.d_nodes, b->w8.d_nodes})
            if (p && p != m->d_nodes) ilm_free(b->ctx, p);
        ilm_scene_destroy(b->merged);
    }
    delete b;
    return ILM_OK;
}

int ilm_batch_create(ilm_ctx* ctx, ilm_scene* const* scenes, int n, ilm_batch** out) {
    if (!ctx || !scenes || n <= 0 || !out) return fail(ILM_ERR_INVALID, "bad arguments");
    *out = nullptr;
    // counts and bases
    int64_t nv = 0, nt = 0, slots = 0, nodes = 0;
    int nmat = 0, nlight = 0;
    int stack_need = 0, depth = 0, max_leaf = 0;
    double sah = 0;
    for (int i = 0; i < n; ++i) {
        ilm_scene* s = scenes[i];
        if (!s || s->ctx != ctx) return fail(ILM_ERR_INVALID, "scene missing or from another context");
        ilm_mesh* m = s->mesh;
        if (m->node_format != ILM_NODES_FLOAT || m->width != 4)
            return fail(ILM_ERR_INVALID, "batched scenes need meshes built with float BVH4 nodes");
        nv += m->nv;
        nt += m->nt;
        slots += (int64_t)m->h_order.size();
        nodes += m->info.nodes;
        nmat += s->nmat;
        nlight += s->nlights;
        stack_need = std::max(stack_need, m->info.stack_need);
        depth = std::max(depth, m->info.depth);
        max_leaf = std::max(max_leaf, m->info.max_leaf);
        sah += m->info.sah_cost;
    }
    if (slots >= (int64_t(1) << 26) || nodes >= (int64_t(1) << 31) || nt >= (int64_t(1) << 31))
        return fail(ILM_ERR_INVALID, "batch exceeds 2^26 triangle slots");
    ilm_batch* b = new ilm_batch();
    b->ctx = ctx;
    b->nodes = nodes;
    std::vector<float> pos;
    std::vector<int32_t> idx, order;
    std::vector<ilm::WideNode<4>> fnodes;
    pos.reserve(3 * nv);
    idx.reserve(3 * nt);
    order.reserve(slots);
    fnodes.reserve(nodes);
    ilm_scene* ms = new ilm_scene();
    ms->ctx = ctx;
    int64_t vbase = 0, pbase = 0, sbase = 0, nbase = 0;
    int mbase = 0, lb = 0, ltbase = 0;
    for (int i = 0; i < n; ++i) {
        ilm_scene* s = scenes[i];
        ilm_mesh* m = s->mesh;
        b->root.push_back((uint32_t)nbase);
        b->root4.push_back((uint32_t)nbase);
        b->lbase.push_back(lb);
        b->nl.push_back(s->nlights);
        pos.insert(pos.end(), m->pos.begin(), m->pos.end());
        for (int32_t v : m->idx) idx.push_back((int32_t)(v + vbase));
        for (int32_t o : m->h_order) order.push_back((int32_t)(o + pbase));
        const auto* src = reinterpret_cast<const ilm::WideNode<4>*>(m->h_nodes.data());
        for (int64_t k = 0; k < m->info.nodes; ++k) {
            ilm::WideNode<4> w = src[k];
            rebase_children(w.child, (uint32_t)nbase, (uint32_t)sbase);
            fnodes.push_back(w);
        }
        // shading: rebase material and light ids
        for (int64_t t = 0; t < m->nt; ++t) {
            float rec[16];
            memcpy(rec, s->h_shade.data() + 16 * t, sizeof rec);
            int32_t mm, ll;
            memcpy(&mm, &rec[3], 4);
            memcpy(&ll, &rec[7], 4);
            if (mm >= 0) mm += mbase;
            if (ll >= 0) ll += lb;
            memcpy(&rec[3], &mm, 4);
            memcpy(&rec[7], &ll, 4);
            ms->h_shade.insert(ms->h_shade.end(), rec, rec + 16);
        }
        ms->h_mats.insert(ms->h_mats.end(), s->h_mats.begin(), s->h_mats.end());
        for (int l = 0; l < s->nlights; ++l) {
            ms->h_lrange.push_back(s->h_lrange[2 * l] + ltbase);
            ms->h_lrange.push_back(s->h_lrange[2 * l + 1]);
            ms->h_lrec.insert(ms->h_lrec.end(), s->h_lrec.begin() + 4 * l, s->h_lrec.begin() + 4 * l + 4);
        }
        if (s->nlights > 0) {
            int64_t nlt = 0;
            for (int l = 0; l < s->nlights; ++l) nlt += s->h_lrange[2 * l + 1];
            ms->h_lt.insert(ms->h_lt.end(), s->h_lt.begin(), s->h_lt.begin() + 12 * nlt);
            ltbase += (int)nlt;
        }
        vbase += m->nv;
        pbase += m->nt;
        sbase += (int64_t)m->h_order.size();
        nbase += m->info.nodes;
        mbase += s->nmat;
        lb += s->nlights;
    }
    if (ms->h_lt.empty()) ms->h_lt.assign(12, 0.f);
    if (ms->h_lrec.empty()) ms->h_lrec.assign(4, 0.f);
    if (ms->h_lrange.empty()) ms->h_lrange.assign(2, 0);
    ms->nmat = nmat;
    ms->nlights = nlight;
    // the packed mesh: positions and indices for export, nodes and triangles on the device
    ilm_mesh* pm = nullptr;
    int rc = ilm_mesh_create(ctx, pos.data(), nv, idx.data(), nt, &pm);
    if (rc) {
        delete ms;
        delete b;
        return rc;
    }
    std::vector<float> tris;
    ilm::pack_triangles(pos.data(), idx.data(), order, tris);
    std::vector<ilm::QNode4> qnodes(nodes);
    for (int64_t k = 0; k < nodes; ++k) ilm::quantize_node(fnodes[k], qnodes[k]);
    rc = ilm_alloc(ctx, (size_t)nodes * sizeof(ilm::WideNode<4>), &b->d_nodes_float);
    if (!rc) rc = ilm_alloc(ctx, (size_t)nodes * sizeof(ilm::QNode4), &b->d_nodes_quant);
    if (!rc) rc = ilm_alloc(ctx, tris.size() * sizeof(float), (void**)&pm->d_tris);
    if (!rc) rc = ilm_upload(ctx, b->d_nodes_float, fnodes.data(), (size_t)nodes * sizeof(ilm::WideNode<4>));
    if (!rc) rc = ilm_upload(ctx, b->d_nodes_quant, qnodes.data(), (size_t)nodes * sizeof(ilm::QNode4));
    if (!rc) rc = ilm_upload(ctx, pm->d_tris, tris.data(), tris.size() * sizeof(float));
    if (rc) {
        if (b->d_nodes_float) ilm_free(ctx, b->d_nodes_float);
        if (b->d_nodes_quant) ilm_free(ctx, b->d_nodes_quant);
        ilm_mesh_destroy(pm);
        delete ms;
        delete b;
        return rc;
    }
    pm->d_nodes = b->d_nodes_float;
    pm->width = 4;
    pm->h_nodes.assign(reinterpret_cast<const uint8_t*>(fnodes.data()),
                       reinterpret_cast<const uint8_t*>(fnodes.data()) + (size_t)nodes * sizeof(ilm::WideNode<4>));
    pm->h_order.swap(order);
    pm->info.nodes = nodes;
    pm->info.node_stride = (int32_t)sizeof(ilm::WideNode<4>);
    pm->info.node_bytes = nodes * (int64_t)sizeof(ilm::WideNode<4>);
    pm->info.tri_bytes = (int64_t)(tris.size() * sizeof(float));
    pm->info.width = 4;
    pm->info.depth = depth;
    pm->info.max_leaf = max_leaf;
    pm->info.stack_need = stack_need;
    pm->info.sah_cost = sah;
    b->need4 = stack_need;
    b->depth4 = depth;
    pm->built = true;
    ms->mesh = pm;
    ms->owns_mesh = true;
    rc = scene_upload(ms);
    if (rc) {
        ilm_free(ctx, b->d_nodes_quant);
        ilm_mesh_destroy(pm);  // frees the float nodes and the triangles
        delete ms;
        delete b;
        return rc;
    }
    b->merged = ms;
    batch_use_format(b, ILM_NODES_FLOAT);
    *out = b;
    return ILM_OK;
}

int ilm_batch_tree(ilm_batch* b, ilm_mesh** mesh, uint32_t* roots) {
    if (!b || !mesh) return fail(ILM_ERR_INVALID, "null argument");
    if (b->node_format != ILM_NODES_FLOAT || b->width != 4)
        return fail(ILM_ERR_STATE, "the batch's tree is exported while it uses float BVH4 nodes");
    *mesh = b->merged->mesh;
    if (roots) memcpy(roots, b->root4.data(), b->root4.size() * sizeof(uint32_t));
    return ILM_OK;
}

int ilm_batch_set_graph(ilm_batch* b, int enable) {
    if (!b) return fail(ILM_ERR_INVALID, "null batch");
    return ilm_scene_set_graph(b->merged, enable);
}

int ilm_batch_configure(ilm_batch* b, int node_format, int width, int64_t batch_paths) {
    if (!b || batch_paths < 0) return fail(ILM_ERR_INVALID, "bad arguments");
    if (node_format != ILM_NODES_FLOAT && node_format != ILM_NODES_QUANTIZED)
        return fail(ILM_ERR_INVALID, "unknown node format");
    int rc = batch_use_format(b, node_format, width == 0 ? 4 : width);
    if (rc) return rc;
    b->batch_paths = batch_paths;
    return ILM_OK;
}

}  // extern "C"

namespace {

int batch_views(ilm_batch* b, const ilm_batch_view* bv, int nviews, std::vector<ilm::View>& views) {
    views.clear();
    for (int i = 0; i < nviews; ++i) {
        const ilm_batch_view& q = bv[i];
        if (q.scene < 0 || q.scene >= (int)b->root.size()) return fail(ILM_ERR_INVALID, "view names no scene of the batch");
        ilm_render_params p{};
        p.cam = q.cam;
        p.width = q.width;
        p.height = q.height;
        p.spp = q.spp;
        ilm::View v = view_of(p, 0);
        v.l0 = b->lbase[q.scene];
        v.l1 = b->lbase[q.scene] + b->nl[q.scene];
        v.root = b->root[q.scene];
        views.push_back(v);
    }
    return ILM_OK;
}

}  // namespace

extern "C" {

int ilm_render_batch(ilm_batch* b, const ilm_batch_view* bv, int nviews, const ilm_render_params* prm, double* accum,
                     ilm_render_stats* stats) {
    if (!b || !bv || nviews <= 0 || !prm) return fail(ILM_ERR_INVALID, "bad arguments");
    std::vector<ilm::View> views;
    int rc = batch_views(b, bv, nviews, views);
    if (rc) return rc;
    ilm_render_params p = *prm;
    if (p.batch_paths <= 0 && b->batch_paths > 0) p.batch_paths = (int32_t)std::min<int64_t>(b->batch_paths, INT32_MAX);
    return render_views(b->merged, nullptr, views, p, nullptr, accum, stats);
}

int ilm_render_batch_nrc(ilm_batch* b, ilm_nrc* nrc, const ilm_batch_view* bv, int nviews,
                         const ilm_render_params* prm, const ilm_nrc_render_params* np, double* accum,
                         ilm_render_stats* stats) {
    if (!b || !nrc || !bv || nviews <= 0 || !prm || !np) return fail(ILM_ERR_INVALID, "bad arguments");
    std::vector<ilm::View> views;
    int rc = batch_views(b, bv, nviews, views);
    if (rc) return rc;
    ilm_render_params p = *prm;
    if (p.batch_paths <= 0 && b->batch_paths > 0) p.batch_paths = (int32_t)std::min<int64_t>(b->batch_paths, INT32_MAX);
    return render_views(b->merged, nrc, views, p, np, accum, stats);
}

int ilm_batch_autotune(ilm_batch* b, const ilm_batch_view* bv, int nviews, const ilm_render_params* prm,
                       ilm_batch_tuning* out) {
    if (!b || !bv || nviews <= 0 || !prm) return fail(ILM_ERR_INVALID, "bad arguments");
    // timing frames at (at most) 2 spp per pixel
    std::vector<ilm_batch_view> tv(bv, bv + nviews);
    int64_t npix = 0;
    for (ilm_batch_view& v : tv) {
        v.spp = std::min(v.spp, 2);
        npix += (int64_t)v.width * v.height;
    }
    double* acc = nullptr;
    int rc = ilm_alloc(b->ctx, (size_t)npix * 3 * sizeof(double), (void**)&acc);
    if (rc) return rc;
    // trees: float BVH4, quantised BVH4, float BVH2, float BVH8
    const int fmts[4][2] = {{ILM_NODES_FLOAT, 4}, {ILM_NODES_QUANTIZED, 4}, {ILM_NODES_FLOAT, 2}, {ILM_NODES_FLOAT, 8}};
    const int64_t sizes[4] = {int64_t(1) << 20, int64_t(1) << 22, int64_t(1) << 24, int64_t(1) << 26};
    static_assert(4 * 4 <= 16, "the tuning table holds 16 candidates");
    ilm_batch_tuning t{};
    t.best_ms = 1e30;
    int k = 0;
    for (const auto& fw : fmts)
        for (int64_t sz : sizes) {
            const int f = fw[0], width = fw[1];
            rc = batch_use_format(b, f, width);
            if (rc) break;
            ilm_render_params p = *prm;
            p.batch_paths = (int32_t)sz;
            double best = 1e30;
            for (int rep = 0; rep < 2 && rc == ILM_OK; ++rep) {
                ilm_render_stats st{};
                p.seed = prm->seed + 7919u * (uint32_t)rep;
                std::vector<ilm::View> views;
                rc = batch_views(b, tv.data(), nviews, views);
                if (rc) break;
                rc = render_views(b->merged, nullptr, views, p, nullptr, acc, &st);
                if (rc) break;
                best = std::min(best, st.ms);
            }
            if (rc) break;
            t.candidate_ms[k] = best;
            t.candidate_format[k] = f;
            t.candidate_width[k] = width;
            t.candidate_batch[k] = sz;
            ++k;
            if (best < t.best_ms) {
                t.best_ms = best;
                t.node_format = f;
                t.width = width;
                t.batch_paths = sz;
            }
        }
    ilm_free(b->ctx, acc);
    if (rc) return rc;
    t.candidates = k;
    rc = batch_use_format(b, t.node_format, t.width);
    if (rc) return rc;
    b->batch_paths = t.batch_paths;
    if (out) *out = t;
    return ILM_OK;
}

}  // extern "C"
