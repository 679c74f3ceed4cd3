// Internal state behind the C ABI: error reporting, contexts, meshes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <chrono>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/ilm_gpu.h"
#include "bvh_build.hpp"

namespace ilm {

// Records the calling thread's last error and returns the code.
int fail(int code, const std::string& msg);
int fail_hip(hipError_t e, const char* what);

#define ILM_HIP(call)                                   \
    do {                                                \
        hipError_t e_ = (call);                         \
        if (e_ != hipSuccess) return ::ilm::fail_hip(e_, #call); \
    } while (0)

}  // namespace ilm

namespace ilm {
constexpr int kTraceSegs = 8;           // ray segments of a traversal launch (one per XCD)
constexpr int kSegCounter0 = 64;        // first segment counter in ilm_ctx::counters
constexpr int kSegCounterStride = 64;   // 512 bytes between segment counters
constexpr int kCounterWords = kSegCounter0 + kTraceSegs * kSegCounterStride;
}  // namespace ilm

struct ilm_ctx {
    int device = 0;
    int cu_count = 0;
    std::string name;
    hipStream_t stream = nullptr;
    std::set<void*> allocs;
    std::vector<hipEvent_t> events;  // timer slots
    int next_event = 0;
    // per-context scratch for traversal: ray-fetch counters and the
    // traversal-stack spill area, grown on demand
    // [1] finished waves, [2] error flags, [4] post-process scratch; the
    // ray-fetch counters of the kTraceSegs ray segments at
    // kSegCounter0 + k * kSegCounterStride (each on its own 512-byte line)
    unsigned long long* counters = nullptr;
    uint32_t* spill = nullptr;
    size_t spill_words = 0;
    uint64_t spill_generation = 0;  // counts reallocations of spill (graphs key on it)
    int trace_refill = 24;  // idle lanes of a wave that trigger its ray refill
    std::mutex mu;
};

struct ilm_mesh {
    ilm_ctx* ctx = nullptr;
    std::vector<float> pos;
    std::vector<int32_t> idx;
    int64_t nv = 0, nt = 0;
    // device data
    void* d_nodes = nullptr;
    float* d_tris = nullptr;  // 12 floats per triangle, leaf order
    int32_t* d_slot_of = nullptr;  // original triangle index -> slot (built on demand)
    int width = 0;
    int node_format = ILM_NODES_FLOAT;
    ilm_bvh_info info{};
    std::vector<uint8_t> h_nodes;
    std::vector<int32_t> h_order;
    bool built = false;
};

namespace ilm {
// Ensures ctx->spill holds at least `words` 32-bit words.
int ensure_spill(ilm_ctx* ctx, size_t words);
// Takes this device's address of the fixed scratch of ilm_trace_closest_noalloc
// and zeroes its fetch counters (once per device; called at context creation).
int noalloc_setup(int device, hipStream_t stream);
hipStream_t pick_stream(ilm_ctx* ctx, void* stream);
// Two-level traversal of an instance table: tlas = top-level float BVH4 over
// the instances' world boxes, blas = the meshes' float BVH4 trees in one node
// array, tris = their triangles (3 float4 each, object space), inst = the
// instance records (4 float4 each: world-to-object rows 0..2, then (mesh
// root, instance index, mesh, 0)). Closest hits (hits, inst_out) or any hits (occ).
int launch_trace_inst(ilm_ctx* ctx, const void* tlas, const void* blas, const float* tris, const float4* inst,
                      int stack_need, const ilm_ray* rays, int64_t n, ilm_hit* hits, int32_t* inst_out, uint8_t* occ,
                      bool any, void* stream);
// PerspectiveCamera frame: out = centre, corner, across, up (12 floats).
void camera_frame(const ilm_camera& c, float out[12]);
// Closest hit (occ == null) or any hit over a ray queue whose length lives in
// device memory (*n_dev, at most `capacity`), without a host round trip.
// rooted: each ray's tmin field holds (as bits) the BVH node to start from,
// so one launch can trace rays of several BVHs packed in one node array.
// Collapses b to a BVH of the given width (2, 4, 8), packs the triangles in leaf order, quantises the
// nodes when asked, uploads (device meshes) and fills m's info and host
// copies, as every host builder ends. t0 = when the build started.
int install_bvh(ilm_mesh* m, Bvh2& b, int node_format, std::chrono::steady_clock::time_point t0, int width = 4);
// Variants of a frame of the wavefront path integrator used by the
// irradiance cache (one view of width x height x spp paths, paths numbered
// pixel-major): gather paths start from given rays; first_t and gbuf record
// per path the first hit's distance and its shading point (3 float4: P and
// kind 0 hit without material / 1 reflector / 2 emitter / 3 miss, Ns, albedo).
struct PathMode {
    const ilm_ray* gather = nullptr;
    float* first_t = nullptr;
    float* gbuf = nullptr;
    int rr_rule = 0;    // 0 PathIntegrator's roulette, 1 ICIntegrator::PathLi's
    bool emit0 = true;  // count emission at depth 0
};
ilm_ctx* scene_ctx(ilm_scene* s);
ilm_mesh* scene_mesh(ilm_scene* s);
// One frame of p (camera, size, spp, depth, roulette depth, eps, seed,
// batch size) with the given mode; adds the paths' radiance sums per pixel
// to accum (width*height*3 doubles, may be null only if nothing is accumulated).
int render_paths(ilm_scene* s, const ilm_render_params& p, const PathMode& mode, double* accum,
                 ilm_render_stats* stats);
int trace_queue(ilm_mesh* m, const ilm_ray* rays, const unsigned long long* n_dev, int64_t capacity, ilm_hit* hits,
                uint8_t* occ, hipStream_t stream, bool rooted = false, int refill = 0);
// One launch over two queues: closest hits of (rays, n_dev) into hits and any
// hits of (rays2, n2_dev) into occ2 (a bounce's shadow segments traced
// together with the next bounce's rays).
int trace_queue_dual(ilm_mesh* m, const ilm_ray* rays, const unsigned long long* n_dev, int64_t capacity,
                     ilm_hit* hits, const ilm_ray* rays2, const unsigned long long* n2_dev, int64_t capacity2,
                     uint8_t* occ2, hipStream_t stream, bool rooted = false, int refill = 0);
// Builds m's BVH4 on its device (linear BVH + collapse); fills the device
// arrays, info and host copies as ilm_bvh_build does.
int build_lbvh(ilm_mesh* m, int max_leaf, int node_format);
// Cache inference over a query queue whose length lives in device memory.
int nrc_infer_queue(ilm_nrc* r, const float* queries, const unsigned long long* n_dev, int64_t capacity, float* out,
                    hipStream_t stream);
}  // namespace ilm
