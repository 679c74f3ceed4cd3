// C ABI: contexts, device buffers, timers, meshes and BVH construction.
#include <stdio.h>
#include <string.h>

#include <chrono>
#include <exception>
#include <new>

#include "bvh_build.hpp"
#include "geom.hpp"
#include "internal.hpp"

namespace ilm {

static thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int fail_hip(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return ILM_ERR_HIP;
}

hipStream_t pick_stream(ilm_ctx* ctx, void* stream) {
    return stream ? (hipStream_t)stream : ctx->stream;
}

int ensure_spill(ilm_ctx* ctx, size_t words) {
    if (ctx->spill_words >= words) return ILM_OK;
    if (ctx->spill) {
        ILM_HIP(hipStreamSynchronize(ctx->stream));
        ILM_HIP(hipFree(ctx->spill));
        ctx->spill = nullptr;
        ctx->spill_words = 0;
    }
    ILM_HIP(hipMalloc(&ctx->spill, words * sizeof(uint32_t)));
    ctx->spill_words = words;
    ++ctx->spill_generation;
    return ILM_OK;
}

}  // namespace ilm

using ilm::fail;

extern "C" {

const char* ilm_last_error(void) { return ilm::g_last_error.c_str(); }

int ilm_device_count(int* count) {
    if (!count) return fail(ILM_ERR_INVALID, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = (e == hipSuccess) ? n : 0;
    return ILM_OK;
}

int ilm_ctx_create(ilm_ctx** out, int device) {
    if (!out) return fail(ILM_ERR_INVALID, "null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(ILM_ERR_NODEV, "no HIP device visible: the MI355X path has no CPU fallback");
    if (device < 0 || device >= n) return fail(ILM_ERR_INVALID, "device index out of range");
    ILM_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    ILM_HIP(hipGetDeviceProperties(&prop, device));
    ilm_ctx* c = new (std::nothrow) ilm_ctx();
    if (!c) return fail(ILM_ERR_NOMEM, "context allocation");
    c->device = device;
    c->cu_count = prop.multiProcessorCount;
    c->name = prop.gcnArchName;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->counters, ilm::kCounterWords * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->counters, 0, ilm::kCounterWords * sizeof(unsigned long long));
    int rc = e != hipSuccess ? ilm::fail_hip(e, "context setup") : ilm::noalloc_setup(device, c->stream);
    if (rc != ILM_OK) {
        if (c->counters) (void)hipFree(c->counters);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return ILM_OK;
}

int ilm_ctx_destroy(ilm_ctx* ctx) {
    if (!ctx) return ILM_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (void* p : ctx->allocs) (void)hipFree(p);
    for (hipEvent_t ev : ctx->events) (void)hipEventDestroy(ev);
    if (ctx->spill) (void)hipFree(ctx->spill);
    if (ctx->counters) (void)hipFree(ctx->counters);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return ILM_OK;
}

int ilm_device_info(ilm_ctx* ctx, char* name, int name_len, int* cu_count) {
    if (!ctx) return fail(ILM_ERR_INVALID, "null ctx");
    if (name && name_len > 0) {
        strncpy(name, ctx->name.c_str(), (size_t)name_len - 1);
        name[name_len - 1] = 0;
    }
    if (cu_count) *cu_count = ctx->cu_count;
    return ILM_OK;
}

int ilm_device_memory(ilm_ctx* ctx, uint64_t* free_bytes, uint64_t* total_bytes) {
    if (!ctx) return fail(ILM_ERR_INVALID, "null ctx");
    ILM_HIP(hipSetDevice(ctx->device));
    size_t fr = 0, tot = 0;
    ILM_HIP(hipMemGetInfo(&fr, &tot));
    if (free_bytes) *free_bytes = fr;
    if (total_bytes) *total_bytes = tot;
    return ILM_OK;
}

int ilm_alloc(ilm_ctx* ctx, size_t bytes, void** dev) {
    if (!ctx || !dev) return fail(ILM_ERR_INVALID, "null argument");
    *dev = nullptr;
    if (bytes == 0) bytes = 16;
    ILM_HIP(hipSetDevice(ctx->device));
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory) return fail(ILM_ERR_NOMEM, "device allocation of " + std::to_string(bytes) + " bytes");
    if (e != hipSuccess) return ilm::fail_hip(e, "hipMalloc");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->allocs.insert(p);
    *dev = p;
    return ILM_OK;
}

int ilm_free(ilm_ctx* ctx, void* dev) {
    if (!ctx) return fail(ILM_ERR_INVALID, "null ctx");
    if (!dev) return ILM_OK;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        auto it = ctx->allocs.find(dev);
        if (it == ctx->allocs.end()) return fail(ILM_ERR_INVALID, "pointer not allocated by this context");
        ctx->allocs.erase(it);
    }
    ILM_HIP(hipStreamSynchronize(ctx->stream));
    ILM_HIP(hipFree(dev));
    return ILM_OK;
}

int ilm_upload(ilm_ctx* ctx, void* dev, const void* host, size_t bytes) {
    if (!ctx || (!dev && bytes) || (!host && bytes)) return fail(ILM_ERR_INVALID, "null argument");
    ILM_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    ILM_HIP(hipStreamSynchronize(ctx->stream));
    return ILM_OK;
}

int ilm_download(ilm_ctx* ctx, void* host, const void* dev, size_t bytes) {
    if (!ctx || (!dev && bytes) || (!host && bytes)) return fail(ILM_ERR_INVALID, "null argument");
    ILM_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
    ILM_HIP(hipStreamSynchronize(ctx->stream));
    return ILM_OK;
}

int ilm_memset(ilm_ctx* ctx, void* dev, int value, size_t bytes) {
    if (!ctx || (!dev && bytes)) return fail(ILM_ERR_INVALID, "null argument");
    ILM_HIP(hipMemsetAsync(dev, value, bytes, ctx->stream));
    ILM_HIP(hipStreamSynchronize(ctx->stream));
    return ILM_OK;
}

int ilm_sync(ilm_ctx* ctx) {
    if (!ctx) return fail(ILM_ERR_INVALID, "null ctx");
    ILM_HIP(hipStreamSynchronize(ctx->stream));
    unsigned long long flags = 0;
    ILM_HIP(hipMemcpy(&flags, ctx->counters + 2, sizeof(flags), hipMemcpyDeviceToHost));
    if (flags) {
        // reported once: cleared for the launches that follow
        const unsigned long long zero = 0;
        ILM_HIP(hipMemcpy(ctx->counters + 2, &zero, sizeof(zero), hipMemcpyHostToDevice));
        return fail(ILM_ERR_STATE, "a traversal exceeded its stack bound; results are incomplete");
    }
    return ILM_OK;
}

int ilm_timer_mark(ilm_ctx* ctx, int* slot) {
    if (!ctx || !slot) return fail(ILM_ERR_INVALID, "null argument");
    const int kSlots = 64;
    if ((int)ctx->events.size() < kSlots) {
        hipEvent_t ev;
        ILM_HIP(hipEventCreate(&ev));
        ctx->events.push_back(ev);
        *slot = (int)ctx->events.size() - 1;
    } else {
        *slot = ctx->next_event;
        ctx->next_event = (ctx->next_event + 1) % kSlots;
    }
    ILM_HIP(hipEventRecord(ctx->events[*slot], ctx->stream));
    return ILM_OK;
}

int ilm_timer_elapsed(ilm_ctx* ctx, int s0, int s1, float* ms) {
    if (!ctx || !ms || s0 < 0 || s1 < 0 || s0 >= (int)ctx->events.size() || s1 >= (int)ctx->events.size())
        return fail(ILM_ERR_INVALID, "bad timer slot");
    ILM_HIP(hipEventSynchronize(ctx->events[s1]));
    ILM_HIP(hipEventElapsedTime(ms, ctx->events[s0], ctx->events[s1]));
    return ILM_OK;
}

// ---------------------------------------------------------------- meshes

int ilm_mesh_create(ilm_ctx* ctx, const float* pos, int64_t nv, const int32_t* idx, int64_t nt, ilm_mesh** out) {
    if (!pos || !idx || !out) return fail(ILM_ERR_INVALID, "null argument");
    if (nv <= 0 || nt <= 0) return fail(ILM_ERR_INVALID, "empty mesh");
    if (nt >= (int64_t(1) << 26)) return fail(ILM_ERR_INVALID, "more than 2^26 - 1 triangles");
    for (int64_t i = 0; i < 3 * nt; ++i)
        if (idx[i] < 0 || idx[i] >= nv) return fail(ILM_ERR_INVALID, "vertex index out of range");
    ilm_mesh* m = new (std::nothrow) ilm_mesh();
    if (!m) return fail(ILM_ERR_NOMEM, "mesh allocation");
    m->ctx = ctx;
    m->nv = nv;
    m->nt = nt;
    m->pos.assign(pos, pos + 3 * nv);
    m->idx.assign(idx, idx + 3 * nt);
    *out = m;
    return ILM_OK;
}

int ilm_mesh_destroy(ilm_mesh* m) {
    if (!m) return ILM_OK;
    if (!m->ctx) { delete m; return ILM_OK; }
    if (m->d_nodes) ilm_free(m->ctx, m->d_nodes);
    if (m->d_tris) ilm_free(m->ctx, m->d_tris);
    if (m->d_slot_of) ilm_free(m->ctx, m->d_slot_of);
    delete m;
    return ILM_OK;
}

}  // extern "C"

namespace ilm {

int install_bvh(ilm_mesh* m, Bvh2& b, int node_format, std::chrono::steady_clock::time_point t0, int width) {
    WideBvh w;
    if (node_format == ILM_NODES_QUANTIZED && width != 4) return fail(ILM_ERR_INVALID, "quantised nodes are BVH4");
    try {
        w = collapse(b, width);
    } catch (const std::exception& e) {
        return fail(ILM_ERR_INVALID, std::string("BVH build failed: ") + e.what());
    }
    std::vector<float> tris;
    pack_triangles(m->pos.data(), m->idx.data(), w.order, tris);
    if (node_format == ILM_NODES_QUANTIZED) {
        std::vector<uint8_t> q((size_t)w.nnodes * sizeof(QNode4));
        const auto* src = reinterpret_cast<const WideNode<4>*>(w.nodes.data());
        auto* dst = reinterpret_cast<QNode4*>(q.data());
        for (int64_t i = 0; i < w.nnodes; ++i) quantize_node(src[i], dst[i]);
        w.nodes.swap(q);
    }
    m->node_format = node_format;
    m->info.node_format = node_format;
    m->info.node_stride = node_format == ILM_NODES_QUANTIZED ? (int32_t)sizeof(QNode4)
                          : width == 2                      ? (int32_t)sizeof(WideNode<2>)
                          : width == 8                      ? (int32_t)sizeof(WideNode<8>)
                                                            : (int32_t)sizeof(WideNode<4>);
    auto t1 = std::chrono::steady_clock::now();
    if (m->ctx) {  // a host-only mesh keeps the flattened tree for export only
        if (m->d_nodes) { ilm_free(m->ctx, m->d_nodes); m->d_nodes = nullptr; }
        if (m->d_tris) { ilm_free(m->ctx, m->d_tris); m->d_tris = nullptr; }
        if (m->d_slot_of) { ilm_free(m->ctx, m->d_slot_of); m->d_slot_of = nullptr; }
        int rc = ilm_alloc(m->ctx, w.nodes.size(), &m->d_nodes);
        if (rc) return rc;
        rc = ilm_alloc(m->ctx, tris.size() * sizeof(float), (void**)&m->d_tris);
        if (rc) return rc;
        rc = ilm_upload(m->ctx, m->d_nodes, w.nodes.data(), w.nodes.size());
        if (rc) return rc;
        rc = ilm_upload(m->ctx, m->d_tris, tris.data(), tris.size() * sizeof(float));
        if (rc) return rc;
    }
    m->width = w.width;
    m->info.nodes = w.nnodes;
    m->info.node_bytes = (int64_t)w.nodes.size();
    m->info.tri_bytes = (int64_t)(tris.size() * sizeof(float));
    m->info.width = w.width;
    m->info.depth = w.depth;
    m->info.max_leaf = w.max_leaf;
    m->info.stack_need = w.stack_need;
    m->info.build_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    m->info.sah_cost = w.sah_cost;
    m->h_nodes.swap(w.nodes);
    m->h_order.swap(w.order);
    m->built = true;
    return ILM_OK;
}

}  // namespace ilm

extern "C" {

int ilm_bvh_build(ilm_mesh* m, const ilm_bvh_params* params) {
    if (!m) return fail(ILM_ERR_INVALID, "null mesh");
    ilm_bvh_params p{};
    if (params) p = *params;
    if (p.width == 0) p.width = 4;
    if (p.width != 2 && p.width != 4 && p.width != 8) return fail(ILM_ERR_INVALID, "BVH width must be 2, 4 or 8");
    if (p.node_format != ILM_NODES_FLOAT && p.node_format != ILM_NODES_QUANTIZED)
        return fail(ILM_ERR_INVALID, "unknown node format");
    if (p.node_format == ILM_NODES_QUANTIZED && p.width != 4)
        return fail(ILM_ERR_INVALID, "quantised nodes are BVH4");
    if (p.builder == ILM_BUILDER_LBVH) {
        if (p.width != 4) return fail(ILM_ERR_INVALID, "the device builder makes BVH4 trees");
        if (!m->ctx) return fail(ILM_ERR_STATE, "the GPU builder needs a device context");
        return ilm::build_lbvh(m, p.max_leaf > 0 ? p.max_leaf : 4, p.node_format);
    }
    auto t0 = std::chrono::steady_clock::now();
    ilm::Bvh2 b;
    try {
        if (p.builder == ILM_BUILDER_REFERENCE) {
            b = ilm::build_reference(m->pos.data(), m->idx.data(), m->nt, p.max_leaf > 0 ? p.max_leaf : 10,
                                     p.max_depth > 0 ? p.max_depth : 35);
        } else if (p.builder == ILM_BUILDER_SAH) {
            b = ilm::build_sah(m->pos.data(), m->idx.data(), m->nt, p.max_leaf > 0 ? p.max_leaf : 4, p.threads);
        } else if (p.builder == ILM_BUILDER_KDTREE || p.builder == ILM_BUILDER_KDTREE_EX) {
            const int leaf = p.max_leaf > 0 ? p.max_leaf : 10;
            // KDTreeMeshEx's constructors take the depth limit from the leaf size
            // (KDTreeMeshEx.cpp:36-44).
            const int depth = p.builder == ILM_BUILDER_KDTREE_EX ? leaf : (p.max_depth > 0 ? p.max_depth : 35);
            if (depth > 40) return fail(ILM_ERR_INVALID, "kd-tree depth limit above 40");
            b = ilm::build_kdtree(m->pos.data(), m->idx.data(), m->nt, leaf, depth,
                                  p.builder == ILM_BUILDER_KDTREE_EX);
        } else {
            return fail(ILM_ERR_INVALID, "unknown builder");
        }
    } catch (const std::exception& e) {
        return fail(ILM_ERR_INVALID, std::string("BVH build failed: ") + e.what());
    }
    return ilm::install_bvh(m, b, p.node_format, t0, p.width);
}

int ilm_bvh_get_info(ilm_mesh* m, ilm_bvh_info* info) {
    if (!m || !info) return fail(ILM_ERR_INVALID, "null argument");
    if (!m->built) return fail(ILM_ERR_STATE, "BVH not built");
    *info = m->info;
    return ILM_OK;
}

int ilm_bvh_export(ilm_mesh* m, void* nodes, int32_t* order) {
    if (!m) return fail(ILM_ERR_INVALID, "null mesh");
    if (!m->built) return fail(ILM_ERR_STATE, "BVH not built");
    if (nodes) memcpy(nodes, m->h_nodes.data(), m->h_nodes.size());
    if (order) memcpy(order, m->h_order.data(), m->h_order.size() * sizeof(int32_t));
    return ILM_OK;
}

}  // extern "C"
