// GPU BVH construction: a linear BVH (Karras, "Maximizing parallelism in
// the construction of BVHs, octrees, and k-d trees", HPG 2012) over 63-bit
// Morton codes of triangle centroids, refitted bottom-up, then collapsed to
// the BVH4 layout the traversal kernels read, all on the device.
// Replaces the host BVHMesh::Compile (Shape/BVHMesh.cpp:78-107,203-269) where
// build time matters more than tree quality (many or large scenes).
//
// Steps (one stream, no host round trips except the per-level counts of the
// collapse):
//   1. triangle boxes; centroid bounds by block reduction + ordered-int atomics
//   2. Morton codes (21 bits per axis) and radix sort of (code, triangle)
//   3. Karras hierarchy: n-1 internal nodes, n leaves, parents and ranges
//   4. refit: each leaf climbs; the second child to arrive unites the boxes
//      (agent-scope acquire/release on the arrival counter)
//   5. collapse, level by level: each wide node opens its largest-area
//      internal children until it has four; a subtree of at most max_leaf
//      triangles becomes one leaf (its triangles are contiguous in Morton order)
//   6. triangles packed in Morton order
#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>

#include <algorithm>
#include <cfloat>
#include <cstring>
#include <vector>

#include "geom.hpp"
#include "internal.hpp"

namespace ilm {
namespace {

constexpr int kBlock = 256;

struct Box6 {
    float lo[3], hi[3];
};

__device__ __forceinline__ uint32_t ordered(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered(uint32_t u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7FFFFFFFu) : ~u);
}

// 1. triangle boxes and the bounds of their centroids
__global__ void __launch_bounds__(kBlock) k_tri_boxes(const float* __restrict__ pos, const int32_t* __restrict__ idx,
                                                      int64_t nt, Box6* __restrict__ boxes,
                                                      uint32_t* __restrict__ cbounds) {
    __shared__ float red[6][kBlock];
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    float c[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, C[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    if (i < nt) {
        Box6 b;
        for (int k = 0; k < 3; ++k) { b.lo[k] = FLT_MAX; b.hi[k] = -FLT_MAX; }
        for (int v = 0; v < 3; ++v) {
            const float* p = pos + 3 * (int64_t)idx[3 * i + v];
            for (int k = 0; k < 3; ++k) { b.lo[k] = fminf(b.lo[k], p[k]); b.hi[k] = fmaxf(b.hi[k], p[k]); }
        }
        boxes[i] = b;
        for (int k = 0; k < 3; ++k) c[k] = C[k] = 0.5f * (b.lo[k] + b.hi[k]);
    }
    for (int k = 0; k < 3; ++k) { red[k][threadIdx.x] = c[k]; red[3 + k][threadIdx.x] = C[k]; }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = fminf(red[k][threadIdx.x], red[k][threadIdx.x + s]);
                red[3 + k][threadIdx.x] = fmaxf(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k) {
            atomicMin(&cbounds[k], ordered(red[k][0]));
            atomicMax(&cbounds[3 + k], ordered(red[3 + k][0]));
        }
    }
}

__device__ __forceinline__ uint64_t spread21(uint64_t x) {
    x &= 0x1FFFFFull;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// 2. Morton codes of centroids in the centroid bounds
__global__ void __launch_bounds__(kBlock) k_morton(const Box6* __restrict__ boxes, int64_t nt,
                                                   const uint32_t* __restrict__ cbounds, uint64_t* __restrict__ codes,
                                                   int32_t* __restrict__ ids) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nt) return;
    const Box6 b = boxes[i];
    uint64_t q[3];
    for (int k = 0; k < 3; ++k) {
        float lo = unordered(cbounds[k]), hi = unordered(cbounds[3 + k]);
        float c = 0.5f * (b.lo[k] + b.hi[k]);
        float e = hi - lo;
        float u = e > 0.0f ? (c - lo) / e : 0.5f;
        u = fminf(fmaxf(u, 0.0f), 1.0f);
        q[k] = (uint64_t)fminf(u * 2097152.0f, 2097151.0f);
    }
    codes[i] = (spread21(q[0]) << 2) | (spread21(q[1]) << 1) | spread21(q[2]);
    ids[i] = (int32_t)i;
}

// common-prefix length of sorted keys i and j; equal codes are told apart by
// their positions (Karras, section 4)
__device__ __forceinline__ int delta(const uint64_t* __restrict__ k, int32_t n, int32_t i, int32_t j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = k[i], b = k[j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz((int)((uint32_t)i ^ (uint32_t)j));
}

// 3. internal node i of n-1; nodes 0..n-2 internal, n-1.. leaves (leaf k = n-1+k)
__global__ void __launch_bounds__(kBlock) k_karras(const uint64_t* __restrict__ keys, int32_t n,
                                                   int32_t* __restrict__ left, int32_t* __restrict__ right,
                                                   int32_t* __restrict__ parent, int32_t* __restrict__ range_lo,
                                                   int32_t* __restrict__ range_hi) {
    int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n - 1) return;
    int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, n, i, i - d);
    int32_t lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int32_t l = 0;
    for (int32_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    int32_t j = i + l * d;
    int dnode = delta(keys, n, i, j);
    // l >= 1 because keys extended by their positions are distinct; the loop
    // also ends for l == 0
    int32_t s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int32_t gamma = i + s * d + min(d, 0);
    int32_t lo = min(i, j), hi = max(i, j);
    int32_t lc = (lo == gamma) ? (n - 1 + gamma) : gamma;
    int32_t rc = (hi == gamma + 1) ? (n - 1 + gamma + 1) : (gamma + 1);
    left[i] = lc;
    right[i] = rc;
    parent[lc] = i;
    parent[rc] = i;
    range_lo[i] = lo;
    range_hi[i] = hi;
}

// 4. bottom-up boxes; flags start at zero
__global__ void __launch_bounds__(kBlock) k_refit(const Box6* __restrict__ tri_boxes, const int32_t* __restrict__ ids,
                                                  int32_t n, const int32_t* __restrict__ left,
                                                  const int32_t* __restrict__ right, const int32_t* __restrict__ parent,
                                                  Box6* nodes, uint32_t* flags) {
    int32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    int32_t node = n - 1 + k;
    nodes[node] = tri_boxes[ids[k]];
    int32_t p = parent[node];
    while (p >= 0) {
        // release this subtree's box; the second arrival acquires its sibling's
        uint32_t prev = __hip_atomic_fetch_add(&flags[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (prev == 0) return;
        Box6 a = nodes[left[p]], b = nodes[right[p]], u;
        for (int c = 0; c < 3; ++c) { u.lo[c] = fminf(a.lo[c], b.lo[c]); u.hi[c] = fmaxf(a.hi[c], b.hi[c]); }
        nodes[p] = u;
        p = parent[p];
    }
}

__device__ __forceinline__ float area6(const Box6& b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}

struct CollapseArgs {
    int32_t n, max_leaf;
    const int32_t *left, *right, *range_lo, *range_hi;
    const Box6* nodes;
    WideNode<4>* wide;
    const int2* qin;           // (bvh2 node, wide node)
    const uint32_t* qin_count;
    int2* qout;
    uint32_t* qout_count;
    uint32_t* wide_count;
    uint32_t* max_leaf_seen;
    double* cost;
    float inv_root_area;
};

__device__ __forceinline__ int32_t subtree_size(const CollapseArgs& a, int32_t node) {
    return node >= a.n - 1 ? 1 : a.range_hi[node] - a.range_lo[node] + 1;
}
__device__ __forceinline__ bool leafable(const CollapseArgs& a, int32_t node) {
    return node >= a.n - 1 || subtree_size(a, node) <= a.max_leaf;
}

// 5. one level of the collapse: one thread per wide node of this level
__global__ void __launch_bounds__(kBlock) k_collapse(CollapseArgs a) {
    uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    if (t >= *a.qin_count) return;
    const int2 item = a.qin[t];
    int32_t kids[4] = {a.left[item.x], a.right[item.x], -1, -1};
    int nk = 2;
    while (nk < 4) {
        int best = -1;
        float best_a = -1.0f;
        for (int k = 0; k < nk; ++k) {
            if (leafable(a, kids[k])) continue;
            float ar = area6(a.nodes[kids[k]]);
            if (ar > best_a) { best_a = ar; best = k; }
        }
        if (best < 0) break;
        int32_t open = kids[best];
        kids[best] = a.left[open];
        kids[nk++] = a.right[open];
    }
    WideNode<4> w;
    double cost = (double)area6(a.nodes[item.x]) * a.inv_root_area;
    for (int k = 0; k < 4; ++k) {
        if (k >= nk) {
            w.lox[k] = w.loy[k] = w.loz[k] = w.hix[k] = w.hiy[k] = w.hiz[k] = FLT_MAX;
            w.child[k] = make_leaf(0, 0);
            continue;
        }
        const int32_t c = kids[k];
        const Box6 b = a.nodes[c];
        w.lox[k] = b.lo[0]; w.hix[k] = b.hi[0];
        w.loy[k] = b.lo[1]; w.hiy[k] = b.hi[1];
        w.loz[k] = b.lo[2]; w.hiz[k] = b.hi[2];
        if (leafable(a, c)) {
            int32_t first = c >= a.n - 1 ? c - (a.n - 1) : a.range_lo[c];
            int32_t cnt = subtree_size(a, c);
            w.child[k] = make_leaf((uint32_t)first, (uint32_t)cnt);
            atomicMax(a.max_leaf_seen, (uint32_t)cnt);
            cost += (double)area6(b) * a.inv_root_area * cnt;
        } else {
            uint32_t slot = atomicAdd(a.wide_count, 1u);
            uint32_t q = atomicAdd(a.qout_count, 1u);
            a.qout[q] = make_int2(c, (int)slot);
            w.child[k] = slot;
        }
    }
    a.wide[item.y] = w;
    atomicAdd(a.cost, cost);
}

// 6. triangles in Morton order: (v0, original index), (v1, 0), (v2, 0)
__global__ void __launch_bounds__(kBlock) k_pack(const float* __restrict__ pos, const int32_t* __restrict__ idx,
                                                 const int32_t* __restrict__ ids, int32_t n, float4* __restrict__ out) {
    int32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    int32_t t = ids[k];
    for (int v = 0; v < 3; ++v) {
        const float* p = pos + 3 * (int64_t)idx[3 * (int64_t)t + v];
        out[3 * (int64_t)k + v] = make_float4(p[0], p[1], p[2], v == 0 ? __int_as_float(t) : 0.0f);
    }
}

// float BVH4 nodes -> quantised nodes (same quantiser as the host builders)
__global__ void __launch_bounds__(kBlock) k_quantize(const WideNode<4>* __restrict__ in, int64_t n,
                                                     QNode4* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    QNode4 q;
    quantize_node(in[i], q);
    out[i] = q;
}

__global__ void k_init_u32(uint32_t* p, int64_t n, uint32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

// Builds the BVH4 of m on its context's device. On success m->d_nodes,
// m->d_tris, m->info, m->h_nodes and m->h_order hold the result.
int build_lbvh(ilm_mesh* m, int max_leaf, int node_format) {
    ilm_ctx* ctx = m->ctx;
    const int64_t nt = m->nt, nv = m->nv;
    if (nt > (1 << 26) - 1) return fail(ILM_ERR_INVALID, "more than 2^26 - 1 triangles");
    max_leaf = std::max(1, std::min(max_leaf, kMaxLeafTris));
    const int32_t n = (int32_t)nt;
    hipStream_t st = ctx->stream;

    // scratch (freed at the end)
    std::vector<void*> tmp;
    auto alloc = [&](size_t bytes, void** p) -> int {
        int rc = ilm_alloc(ctx, std::max<size_t>(bytes, 16), p);
        if (rc == ILM_OK) tmp.push_back(*p);
        return rc;
    };
    auto cleanup = [&]() {
        for (void* p : tmp) ilm_free(ctx, p);
        tmp.clear();
    };
    float *d_pos = nullptr;
    int32_t *d_idx = nullptr, *ids_in = nullptr, *ids = nullptr, *left = nullptr, *right = nullptr,
            *parent = nullptr, *rlo = nullptr, *rhi = nullptr;
    uint64_t *codes_in = nullptr, *codes = nullptr;
    Box6 *tri_boxes = nullptr, *nodes = nullptr;
    uint32_t *cb = nullptr, *flags = nullptr, *counts = nullptr;
    int2 *qa = nullptr, *qb = nullptr;
    double* cost = nullptr;
    const int32_t ninner = std::max(n - 1, 1);
    int rc = ILM_OK;
    rc |= alloc(nv * 3 * sizeof(float), (void**)&d_pos);
    rc |= alloc(nt * 3 * sizeof(int32_t), (void**)&d_idx);
    rc |= alloc(nt * sizeof(Box6), (void**)&tri_boxes);
    rc |= alloc(6 * sizeof(uint32_t), (void**)&cb);
    rc |= alloc(nt * sizeof(uint64_t), (void**)&codes_in);
    rc |= alloc(nt * sizeof(uint64_t), (void**)&codes);
    rc |= alloc(nt * sizeof(int32_t), (void**)&ids_in);
    rc |= alloc(nt * sizeof(int32_t), (void**)&ids);
    rc |= alloc(ninner * sizeof(int32_t), (void**)&left);
    rc |= alloc(ninner * sizeof(int32_t), (void**)&right);
    rc |= alloc(ninner * sizeof(int32_t), (void**)&rlo);
    rc |= alloc(ninner * sizeof(int32_t), (void**)&rhi);
    rc |= alloc((2 * nt) * sizeof(int32_t), (void**)&parent);
    rc |= alloc((2 * nt) * sizeof(Box6), (void**)&nodes);
    rc |= alloc(ninner * sizeof(uint32_t), (void**)&flags);
    rc |= alloc(ninner * sizeof(int2), (void**)&qa);
    rc |= alloc(ninner * sizeof(int2), (void**)&qb);
    rc |= alloc(8 * sizeof(uint32_t), (void**)&counts);  // [0] qa, [1] qb, [2] wide, [3] max leaf
    rc |= alloc(sizeof(double), (void**)&cost);
    if (rc != ILM_OK) { cleanup(); return fail(ILM_ERR_NOMEM, "LBVH scratch allocation failed"); }
    // wide nodes: at most one per internal BVH2 node, plus the root
    void* d_nodes = nullptr;
    float* d_tris = nullptr;
    const int64_t max_wide = std::max<int64_t>(n - 1, 1);
    if (ilm_alloc(ctx, (size_t)max_wide * sizeof(WideNode<4>), &d_nodes) != ILM_OK ||
        ilm_alloc(ctx, (size_t)nt * 12 * sizeof(float), (void**)&d_tris) != ILM_OK) {
        if (d_nodes) ilm_free(ctx, d_nodes);
        cleanup();
        return fail(ILM_ERR_NOMEM, "LBVH output allocation failed");
    }
    auto fail_all = [&](int code, const char* what) {
        ilm_free(ctx, d_nodes);
        ilm_free(ctx, d_tris);
        cleanup();
        return fail(code, what);
    };
#define LBVH_HIP(call)                                              \
    do {                                                            \
        hipError_t e_ = (call);                                     \
        if (e_ != hipSuccess) {                                     \
            fail_hip(e_, #call);                                    \
            std::string msg_ = ilm_last_error();                    \
            return fail_all(ILM_ERR_HIP, msg_.c_str());             \
        }                                                           \
    } while (0)

    hipEvent_t e0, e1;
    LBVH_HIP(hipEventCreate(&e0));
    LBVH_HIP(hipEventCreate(&e1));
    LBVH_HIP(hipMemcpyAsync(d_pos, m->pos.data(), nv * 3 * sizeof(float), hipMemcpyHostToDevice, st));
    LBVH_HIP(hipMemcpyAsync(d_idx, m->idx.data(), nt * 3 * sizeof(int32_t), hipMemcpyHostToDevice, st));
    LBVH_HIP(hipEventRecord(e0, st));
    const uint32_t cb_init[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    // small initialisations by kernel (no memset nodes if a caller captures the stream)
    hipLaunchKernelGGL(k_init_u32, dim3(1), dim3(64), 0, st, cb, 3, cb_init[0]);
    hipLaunchKernelGGL(k_init_u32, dim3(1), dim3(64), 0, st, cb + 3, 3, 0u);
    hipLaunchKernelGGL(k_init_u32, dim3(blocks_for(ninner)), dim3(kBlock), 0, st, flags, (int64_t)ninner, 0u);
    hipLaunchKernelGGL(k_init_u32, dim3(1), dim3(64), 0, st, counts, 8, 0u);
    hipLaunchKernelGGL(k_init_u32, dim3(1), dim3(64), 0, st, (uint32_t*)cost, 2, 0u);
    hipLaunchKernelGGL(k_tri_boxes, dim3(blocks_for(nt)), dim3(kBlock), 0, st, d_pos, d_idx, nt, tri_boxes, cb);
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(nt)), dim3(kBlock), 0, st, tri_boxes, nt, cb, codes_in, ids_in);
    LBVH_HIP(hipGetLastError());
    size_t sort_bytes = 0;
    LBVH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, codes_in, codes, ids_in, ids, n, 0, 63, st));
    void* sort_tmp = nullptr;
    if (alloc(sort_bytes, &sort_tmp) != ILM_OK) return fail_all(ILM_ERR_NOMEM, "LBVH sort scratch");
    LBVH_HIP(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, codes_in, codes, ids_in, ids, n, 0, 63, st));

    int levels = 0;
    uint32_t h_counts[8] = {0};
    double h_cost = 0;
    if (n <= max_leaf) {
        // a single leaf under the root
        WideNode<4> w;
        float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
        for (int64_t t = 0; t < nt; ++t)
            for (int v = 0; v < 3; ++v)
                for (int k = 0; k < 3; ++k) {
                    float c = m->pos[3 * (int64_t)m->idx[3 * t + v] + k];
                    lo[k] = std::min(lo[k], c);
                    hi[k] = std::max(hi[k], c);
                }
        for (int k = 0; k < 4; ++k) {
            w.lox[k] = w.loy[k] = w.loz[k] = w.hix[k] = w.hiy[k] = w.hiz[k] = FLT_MAX;
            w.child[k] = make_leaf(0, 0);
        }
        w.lox[0] = lo[0]; w.loy[0] = lo[1]; w.loz[0] = lo[2];
        w.hix[0] = hi[0]; w.hiy[0] = hi[1]; w.hiz[0] = hi[2];
        w.child[0] = make_leaf(0, (uint32_t)n);
        LBVH_HIP(hipMemcpyAsync(d_nodes, &w, sizeof w, hipMemcpyHostToDevice, st));
        h_counts[2] = 1;
        h_counts[3] = (uint32_t)n;
        h_cost = (double)n;
        levels = 1;
    } else {
        hipLaunchKernelGGL(k_init_u32, dim3(blocks_for(2 * nt)), dim3(kBlock), 0, st, (uint32_t*)parent, 2 * nt,
                           0xFFFFFFFFu);
        hipLaunchKernelGGL(k_karras, dim3(blocks_for(n - 1)), dim3(kBlock), 0, st, codes, n, left, right, parent,
                           rlo, rhi);
        hipLaunchKernelGGL(k_refit, dim3(blocks_for(n)), dim3(kBlock), 0, st, tri_boxes, ids, n, left, right, parent,
                           nodes, flags);
        LBVH_HIP(hipGetLastError());
        Box6 root;
        LBVH_HIP(hipMemcpyAsync(&root, nodes, sizeof root, hipMemcpyDeviceToHost, st));
        LBVH_HIP(hipStreamSynchronize(st));
        float ra = 2.0f * ((root.hi[0] - root.lo[0]) * (root.hi[1] - root.lo[1]) +
                           (root.hi[1] - root.lo[1]) * (root.hi[2] - root.lo[2]) +
                           (root.hi[2] - root.lo[2]) * (root.hi[0] - root.lo[0]));
        // level 0: internal node 0 -> wide node 0
        const int2 first = make_int2(0, 0);
        const uint32_t one[3] = {1u, 0u, 1u};  // qa count, qb count, wide count
        LBVH_HIP(hipMemcpyAsync(qa, &first, sizeof first, hipMemcpyHostToDevice, st));
        LBVH_HIP(hipMemcpyAsync(counts, one, sizeof one, hipMemcpyHostToDevice, st));
        CollapseArgs a;
        a.n = n;
        a.max_leaf = max_leaf;
        a.left = left; a.right = right; a.range_lo = rlo; a.range_hi = rhi;
        a.nodes = nodes;
        a.wide = reinterpret_cast<WideNode<4>*>(d_nodes);
        a.wide_count = counts + 2;
        a.max_leaf_seen = counts + 3;
        a.cost = cost;
        a.inv_root_area = ra > 0 ? 1.0f / ra : 0.0f;
        uint32_t level_items = 1;
        int2 *qin = qa, *qout = qb;
        uint32_t *cin = counts, *cout = counts + 1;
        while (level_items > 0) {
            if (++levels > 4096) return fail_all(ILM_ERR_STATE, "LBVH collapse did not terminate");
            a.qin = qin; a.qin_count = cin; a.qout = qout; a.qout_count = cout;
            hipLaunchKernelGGL(k_init_u32, dim3(1), dim3(64), 0, st, cout, 1, 0u);
            hipLaunchKernelGGL(k_collapse, dim3(blocks_for(level_items)), dim3(kBlock), 0, st, a);
            LBVH_HIP(hipGetLastError());
            LBVH_HIP(hipMemcpyAsync(&level_items, cout, sizeof level_items, hipMemcpyDeviceToHost, st));
            LBVH_HIP(hipStreamSynchronize(st));
            std::swap(qin, qout);
            std::swap(cin, cout);
        }
        LBVH_HIP(hipMemcpyAsync(h_counts, counts, sizeof h_counts, hipMemcpyDeviceToHost, st));
        LBVH_HIP(hipMemcpyAsync(&h_cost, cost, sizeof h_cost, hipMemcpyDeviceToHost, st));
    }
    hipLaunchKernelGGL(k_pack, dim3(blocks_for(n)), dim3(kBlock), 0, st, d_pos, d_idx, ids, n,
                       reinterpret_cast<float4*>(d_tris));
    LBVH_HIP(hipGetLastError());
    const int64_t nwide = h_counts[2];
    size_t stride = sizeof(WideNode<4>);
    if (node_format == ILM_NODES_QUANTIZED) {
        void* qn = nullptr;
        if (alloc((size_t)nwide * sizeof(QNode4), &qn) != ILM_OK) return fail_all(ILM_ERR_NOMEM, "quantised nodes");
        hipLaunchKernelGGL(k_quantize, dim3(blocks_for(nwide)), dim3(kBlock), 0, st,
                           reinterpret_cast<const WideNode<4>*>(d_nodes), nwide, reinterpret_cast<QNode4*>(qn));
        LBVH_HIP(hipGetLastError());
        // the float nodes become scratch and the quantised ones the output
        tmp.erase(std::find(tmp.begin(), tmp.end(), qn));
        tmp.push_back(d_nodes);
        d_nodes = qn;
        stride = sizeof(QNode4);
    }
    LBVH_HIP(hipEventRecord(e1, st));
    LBVH_HIP(hipStreamSynchronize(st));
    float gpu_ms = 0;
    LBVH_HIP(hipEventElapsedTime(&gpu_ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);

    // host copies for export and for hit-slot lookups
    m->h_nodes.resize((size_t)nwide * stride);
    m->h_order.resize(nt);
    LBVH_HIP(hipMemcpy(m->h_nodes.data(), d_nodes, m->h_nodes.size(), hipMemcpyDeviceToHost));
    LBVH_HIP(hipMemcpy(m->h_order.data(), ids, nt * sizeof(int32_t), hipMemcpyDeviceToHost));
    cleanup();
#undef LBVH_HIP

    if (m->d_nodes) ilm_free(ctx, m->d_nodes);
    if (m->d_tris) ilm_free(ctx, m->d_tris);
    if (m->d_slot_of) { ilm_free(ctx, m->d_slot_of); m->d_slot_of = nullptr; }
    m->d_nodes = d_nodes;
    m->d_tris = d_tris;
    m->width = 4;
    m->node_format = node_format;
    m->info.node_format = node_format;
    m->info.node_stride = (int32_t)stride;
    m->info.nodes = nwide;
    m->info.node_bytes = nwide * (int64_t)stride;
    m->info.tri_bytes = nt * 12 * (int64_t)sizeof(float);
    m->info.width = 4;
    m->info.depth = levels;
    m->info.max_leaf = (int32_t)h_counts[3];
    m->info.stack_need = 3 * levels + 1;  // at most three deferred siblings per level
    m->info.build_ms = gpu_ms;
    m->info.sah_cost = h_cost;
    m->built = true;
    return ILM_OK;
}

}  // namespace ilm
