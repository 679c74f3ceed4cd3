// Ray traversal kernels for CDNA4 (gfx950) and the tracing entry points of
// the C ABI.
//
// Design: persistent workgroups sized to the chip's occupancy; each wave
// refills idle lanes with new rays from a fetch counter (one atomic per wave
// per refill, prefix via ballot/mbcnt), so incoherent rays of very different
// traversal length keep SIMD lanes busy. The rays are split into eight
// contiguous segments with a counter each; a workgroup starts on segment
// blockIdx % 8, which under the round-robin dispatch of workgroups to the
// eight XCDs is its XCD's own, and moves on once a segment is spent. Each
// XCD's L2 then serves a narrow window of neighbouring (coherent) rays, and
// the fetch atomics spread over eight counters. Only the first ceil(n / 64)
// waves of the grid take part, so short launches stay cheap. A wave refills once at
// least refill_min of its lanes are idle (24 by default; 56 for coherent
// camera rays, whose lanes finish together). Per lane: a 16-entry
// traversal stack ring in LDS (lane-strided, conflict-free) that spills its
// oldest entries to a global area only when a node's pushes could overrun
// it. Nodes are wide (BVH2/4/8 with float boxes, 64/112/224 bytes, or
// 64-byte quantised BVH4), loaded as 16-byte vectors. The vector ALU and the
// vector-memory path (address unit, L1) both charge a wave instruction in
// full however few of its lanes are active, so the kernel keeps instructions
// per ray low and lanes busy. A node step avoids branches: the children's
// keys (entry distance bits with the slot in the low bits) are sorted with a
// min/max comparator network, the child words are read back from LDS by slot,
// and every push is written at the top and kept only when its child was hit;
// closest-hit rays visit the nearest child first, any-hit launches keep the
// children in slot order. Traversal is speculative (Aila and Laine,
// "Understanding the Efficiency of Ray Traversal on GPUs", HPG 2009): a lane
// that reaches a leaf parks it and keeps descending until a quarter (closest
// hits) or three eighths (any-hit, dual and two-level launches) of the lanes
// in the node loop hold a leaf; the leaf loop then tests one triangle per lane
// per iteration, each lane moving on to its next leaf on its own, and hands
// the wave back to the node loop once few lanes hold triangles.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <mutex>
#include <string>

#include "geom.hpp"
#include "internal.hpp"

namespace ilm {

constexpr int kBlock = 256;
constexpr int kLds = 16;

struct TraceArgs {
    const float4* __restrict__ nodes;  // wide nodes (Kind<K>::F4 float4 each)
    const float4* __restrict__ tris;   // 3 float4 per triangle
    const ilm_ray* __restrict__ rays;
    int64_t n;                        // ray count, or capacity when n_dev is set
    const unsigned long long* n_dev;  // ray count in device memory (wavefront queues), or null
    ilm_hit* __restrict__ hits;
    uint8_t* __restrict__ occ;
    int32_t* __restrict__ stats;
    uint32_t* __restrict__ spill;
    int32_t spill_depth;
    int32_t rooted;                   // start node of each ray from its tmin bits
    int32_t refill_min;               // idle lanes that trigger a wave's refill (1..64)
    int64_t grid_threads;
    unsigned long long* counters;
    uint32_t total_waves;
    // dual launches: a second queue traced for any hits in the same launch,
    // as ids n1 .. n1 + n2 of the combined range (n1 = the first queue's count)
    const ilm_ray* __restrict__ rays2;
    int64_t n2;                        // its count, or capacity when n2_dev is set
    const unsigned long long* n2_dev;
    uint8_t* __restrict__ occ2;
    // instance tables (two-level traversal): the top-level tree over the
    // instances' world boxes (nodes then holds the meshes' trees), the
    // instance records in its leaf order (world-to-object rows 0..2, then
    // (mesh root node, instance index, mesh, 0)) and the hit instances out
    const float4* __restrict__ tlas;
    const float4* __restrict__ inst;
    int32_t* __restrict__ inst_out;
};

__device__ __forceinline__ float slab1(float lox, float hix, float loy, float hiy, float loz, float hiz,
                                       const RayState& r) {
    float tx0 = fmaf(lox, r.inv.x, -r.oinv.x), tx1 = fmaf(hix, r.inv.x, -r.oinv.x);
    float ty0 = fmaf(loy, r.inv.y, -r.oinv.y), ty1 = fmaf(hiy, r.inv.y, -r.oinv.y);
    float tz0 = fmaf(loz, r.inv.z, -r.oinv.z), tz1 = fmaf(hiz, r.inv.z, -r.oinv.z);
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), r.tlo));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), r.thi));
    return (tn <= tf * 1.00000048f) ? tn : INFINITY;
}

// Node kinds of the traversal kernels: float boxes at width 4, 2 or 8
// (WideNode<W>, 7 / 4 / 14 x 16 B) and the quantised BVH4 node (QNode4, 4 x 16 B).
enum NodeKind { kFloat4 = 0, kQuant4 = 1, kFloat2 = 2, kFloat8 = 3 };
template <int K> struct Kind;
template <> struct Kind<kFloat4> { static constexpr int W = 4, F4 = 7; };
template <> struct Kind<kQuant4> { static constexpr int W = 4, F4 = 4; };
template <> struct Kind<kFloat2> { static constexpr int W = 2, F4 = 4; };
template <> struct Kind<kFloat8> { static constexpr int W = 8, F4 = 14; };

// Child entry distances of a node (INFINITY: missed or empty) and its child words.
template <int K>
__device__ __forceinline__ void node_hits(const float4* __restrict__ nodes, uint32_t cur, const RayState& r,
                                          float (&d)[Kind<K>::W], uint32_t (&c)[Kind<K>::W]) {
    const float4* np = nodes + (int64_t)cur * Kind<K>::F4;
    if (K == kFloat4) {
        float4 lx = np[0], hx = np[1], ly = np[2], hy = np[3], lz = np[4], hz = np[5];
        uint4 ch = reinterpret_cast<const uint4*>(np)[6];
        d[0] = slab1(lx.x, hx.x, ly.x, hy.x, lz.x, hz.x, r);
        d[1] = slab1(lx.y, hx.y, ly.y, hy.y, lz.y, hz.y, r);
        d[2] = slab1(lx.z, hx.z, ly.z, hy.z, lz.z, hz.z, r);
        d[3] = slab1(lx.w, hx.w, ly.w, hy.w, lz.w, hz.w, r);
        c[0] = ch.x; c[1] = ch.y; c[2] = ch.z; c[3] = ch.w;
    } else if (K == kFloat2) {
        // (lox0 lox1 hix0 hix1) (loy.. hiy..) (loz.. hiz..) (child0 child1 pad pad)
        float4 x = np[0], y = np[1], z = np[2];
        uint4 ch = reinterpret_cast<const uint4*>(np)[3];
        d[0] = slab1(x.x, x.z, y.x, y.z, z.x, z.z, r);
        d[1] = slab1(x.y, x.w, y.y, y.w, z.y, z.w, r);
        c[0] = ch.x; c[1] = ch.y;
    } else if (K == kFloat8) {
        float4 lx0 = np[0], lx1 = np[1], hx0 = np[2], hx1 = np[3], ly0 = np[4], ly1 = np[5], hy0 = np[6],
               hy1 = np[7], lz0 = np[8], lz1 = np[9], hz0 = np[10], hz1 = np[11];
        uint4 c0 = reinterpret_cast<const uint4*>(np)[12], c1 = reinterpret_cast<const uint4*>(np)[13];
        d[0] = slab1(lx0.x, hx0.x, ly0.x, hy0.x, lz0.x, hz0.x, r);
        d[1] = slab1(lx0.y, hx0.y, ly0.y, hy0.y, lz0.y, hz0.y, r);
        d[2] = slab1(lx0.z, hx0.z, ly0.z, hy0.z, lz0.z, hz0.z, r);
        d[3] = slab1(lx0.w, hx0.w, ly0.w, hy0.w, lz0.w, hz0.w, r);
        d[4] = slab1(lx1.x, hx1.x, ly1.x, hy1.x, lz1.x, hz1.x, r);
        d[5] = slab1(lx1.y, hx1.y, ly1.y, hy1.y, lz1.y, hz1.y, r);
        d[6] = slab1(lx1.z, hx1.z, ly1.z, hy1.z, lz1.z, hz1.z, r);
        d[7] = slab1(lx1.w, hx1.w, ly1.w, hy1.w, lz1.w, hz1.w, r);
        c[0] = c0.x; c[1] = c0.y; c[2] = c0.z; c[3] = c0.w;
        c[4] = c1.x; c[5] = c1.y; c[6] = c1.z; c[7] = c1.w;
    } else {
        float4 o = np[0];
        uint4 ch = reinterpret_cast<const uint4*>(np)[1];
        uint4 qa = reinterpret_cast<const uint4*>(np)[2];
        uint4 qb = reinterpret_cast<const uint4*>(np)[3];
        const uint32_t ex = __float_as_uint(o.w);
        // child plane t = (p + q 2^e - o) / d = q * (2^e / d) + (p - o) / d
        const float Sx = __uint_as_float((ex & 0xFFu) << 23) * r.inv.x, Bx = fmaf(o.x, r.inv.x, -r.oinv.x);
        const float Sy = __uint_as_float(((ex >> 8) & 0xFFu) << 23) * r.inv.y, By = fmaf(o.y, r.inv.y, -r.oinv.y);
        const float Sz = __uint_as_float(((ex >> 16) & 0xFFu) << 23) * r.inv.z, Bz = fmaf(o.z, r.inv.z, -r.oinv.z);
        const uint32_t cw[4] = {ch.x, ch.y, ch.z, ch.w};
        // near and far byte planes by the direction signs (S has the sign of
        // 1/d, and t = q * S + B is monotone in q): the same distances as the
        // min/max form, in a select per axis instead of two per child
        const bool fx = r.inv.x < 0.f, fy = r.inv.y < 0.f, fz = r.inv.z < 0.f;
        const uint32_t nqx = fx ? qa.y : qa.x, fqx = fx ? qa.x : qa.y;
        const uint32_t nqy = fy ? qa.w : qa.z, fqy = fy ? qa.z : qa.w;
        const uint32_t nqz = fz ? qb.y : qb.x, fqz = fz ? qb.x : qb.y;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int sh = 8 * k;
            const float tn = fmaxf(fmaxf(fmaf((float)((nqx >> sh) & 0xFFu), Sx, Bx), fmaf((float)((nqy >> sh) & 0xFFu), Sy, By)),
                                   fmaxf(fmaf((float)((nqz >> sh) & 0xFFu), Sz, Bz), r.tlo));
            const float tf = fminf(fminf(fmaf((float)((fqx >> sh) & 0xFFu), Sx, Bx), fmaf((float)((fqy >> sh) & 0xFFu), Sy, By)),
                                   fminf(fmaf((float)((fqz >> sh) & 0xFFu), Sz, Bz), r.thi));
            d[k] = (tn <= tf * 1.00000048f && cw[k] != kEmptySlot) ? tn : INFINITY;
            c[k] = cw[k];
        }
    }
}

// Child keys of a node (child_key: entry distance bits with the slot in the
// low bits), and its child words to the lane's W staging words in LDS, from
// which the traversal reads them back by slot once the keys are sorted.
template <int K>
__device__ __forceinline__ void node_keys(const float4* __restrict__ nodes, uint32_t cur, const RayState& r,
                                          uint32_t (&key)[Kind<K>::W], uint32_t* __restrict__ stage) {
    constexpr int W = Kind<K>::W;
    float d[W];
    uint32_t c[W];
    node_hits<K>(nodes, cur, r, d, c);
#pragma unroll
    for (int k = 0; k < W; ++k) key[k] = child_key<W>(d[k], k);
    if (W == 2) {
        *reinterpret_cast<uint2*>(stage) = make_uint2(c[0], c[1]);
    } else {
#pragma unroll
        for (int q = 0; q < W / 4; ++q)
            reinterpret_cast<uint4*>(stage)[q] = make_uint4(c[4 * q], c[4 * q + 1], c[4 * q + 2], c[4 * q + 3]);
    }
}

// Entries a lane keeps in its LDS ring before a node's W - 1 pushes: with at
// most this many live entries the pushes (each written at the top, kept when
// the child was hit) never overwrite a live entry of the 16-entry ring.
template <int W> constexpr int kKeep = kLds - (W - 1);

// Speculation thresholds: the node loop ends once kSpecNum / kSpecDen of its
// lanes hold a parked leaf (a quarter for closest hits; three eighths in the
// instance walk, whose mesh descents are short and re-enter the node loop
// after each instance, in dual launches, whose shadow segments end early, and
// in any-hit launches); the leaf loop ends once fewer than kLeafNum / kLeafDen
// of the wave's active lanes hold triangles (and a lane can descend).
template <bool ANY, bool INST, bool DUAL> constexpr bool kSpecWide = ANY || INST || DUAL;
template <bool ANY, bool INST, bool DUAL> constexpr int kSpecNum = kSpecWide<ANY, INST, DUAL> ? 3 : 1;
template <bool ANY, bool INST, bool DUAL> constexpr int kSpecDen = kSpecWide<ANY, INST, DUAL> ? 8 : 4;
constexpr int kLeafNum = 1, kLeafDen = 4;
static_assert((kLds & (kLds - 1)) == 0, "the LDS ring is indexed by position & (kLds - 1)");

// Waves per SIMD the traversal kernels are compiled for: seven (72 VGPRs)
// measured 4% faster than the six that the unconstrained register count
// allows; BVH8 nodes need ~120 VGPRs and keep the four those allow. The
// statistics variant keeps six (80 VGPRs), where it needs no scratch memory.
// The two-level walk stays at seven although it spills 44 bytes per lane
// there: at six, without the spill, it traced 3% fewer incoherent rays.
template <int K, bool WIDE> constexpr int kTraceWaves = K == kFloat8 ? 4 : WIDE ? 6 : 7;

// Stack marker of the two-level traversal: back to the world-space ray.
constexpr uint32_t kReturn = 0x7FFFFFFEu;

// INST: two-level traversal of an instance table. A lane walks the top-level
// tree in world space; at an instance leaf it pushes kReturn and continues in
// the instance's mesh tree with the ray transformed into object space (the
// direction not normalised, so t is shared); popping kReturn reloads the
// world ray. Leaves are parked only inside a mesh tree, and a lane holding
// one stops at kReturn until the leaf's triangles are tested with the
// object-space ray.
template <bool ANY, bool STATS, int KIND, bool DUAL, bool INST = false>
__global__ void __launch_bounds__(kBlock, (kTraceWaves<KIND, STATS>)) k_trace4(TraceArgs a) {
    static_assert(!INST || (KIND == kFloat4 && !DUAL && !STATS), "instance tables hold float BVH4 trees");
    constexpr int W = Kind<KIND>::W;
    // per lane: a kLds-entry stack ring (lane-strided, conflict-free) and W
    // staging words for the current node's child words
    __shared__ __align__(16) uint32_t lds[(kLds + W) * kBlock];
    const int tid = threadIdx.x;
    uint32_t* const ring = lds + tid;
    uint32_t* const stage = lds + kLds * kBlock + tid * W;
    const uint32_t stage_off = (uint32_t)((kLds * kBlock + tid * W) * sizeof(uint32_t));
    const int64_t gtid = (int64_t)blockIdx.x * kBlock + tid;
    const int64_t n1 = a.n_dev ? (int64_t)min((unsigned long long)a.n, *a.n_dev) : a.n;
    const int64_t n = DUAL ? n1 + (a.n2_dev ? (int64_t)min((unsigned long long)a.n2, *a.n2_dev) : a.n2) : n1;
    // only the first ceil(n / 64) waves can receive rays: the others leave at
    // once, without touching the counters, so that a short launch (the late
    // bounces of a frame) costs few atomics
    const unsigned long long wave_id = (unsigned long long)gtid >> 6;
    const unsigned long long part = min((unsigned long long)a.total_waves, ((unsigned long long)n + 63ull) >> 6);
    if (wave_id >= part) return;
    const unsigned long long seg_len = ((unsigned long long)n + kTraceSegs - 1) / kTraceSegs;

    bool active = false;
    bool exhausted = false;  // uniform across the wave
    // the segment the wave fetches from (it starts on its XCD's) and the
    // number of segments it has found spent; uniform across the wave
    unsigned int seg = blockIdx.x % kTraceSegs, spent = 0;
    uint32_t ray_id = 0;  // launches hold fewer than 2^32 rays (checked on the host)
    bool any_ray = ANY;  // this lane's ray wants any hit (dual launches: the second queue's rays)
    RayState r;
    uint32_t cur = kNoNode;
    uint32_t stash = kNoNode;  // a leaf parked during speculative descent
    uint32_t tk = 0;           // next triangle of the parked leaf
    // stack: entries at positions [0, sp); those at [base, sp) are in the
    // LDS ring (position p in slot p & (kLds - 1)), the older ones in the
    // global spill area (position p at spill[p * grid_threads + gtid])
    int sp = 0, base = 0;
    float best_t = 0.f, best_u = 0.f, best_v = 0.f;
    int32_t best_prim = -1;
    int32_t n_nodes = 0, n_tris = 0;
    bool in_blas = false;              // INST: the lane is inside an instance's mesh tree
    bool world_stale = false;          // INST: r still holds an instance's object-space set-up
    int32_t inst_id = -1, best_inst = -1;

    auto pop = [&]() -> uint32_t {
        if (sp > base) {
            --sp;
            return ring[(sp & (kLds - 1)) * kBlock];
        }
        if (sp == 0) return kNoNode;
        --sp;
        base = sp;
        // a position beyond the spill area was dropped (flagged when it was)
        return sp < a.spill_depth ? a.spill[(int64_t)sp * a.grid_threads + gtid] : kNoNode;
    };
    auto finish = [&]() {
        if (ANY) {
            a.occ[ray_id] = best_prim >= 0 ? 1 : 0;
        } else if (DUAL && any_ray) {
            a.occ2[ray_id - (uint32_t)n1] = best_prim >= 0 ? 1 : 0;
        } else {
            float4 h = make_float4(best_t, __int_as_float(best_prim), best_u, best_v);
            reinterpret_cast<float4*>(a.hits)[ray_id] = h;
            if (INST) a.inst_out[ray_id] = best_prim >= 0 ? best_inst : -1;
        }
        if (STATS) {
            a.stats[2 * (int64_t)ray_id] = n_nodes;
            a.stats[2 * (int64_t)ray_id + 1] = n_tris;
        }
        active = false;
    };
    // one push (INST's instance entries), spilling the oldest ring entry when full
    auto push1 = [&](uint32_t x) {
        while (sp - base >= kLds) {
            if (base < a.spill_depth)
                a.spill[(int64_t)base * a.grid_threads + gtid] = ring[(base & (kLds - 1)) * kBlock];
            else
                atomicOr(&a.counters[2], 1ull);
            ++base;
        }
        ring[(sp & (kLds - 1)) * kBlock] = x;
        ++sp;
    };
    // the lane has a traversal step to take: an inner node (INST: also any
    // top-level entry, and kReturn once no leaf is parked)
    auto can_step = [&]() -> bool {
        if (!active || cur == kNoNode) return false;
        if (INST) {
            if (!in_blas) return true;
            if (cur == kReturn) return stash == kNoNode;
        }
        return !is_leaf(cur);
    };

    for (;;) {
        unsigned long long idle = __ballot(!active);
        if (!exhausted && idle != 0ull && (__popcll(idle) >= a.refill_min || idle == ~0ull)) {
            // idle lanes take the next ray ids of the current segment; a spent
            // segment sends the wave on to the next one
            const unsigned long long need = (unsigned long long)__popcll(idle);
            const unsigned long long rank = (unsigned long long)__popcll(idle & __lanemask_lt());
            unsigned long long got = 0, mine = ~0ull;
            while (got < need && spent < (unsigned)kTraceSegs) {
                const unsigned long long s0 = seg * seg_len;
                const unsigned long long s1 = min(s0 + seg_len, (unsigned long long)n);
                const unsigned long long size = s1 > s0 ? s1 - s0 : 0ull;
                const unsigned long long want = need - got;
                unsigned long long base = 0;
                if (__lane_id() == 0) base = atomicAdd(&a.counters[kSegCounter0 + seg * kSegCounterStride], want);
                base = __shfl(base, 0);
                const unsigned long long take = base < size ? min(want, size - base) : 0ull;
                if (rank >= got && rank < got + take) mine = s0 + base + (rank - got);
                got += take;
                if (base + want >= size) {
                    seg = (seg + 1) % kTraceSegs;
                    ++spent;
                }
            }
            if (spent >= (unsigned)kTraceSegs) exhausted = true;
            if (!active) {
                const int64_t my = (int64_t)mine;
                if (mine < (unsigned long long)n) {
                    ray_id = (uint32_t)my;
                    any_ray = ANY || (DUAL && my >= n1);
                    const float4* rp = reinterpret_cast<const float4*>(DUAL && my >= n1 ? a.rays2 + (my - n1) : a.rays + my);
                    float4 o4 = rp[0], d4 = rp[1];
                    ray_setup(r, mk(o4.x, o4.y, o4.z), mk(d4.x, d4.y, d4.z), d4.w);
                    cur = a.rooted ? __float_as_uint(o4.w) : 0u;
                    sp = 0;
                    base = 0;
                    stash = kNoNode;
                    tk = 0;
                    best_prim = -1;
                    best_t = best_u = best_v = 0.f;
                    n_nodes = n_tris = 0;
                    in_blas = false;
                    world_stale = false;
                    inst_id = best_inst = -1;
                    active = true;
                }
            }
        }
        if (!__any(active)) break;

        // inner nodes; a lane that meets a leaf parks it and keeps descending
        // until kSpecNum / kSpecDen of the lanes in the loop hold a parked leaf
        // (a quarter measured best with the leaf loop below and quantised
        // nodes: 1/8, 3/16, 3/8 and 1/2 were 2-5% slower on incoherent rays,
        // three quarters 9% slower with float nodes; the instance walk was 14%
        // faster at 3/8 than at 1/4, the path tracer's dual launches 1.5%
        // (and as fast as at 1/2), shadow rays 2% (1/8 6% slower))
        while (can_step()) {
            if (INST && in_blas && cur == kReturn) {
                // back in the top-level tree; the world-space set-up waits for
                // its next node (the next instance entry, often the next entry
                // on the stack, sets up its own ray from the stored one)
                in_blas = false;
                world_stale = true;
                cur = pop();
                continue;
            }
            if (INST && is_leaf(cur)) {
                // a top-level leaf: enter its first instance (the rest wait on the stack)
                const uint32_t first = leaf_first(cur), cnt = leaf_count(cur);
                if (cnt == 0) {
                    cur = pop();
                    continue;
                }
                if (cnt > 1) push1(make_leaf(first + 1, cnt - 1));
                const float4* rp = reinterpret_cast<const float4*>(a.rays + ray_id);
                const float4 o4 = rp[0], d4 = rp[1];
                const float4* ip = a.inst + 4 * (int64_t)first;
                const float4 m0 = ip[0], m1 = ip[1], m2 = ip[2];
                const int4 meta = reinterpret_cast<const int4*>(ip)[3];
                const f3 oo = mk(fmaf(m0.x, o4.x, fmaf(m0.y, o4.y, fmaf(m0.z, o4.z, m0.w))),
                                 fmaf(m1.x, o4.x, fmaf(m1.y, o4.y, fmaf(m1.z, o4.z, m1.w))),
                                 fmaf(m2.x, o4.x, fmaf(m2.y, o4.y, fmaf(m2.z, o4.z, m2.w))));
                const f3 od = mk(fmaf(m0.x, d4.x, fmaf(m0.y, d4.y, m0.z * d4.z)),
                                 fmaf(m1.x, d4.x, fmaf(m1.y, d4.y, m1.z * d4.z)),
                                 fmaf(m2.x, d4.x, fmaf(m2.y, d4.y, m2.z * d4.z)));
                const float thi = r.thi;
                ray_setup(r, oo, od, d4.w);
                r.thi = thi;
                push1(kReturn);
                in_blas = true;
                inst_id = meta.y;
                cur = (uint32_t)meta.x;
                continue;
            }
            if (INST && !in_blas && world_stale) {
                // the world-space ray, keeping the window its hits shrank
                const float4* rp = reinterpret_cast<const float4*>(a.rays + ray_id);
                const float4 o4 = rp[0], d4 = rp[1];
                const float thi = r.thi;
                ray_setup(r, mk(o4.x, o4.y, o4.z), mk(d4.x, d4.y, d4.z), d4.w);
                r.thi = fminf(thi, r.thi);
                world_stale = false;
            }
            if (STATS) ++n_nodes;
            uint32_t key[W];
            node_keys<KIND>(INST && !in_blas ? a.tlas : a.nodes, cur, r, key, stage);
            // nearest child first for closest hits; any-hit launches keep the
            // children in slot order (6% more shadow rays per second: no sort,
            // and occluders are found after 6.9 triangle tests per ray instead
            // of 7.8; farthest-first took 5.5 tests but 10% more nodes and
            // was 12% slower)
            if (!ANY) sort_keys<W>(key);
            // make room for W - 1 pushes: the oldest ring entries go to the
            // spill area (rare: stacks deeper than kKeep)
            while (sp - base > kKeep<W>) {
                if (base < a.spill_depth)
                    a.spill[(int64_t)base * a.grid_threads + gtid] = ring[(base & (kLds - 1)) * kBlock];
                else
                    atomicOr(&a.counters[2], 1ull);  // stack bound exceeded (never expected)
                ++base;
            }
            uint32_t cw[W];
#pragma unroll
            for (int k = 0; k < W; ++k)  // stage byte offset | slot * 4 (the offset's low bits are clear)
                cw[k] = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(lds) +
                                                           (stage_off | (key[k] & ((W - 1) << 2))));
            // farthest first; a missed child (sorted last, or anywhere in an
            // any-hit launch) is written at the top and left there for the
            // next write to overwrite
#pragma unroll
            for (int k = W - 1; k >= 1; --k) {
                ring[(sp & (kLds - 1)) * kBlock] = cw[k];
                sp += key[k] < kMissKey ? 1 : 0;
            }
            cur = key[0] < kMissKey ? cw[0] : pop();
            if ((!INST || in_blas) && cur != kNoNode && is_leaf(cur) && stash == kNoNode) {
                stash = cur;
                cur = pop();
            }
            if (__popcll(__ballot(stash != kNoNode)) * kSpecDen<ANY, INST, DUAL> >= __popcll(__ballot(true)) * kSpecNum<ANY, INST, DUAL>) break;
        }
        // parked leaves, one triangle per lane per iteration: a lane moves on
        // to the next leaf on its stack as soon as its current one is done, so
        // lanes with short leaves do not wait for the wave's longest one. The
        // wave goes back to descending once fewer than kLeafNum / kLeafDen of
        // its active lanes hold triangles while some lane has an inner node
        // to visit (a lane keeps its parked leaf and its place in it).
        for (;;) {
            const bool has = active && stash != kNoNode;
            const unsigned long long hm = __ballot(has);
            if (hm == 0ull) break;
            const unsigned long long dm = __ballot(can_step());
            if (dm != 0ull && __popcll(hm) * kLeafDen < __popcll(__ballot(active)) * kLeafNum) break;
            if (has) {
                const float4* tp = a.tris + 3 * (int64_t)(leaf_first(stash) + tk);
                float4 A = tp[0], B = tp[1], C = tp[2];
                if (STATS) ++n_tris;
                float t, u, v;
                bool stop = false;
                if (tri_watertight(r, mk(A.x, A.y, A.z), mk(B.x, B.y, B.z), mk(C.x, C.y, C.z), t, u, v)) {
                    best_t = t;
                    best_u = u;
                    best_v = v;
                    best_prim = __float_as_int(A.w);
                    best_inst = inst_id;
                    r.thi = tmax_window(t);
                    if (ANY || (DUAL && any_ray)) {
                        stash = kNoNode;
                        tk = 0;
                        finish();
                        stop = true;
                    }
                }
                if (!stop && ++tk >= leaf_count(stash)) {
                    tk = 0;
                    if (cur != kNoNode && is_leaf(cur)) {
                        stash = cur;
                        cur = pop();
                    } else {
                        stash = kNoNode;
                    }
                }
            }
        }
        if (active && cur == kNoNode && stash == kNoNode) finish();
    }

    // the last wave of the grid resets the fetch counters for the next launch
    if (__lane_id() == 0) {
        unsigned long long prev = atomicAdd(&a.counters[1], 1ull);
        if (prev == part - 1ull) {
            for (int k = 0; k < kTraceSegs; ++k) atomicExch(&a.counters[kSegCounter0 + k * kSegCounterStride], 0ull);
            atomicExch(&a.counters[1], 0ull);
        }
    }
}

// The second queue of a dual launch (any hits).
struct AnyQueue {
    const ilm_ray* rays;
    int64_t capacity;
    const unsigned long long* n_dev;
    uint8_t* occ;
};

// ---------------------------------------------------------------- fixed scratch
// The entry point that allocates nothing (ilm_trace_closest_noalloc) keeps
// its ray-fetch counters and traversal-stack spill area in this array: the
// first kScratchCounterBytes hold the counters (same layout as
// ilm_ctx::counters), the rest the spill area. Each device holds its own copy
// for the life of the process; its device address is taken once per device
// at context creation (noalloc_setup), and host code only ever passes that
// address, never the array's name.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr size_t kScratchCounterBytes = 8192;
static_assert(kCounterWords * sizeof(unsigned long long) <= kScratchCounterBytes, "counter block too small");
static __device__ __align__(256) uint32_t scratch[kScratchBytes / sizeof(uint32_t)];

constexpr int kMaxDevices = 64;
static void* g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

__global__ void k_zero_u64(unsigned long long* p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

int noalloc_setup(int device, hipStream_t stream) {
    if (device < 0 || device >= kMaxDevices) return fail(ILM_ERR_INVALID, "device index beyond the scratch table");
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch[device]) return ILM_OK;
    void* p = nullptr;
    ILM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(ILM_ERR_HIP, "no device address for the fixed scratch");
    // the fetch counters must start at zero (each launch's last wave resets them)
    hipLaunchKernelGGL(k_zero_u64, dim3((kCounterWords + 255) / 256), dim3(256), 0, stream,
                       reinterpret_cast<unsigned long long*>(p), (int64_t)kCounterWords);
    ILM_HIP(hipGetLastError());
    ILM_HIP(hipStreamSynchronize(stream));
    g_scratch[device] = p;
    return ILM_OK;
}

template <bool ANY, bool STATS, bool DUAL = false>
static int launch_trace(ilm_mesh* m, const ilm_ray* rays, int64_t n, ilm_hit* hits, uint8_t* occ, int32_t* stats,
                        void* stream, const unsigned long long* n_dev = nullptr, bool rooted = false,
                        int refill = 0, const AnyQueue* q2 = nullptr, bool fixed_scratch = false) {
    static_assert(!DUAL || (!ANY && !STATS), "dual launches trace closest hits plus a queue of any hits");
    if (!m) return fail(ILM_ERR_INVALID, "null mesh");
    if (!m->built) return fail(ILM_ERR_STATE, "BVH not built");
    if (!m->ctx) return fail(ILM_ERR_STATE, "mesh was created without a device context");
    if (n < 0 || (n > 0 && !rays)) return fail(ILM_ERR_INVALID, "bad ray buffer");
    if (DUAL && (!q2 || q2->capacity < 0 || (q2->capacity > 0 && (!q2->rays || !q2->occ))))
        return fail(ILM_ERR_INVALID, "bad second queue");
    const int64_t total = n + (DUAL ? q2->capacity : 0);
    if (total == 0) return ILM_OK;
    if (total > (int64_t)UINT32_MAX) return fail(ILM_ERR_INVALID, "a launch traces fewer than 2^32 rays");
    if (!ANY && !hits) return fail(ILM_ERR_INVALID, "null hits");
    if (ANY && !occ) return fail(ILM_ERR_INVALID, "null occlusion output");
    if (STATS && !stats) return fail(ILM_ERR_INVALID, "null stats");
    ilm_ctx* ctx = m->ctx;
    ILM_HIP(hipSetDevice(ctx->device));
    const int kind = m->node_format == ILM_NODES_QUANTIZED ? kQuant4
                     : m->width == 2 ? kFloat2 : m->width == 8 ? kFloat8 : kFloat4;
    if (m->node_format == ILM_NODES_QUANTIZED && m->width != 4) return fail(ILM_ERR_STATE, "quantised nodes are BVH4");
    if (m->width != 2 && m->width != 4 && m->width != 8) return fail(ILM_ERR_STATE, "BVH width must be 2, 4 or 8");
    const void* fn = kind == kQuant4   ? (const void*)k_trace4<ANY, STATS, kQuant4, DUAL>
                     : kind == kFloat2 ? (const void*)k_trace4<ANY, STATS, kFloat2, DUAL>
                     : kind == kFloat8 ? (const void*)k_trace4<ANY, STATS, kFloat8, DUAL>
                                       : (const void*)k_trace4<ANY, STATS, kFloat4, DUAL>;
    int per_cu = 0;
    ILM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0));
    per_cu = std::max(per_cu, 1);
    int64_t blocks = (int64_t)per_cu * ctx->cu_count;
    blocks = std::min<int64_t>(blocks, (total + kBlock - 1) / kBlock);
    blocks = std::max<int64_t>(blocks, 1);
    TraceArgs a;
    a.nodes = reinterpret_cast<const float4*>(m->d_nodes);
    a.tris = reinterpret_cast<const float4*>(m->d_tris);
    a.rays = rays;
    a.n = n;
    a.n_dev = n_dev;
    a.hits = hits;
    a.occ = occ;
    a.stats = stats;
    a.grid_threads = blocks * kBlock;
    a.rooted = rooted ? 1 : 0;
    a.refill_min = refill > 0 ? refill : ctx->trace_refill;
    if (fixed_scratch) {
        // the spill area holds the positions below the ring's: a stack of at
        // most stack_need entries keeps at least kKeep<W> of them in LDS, so
        // it spills at most its stack_need - kKeep<W> oldest
        void* base = ctx->device < kMaxDevices ? g_scratch[ctx->device] : nullptr;
        if (!base) return fail(ILM_ERR_STATE, "fixed scratch not set up on this device");
        const int keep = m->width == 2 ? kKeep<2> : m->width == 8 ? kKeep<8> : kKeep<4>;
        a.spill_depth = std::max(1, m->info.stack_need + 1 - keep);
        const int64_t per_block = (int64_t)a.spill_depth * kBlock * (int64_t)sizeof(uint32_t);
        const int64_t fit = (int64_t)(kScratchBytes - kScratchCounterBytes) / per_block;
        if (fit < std::min<int64_t>(blocks, ctx->cu_count))
            return fail(ILM_ERR_INVALID, "tree too deep for the fixed scratch (stack need " +
                                             std::to_string(m->info.stack_need) + ", 64 MiB scratch)");
        blocks = std::min<int64_t>(blocks, fit);
        a.grid_threads = blocks * kBlock;
        a.counters = reinterpret_cast<unsigned long long*>(base);
        a.spill = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(base) + kScratchCounterBytes);
    } else {
        a.spill_depth = std::max(1, m->info.stack_need + 1);
        int rc = ensure_spill(ctx, (size_t)a.spill_depth * (size_t)a.grid_threads);
        if (rc) return rc;
        a.spill = ctx->spill;
        a.counters = ctx->counters;
    }
    a.total_waves = (uint32_t)(blocks * (kBlock / 64));
    a.rays2 = DUAL ? q2->rays : nullptr;
    a.n2 = DUAL ? q2->capacity : 0;
    a.n2_dev = DUAL ? q2->n_dev : nullptr;
    a.occ2 = DUAL ? q2->occ : nullptr;
    hipStream_t st = pick_stream(ctx, stream);
    if (kind == kQuant4)
        hipLaunchKernelGGL((k_trace4<ANY, STATS, kQuant4, DUAL>), dim3((unsigned)blocks), dim3(kBlock), 0, st, a);
    else if (kind == kFloat2)
        hipLaunchKernelGGL((k_trace4<ANY, STATS, kFloat2, DUAL>), dim3((unsigned)blocks), dim3(kBlock), 0, st, a);
    else if (kind == kFloat8)
        hipLaunchKernelGGL((k_trace4<ANY, STATS, kFloat8, DUAL>), dim3((unsigned)blocks), dim3(kBlock), 0, st, a);
    else
        hipLaunchKernelGGL((k_trace4<ANY, STATS, kFloat4, DUAL>), dim3((unsigned)blocks), dim3(kBlock), 0, st, a);
    ILM_HIP(hipGetLastError());
    return ILM_OK;
}

// Two-level traversal of an instance table (k_trace4 in instance mode): the
// persistent grid, XCD-aware fetch and stack machinery of the flat kernel.
int launch_trace_inst(ilm_ctx* ctx, const void* tlas, const void* blas, const float* tris, const float4* inst,
                      int stack_need, const ilm_ray* rays, int64_t n, ilm_hit* hits, int32_t* inst_out, uint8_t* occ,
                      bool any, void* stream) {
    if (!ctx) return fail(ILM_ERR_STATE, "instance table without a device context");
    if (n < 0 || (n > 0 && !rays)) return fail(ILM_ERR_INVALID, "bad ray buffer");
    if (n == 0) return ILM_OK;
    if (n > (int64_t)UINT32_MAX) return fail(ILM_ERR_INVALID, "a launch traces fewer than 2^32 rays");
    if (any ? !occ : (!hits || !inst_out)) return fail(ILM_ERR_INVALID, "null output");
    ILM_HIP(hipSetDevice(ctx->device));
    const void* fn = any ? (const void*)k_trace4<true, false, kFloat4, false, true>
                         : (const void*)k_trace4<false, false, kFloat4, false, true>;
    int per_cu = 0;
    ILM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kBlock, 0));
    per_cu = std::max(per_cu, 1);
    int64_t blocks = std::min<int64_t>((int64_t)per_cu * ctx->cu_count, (n + kBlock - 1) / kBlock);
    blocks = std::max<int64_t>(blocks, 1);
    TraceArgs a{};
    a.nodes = reinterpret_cast<const float4*>(blas);
    a.tris = reinterpret_cast<const float4*>(tris);
    a.rays = rays;
    a.n = n;
    a.hits = hits;
    a.occ = occ;
    a.grid_threads = blocks * kBlock;
    a.refill_min = ctx->trace_refill;
    a.spill_depth = std::max(1, stack_need + 1);
    int rc = ensure_spill(ctx, (size_t)a.spill_depth * (size_t)a.grid_threads);
    if (rc) return rc;
    a.spill = ctx->spill;
    a.counters = ctx->counters;
    a.total_waves = (uint32_t)(blocks * (kBlock / 64));
    a.tlas = reinterpret_cast<const float4*>(tlas);
    a.inst = inst;
    a.inst_out = inst_out;
    hipStream_t st = pick_stream(ctx, stream);
    if (any)
        hipLaunchKernelGGL((k_trace4<true, false, kFloat4, false, true>), dim3((unsigned)blocks), dim3(kBlock), 0, st, a);
    else
        hipLaunchKernelGGL((k_trace4<false, false, kFloat4, false, true>), dim3((unsigned)blocks), dim3(kBlock), 0, st, a);
    ILM_HIP(hipGetLastError());
    return ILM_OK;
}

// ---------------------------------------------------------------- ray generation

struct CamConsts {
    f3 centre, corner, across, up;
};

// PerspectiveCamera frame (Camera/PerspectiveCamera.cpp:15-25, Camera.cpp:30-44,
// Geometry/Basis.cpp:28-44) with the factory's distance 1 (CameraFactories.h:46-49).
static CamConsts camera_consts(const ilm_camera& c) {
    f3 pos = mk(c.pos[0], c.pos[1], c.pos[2]), tgt = mk(c.target[0], c.target[1], c.target[2]);
    f3 upv = mk(c.up[0], c.up[1], c.up[2]);
    f3 W = normalize(normalize(tgt - pos));
    f3 V = normalize(upv);
    V = V - dot(V, W) * W;
    if (dot(V, V) <= FLT_EPSILON) {
        V = mk(W.z, W.x, W.y);
        V = normalize(cross(V, W));
        V = normalize(V - dot(V, W) * W);
    }
    f3 U = cross(V, W);
    float deg2rad = (float)atan(1.0) / 45.0f;
    float half = c.fov_deg * deg2rad * 0.5f;
    float v1 = (float)tan((double)half), v0 = -v1;
    float u0 = v0 * c.aspect, u1 = v1 * c.aspect;
    CamConsts k;
    k.centre = pos;
    k.corner = pos + u0 * U + v0 * V + 1.0f * W;
    k.across = (u1 - u0) * U;
    k.up = (v1 - v0) * V;
    return k;
}

void camera_frame(const ilm_camera& c, float out[12]) {
    CamConsts k = camera_consts(c);
    const f3 v[4] = {k.centre, k.corner, k.across, k.up};
    for (int i = 0; i < 4; ++i) {
        out[3 * i] = v[i].x;
        out[3 * i + 1] = v[i].y;
        out[3 * i + 2] = v[i].z;
    }
}

__device__ __forceinline__ void camera_ray(const CamConsts& c, int W, int H, int x, int y, float u, float v,
                                           ilm_ray& out) {
    float rcpW = 1.f / (float)W, rcpH = 1.f / (float)H;
    float nx = (float)x * rcpW, ny = (float)y * rcpH;
    float xi1 = (u - 0.5f) * rcpW, xi2 = (v - 0.5f) * rcpH;
    f3 t = c.corner + c.across * (nx + xi1) + c.up * (ny + xi2);
    f3 d = normalize(t - c.centre);
    out.ox = c.centre.x; out.oy = c.centre.y; out.oz = c.centre.z; out.tmin = 0.0f;
    out.dx = d.x; out.dy = d.y; out.dz = d.z; out.tmax = FLT_MAX;
}

__global__ void k_camera(CamConsts c, int W, int H, int jitter, uint32_t seed, uint32_t sample, ilm_ray* out) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)W * H) return;
    int x = (int)(p % W), y = (int)(p / W);
    float u = 0.5f, v = 0.5f;
    if (jitter) {
        Block8 b = sample_block(seed, (uint64_t)p, sample, 0, 0);
        u = b.v[0];
        v = b.v[1];
    }
    ilm_ray r;
    camera_ray(c, W, H, x, y, u, v, r);
    out[p] = r;
}

// out[i] = camera ray through pixel pixels[i % npix]; the first npix rays use the
// pixel's own sample position, later ones counter-based jitter (sample i / npix).
__global__ void k_camera_list(CamConsts c, int W, int H, const int32_t* pixels, int64_t npix, int64_t n,
                              uint32_t seed, ilm_ray* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t p = pixels[i % npix];
    uint32_t s = (uint32_t)(i / npix);
    float u = 0.5f, v = 0.5f;
    if (s > 0) {
        Block8 b = sample_block(seed, (uint64_t)p, s, 0, 0);
        u = b.v[0];
        v = b.v[1];
    }
    ilm_ray r;
    camera_ray(c, W, H, (int)(p % W), (int)(p / W), u, v, r);
    out[i] = r;
}

struct SpawnArgs {
    const ilm_ray* rays;
    const ilm_hit* hits;
    const int32_t* src;
    int64_t n;
    const float4* tris;
    const int32_t* slot_of;  // original triangle index -> triangle slot
    int kind;
    f3 lc, lu, lv;
    float eps;
    uint32_t seed;
    ilm_ray* out;
};

__global__ void k_spawn(SpawnArgs a) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    int64_t s = a.src ? (int64_t)a.src[i] : i;
    ilm_hit h = a.hits[s];
    ilm_ray r = a.rays[s];
    ilm_ray o;
    if (h.prim < 0) {  // no hit: a ray that cannot hit anything
        o.ox = r.ox; o.oy = r.oy; o.oz = r.oz; o.tmin = 0.0f;
        o.dx = r.dx; o.dy = r.dy; o.dz = r.dz; o.tmax = -1.0f;
        a.out[i] = o;
        return;
    }
    f3 ro = mk(r.ox, r.oy, r.oz), rd = mk(r.dx, r.dy, r.dz);
    f3 P = ro + h.t * rd;  // Ray::PointAlongRay (Geometry/Ray.inl:32-34)
    Block8 b = sample_block(a.seed, (uint64_t)i, 0, 1, (uint32_t)a.kind);
    if (a.kind == ILM_SPAWN_SHADOW) {
        f3 Q = a.lc + b.v[0] * a.lu + b.v[1] * a.lv;
        f3 d = Q - P;
        o.ox = P.x; o.oy = P.y; o.oz = P.z; o.tmin = 1e-4f;
        o.dx = d.x; o.dy = d.y; o.dz = d.z; o.tmax = 1.0f - 1e-4f;
    } else {
        int32_t slot = a.slot_of[h.prim];
        const float4* tp = a.tris + 3 * (int64_t)slot;
        float4 A = tp[0], B = tp[1], C = tp[2];
        f3 ng = normalize(cross(mk(B.x - A.x, B.y - A.y, B.z - A.z), mk(C.x - A.x, C.y - A.y, C.z - A.z)));
        if (dot(ng, rd) > 0.0f) ng = -ng;  // face the incoming ray
        f3 org = P + ng * a.eps;
        Basis bs = basis_from_w(ng);
        f3 d = to_world(bs, cosine_hemisphere(b.v[0], b.v[1]));
        o.ox = org.x; o.oy = org.y; o.oz = org.z; o.tmin = a.eps;
        o.dx = d.x; o.dy = d.y; o.dz = d.z; o.tmax = FLT_MAX;
    }
    a.out[i] = o;
}

// ---------------------------------------------------------------- compaction
// Deterministic three-pass compaction of valid hits: per-block counts with
// wave ballots, one scan block, then ordered writes.
constexpr int kCompactItems = 4;  // items per thread per block pass

__global__ void k_compact_count(const ilm_hit* hits, int64_t n, int32_t* block_counts) {
    __shared__ int wave_sums[kBlock / 64];
    int64_t base = (int64_t)blockIdx.x * kBlock * kCompactItems;
    int cnt = 0;
    for (int k = 0; k < kCompactItems; ++k) {
        int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
        bool ok = i < n && hits[i].prim >= 0;
        cnt += __popcll(__ballot(ok));
    }
    if (__lane_id() == 0) wave_sums[threadIdx.x / 64] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int s = 0;
        for (int w = 0; w < kBlock / 64; ++w) s += wave_sums[w];
        block_counts[blockIdx.x] = s;
    }
}

__global__ void k_compact_scan(int32_t* block_counts, int64_t nblocks, int64_t* total) {
    // single block: each thread scans a contiguous chunk, then the chunk sums
    __shared__ int64_t part[kBlock];
    int64_t per = (nblocks + kBlock - 1) / kBlock;
    int64_t s0 = threadIdx.x * per, s1 = std::min<int64_t>(nblocks, s0 + per);
    int64_t sum = 0;
    for (int64_t i = s0; i < s1; ++i) sum += block_counts[i];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = 0;
        for (int t = 0; t < kBlock; ++t) { int64_t v = part[t]; part[t] = run; run += v; }
        *total = run;
    }
    __syncthreads();
    int64_t run = part[threadIdx.x];
    for (int64_t i = s0; i < s1; ++i) { int32_t v = block_counts[i]; block_counts[i] = (int32_t)run; run += v; }
}

__global__ void k_compact_write(const ilm_hit* hits, int64_t n, const int32_t* block_offsets, int32_t* out) {
    __shared__ int cnt[kCompactItems][kBlock / 64];
    const int wave = threadIdx.x / 64;
    int64_t base = (int64_t)blockIdx.x * kBlock * kCompactItems;
    unsigned long long masks[kCompactItems];
    bool oks[kCompactItems];
    for (int k = 0; k < kCompactItems; ++k) {
        int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
        oks[k] = i < n && hits[i].prim >= 0;
        masks[k] = __ballot(oks[k]);
        if (__lane_id() == 0) cnt[k][wave] = __popcll(masks[k]);
    }
    __syncthreads();
    // slots follow the index order: item k, then wave, then lane
    int off = block_offsets[blockIdx.x];
    for (int k = 0; k < kCompactItems; ++k) {
        int b = off;
        for (int w = 0; w < wave; ++w) b += cnt[k][w];
        if (oks[k]) out[b + __popcll(masks[k] & __lanemask_lt())] = (int32_t)(base + (int64_t)k * kBlock + threadIdx.x);
        for (int w = 0; w < kBlock / 64; ++w) off += cnt[k][w];
    }
}

// Plain streaming copy (16 bytes per lane per iteration, grid-stride): the
// bandwidth probe used to re-measure HBM throughput and to calibrate the
// byte counters of the profiler.
__global__ void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n16) {
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) dst[i] = src[i];
}

// Cache-bandwidth probe: every workgroup streams the whole (small) buffer
// `passes` times with 16-byte loads, starting 64 KiB after the previous
// workgroup and wrapping around. A buffer that fits each XCD's L2 (4 MiB)
// but not a CU's L1 (32 KiB) measures the aggregate L2-to-CU read bandwidth.
// The loaded words are folded into a value that is stored only if it equals
// an impossible pattern, so the loads cannot be dropped.
__global__ void __launch_bounds__(256) k_read_probe(const uint4* __restrict__ src, int64_t n16, int passes,
                                                    uint4* __restrict__ sink) {
    uint4 acc = make_uint4(0u, 0u, 0u, 0u);
    const int64_t start = ((int64_t)blockIdx.x * 4096) % n16;
    for (int p = 0; p < passes; ++p) {
#pragma unroll 8
        for (int64_t i = threadIdx.x; i < n16; i += blockDim.x) {
            int64_t j = start + i;
            if (j >= n16) j -= n16;
            const uint4 v = src[j];
            acc.x ^= v.x; acc.y += v.y; acc.z ^= v.z; acc.w += v.w;
        }
    }
    if (acc.x == 0x9E3779B9u && acc.y == 0x7F4A7C15u && acc.z == 0xF39CC060u && acc.w == 0x5CEDC834u) sink[0] = acc;
}

}  // namespace ilm

namespace ilm {
int trace_queue(ilm_mesh* m, const ilm_ray* rays, const unsigned long long* n_dev, int64_t capacity, ilm_hit* hits,
                uint8_t* occ, hipStream_t stream, bool rooted, int refill) {
    if (occ) return launch_trace<true, false>(m, rays, capacity, nullptr, occ, nullptr, stream, n_dev, rooted, refill);
    return launch_trace<false, false>(m, rays, capacity, hits, nullptr, nullptr, stream, n_dev, rooted, refill);
}

int trace_queue_dual(ilm_mesh* m, const ilm_ray* rays, const unsigned long long* n_dev, int64_t capacity,
                     ilm_hit* hits, const ilm_ray* rays2, const unsigned long long* n2_dev, int64_t capacity2,
                     uint8_t* occ2, hipStream_t stream, bool rooted, int refill) {
    const AnyQueue q2{rays2, capacity2, n2_dev, occ2};
    return launch_trace<false, false, true>(m, rays, capacity, hits, nullptr, nullptr, stream, n_dev, rooted, refill,
                                            &q2);
}
}  // namespace ilm

using namespace ilm;

extern "C" {

int ilm_ctx_set_trace_refill(ilm_ctx* ctx, int idle_lanes) {
    if (!ctx) return fail(ILM_ERR_INVALID, "null ctx");
    if (idle_lanes < 1 || idle_lanes > 64) return fail(ILM_ERR_INVALID, "idle_lanes must be in 1..64");
    ctx->trace_refill = idle_lanes;
    return ILM_OK;
}

int ilm_trace_closest(ilm_mesh* mesh, const ilm_ray* rays, int64_t n, ilm_hit* hits, void* stream) {
    return launch_trace<false, false>(mesh, rays, n, hits, nullptr, nullptr, stream);
}

int ilm_trace_any(ilm_mesh* mesh, const ilm_ray* rays, int64_t n, uint8_t* occluded, void* stream) {
    return launch_trace<true, false>(mesh, rays, n, nullptr, occluded, nullptr, stream);
}

int ilm_trace_closest_noalloc(ilm_mesh* mesh, const ilm_ray* rays, int64_t n, ilm_hit* hits, void* stream) {
    return launch_trace<false, false>(mesh, rays, n, hits, nullptr, nullptr, stream, nullptr, false, 0, nullptr,
                                      true);
}

int ilm_trace_stats(ilm_mesh* mesh, const ilm_ray* rays, int64_t n, int any_hit, int32_t* visits, void* stream) {
    if (!mesh) return fail(ILM_ERR_INVALID, "null mesh");
    if (any_hit) {
        uint8_t* occ = nullptr;
        int rc = ilm_alloc(mesh->ctx, (size_t)std::max<int64_t>(n, 1), (void**)&occ);
        if (rc) return rc;
        rc = launch_trace<true, true>(mesh, rays, n, nullptr, occ, visits, stream);
        int rc2 = ilm_free(mesh->ctx, occ);
        return rc ? rc : rc2;
    }
    ilm_hit* h = nullptr;
    int rc = ilm_alloc(mesh->ctx, (size_t)std::max<int64_t>(n, 1) * sizeof(ilm_hit), (void**)&h);
    if (rc) return rc;
    rc = launch_trace<false, true>(mesh, rays, n, h, nullptr, visits, stream);
    int rc2 = ilm_free(mesh->ctx, h);
    return rc ? rc : rc2;
}

int ilm_raygen_camera(ilm_ctx* ctx, const ilm_camera* cam, int width, int height, int jitter, uint32_t seed,
                      uint32_t sample_index, ilm_ray* rays, void* stream) {
    if (!ctx || !cam || !rays || width <= 0 || height <= 0) return fail(ILM_ERR_INVALID, "bad camera arguments");
    CamConsts c = camera_consts(*cam);
    int64_t n = (int64_t)width * height;
    hipLaunchKernelGGL(k_camera, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, pick_stream(ctx, stream), c, width,
                       height, jitter, seed, sample_index, rays);
    ILM_HIP(hipGetLastError());
    return ILM_OK;
}

int ilm_raygen_camera_list(ilm_ctx* ctx, const ilm_camera* cam, int width, int height, const int32_t* pixels,
                           int64_t npix, int64_t n, uint32_t seed, ilm_ray* rays, void* stream) {
    if (!ctx || !cam || !rays || !pixels || npix <= 0 || n < 0) return fail(ILM_ERR_INVALID, "bad arguments");
    if (n == 0) return ILM_OK;
    CamConsts c = camera_consts(*cam);
    hipLaunchKernelGGL(k_camera_list, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, pick_stream(ctx, stream), c,
                       width, height, pixels, npix, n, seed, rays);
    ILM_HIP(hipGetLastError());
    return ILM_OK;
}

int ilm_raygen_secondary(ilm_mesh* mesh, const ilm_ray* rays_in, const ilm_hit* hits, const int32_t* src, int64_t n,
                         int kind, const float* light, float eps, uint32_t seed, ilm_ray* rays_out, void* stream) {
    if (!mesh || !rays_in || !hits || !rays_out || n < 0) return fail(ILM_ERR_INVALID, "bad arguments");
    if (!mesh->built) return fail(ILM_ERR_STATE, "BVH not built");
    if (!mesh->ctx) return fail(ILM_ERR_STATE, "mesh was created without a device context");
    if (kind != ILM_SPAWN_SHADOW && kind != ILM_SPAWN_COSINE) return fail(ILM_ERR_INVALID, "unknown spawn kind");
    if (kind == ILM_SPAWN_SHADOW && !light) return fail(ILM_ERR_INVALID, "shadow rays need a light");
    if (n == 0) return ILM_OK;
    ilm_ctx* ctx = mesh->ctx;
    if (!mesh->d_slot_of) {
        // kd-tree leaves may repeat a triangle: any of its slots holds it
        std::vector<int32_t> slot(mesh->nt);
        for (size_t k = 0; k < mesh->h_order.size(); ++k) slot[mesh->h_order[k]] = (int32_t)k;
        int rc = ilm_alloc(ctx, slot.size() * sizeof(int32_t), (void**)&mesh->d_slot_of);
        if (rc) return rc;
        rc = ilm_upload(ctx, mesh->d_slot_of, slot.data(), slot.size() * sizeof(int32_t));
        if (rc) return rc;
    }
    SpawnArgs a;
    a.rays = rays_in;
    a.hits = hits;
    a.src = src;
    a.n = n;
    a.tris = reinterpret_cast<const float4*>(mesh->d_tris);
    a.slot_of = mesh->d_slot_of;
    a.kind = kind;
    a.lc = light ? mk(light[0], light[1], light[2]) : mk(0, 0, 0);
    a.lu = light ? mk(light[3], light[4], light[5]) : mk(0, 0, 0);
    a.lv = light ? mk(light[6], light[7], light[8]) : mk(0, 0, 0);
    a.eps = eps;
    a.seed = seed;
    a.out = rays_out;
    hipLaunchKernelGGL(k_spawn, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, pick_stream(ctx, stream), a);
    ILM_HIP(hipGetLastError());
    return ILM_OK;
}

int ilm_device_copy(ilm_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream) {
    if (!ctx || !dst || !src || (bytes % 16) != 0) return fail(ILM_ERR_INVALID, "copy needs 16-byte multiples");
    if (bytes == 0) return ILM_OK;
    int64_t n16 = (int64_t)(bytes / 16);
    int64_t blocks = std::min<int64_t>((n16 + 255) / 256, (int64_t)ctx->cu_count * 32);
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, pick_stream(ctx, stream),
                       reinterpret_cast<const float4*>(src), reinterpret_cast<float4*>(dst), n16);
    ILM_HIP(hipGetLastError());
    return ILM_OK;
}

int ilm_read_probe(ilm_ctx* ctx, const void* src, size_t bytes, int passes, int blocks_per_cu, void* sink,
                   void* stream) {
    if (!ctx || !src || !sink || bytes < 16 || (bytes % 16) != 0 || passes <= 0 || blocks_per_cu <= 0)
        return fail(ILM_ERR_INVALID, "read probe needs a buffer of 16-byte multiples, passes and blocks");
    const int64_t blocks = (int64_t)ctx->cu_count * blocks_per_cu;
    hipLaunchKernelGGL(k_read_probe, dim3((unsigned)blocks), dim3(256), 0, pick_stream(ctx, stream),
                       reinterpret_cast<const uint4*>(src), (int64_t)(bytes / 16), passes, reinterpret_cast<uint4*>(sink));
    ILM_HIP(hipGetLastError());
    return ILM_OK;
}

int ilm_compact_hits(ilm_ctx* ctx, const ilm_hit* hits, int64_t n, int32_t* idx_out, int64_t* count_dev,
                     void* stream) {
    if (!ctx || !hits || !idx_out || !count_dev || n < 0) return fail(ILM_ERR_INVALID, "bad arguments");
    if (n >= (int64_t(1) << 31)) return fail(ILM_ERR_INVALID, "too many hits for int32 indices");
    int64_t per_block = (int64_t)kBlock * kCompactItems;
    int64_t nb = std::max<int64_t>(1, (n + per_block - 1) / per_block);
    int32_t* bc = nullptr;
    int rc = ilm_alloc(ctx, (size_t)nb * sizeof(int32_t), (void**)&bc);
    if (rc) return rc;
    hipStream_t s = pick_stream(ctx, stream);
    hipLaunchKernelGGL(k_compact_count, dim3((unsigned)nb), dim3(kBlock), 0, s, hits, n, bc);
    hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(kBlock), 0, s, bc, nb, count_dev);
    hipLaunchKernelGGL(k_compact_write, dim3((unsigned)nb), dim3(kBlock), 0, s, hits, n, bc, idx_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    int rc2 = ilm_free(ctx, bc);
    if (e != hipSuccess) return fail_hip(e, "compaction");
    return rc2;
}

}  // extern "C"
