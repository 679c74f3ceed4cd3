// Math shared by the traversal and shading kernels (gfx950). A few helpers
// (node layout, camera frame) are also used by host code that builds the
// inputs of those kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cfloat>
#include <cmath>
#include <cstring>

#define ILM_HD __host__ __device__ __forceinline__

namespace ilm {

struct f3 {
    float x, y, z;
};
ILM_HD f3 mk(float x, float y, float z) { return f3{x, y, z}; }
ILM_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
ILM_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
ILM_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
ILM_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
ILM_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
ILM_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
ILM_HD f3 mul(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
ILM_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
ILM_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
ILM_HD float length(f3 a) { return sqrtf(dot(a, a)); }
ILM_HD f3 normalize(f3 a) { float l = length(a); return f3{a.x / l, a.y / l, a.z / l}; }
ILM_HD float comp(f3 v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// ------------------------------------------------------------------ BVH layout
// Wide node, W children, SoA boxes then child words: 7 x 16 B for W = 4.
// child word: bit 31 clear -> internal node index; bit 31 set -> leaf with
// first triangle (bits 5..30) and triangle count (bits 0..4). A leaf with
// count 0 is an empty slot (its box is a far-away point).
template <int W>
struct alignas(16) WideNode {
    float lox[W], hix[W], loy[W], hiy[W], loz[W], hiz[W];
    uint32_t child[W];
};
static_assert(sizeof(WideNode<4>) == 112, "BVH4 node is 112 bytes");
static_assert(sizeof(WideNode<8>) == 224, "BVH8 node is 224 bytes");
static_assert(sizeof(WideNode<2>) == 64, "BVH2 node is 64 bytes (padded)");

constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kNoNode = 0x7FFFFFFFu;  // stack sentinel
constexpr int kLeafCountBits = 5;
constexpr int kMaxLeafTris = 31;
ILM_HD bool is_leaf(uint32_t c) { return (c & kLeafBit) != 0; }
ILM_HD uint32_t leaf_first(uint32_t c) { return (c & ~kLeafBit) >> kLeafCountBits; }
ILM_HD uint32_t leaf_count(uint32_t c) { return c & ((1u << kLeafCountBits) - 1u); }
ILM_HD uint32_t make_leaf(uint32_t first, uint32_t count) { return kLeafBit | (first << kLeafCountBits) | count; }
constexpr uint32_t kEmptySlot = kLeafBit;  // make_leaf(0, 0)

// Quantised BVH4 node, 64 bytes (four 16-byte loads, half a 128-byte line).
// Child boxes are stored as bytes relative to an origin p (the minimum corner
// of the children) in steps of 2^e per axis:
//   lo = p + qlo * 2^e,  hi = p + qhi * 2^e,
// rounded outward so that every decoded box contains its child. Bytes k of
// qlo*/qhi* belong to child k; exps holds e + 127 for x, y, z in bytes 0-2
// (the bit pattern of 2^e as a float is that byte shifted left by 23).
// Empty slots have child word kEmptySlot and are skipped by the traversal.
struct alignas(16) QNode4 {
    float px, py, pz;
    uint32_t exps;
    uint32_t child[4];
    uint32_t qlox, qhix, qloy, qhiy;
    uint32_t qloz, qhiz, pad0, pad1;
};
static_assert(sizeof(QNode4) == 64, "quantised BVH4 node is 64 bytes");

ILM_HD uint32_t float_bits(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(f);
#else
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
#endif
}

ILM_HD float pow2_from_biased(uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(b << 23);
#else
    uint32_t u = b << 23;
    float f;
    memcpy(&f, &u, 4);
    return f;
#endif
}

// Quantises one float BVH4 node. The step per axis is the smallest power of
// two with (max hi - p) / step <= 254; each byte is rounded outward and then
// checked in float (p + q * step) so that the decoded box contains the child.
ILM_HD void quantize_node(const WideNode<4>& n, QNode4& q) {
    float p[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, top[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    bool any = false;
    for (int k = 0; k < 4; ++k) {
        if (n.child[k] == kEmptySlot) continue;
        any = true;
        const float lo[3] = {n.lox[k], n.loy[k], n.loz[k]}, hi[3] = {n.hix[k], n.hiy[k], n.hiz[k]};
        for (int a = 0; a < 3; ++a) { p[a] = fminf(p[a], lo[a]); top[a] = fmaxf(top[a], hi[a]); }
    }
    if (!any)
        for (int a = 0; a < 3; ++a) p[a] = top[a] = 0.0f;
    float step[3];
    uint32_t exps = 0;
    for (int a = 0; a < 3; ++a) {
        float ext = top[a] - p[a];
        int e = -126;
        if (ext > 0.0f) {
            int x;
            frexpf(ext / 254.0f, &x);  // ext / 254 < 2^x
            e = x < -126 ? -126 : (x > 127 ? 127 : x);
        }
        step[a] = ldexpf(1.0f, e);
        exps |= (uint32_t)(e + 127) << (8 * a);
    }
    q.px = p[0]; q.py = p[1]; q.pz = p[2];
    q.exps = exps;
    q.pad0 = q.pad1 = 0;
    uint32_t ql[3] = {0, 0, 0}, qh[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        q.child[k] = n.child[k];
        if (n.child[k] == kEmptySlot) continue;
        const float lo[3] = {n.lox[k], n.loy[k], n.loz[k]}, hi[3] = {n.hix[k], n.hiy[k], n.hiz[k]};
        for (int a = 0; a < 3; ++a) {
            float fl = floorf((lo[a] - p[a]) / step[a]);
            int l = fl < 0.0f ? 0 : (fl > 255.0f ? 255 : (int)fl);
            while (l > 0 && p[a] + (float)l * step[a] > lo[a]) --l;
            float fh = ceilf((hi[a] - p[a]) / step[a]);
            int h = fh < 0.0f ? 0 : (fh > 255.0f ? 255 : (int)fh);
            while (h < 255 && p[a] + (float)h * step[a] < hi[a]) ++h;
            ql[a] |= (uint32_t)l << (8 * k);
            qh[a] |= (uint32_t)h << (8 * k);
        }
    }
    q.qlox = ql[0]; q.qhix = qh[0];
    q.qloy = ql[1]; q.qhiy = qh[1];
    q.qloz = ql[2]; q.qhiz = qh[2];
}

// ------------------------------------------------------------------ t window
// The reference accepts a triangle hit when 1e-5 <= t <= Max - 1e-5, both
// compared in double (Shape/IndexedTriangle.cpp:114-115). For float t these
// are t >= kTLo and t <= round_down((double)Max - 1e-5).
constexpr float kTLo = 1.00000008e-05f;  // smallest float >= 1e-5 (double)

ILM_HD float round_down_f(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __double2float_rd(x);
#else
    float f = (float)x;
    if ((double)f > x) f = nextafterf(f, -INFINITY);
    return f;
#endif
}
ILM_HD float tmax_window(float tmax) { return round_down_f((double)tmax - 1e-5); }

// ------------------------------------------------------------------ rays
struct RayState {
    f3 o, d;
    f3 inv;   // 1/d with zero components replaced by +-1e-18
    f3 oinv;  // o * inv
    int kp;   // axis permutation of the watertight test: kx | ky << 2 | kz << 4 (one register)
    float Sx, Sy, Sz;
    float tlo, thi;  // accepted t window (thi shrinks with closest hits)
};

ILM_HD float safe_rcp(float d) {
    const float tiny = 1e-18f;
    return 1.0f / (fabsf(d) > tiny ? d : copysignf(tiny, d));
}

ILM_HD void ray_setup(RayState& r, f3 o, f3 d, float tmax) {
    r.o = o;
    r.d = d;
    r.inv = mk(safe_rcp(d.x), safe_rcp(d.y), safe_rcp(d.z));
    r.oinv = mk(o.x * r.inv.x, o.y * r.inv.y, o.z * r.inv.z);
    // Watertight test setup (Woop, Benthin, Wald, JCGT 2013): kz = dominant axis.
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    float dkz = comp(d, kz);
    if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kp = kx | ky << 2 | kz << 4;
    r.Sx = comp(d, kx) / dkz;
    r.Sy = comp(d, ky) / dkz;
    r.Sz = 1.0f / dkz;
    r.tlo = kTLo;
    r.thi = tmax_window(tmax);
}

// ------------------------------------------------------------------ box test
// Slab test of child i; returns entry distance, or +inf on a miss. The exit
// distance is relaxed by 2^-21 relative so that rounding of the fma form
// never culls a box that the ray touches.
template <int W>
ILM_HD float slab(const WideNode<W>& n, int i, const RayState& r) {
    float tx0 = fmaf(n.lox[i], r.inv.x, -r.oinv.x), tx1 = fmaf(n.hix[i], r.inv.x, -r.oinv.x);
    float ty0 = fmaf(n.loy[i], r.inv.y, -r.oinv.y), ty1 = fmaf(n.hiy[i], r.inv.y, -r.oinv.y);
    float tz0 = fmaf(n.loz[i], r.inv.z, -r.oinv.z), tz1 = fmaf(n.hiz[i], r.inv.z, -r.oinv.z);
    float tn = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), r.tlo));
    float tf = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), r.thi));
    return (tn <= tf * 1.00000048f) ? tn : INFINITY;
}

ILM_HD void cswap(float& da, float& db, uint32_t& ca, uint32_t& cb) {
    if (db < da) {
        float td = da; da = db; db = td;
        uint32_t tc = ca; ca = cb; cb = tc;
    }
}

// Sorts children by entry distance, nearest first (comparator networks:
// 1, 5 and Batcher's 19 exchanges).
template <int W>
ILM_HD void sort_children(float (&d)[W], uint32_t (&c)[W]) {
    if (W == 2) {
        cswap(d[0], d[1], c[0], c[1]);
    } else if (W == 4) {
        cswap(d[0], d[1], c[0], c[1]); cswap(d[2], d[3], c[2], c[3]); cswap(d[0], d[2], c[0], c[2]);
        cswap(d[1], d[3], c[1], c[3]); cswap(d[1], d[2], c[1], c[2]);
    } else {
        constexpr int net[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {1, 2}, {5, 6},
                                    {0, 4}, {3, 7}, {1, 5}, {2, 6}, {1, 4}, {3, 6}, {2, 4}, {3, 5}, {3, 4}};
#pragma unroll
        for (int k = 0; k < 19; ++k) cswap(d[net[k][0]], d[net[k][1]], c[net[k][0]], c[net[k][1]]);
    }
}

// Child keys of the ordered traversal: a child's entry distance as bits
// (non-negative floats order as unsigned integers) with its slot in the low
// bits, as slot * 4 (the byte offset of its child word); the distance loses
// its lowest mantissa bits (2^-19 relative), which only reorders children
// whose entry distances nearly tie, and equal keys cannot occur. Sorting the
// keys as integers orders the children nearest first, ties by slot; a missed
// child (distance +inf) has a key of at least kMissKey and sorts last.
constexpr uint32_t kMissKey = 0x7F800000u;
template <int W> constexpr uint32_t kSlotMask = W == 8 ? 31u : W == 4 ? 15u : 7u;
template <int W>
ILM_HD uint32_t child_key(float d, int slot) {
    return (float_bits(d) & ~kSlotMask<W>) | (uint32_t)(slot << 2);
}
template <int W>
ILM_HD int key_slot(uint32_t key) { return (int)((key >> 2) & (uint32_t)(W - 1)); }

ILM_HD void kswap(uint32_t& a, uint32_t& b) {
    const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    a = lo;
    b = hi;
}

// Sorts child keys ascending with the exchange networks of sort_children
// (branch-free: a minimum and a maximum per exchange).
template <int W>
ILM_HD void sort_keys(uint32_t (&k)[W]) {
    if (W == 2) {
        kswap(k[0], k[1]);
    } else if (W == 4) {
        kswap(k[0], k[1]); kswap(k[2], k[3]); kswap(k[0], k[2]); kswap(k[1], k[3]); kswap(k[1], k[2]);
    } else {
        constexpr int net[19][2] = {{0, 1}, {2, 3}, {4, 5}, {6, 7}, {0, 2}, {1, 3}, {4, 6}, {5, 7}, {1, 2}, {5, 6},
                                    {0, 4}, {3, 7}, {1, 5}, {2, 6}, {1, 4}, {3, 6}, {2, 4}, {3, 5}, {3, 4}};
#pragma unroll
        for (int e = 0; e < 19; ++e) kswap(k[net[e][0]], k[net[e][1]]);
    }
}

// ------------------------------------------------------------------ triangle test
// Watertight ray/triangle test (Woop, Benthin, Wald 2013). The edge functions
// are evaluated without fused multiply-add so that the two triangles sharing
// an edge compute exactly negated values; zero edge functions are recomputed
// in double. Accepts tlo <= t <= thi. u, v = barycentric weights of v1, v2.
ILM_HD float mul_rn(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __fmul_rn(a, b);
#else
    return a * b;
#endif
}
ILM_HD float sub_rn(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __fsub_rn(a, b);
#else
    return a - b;
#endif
}

ILM_HD bool tri_watertight(const RayState& r, f3 v0, f3 v1, f3 v2, float& t_out, float& u_out, float& v_out) {
    f3 A = v0 - r.o, B = v1 - r.o, C = v2 - r.o;
    const int kx = r.kp & 3, ky = (r.kp >> 2) & 3, kz = r.kp >> 4;
    float Akz = comp(A, kz), Bkz = comp(B, kz), Ckz = comp(C, kz);
    float Ax = fmaf(-r.Sx, Akz, comp(A, kx)), Ay = fmaf(-r.Sy, Akz, comp(A, ky));
    float Bx = fmaf(-r.Sx, Bkz, comp(B, kx)), By = fmaf(-r.Sy, Bkz, comp(B, ky));
    float Cx = fmaf(-r.Sx, Ckz, comp(C, kx)), Cy = fmaf(-r.Sy, Ckz, comp(C, ky));
    float U = sub_rn(mul_rn(Cx, By), mul_rn(Cy, Bx));
    float V = sub_rn(mul_rn(Ax, Cy), mul_rn(Ay, Cx));
    float Wf = sub_rn(mul_rn(Bx, Ay), mul_rn(By, Ax));
    if (U == 0.0f || V == 0.0f || Wf == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        U = (float)(CxBy - CyBx);
        V = (float)(AxCy - AyCx);
        Wf = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || Wf < 0.0f) && (U > 0.0f || V > 0.0f || Wf > 0.0f)) return false;
    float det = U + V + Wf;
    if (det == 0.0f) return false;
    float T = U * (r.Sz * Akz) + V * (r.Sz * Bkz) + Wf * (r.Sz * Ckz);
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= r.tlo && t <= r.thi)) return false;
    t_out = t;
    u_out = V * rcp;
    v_out = Wf * rcp;
    return true;
}

// ------------------------------------------------------------------ RNG
// Philox4x64-10 (Salmon, Moraes, Dror, Shaw, SC'11). Counter-based: the
// sample stream of a path is a pure function of (seed, pixel, sample,
// bounce, stream), so renders are reproducible and need no shared state.
ILM_HD void mulhilo64(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
#if defined(__HIP_DEVICE_COMPILE__)
    // the 128-bit product from four 32 x 32 -> 64 multiply-adds (the carries
    // of the middle terms folded into them), where a * b and __umul64hi
    // separately take six multiplies
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
    const uint64_t p00 = (uint64_t)a0 * b0;
    const uint64_t t = (uint64_t)a0 * b1 + (p00 >> 32);          // < 2^64
    const uint64_t u = (uint64_t)a1 * b0 + (uint32_t)t;          // < 2^64
    lo = ((uint64_t)(uint32_t)u << 32) | (uint32_t)p00;
    hi = (uint64_t)a1 * b1 + (t >> 32) + (u >> 32);
#else
    unsigned __int128 p = (unsigned __int128)a * b;
    hi = (uint64_t)(p >> 64);
    lo = (uint64_t)p;
#endif
}

ILM_HD void philox4x64(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0, uint64_t k1, uint64_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B97F4A7C15ull; k1 += 0xBB67AE8584CAA73Bull; }
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ull, c0, hi0, lo0);
        mulhilo64(0xCA5A826395121157ull, c2, hi1, lo1);
        uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// u32 -> [0, 1): the reference's Random::NextFloat mapping (Maths/Random.cpp:117-120).
ILM_HD float u32_to_unit(uint32_t u) {
    return ((float)((double)u * 2.3283064370807974e-10)) * (1.0f - FLT_EPSILON);
}

// 8 uniform floats for (seed, pixel, sample, bounce, stream).
struct Block8 { float v[8]; };
ILM_HD Block8 sample_block(uint32_t seed, uint64_t pixel, uint32_t sample, uint32_t bounce, uint32_t stream) {
    uint64_t o[4];
    philox4x64(pixel, ((uint64_t)bounce << 32) | sample, stream, 0, seed, 0, o);
    Block8 b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        b.v[2 * i] = u32_to_unit((uint32_t)o[i]);
        b.v[2 * i + 1] = u32_to_unit((uint32_t)(o[i] >> 32));
    }
    return b;
}

// ------------------------------------------------------------------ frames and warps
struct Basis { f3 U, V, W; };
// OrthonormalBasis::InitFromW (Geometry/Basis.cpp:85-99).
ILM_HD Basis basis_from_w(f3 w) {
    Basis b;
    b.W = normalize(w);
    if (fabsf(b.W.x) > fabsf(b.W.y)) {
        float inv = 1.f / sqrtf(b.W.x * b.W.x + b.W.z * b.W.z);
        b.V = mk(-b.W.z * inv, 0.f, b.W.x * inv);
    } else {
        float inv = 1.f / sqrtf(b.W.y * b.W.y + b.W.z * b.W.z);
        b.V = mk(0.f, b.W.z * inv, -b.W.y * inv);
    }
    b.U = cross(b.W, b.V);
    return b;
}
ILM_HD f3 to_local(const Basis& b, f3 v) { return mk(dot(b.U, v), dot(b.V, v), dot(b.W, v)); }
ILM_HD f3 to_world(const Basis& b, f3 v) { return v.x * b.U + v.y * b.V + v.z * b.W; }

constexpr float kPi = 3.14159265358979f;
constexpr float kInvPi = 0.318309886183791f;

// Montecarlo::CosineSampleHemisphere (Maths/Montecarlo.h:30-37).
ILM_HD f3 cosine_hemisphere(float u, float v) {
    float a = 2.0f * kPi * v;
    float b = sqrtf(u);
    return mk(b * cosf(a), b * sinf(a), sqrtf(1.0f - u));
}

}  // namespace ilm
