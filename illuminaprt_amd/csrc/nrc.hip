// Neural radiance cache: a small MLP trained online on radiance that the
// path tracer's own extended paths produce, with hand-written bf16 MFMA
// kernels (FP32 accumulation, FP32 master weights). No reference counterpart:
// it stands in for the irradiance-cache integrator
// (Integrator/ICIntegrator.cpp:29-668) on the far end of each path.
//
// Network: 96 inputs (83 used) -> 5 hidden layers of 64 (ReLU) -> 16 outputs
// (RGB used), no biases. Encoding of a query (16 floats: position xyz,
// direction theta/phi, normal theta/phi, roughness, kd rgb, ks rgb, 2 unused):
//   72 = sin/cos(2^i pi p_d) for 12 frequencies of each normalised coordinate,
//   then direction (2), normal (2), roughness (1), kd (3), ks (3), zeros (13).
//
// MFMA data flow (v_mfma_f32_16x16x32_bf16; 16 samples per 16x16 tile):
//   A operand (16x32): lane l holds A[l%16][8(l/16) .. 8(l/16)+7]
//   B operand (32x16): lane l holds B[8(l/16) .. +7][l%16]
//   D (16x16 f32):     lane l holds D[4(l/16) .. +3][l%16]
// Layers are computed as H' = W H (features x samples), weights as A. The D
// tiles of one layer become the next layer's B operand in place: chunk c of
// the next layer takes D tiles 2c and 2c+1, i.e. K slot 8g+e of chunk c holds
// feature perm(c,g,e) = 16(2c + e/4) + 4g + e%4; the packed weights use the
// same K order. The backward pass chains the same way with W^T. Weight
// gradients dW = G H^T reduce over samples in a separate kernel that reads
// activations and gradients stored feature-major.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "geom.hpp"
#include "internal.hpp"

namespace ilm {
namespace nrc {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4;
typedef __attribute__((ext_vector_type(4))) float f32x4;
typedef __attribute__((ext_vector_type(2))) __bf16 bf16x2;
typedef __attribute__((ext_vector_type(2))) float f32x2;
typedef __attribute__((ext_vector_type(2))) short i16x2;
typedef __attribute__((ext_vector_type(4))) unsigned u32x4;

// Two floats rounded to bf16 (nearest even) by one v_cvt_pk_bf16_f32; wider
// conversions compile to one conversion per value and a pack per pair.
__device__ __forceinline__ unsigned bf16_pair(float x, float y) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){x, y}, bf16x2));
}
// The same pair with ReLU applied after rounding, on the bits: a negative
// value (sign bit set, -0 included) is a negative 16-bit integer, so a signed
// max with 0 makes it +0 and keeps every other value. For finite inputs this
// equals rounding max(x, 0).
__device__ __forceinline__ unsigned relu_pair(float x, float y) {
    const i16x2 h = __builtin_bit_cast(i16x2, bf16_pair(x, y));
    return __builtin_bit_cast(unsigned, __builtin_elementwise_max(h, (i16x2)(short)0));
}

constexpr int kIn = 96, kHid = 64, kOut = 16;
constexpr int kQuery = 16;                       // floats per query record
// master weights, row-major [out][in]
constexpr int kW0 = 0;                           // 64 x 96
constexpr int kWh = kW0 + kHid * kIn;            // 4 x (64 x 64)
constexpr int kW5 = kWh + 4 * kHid * kHid;       // 16 x 64
constexpr int kParams = kW5 + kOut * kHid;       // 23552
// packed forward fragments (bf16): 46 fragments of 512
constexpr int kF0 = 0, kFh = 12 * 512, kF5 = kFh + 32 * 512, kFwdElems = kF5 + 2 * 512;
// packed backward (transposed) fragments: layers 1..4 (x32) and layer 5 (x16)
constexpr int kBh = 0, kB5 = 32 * 512, kBwdElems = kB5 + 4 * 256;
static_assert(kFwdElems == kParams, "every weight appears once in the forward pack");

__host__ __device__ __forceinline__ int perm(int c, int g, int e) { return 16 * (2 * c + (e >> 2)) + 4 * g + (e & 3); }
// inverse of perm for an input feature i < 64
__host__ __device__ __forceinline__ void iperm(int i, int& c, int& g, int& e) {
    int q = i >> 4, r = i & 15;
    c = q >> 1;
    g = r >> 2;
    e = 4 * (q & 1) + (r & 3);
}

// Position in the forward pack of master weight (layer, o, i).
__host__ __device__ __forceinline__ int fwd_pos(int layer, int o, int i) {
    if (layer == 0) {
        int t = o >> 4, m = o & 15, c = i >> 5, k = i & 31, g = k >> 3, e = k & 7;
        return kF0 + ((t * 3 + c) * 64 + m + 16 * g) * 8 + e;
    }
    int c, g, e;
    iperm(i, c, g, e);
    if (layer <= 4) {
        int t = o >> 4, m = o & 15;
        return kFh + (layer - 1) * 8 * 512 + ((t * 2 + c) * 64 + m + 16 * g) * 8 + e;
    }
    return kF5 + (c * 64 + o + 16 * g) * 8 + e;  // layer 5: one output tile
}
// Position in the backward pack of master weight (layer, o, i), layers 1..5.
__host__ __device__ __forceinline__ int bwd_pos(int layer, int o, int i) {
    int tp = i >> 4, mp = i & 15;
    if (layer <= 4) {
        int c, g, e;
        iperm(o, c, g, e);
        return kBh + (layer - 1) * 8 * 512 + ((tp * 2 + c) * 64 + mp + 16 * g) * 8 + e;
    }
    int g = o >> 2, e = o & 3;
    return kB5 + (tp * 64 + mp + 16 * g) * 4 + e;
}

__host__ __device__ __forceinline__ int layer_of(int p, int& o, int& i) {
    if (p < kWh) { o = p / kIn; i = p % kIn; return 0; }
    if (p < kW5) { int q = p - kWh; int l = q / (kHid * kHid); q %= kHid * kHid; o = q / kHid; i = q % kHid; return 1 + l; }
    int q = p - kW5;
    o = q / kHid;
    i = q % kHid;
    return 5;
}

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// The 24 features lane group g (0..3) holds for one sample, as three 8-feature
// chunks: chunk c holds features 32c + 8g .. 32c + 8g + 7. A chunk below 72
// is four (sin, cos) pairs of consecutive frequencies of one coordinate; the
// first pair is evaluated directly and the double-angle identities
// sin 2a = 2 sin a cos a, cos 2a = cos^2 a - sin^2 a give the next three
// (three doublings multiply its error by at most 8, far below bf16 rounding).
// The first pair: sin(pi x) and cos(pi x) for x = p 2^fr (0 <= p <= 1,
// fr <= 8) are those of x / 2 revolutions, reduced exactly to [0, 1) and
// evaluated by v_sin_f32 and v_cos_f32, whose argument is in revolutions
// (measured on gfx950 over that range: at most 1.25e-7 from the double
// result, against 5e-8 for sincospif, at a tenth of its instructions).
__device__ __forceinline__ void freq_chunk(float p, int fr, bf16x8& x) {
    const float r = __builtin_amdgcn_fractf(0.5f * p * (float)(1 << fr));
    float sn = __builtin_amdgcn_sinf(r), cs = __builtin_amdgcn_cosf(r);
    u32x4 u;
    for (int k = 0; k < 4; ++k) {
        u[k] = bf16_pair(sn, cs);
        const float s2 = 2.0f * sn * cs, c2 = cs * cs - sn * sn;
        sn = s2;
        cs = c2;
    }
    x = __builtin_bit_cast(bf16x8, u);
}

// Chunks 0 and 1 (features 0..63) are frequency chunks for every g; chunk 2 is
// one for g = 0 (features 64..71: coordinate 2 from frequency 8) and holds the
// query's attributes for g = 1 (features 72..79: direction, normal, roughness,
// kd) and g = 2 (80..82: ks, then zeros), zeros for g = 3. Every index into q
// is a constant: an index that depends on the lane would compile to a compare
// and a select over all 16 entries of q for each feature.
// Chunk 0 (features 8g..) is coordinate 0 from frequency 4g for g < 3 and
// coordinate 1 from frequency 0 for g = 3; chunk 1 (features 32 + 8g..) is
// coordinate 1 from frequency 4 + 4g for g < 2 and coordinate 2 from frequency
// 4(g - 2) for g >= 2. The normalised coordinates p0, p1, p2 come as scalars:
// picking one from an array by a lane-dependent index puts the array in
// scratch memory.
__device__ __forceinline__ void lane_features(const float* q, float p0, float p1, float p2, int g, bf16x8 x[3]) {
    freq_chunk(g == 3 ? p1 : p0, g == 3 ? 0 : 4 * g, x[0]);
    freq_chunk(g < 2 ? p1 : p2, g < 2 ? 4 + 4 * g : 4 * (g - 2), x[1]);
    if (g == 0) {
        freq_chunk(p2, 8, x[2]);
    } else {
        float v[8];
        for (int e = 0; e < 8; ++e) v[e] = g == 1 ? q[3 + e] : (g == 2 && e < 3) ? q[11 + e] : 0.0f;
        const u32x4 u = {bf16_pair(v[0], v[1]), bf16_pair(v[2], v[3]), bf16_pair(v[4], v[5]), bf16_pair(v[6], v[7])};
        x[2] = __builtin_bit_cast(bf16x8, u);
    }
}

struct Bounds { float lo[3], inv[3]; };

__device__ __forceinline__ void load_query(const float* queries, int64_t s, const Bounds& bd, float q[kQuery], float pn[3]) {
    const float4* p4 = reinterpret_cast<const float4*>(queries + s * kQuery);
    float4 a = p4[0], b = p4[1], c = p4[2], d = p4[3];
    q[0] = a.x; q[1] = a.y; q[2] = a.z; q[3] = a.w; q[4] = b.x; q[5] = b.y; q[6] = b.z; q[7] = b.w;
    q[8] = c.x; q[9] = c.y; q[10] = c.z; q[11] = c.w; q[12] = d.x; q[13] = d.y; q[14] = d.z; q[15] = d.w;
    for (int k = 0; k < 3; ++k) pn[k] = fminf(fmaxf((q[k] - bd.lo[k]) * bd.inv[k], 0.f), 1.f);
}

// ReLU of two D tiles' values as one B chunk of bf16.
__device__ __forceinline__ bf16x8 relu_pack(f32x4 a, f32x4 b) {
    const u32x4 u = {relu_pair(a[0], a[1]), relu_pair(a[2], a[3]), relu_pair(b[0], b[1]), relu_pair(b[2], b[3])};
    return __builtin_bit_cast(bf16x8, u);
}

// Makes `off` depend on `dep` in a way the compiler cannot see through. Weight
// fragments addressed with `off` are then read from LDS only after `dep` (the
// previous layer's result) exists, so one layer's fragments are live in
// registers at a time instead of the whole network's.
__device__ __forceinline__ void order_after(int& off, float dep) { asm volatile("" : "+v"(off) : "v"(dep)); }

// Copies `elems` bf16 (a multiple of 8) from global memory to LDS with the
// whole workgroup. Each thread issues B loads of 16 B, then an empty asm
// statement takes every loaded value (so the compiler cannot move each load
// down to its store), then the LDS stores follow. Copied one at a time, every
// load waited for memory before its store: 39 round trips per thread for the
// 80 KB of training weights (batched, the training kernel takes 6-8% less).
// B = 6 covers the inference weights with 512 threads in one batch without
// raising the kernel's registers.
template <int B>
__device__ __forceinline__ void lds_copy(__bf16* dst, const __bf16* src, int elems) {
    static_assert(B == 6 || B == 16, "the asm statement below lists 6 or 16 values");
    const u32x4* s = reinterpret_cast<const u32x4*>(src);
    u32x4* d = reinterpret_cast<u32x4*>(dst);
    const int n16 = elems / 8, nt = (int)blockDim.x;
    for (int k0 = (int)threadIdx.x; k0 < n16; k0 += B * nt) {
        u32x4 v[B];
        // past the end a lane re-reads the last element (its store is skipped)
#pragma unroll
        for (int u = 0; u < B; ++u) v[u] = s[min(k0 + u * nt, n16 - 1)];
        if constexpr (B == 6) {
            asm volatile("" ::"v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]), "v"(v[4]), "v"(v[5]));
        } else {
            asm volatile("" ::"v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]), "v"(v[4]), "v"(v[5]), "v"(v[6]), "v"(v[7]),
                         "v"(v[8]), "v"(v[9]), "v"(v[10]), "v"(v[11]), "v"(v[12]), "v"(v[13]), "v"(v[14]), "v"(v[15]));
        }
#pragma unroll
        for (int u = 0; u < B; ++u)
            if (k0 + u * nt < n16) d[k0 + u * nt] = v[u];
    }
}

// ------------------------------------------------------------------ inference
struct InferArgs {
    const __bf16* fwd;
    const float* queries;
    int64_t n;                        // query count, or capacity when n_dev is set
    const unsigned long long* n_dev;  // query count in device memory, or null
    Bounds bd;
    float* out;  // 4 floats per query (rgb, 0)
};

// Inference: two sample tiles of 16 per wave iteration and eight waves per
// workgroup sharing one copy of the weights in LDS, compiled for six waves per
// SIMD (at most 80 VGPRs; three workgroups per CU fit the LDS). Four tiles
// with four waves (134 VGPRs, three waves per SIMD) took 8% longer, three
// tiles with eight waves 5%.
constexpr int kTiles = 2;
constexpr int kInferWaves = 8;

__global__ void __launch_bounds__(64 * kInferWaves, 6) k_infer(InferArgs a) {
    __shared__ __bf16 w[kFwdElems];
    lds_copy<6>(w, a.fwd, kFwdElems);
    __syncthreads();
    const int lane = (int)__lane_id(), g = lane >> 4, m = lane & 15;
    const int wave = threadIdx.x >> 6;
    const int64_t n = a.n_dev ? (int64_t)min((unsigned long long)a.n, *a.n_dev) : a.n;
    const int64_t step = (int64_t)gridDim.x * kInferWaves * 16 * kTiles;
    for (int64_t base = ((int64_t)blockIdx.x * kInferWaves + wave) * 16 * kTiles; base < n; base += step) {
        // An opaque per-iteration copy of the lane's fragment offset keeps the
        // compiler from hoisting all 46 weight fragments out of the loop into
        // registers (184 VGPRs, one wave per SIMD); they are re-read from LDS.
        int lo8 = lane * 8;
        asm volatile("" : "+v"(lo8));
        bf16x8 x[kTiles][3];
        for (int s = 0; s < kTiles; ++s) {
            int64_t smp = min(base + 16 * s + m, n - 1);
            float q[kQuery], pn[3];
            load_query(a.queries, smp, a.bd, q, pn);
            lane_features(q, pn[0], pn[1], pn[2], g, x[s]);
        }
        f32x4 acc[4][kTiles];
        // layer 0: 96 -> 64
        for (int t = 0; t < 4; ++t) {
            for (int s = 0; s < kTiles; ++s) acc[t][s] = f32x4{0.f, 0.f, 0.f, 0.f};
            for (int c = 0; c < 3; ++c) {
                bf16x8 A = *reinterpret_cast<const bf16x8*>(w + kF0 + (t * 3 + c) * 512 + lo8);
                for (int s = 0; s < kTiles; ++s) acc[t][s] = mfma32(A, x[s][c], acc[t][s]);
            }
        }
        bf16x8 h[kTiles][2];
        for (int s = 0; s < kTiles; ++s)
            for (int c = 0; c < 2; ++c) h[s][c] = relu_pack(acc[2 * c][s], acc[2 * c + 1][s]);
        order_after(lo8, acc[3][kTiles - 1][3]);
        // hidden layers 1..4: 64 -> 64
        for (int l = 0; l < 4; ++l) {
            for (int t = 0; t < 4; ++t) {
                for (int s = 0; s < kTiles; ++s) acc[t][s] = f32x4{0.f, 0.f, 0.f, 0.f};
                for (int c = 0; c < 2; ++c) {
                    bf16x8 A = *reinterpret_cast<const bf16x8*>(w + kFh + l * 8 * 512 + (t * 2 + c) * 512 + lo8);
                    for (int s = 0; s < kTiles; ++s) acc[t][s] = mfma32(A, h[s][c], acc[t][s]);
                }
            }
            for (int s = 0; s < kTiles; ++s)
                for (int c = 0; c < 2; ++c) h[s][c] = relu_pack(acc[2 * c][s], acc[2 * c + 1][s]);
            order_after(lo8, acc[3][kTiles - 1][3]);
        }
        // output layer: 64 -> 16
        for (int s = 0; s < kTiles; ++s) {
            f32x4 o = f32x4{0.f, 0.f, 0.f, 0.f};
            for (int c = 0; c < 2; ++c) {
                bf16x8 A = *reinterpret_cast<const bf16x8*>(w + kF5 + c * 512 + lo8);
                o = mfma32(A, h[s][c], o);
            }
            int64_t smp = base + 16 * s + m;
            if (g == 0 && smp < n)
                reinterpret_cast<float4*>(a.out)[smp] = make_float4(o[0], o[1], o[2], 0.f);
        }
    }
}

// ------------------------------------------------------------------ training: forward + backward
struct TrainArgs {
    const __bf16* fwd;
    const __bf16* bwd;
    const float* queries;
    const float* targets;  // 4 floats per sample (rgb, weight)
    int64_t n;
    Bounds bd;
    __bf16* H;  // activations, feature-major: H0 [96][n] then H1..H5 [64][n]
    __bf16* G;  // gradients, feature-major: G1..G5 [64][n] then G6 [16][n]
    int64_t ld;  // leading dimension of H and G (n rounded up to 32)
    float* loss;  // per-sample loss (may be null)
    float grad_scale;  // 1 / (3 n)
};

// One bf16 store at a row pointer plus a 32-bit byte offset.
__device__ __forceinline__ void st_bf16(__bf16* row, uint32_t byte_off, __bf16 v) {
    *reinterpret_cast<__bf16*>(reinterpret_cast<char*>(row) + byte_off) = v;
}
__device__ __forceinline__ int64_t hoff(int layer, int64_t ld) { return layer == 0 ? 0 : (int64_t)kIn * ld + (int64_t)(layer - 1) * kHid * ld; }
__device__ __forceinline__ int64_t goff(int layer, int64_t ld) { return (int64_t)(layer - 1) * kHid * ld; }  // layer 1..6

// Two 256-thread workgroups per CU (80 KB of packed weights each in LDS).
// 1024-thread workgroups (4 waves per SIMD) measured no faster: the kernel's
// time goes to its feature-major activation and gradient stores.
constexpr int kFbBlock = 256;
__global__ void __launch_bounds__(kFbBlock) k_train_fb(TrainArgs a) {
    __shared__ __bf16 wf[kFwdElems];
    __shared__ __bf16 wb[kBwdElems];
    lds_copy<16>(wf, a.fwd, kFwdElems);
    lds_copy<16>(wb, a.bwd, kBwdElems);
    __syncthreads();
    const int lane = (int)__lane_id(), g = lane >> 4, m = lane & 15;
    const int wave = threadIdx.x >> 6;
    constexpr int kWaves = kFbBlock / 64;
    const int64_t n = a.n, ld = a.ld;
    const int64_t step = (int64_t)gridDim.x * kWaves * 16;
    for (int64_t base = ((int64_t)blockIdx.x * kWaves + wave) * 16; base < n; base += step) {
        int lo8 = lane * 8, lo4 = lane * 4;  // opaque offsets: weights stay in LDS (see k_infer)
        asm volatile("" : "+v"(lo8), "+v"(lo4));
        const int64_t smp = base + m;
        const bool valid = smp < n;
        const int64_t sl = valid ? smp : n - 1;
        float q[kQuery], pn[3];
        load_query(a.queries, sl, a.bd, q, pn);
        bf16x8 x[3];
        lane_features(q, pn[0], pn[1], pn[2], g, x);
        // Each store is a row pointer plus the lane's 32-bit byte offset (the
        // lane's sample, shifted by its group's 4 or 8 features; below 2^32
        // for n <= 2^26). Every lane stores, with no mask per store: a wave's
        // columns (base .. base + 15) lie below ld, and the weight-gradient
        // reduction reads only columns below n.
        const uint32_t off4 = 2u * (uint32_t)(4 * g * ld + smp), off8 = 2u * (uint32_t)(8 * g * ld + smp);
        for (int c = 0; c < 3; ++c)
            for (int e = 0; e < 8; ++e) st_bf16(a.H + hoff(0, ld) + (int64_t)(32 * c + e) * ld, off8, x[c][e]);
        // hc = the current layer's output, as B chunks. Each layer's output is
        // stored (H_l, for the weight gradients) as soon as it exists, and
        // the backward pass keeps only its ReLU derivative: bit 8c + e of
        // relu_on[l] is set when feature (c, e) of H_l is positive. Holding
        // the masks instead of the five layers' outputs saves ~40 registers.
        bf16x8 hc[2];
        uint32_t relu_on[6];
        f32x4 acc[4];
        for (int t = 0; t < 4; ++t) {
            acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
            for (int c = 0; c < 3; ++c)
                acc[t] = mfma32(*reinterpret_cast<const bf16x8*>(wf + kF0 + (t * 3 + c) * 512 + lo8), x[c], acc[t]);
        }
#pragma unroll
        for (int l = 1; l <= 5; ++l) {
            if (l > 1) {
                for (int t = 0; t < 4; ++t) {
                    acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
                    for (int c = 0; c < 2; ++c)
                        acc[t] = mfma32(*reinterpret_cast<const bf16x8*>(wf + kFh + (l - 2) * 8 * 512 + (t * 2 + c) * 512 + lo8),
                                        hc[c], acc[t]);
                }
            }
            for (int c = 0; c < 2; ++c) hc[c] = relu_pack(acc[2 * c], acc[2 * c + 1]);
            order_after(lo8, acc[3][3]);
            // store H_l (feature of K slot (c, g, e) is perm(c, g, e))
            uint32_t on = 0;
            for (int c = 0; c < 2; ++c)
                for (int e = 0; e < 8; ++e) {
                    st_bf16(a.H + hoff(l, ld) + (int64_t)perm(c, 0, e) * ld, off4, hc[c][e]);
                    on |= ((float)hc[c][e] > 0.f ? 1u : 0u) << (8 * c + e);
                }
            relu_on[l] = on;
        }
        f32x4 o = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int c = 0; c < 2; ++c) o = mfma32(*reinterpret_cast<const bf16x8*>(wf + kF5 + c * 512 + lo8), hc[c], o);
        // relative L2 loss on rgb: (y - t)^2 / (sg(y)^2 + 0.01); lanes g == 0 hold outputs 0..3
        bf16x4 g6;
        float lsum = 0.f;
        float4 tg = reinterpret_cast<const float4*>(a.targets)[sl];
        for (int r = 0; r < 4; ++r) {
            float gr = 0.f;
            if (g == 0 && r < 3 && valid) {
                float y = o[r], t = (r == 0 ? tg.x : (r == 1 ? tg.y : tg.z));
                float den = y * y + 0.01f;
                gr = 2.f * (y - t) / den * a.grad_scale * tg.w;
                lsum += (y - t) * (y - t) / den * tg.w;
            }
            g6[r] = (__bf16)gr;
        }
        if (a.loss && g == 0 && valid) a.loss[smp] = lsum;
        for (int r = 0; r < 4; ++r) st_bf16(a.G + goff(6, ld) + (int64_t)r * ld, off4, g6[r]);
        // backward through layer 5 (x16 fragments): dH5 = W5^T g6
        f32x4 d[4];
        order_after(lo4, (float)g6[0]);
        for (int t = 0; t < 4; ++t)
            d[t] = mfma16(*reinterpret_cast<const bf16x4*>(wb + kB5 + t * 256 + lo4), g6, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
        for (int l = 5; l >= 1; --l) {
            // G_l = dH_l * relu'(H_l), back in B-chunk form
            bf16x8 gl[2];
            for (int c = 0; c < 2; ++c) {
                float v[8];
                for (int e = 0; e < 8; ++e) v[e] = ((relu_on[l] >> (8 * c + e)) & 1u) ? d[2 * c + (e >> 2)][e & 3] : 0.f;
                const u32x4 u = {bf16_pair(v[0], v[1]), bf16_pair(v[2], v[3]), bf16_pair(v[4], v[5]), bf16_pair(v[6], v[7])};
                gl[c] = __builtin_bit_cast(bf16x8, u);
            }
            for (int c = 0; c < 2; ++c)
                for (int e = 0; e < 8; ++e) st_bf16(a.G + goff(l, ld) + (int64_t)perm(c, 0, e) * ld, off4, gl[c][e]);
            if (l == 1) break;
            order_after(lo8, d[3][3]);
            // dH_{l-1} = W_{l-1}^T G_l  (layer index of the weights is l-1, 1..4)
            for (int t = 0; t < 4; ++t) {
                d[t] = f32x4{0.f, 0.f, 0.f, 0.f};
                for (int c = 0; c < 2; ++c)
                    d[t] = mfma32(*reinterpret_cast<const bf16x8*>(wb + kBh + (l - 2) * 8 * 512 + (t * 2 + c) * 512 + lo8),
                                  gl[c], d[t]);
            }
        }
    }
}

// ------------------------------------------------------------------ weight gradients
// 92 output tiles of 16x16 in 21 strips of tiles sharing their output rows:
// layer 0 (4 strips of 6 tiles), layers 1..4 (4 strips of 4 tiles each),
// layer 5 (1 strip of 4 tiles). A wave computes one strip over its
// workgroup's chunk of samples, so the strip's G rows are read once for all
// of its tiles.
constexpr int kStrips = 21;
constexpr int kStripWaves = 24;  // six workgroups of four waves per chunk
struct DwArgs {
    const __bf16* H;
    const __bf16* G;
    int64_t n;         // samples
    int64_t ld;        // leading dimension of H and G (multiple of 32)
    int64_t chunk;     // samples per block (multiple of 32)
    float* partial;    // [blocks][kParams]
};

__device__ __forceinline__ void strip_of(int strip, int& layer, int& to, int& ntiles) {
    if (strip < 4) { layer = 0; to = strip; ntiles = 6; return; }
    if (strip < 20) { layer = 1 + (strip - 4) / 4; to = (strip - 4) % 4; ntiles = 4; return; }
    layer = 5; to = 0; ntiles = 4;
}

__global__ void __launch_bounds__(256, 3) k_dw(DwArgs a) {
    const int lane = (int)__lane_id(), g = lane >> 4, m = lane & 15;
    const int wave = threadIdx.x >> 6;
    const int strip = (int)blockIdx.y * 4 + wave;
    if (strip >= kStrips) return;
    const int64_t n = a.n, ld = a.ld;
    const int64_t s0 = (int64_t)blockIdx.x * a.chunk, s1 = min(n, s0 + a.chunk);
    float* part = a.partial + (int64_t)blockIdx.x * kParams;
    int layer, to, ntiles;
    strip_of(strip, layer, to, ntiles);
    // A rows: G_{layer+1} output features; B rows: H_layer input features
    const __bf16* gA = a.G + (int64_t)layer * kHid * ld + (int64_t)(16 * to + m) * ld;
    const __bf16* hB = a.H + (layer == 0 ? 0 : (int64_t)kIn * ld + (int64_t)(layer - 1) * kHid * ld) + (int64_t)m * ld;
    f32x4 acc[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // Three steps of 32 samples at a time: their loads are all issued before
    // the first product, so a wave waits for memory once per 96 samples (four
    // steps need more than the 168 registers of three waves per SIMD).
    constexpr int kSteps = 3;
    int64_t s = s0;
    for (; s + 32 * kSteps <= s1; s += 32 * kSteps) {
        bf16x8 A[kSteps], B[kSteps][6];
#pragma unroll
        for (int u = 0; u < kSteps; ++u) {
            const int64_t k0 = s + 32 * u + 8 * g;
            A[u] = *reinterpret_cast<const bf16x8*>(gA + k0);
#pragma unroll
            for (int t = 0; t < 6; ++t)
                if (t < ntiles) B[u][t] = *reinterpret_cast<const bf16x8*>(hB + (int64_t)(16 * t) * ld + k0);
        }
#pragma unroll
        for (int u = 0; u < kSteps; ++u)
#pragma unroll
            for (int t = 0; t < 6; ++t)
                if (t < ntiles) acc[t] = mfma32(A[u], B[u][t], acc[t]);
    }
    for (; s < s1; s += 32) {
        const int64_t k0 = s + 8 * g;
        const bool full = k0 + 8 <= s1;
        bf16x8 A;
        if (full) {
            A = *reinterpret_cast<const bf16x8*>(gA + k0);
        } else {
            for (int e = 0; e < 8; ++e) A[e] = k0 + e < s1 ? gA[k0 + e] : (__bf16)0.f;
        }
#pragma unroll
        for (int t = 0; t < 6; ++t) {
            if (t >= ntiles) break;
            const __bf16* hp = hB + (int64_t)(16 * t) * ld;
            bf16x8 B;
            if (full) {
                B = *reinterpret_cast<const bf16x8*>(hp + k0);
            } else {
                for (int e = 0; e < 8; ++e) B[e] = k0 + e < s1 ? hp[k0 + e] : (__bf16)0.f;
            }
            acc[t] = mfma32(A, B, acc[t]);
        }
    }
    const int in_dim = layer == 0 ? kIn : kHid;
    const int base = layer == 0 ? kW0 : (layer <= 4 ? kWh + (layer - 1) * kHid * kHid : kW5);
#pragma unroll
    for (int t = 0; t < 6; ++t) {
        if (t >= ntiles) break;
        for (int r = 0; r < 4; ++r) {
            const int o = 16 * to + 4 * g + r, i = 16 * t + m;
            part[base + o * in_dim + i] = acc[t][r];
        }
    }
}

// ------------------------------------------------------------------ Adam + repack
struct AdamArgs {
    const float* partial;
    int nblocks;
    float* w;
    float* mom1;
    float* mom2;
    float* grad;  // summed gradient of the step (kept for validation)
    __bf16* fwd;
    __bf16* bwd;
    float lr, b1, b2, eps, c1, c2;  // c1 = 1 - b1^t, c2 = 1 - b2^t
    int do_update;
};

// 64 parameters per workgroup: its four waves sum every fourth sample
// chunk's partial gradients (a fixed order, so the sum is deterministic),
// the first wave adds the four sums and updates.
__global__ void __launch_bounds__(256) k_adam(AdamArgs a) {
    __shared__ float quarter[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    float part = 0.f;
    if (a.do_update && p < kParams) {
        // eight rows in flight per wave (independent loads; the sum keeps its order)
#pragma unroll 8
        for (int b = wave; b < a.nblocks; b += 4) part += a.partial[(int64_t)b * kParams + p];
    }
    quarter[wave][lane] = part;
    __syncthreads();
    if (wave != 0 || p >= kParams) return;
    float wv = a.w[p];
    if (a.do_update) {
        const float gsum = (quarter[0][lane] + quarter[1][lane]) + (quarter[2][lane] + quarter[3][lane]);
        a.grad[p] = gsum;
        float m1 = a.b1 * a.mom1[p] + (1.f - a.b1) * gsum;
        float m2 = a.b2 * a.mom2[p] + (1.f - a.b2) * gsum * gsum;
        a.mom1[p] = m1;
        a.mom2[p] = m2;
        wv -= a.lr * (m1 / a.c1) / (sqrtf(m2 / a.c2) + a.eps);
        a.w[p] = wv;
    }
    int o, i;
    int layer = layer_of(p, o, i);
    __bf16 b = (__bf16)wv;
    a.fwd[fwd_pos(layer, o, i)] = b;
    if (layer >= 1) a.bwd[bwd_pos(layer, o, i)] = b;
}

// ------------------------------------------------------------------ layout self-test
// D = A B for one 16x16x32 and one 16x16x16 product with integer-valued
// operands (exact in bf16 and f32), compared on the host against the
// operand layouts the kernels above assume.
__global__ void k_layout_probe(float* out32, float* out16) {
    int l = (int)__lane_id(), g = l >> 4, m = l & 15;
    bf16x8 A, B;
    for (int e = 0; e < 8; ++e) {
        int k = 8 * g + e;
        A[e] = (__bf16)(float)((m * 3 + k) % 7 - 3);   // A[m][k]
        B[e] = (__bf16)(float)((k * 5 + m) % 5 - 2);   // B[k][n = m]
    }
    f32x4 d = mfma32(A, B, f32x4{0.f, 0.f, 0.f, 0.f});
    for (int r = 0; r < 4; ++r) out32[(4 * g + r) * 16 + m] = d[r];
    bf16x4 A4, B4;
    for (int e = 0; e < 4; ++e) {
        int k = 4 * g + e;
        A4[e] = (__bf16)(float)((m * 3 + k) % 7 - 3);
        B4[e] = (__bf16)(float)((k * 5 + m) % 5 - 2);
    }
    f32x4 d4 = mfma16(A4, B4, f32x4{0.f, 0.f, 0.f, 0.f});
    for (int r = 0; r < 4; ++r) out16[(4 * g + r) * 16 + m] = d4[r];
}

}  // namespace nrc
}  // namespace ilm

// ------------------------------------------------------------------ C ABI
struct ilm_nrc {
    ilm_ctx* ctx = nullptr;
    ilm::nrc::Bounds bd{};
    float lr = 1e-3f, b1 = 0.9f, b2 = 0.99f, eps = 1e-8f;
    int64_t step = 0;
    float* w = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    float* grad = nullptr;
    __bf16* fwd = nullptr;
    __bf16* bwd = nullptr;
    // training scratch, grown on demand
    int64_t cap = 0;
    __bf16* H = nullptr;
    __bf16* G = nullptr;
    float* partial = nullptr;
    int64_t partial_blocks = 0;
};

namespace ilm {
int nrc_infer_queue(ilm_nrc* r, const float* queries, const unsigned long long* n_dev, int64_t capacity, float* out,
                    hipStream_t stream) {
    if (capacity <= 0) return ILM_OK;
    nrc::InferArgs a;
    a.fwd = r->fwd;
    a.queries = queries;
    a.n = capacity;
    a.n_dev = n_dev;
    a.bd = r->bd;
    a.out = out;
    // One round of resident workgroups (the 47 KB of weights each holds in LDS
    // allow three per CU): a larger grid runs a second, partial round after
    // the first.
    int per_cu = 0;
    ILM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)nrc::k_infer, 64 * nrc::kInferWaves, 0));
    per_cu = std::max(per_cu, 1);
    int64_t per_block = nrc::kInferWaves * 16 * nrc::kTiles;
    int64_t blocks = std::min<int64_t>((capacity + per_block - 1) / per_block, (int64_t)r->ctx->cu_count * per_cu);
    hipLaunchKernelGGL(nrc::k_infer, dim3((unsigned)blocks), dim3(64 * nrc::kInferWaves), 0, stream, a);
    ILM_HIP(hipGetLastError());
    return ILM_OK;
}
}  // namespace ilm

namespace {
using namespace ilm::nrc;

int nrc_repack(ilm_nrc* r, hipStream_t st, int update, float c1, float c2, int nblocks) {
    AdamArgs a;
    a.partial = r->partial;
    a.nblocks = nblocks;
    a.w = r->w;
    a.mom1 = r->m1;
    a.mom2 = r->m2;
    a.grad = r->grad;
    a.fwd = r->fwd;
    a.bwd = r->bwd;
    a.lr = r->lr;
    a.b1 = r->b1;
    a.b2 = r->b2;
    a.eps = r->eps;
    a.c1 = c1;
    a.c2 = c2;
    a.do_update = update;
    hipLaunchKernelGGL(k_adam, dim3((kParams + 63) / 64), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? ILM_OK : ilm::fail(ILM_ERR_HIP, "adam launch");
}

int layout_selftest(ilm_ctx* ctx) {
    static int verified = 0;
    if (verified) return ILM_OK;
    float* d = nullptr;
    ILM_HIP(hipMalloc(&d, 512 * sizeof(float)));
    hipLaunchKernelGGL(k_layout_probe, dim3(1), dim3(64), 0, ctx->stream, d, d + 256);
    float h[512];
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return ilm::fail_hip(e, "MFMA layout probe");
    for (int K = 32, off = 0; K >= 16; K -= 16, off += 256)
        for (int i = 0; i < 16; ++i)
            for (int j = 0; j < 16; ++j) {
                float want = 0.f;
                for (int k = 0; k < K; ++k) want += (float)((i * 3 + k) % 7 - 3) * (float)((k * 5 + j) % 5 - 2);
                if (h[off + i * 16 + j] != want)
                    return ilm::fail(ILM_ERR_STATE, "MFMA operand layout differs from the one the cache kernels assume");
            }
    verified = 1;
    return ILM_OK;
}

int ensure_train(ilm_nrc* r, int64_t n, int64_t blocks) {
    n = (n + 31) / 32 * 32;
    if (r->cap < n) {
        if (r->H) ilm_free(r->ctx, r->H);
        if (r->G) ilm_free(r->ctx, r->G);
        r->H = r->G = nullptr;
        r->cap = 0;
        int rc = ilm_alloc(r->ctx, (size_t)n * (kIn + 5 * kHid) * 2, (void**)&r->H);
        if (!rc) rc = ilm_alloc(r->ctx, (size_t)n * (5 * kHid + kOut) * 2, (void**)&r->G);
        if (rc) return rc;
        r->cap = n;
    }
    if (r->partial_blocks < blocks) {
        if (r->partial) ilm_free(r->ctx, r->partial);
        r->partial = nullptr;
        int rc = ilm_alloc(r->ctx, (size_t)blocks * kParams * 4, (void**)&r->partial);
        if (rc) return rc;
        r->partial_blocks = blocks;
    }
    return ILM_OK;
}
}  // namespace

extern "C" {

int ilm_nrc_create(ilm_ctx* ctx, const ilm_nrc_params* p, ilm_nrc** out) {
    if (!ctx || !p || !out) return ilm::fail(ILM_ERR_INVALID, "null argument");
    *out = nullptr;
    int rc = layout_selftest(ctx);
    if (rc) return rc;
    ilm_nrc* r = new ilm_nrc();
    r->ctx = ctx;
    for (int k = 0; k < 3; ++k) {
        float ext = p->bounds_hi[k] - p->bounds_lo[k];
        r->bd.lo[k] = p->bounds_lo[k];
        r->bd.inv[k] = ext > 0 ? 1.f / ext : 0.f;
    }
    if (p->lr > 0) r->lr = p->lr;
    if (p->beta1 > 0) r->b1 = p->beta1;
    if (p->beta2 > 0) r->b2 = p->beta2;
    if (p->eps > 0) r->eps = p->eps;
    rc = ilm_alloc(ctx, kParams * 4, (void**)&r->w);
    if (!rc) rc = ilm_alloc(ctx, kParams * 4, (void**)&r->m1);
    if (!rc) rc = ilm_alloc(ctx, kParams * 4, (void**)&r->m2);
    if (!rc) rc = ilm_alloc(ctx, kParams * 4, (void**)&r->grad);
    if (!rc) rc = ilm_alloc(ctx, kFwdElems * 2, (void**)&r->fwd);
    if (!rc) rc = ilm_alloc(ctx, kBwdElems * 2, (void**)&r->bwd);
    if (rc) { ilm_nrc_destroy(r); return rc; }
    // He-uniform initialisation of the used weights (seeded splitmix), zeros in padding
    std::vector<float> w(kParams, 0.f);
    uint64_t s = 0x5EEDull ^ ((uint64_t)p->seed << 17);
    auto uni = [&]() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    for (int q = 0; q < kParams; ++q) {
        int o, i;
        int l = layer_of(q, o, i);
        int fan_in = l == 0 ? 83 : kHid;
        bool used = (l == 0 ? i < 83 : true) && (l == 5 ? o < 3 : true);
        double bound = std::sqrt(6.0 / fan_in) * (l == 5 ? 0.1 : 1.0);
        if (used) w[q] = (float)((2.0 * uni() - 1.0) * bound);
    }
    rc = ilm_upload(ctx, r->w, w.data(), kParams * 4);
    if (!rc) rc = ilm_memset(ctx, r->m1, 0, kParams * 4);
    if (!rc) rc = ilm_memset(ctx, r->m2, 0, kParams * 4);
    if (!rc) rc = ilm_memset(ctx, r->grad, 0, kParams * 4);
    if (!rc) rc = ilm_memset(ctx, r->bwd, 0, kBwdElems * 2);
    if (!rc) rc = nrc_repack(r, ctx->stream, 0, 1.f, 1.f, 0);
    if (!rc) rc = ilm_sync(ctx);
    if (rc) { ilm_nrc_destroy(r); return rc; }
    *out = r;
    return ILM_OK;
}

int ilm_nrc_destroy(ilm_nrc* r) {
    if (!r) return ILM_OK;
    for (void* p : {(void*)r->w, (void*)r->m1, (void*)r->m2, (void*)r->grad, (void*)r->fwd, (void*)r->bwd, (void*)r->H, (void*)r->G,
                    (void*)r->partial})
        if (p) ilm_free(r->ctx, p);
    delete r;
    return ILM_OK;
}

int ilm_nrc_param_count(int* n) {
    if (!n) return ilm::fail(ILM_ERR_INVALID, "null argument");
    *n = kParams;
    return ILM_OK;
}

int ilm_nrc_set_weights(ilm_nrc* r, const float* w) {
    if (!r || !w) return ilm::fail(ILM_ERR_INVALID, "null argument");
    int rc = ilm_upload(r->ctx, r->w, w, kParams * 4);
    if (!rc) rc = ilm_memset(r->ctx, r->m1, 0, kParams * 4);
    if (!rc) rc = ilm_memset(r->ctx, r->m2, 0, kParams * 4);
    r->step = 0;
    if (!rc) rc = nrc_repack(r, r->ctx->stream, 0, 1.f, 1.f, 0);
    if (!rc) rc = ilm_sync(r->ctx);
    return rc;
}

int ilm_nrc_get_gradients(ilm_nrc* r, float* g) {
    if (!r || !g) return ilm::fail(ILM_ERR_INVALID, "null argument");
    return ilm_download(r->ctx, g, r->grad, kParams * 4);
}

int ilm_nrc_get_weights(ilm_nrc* r, float* w) {
    if (!r || !w) return ilm::fail(ILM_ERR_INVALID, "null argument");
    return ilm_download(r->ctx, w, r->w, kParams * 4);
}

int ilm_nrc_infer(ilm_nrc* r, const float* queries, int64_t n, float* out, void* stream) {
    if (!r || (n > 0 && (!queries || !out)) || n < 0) return ilm::fail(ILM_ERR_INVALID, "bad arguments");
    if (n == 0) return ILM_OK;
    return ilm::nrc_infer_queue(r, queries, nullptr, n, out, ilm::pick_stream(r->ctx, stream));
}

int ilm_nrc_train_step(ilm_nrc* r, const float* queries, const float* targets, int64_t n, float* loss, void* stream) {
    if (!r || n <= 0 || !queries || !targets) return ilm::fail(ILM_ERR_INVALID, "bad arguments");
    // the training kernel addresses activations with 32-bit lane offsets
    if (n > (int64_t(1) << 26)) return ilm::fail(ILM_ERR_INVALID, "training batch above 2^26 samples");
    const int64_t chunk = 512;
    const int64_t dw_blocks = (n + chunk - 1) / chunk;
    int rc = ensure_train(r, n, dw_blocks);
    if (rc) return rc;
    hipStream_t st = ilm::pick_stream(r->ctx, stream);
    TrainArgs t;
    t.fwd = r->fwd;
    t.bwd = r->bwd;
    t.queries = queries;
    t.targets = targets;
    t.n = n;
    t.bd = r->bd;
    t.H = r->H;
    t.G = r->G;
    t.ld = (n + 31) / 32 * 32;
    t.loss = loss;
    t.grad_scale = 1.f / (3.f * (float)n);
    const int64_t fb_per_block = kFbBlock / 64 * 16;  // 16 samples per wave
    int64_t fb_blocks = std::min<int64_t>((n + fb_per_block - 1) / fb_per_block, (int64_t)r->ctx->cu_count * 2);
    hipLaunchKernelGGL(k_train_fb, dim3((unsigned)fb_blocks), dim3(kFbBlock), 0, st, t);
    DwArgs d;
    d.H = r->H;
    d.G = r->G;
    d.n = n;
    d.ld = (n + 31) / 32 * 32;
    d.chunk = chunk;
    d.partial = r->partial;
    // six workgroups per chunk of samples, one wave per strip of output tiles
    hipLaunchKernelGGL(k_dw, dim3((unsigned)dw_blocks, kStripWaves / 4), dim3(256), 0, st, d);
    ILM_HIP(hipGetLastError());
    r->step += 1;
    float c1 = 1.f - powf(r->b1, (float)r->step), c2 = 1.f - powf(r->b2, (float)r->step);
    return nrc_repack(r, st, 1, c1, c2, (int)dw_blocks);
}

}  // extern "C"
