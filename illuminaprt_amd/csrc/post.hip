// Frame resolve, tone mapping and image output: the step after the path's
// accumulation. Replaces AccumulationBuffer::Apply (Postproc/AccumulationBuffer.h:59-88),
// the GlobalTone and DragoTone post-processes (Postproc/GlobalTone.h:37-53,
// Postproc/DragoTone.h:37-80), ImageDevice's mirrored write-out
// (Device/ImageDevice.cpp:137-148) and the PFM / PPM writers
// (Image/ImagePFM.h:89-111, Image/ImagePPM.h:85-110).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "ilm_gpu.h"
#include "internal.hpp"

namespace ilm {
namespace {

constexpr int kBlock = 256;

// out pixel (x, y) = accum / samples; with flip, pixel (x, y) of the frame is
// written to (W-1-x, H-1-y) as ImageDevice does.
__global__ void k_resolve(const double* __restrict__ acc, int w, int h, double inv, int flip, float* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    int x = (int)(i % w), y = (int)(i / w);
    int64_t o = flip ? (int64_t)(h - 1 - y) * w + (w - 1 - x) : i;
    out[3 * o + 0] = (float)(acc[3 * i + 0] * inv);
    out[3 * o + 1] = (float)(acc[3 * i + 1] * inv);
    out[3 * o + 2] = (float)(acc[3 * i + 2] * inv);
}

__global__ void k_global_tone(const float* __restrict__ in, int64_t n, float* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float c = in[i];
    out[i] = c / (c + 1.0f);
}

__global__ void k_zero3(uint32_t* p) {
    if (threadIdx.x < 3) p[threadIdx.x] = 0u;
}

// per-channel maximum, starting from 0 as DragoTone's Lwmax does; values
// that are not above 0 never raise it, so float bits order as unsigned
__global__ void k_channel_max(const float* __restrict__ in, int64_t npix, uint32_t* __restrict__ mx) {
    __shared__ float red[3][kBlock];
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = (i < npix) ? fmaxf(in[3 * i + c], 0.0f) : 0.0f;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; ++c) red[c][threadIdx.x] = fmaxf(red[c][threadIdx.x], red[c][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int c = 0; c < 3; ++c) atomicMax(&mx[c], __float_as_uint(red[c][0]));
}

// DragoTone with its fixed bias 0.8 and scale factor 95%:
// v = 0.95 / log10(Lmax + 1) * ln(L + 1) / ln(2 + 8 (L / Lmax)^(ln 0.8 / ln 0.5)),
// clamped to [0, 1]
__global__ void k_drago(const float* __restrict__ in, int64_t npix, const uint32_t* __restrict__ mx,
                        float* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= npix) return;
    const float bias = logf(0.8f) / logf(0.5f);
    for (int c = 0; c < 3; ++c) {
        float lmax = __uint_as_float(mx[c]);
        float L = in[3 * i + c];
        float v = 0.95f / (logf(lmax + 1.0f) / logf(10.0f)) * logf(L + 1.0f) / logf(2.0f + powf(L / lmax, bias) * 8.0f);
        if (v < 0.0f) v = 0.0f;
        else if (v > 1.0f) v = 1.0f;
        out[3 * i + c] = v;
    }
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace
}  // namespace ilm

using ilm::fail;

extern "C" {

int ilm_image_resolve(ilm_ctx* ctx, const double* accum, int width, int height, double inv_samples, int flip,
                      float* out, void* stream) {
    if (!ctx || !accum || !out || width <= 0 || height <= 0) return fail(ILM_ERR_INVALID, "bad arguments");
    hipStream_t st = ilm::pick_stream(ctx, stream);
    const int64_t n = (int64_t)width * height;
    hipLaunchKernelGGL(ilm::k_resolve, dim3(ilm::blocks_for(n)), dim3(ilm::kBlock), 0, st, accum, width, height,
                       inv_samples, flip ? 1 : 0, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? ILM_OK : ilm::fail_hip(e, "resolve launch");
}

int ilm_image_tonemap(ilm_ctx* ctx, const float* in, int width, int height, int op, float* out, void* stream) {
    if (!ctx || !in || !out || width <= 0 || height <= 0) return fail(ILM_ERR_INVALID, "bad arguments");
    if (op != ILM_TONE_GLOBAL && op != ILM_TONE_DRAGO) return fail(ILM_ERR_INVALID, "unknown tone operator");
    hipStream_t st = ilm::pick_stream(ctx, stream);
    const int64_t npix = (int64_t)width * height;
    if (op == ILM_TONE_GLOBAL) {
        hipLaunchKernelGGL(ilm::k_global_tone, dim3(ilm::blocks_for(3 * npix)), dim3(ilm::kBlock), 0, st, in,
                           3 * npix, out);
    } else {
        // three 32-bit maxima in the context's spare counter words
        uint32_t* mx = reinterpret_cast<uint32_t*>(ctx->counters + 4);
        hipLaunchKernelGGL(ilm::k_zero3, dim3(1), dim3(64), 0, st, mx);
        hipLaunchKernelGGL(ilm::k_channel_max, dim3(ilm::blocks_for(npix)), dim3(ilm::kBlock), 0, st, in, npix, mx);
        hipLaunchKernelGGL(ilm::k_drago, dim3(ilm::blocks_for(npix)), dim3(ilm::kBlock), 0, st, in, npix,
                           (const uint32_t*)mx, out);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? ILM_OK : ilm::fail_hip(e, "tone map launch");
}

int ilm_image_write(const char* path, const float* rgb, int width, int height, int format) {
    if (!path || !rgb || width <= 0 || height <= 0) return fail(ILM_ERR_INVALID, "bad arguments");
    if (format != ILM_IMAGE_PFM && format != ILM_IMAGE_PPM) return fail(ILM_ERR_INVALID, "unknown image format");
    FILE* f = fopen(path, "wb");
    if (!f) return fail(ILM_ERR_INVALID, std::string("cannot open ") + path);
    const int64_t n = (int64_t)width * height;
    bool ok;
    if (format == ILM_IMAGE_PFM) {
        ok = fprintf(f, "PF %d %d -1 ", width, height) > 0 && fwrite(rgb, sizeof(float), 3 * n, f) == (size_t)(3 * n);
    } else {
        ok = fprintf(f, "P6 %d %d 255 ", width, height) > 0;
        std::vector<unsigned char> px(3 * n);
        for (int64_t i = 0; i < 3 * n; ++i) {
            int v = (int)(256.0f * rgb[i]);
            px[i] = (unsigned char)(v > 255 ? 255 : (v < 0 ? 0 : v));
        }
        ok = ok && fwrite(px.data(), 1, px.size(), f) == px.size();
    }
    if (fclose(f) != 0) ok = false;
    return ok ? ILM_OK : fail(ILM_ERR_INVALID, std::string("write failed: ") + path);
}

}  // extern "C"
