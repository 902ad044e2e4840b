// The convergence meter (crates/gpu/src/stats.rs:27-173,
// shaders/stats/convergence.wgsl): a film's accumulated sums and sums of
// squares reduced to one number, the mean over pixels of the relative
// standard error of the pixel's mean, sqrt(luminance(Var[sample]) / N) /
// luminance(mean). Pixels with fewer than two samples or a mean luminance of
// at most 1e-3 are left out; -1 reports that no pixel qualified.
//
// Two launches: a grid-stride pass with one block per CU (each thread
// reads whole 64-byte Accum records, consecutive threads consecutive pixels)
// whose blocks reduce their (error, count) pairs through wave shuffles and
// LDS into one partial each, then one block folding the partials.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "context.hpp"

namespace lucida {
namespace {

constexpr int CONV_BLOCK = 256;
constexpr float MIN_LUMINANCE = 1.0e-3f;

__device__ inline float lum(float r, float g, float b) { return r * 0.2126f + g * 0.7152f + b * 0.0722f; }

__device__ inline float2 block_sum(float2 v, float2* red) {
    for (int off = 32; off > 0; off >>= 1) {
        v.x += __shfl_down(v.x, off, 64);
        v.y += __shfl_down(v.y, off, 64);
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) red[wave] = v;
    __syncthreads();
    float2 s = make_float2(0.0f, 0.0f);
    if (threadIdx.x == 0)
        for (int w = 0; w < CONV_BLOCK / 64; ++w) {
            s.x += red[w].x;
            s.y += red[w].y;
        }
    return s;
}

__global__ __launch_bounds__(CONV_BLOCK) void k_conv_reduce(const GpuAccum* __restrict__ accum, uint32_t pixels,
                                                            float2* __restrict__ partials) {
    __shared__ float2 red[CONV_BLOCK / 64];
    float2 acc = make_float2(0.0f, 0.0f);
    for (uint32_t i = blockIdx.x * CONV_BLOCK + threadIdx.x; i < pixels; i += gridDim.x * CONV_BLOCK) {
        const float4* a = reinterpret_cast<const float4*>(accum + i);
        const float4 a0 = a[0], a3 = a[3];
        const float n = a0.w;
        if (!(n >= 2.0f)) continue;
        const float inv = 1.0f / n;
        const float mr = a0.x * inv, mg = a0.y * inv, mb = a0.z * inv;
        const float mean_l = lum(mr, mg, mb);
        if (!(mean_l > MIN_LUMINANCE)) continue;
        const float vr = fmaxf(a3.x * inv - mr * mr, 0.0f), vg = fmaxf(a3.y * inv - mg * mg, 0.0f),
                    vb = fmaxf(a3.z * inv - mb * mb, 0.0f);
        acc.x += sqrtf(lum(vr, vg, vb) * inv) / mean_l;
        acc.y += 1.0f;
    }
    const float2 s = block_sum(acc, red);
    if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

__global__ __launch_bounds__(CONV_BLOCK) void k_conv_finish(const float2* __restrict__ partials, uint32_t count,
                                                            float* __restrict__ result) {
    __shared__ float2 red[CONV_BLOCK / 64];
    float2 acc = make_float2(0.0f, 0.0f);
    for (uint32_t i = threadIdx.x; i < count; i += CONV_BLOCK) {
        acc.x += partials[i].x;
        acc.y += partials[i].y;
    }
    const float2 s = block_sum(acc, red);
    if (threadIdx.x == 0) result[0] = s.y > 0.0f ? s.x / s.y : -1.0f;
}

}  // namespace
}  // namespace lucida

using namespace lucida;

extern "C" {

lucida_status lucida_film_convergence(lucida_film* film, float* out) {
    if (!film || !out) return set_error(LUCIDA_INVALID_ARG, "lucida_film_convergence: null argument");
    clear_error();
    lucida_ctx* ctx = film->ctx;
    const uint32_t pixels = film->width * film->height;
    const uint32_t blocks = (uint32_t)std::max(1, std::min<int>(ctx->cu_count, (int)((pixels + 255) / 256)));
    LUCIDA_HIP(ctx->conv_buf.ensure((size_t)blocks * sizeof(float2) + 16), LUCIDA_OOM);
    float2* partials = ctx->conv_buf.as<float2>();
    float* result = reinterpret_cast<float*>(partials + blocks);
    hipLaunchKernelGGL(k_conv_reduce, dim3(blocks), dim3(CONV_BLOCK), 0, ctx->stream, film->accum, pixels, partials);
    hipLaunchKernelGGL(k_conv_finish, dim3(1), dim3(CONV_BLOCK), 0, ctx->stream, partials, blocks, result);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    LUCIDA_HIP(hipMemcpyAsync(out, result, sizeof(float), hipMemcpyDeviceToHost, ctx->stream), LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(ctx->stream), LUCIDA_READBACK);
    return LUCIDA_OK;
}

}  // extern "C"
