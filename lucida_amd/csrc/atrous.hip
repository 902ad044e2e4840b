// Edge-avoiding a-trous denoiser for gfx950.
//
// Reference: shaders/denoise/atrous.wgsl:73-198 with host driver
// crates/gpu/src/denoise.rs:301-408; CPU twin crates/core/src/denoise.rs:134.
// Same algorithm: demodulate by albedo (floored at 1e-3), a 3x3 luminance
// deviation measured once, `iterations` passes of the 5x5 B-spline kernel at
// stride 2^level weighted by normal, depth, luminance and albedo stops, then
// remodulate. Changes for the GPU: the guides are packed once into two
// 16-byte planes (normal+depth, albedo+luminance) so each tap reads 48 bytes
// instead of a 64-byte Accum plus the colour, the colour plane carries its
// luminance in .w, the deviation is fused into the demodulate pass's
// successor, and the remodulation is fused into the last filter pass.
#include <hip/hip_runtime.h>

#include "context.hpp"
#include "shading.hpp"

namespace lucida {
namespace {

using namespace dev;

constexpr int TX = 16, TY = 8;

__device__ inline float lum3(float x, float y, float z) { return x * 0.2126f + y * 0.7152f + z * 0.0722f; }

// Means of the accumulated channels, and the demodulated colour.
__global__ __launch_bounds__(TX* TY) void k_dn_prepare(const GpuAccum* __restrict__ in, uint32_t w, uint32_t h,
                                                       float4* __restrict__ color, float4* __restrict__ gnd,
                                                       float4* __restrict__ galb) {
    const uint32_t x = blockIdx.x * TX + threadIdx.x, y = blockIdx.y * TY + threadIdx.y;
    if (x >= w || y >= h) return;
    const uint32_t i = y * w + x;
    const float4* a = reinterpret_cast<const float4*>(in + i);
    const float4 a0 = a[0], a1 = a[1], a2 = a[2];
    const float inv = 1.0f / fmaxf(a0.w, 1.0f);
    const V3 alb = vmax(V3{a1.x * inv, a1.y * inv, a1.z * inv}, v3s(1.0e-3f));
    const V3 c = V3{a0.x * inv, a0.y * inv, a0.z * inv} / alb;
    color[i] = make_float4(c.x, c.y, c.z, luminance(c));
    gnd[i] = make_float4(a2.x * inv, a2.y * inv, a2.z * inv, a1.w * inv);
    galb[i] = make_float4(alb.x, alb.y, alb.z, luminance(alb));
}

// denoise.rs:111 local_deviation: 3x3 luminance standard deviation.
__global__ __launch_bounds__(TX* TY) void k_dn_deviation(const float4* __restrict__ color, uint32_t w, uint32_t h,
                                                         float* __restrict__ dev) {
    const int x = blockIdx.x * TX + threadIdx.x, y = blockIdx.y * TY + threadIdx.y;
    if (x >= (int)w || y >= (int)h) return;
    float sum = 0.0f, sum_sq = 0.0f, n = 0.0f;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            const int px = x + dx, py = y + dy;
            if (px < 0 || py < 0 || px >= (int)w || py >= (int)h) continue;
            const float l = color[py * w + px].w;
            sum += l;
            sum_sq += l * l;
            n += 1.0f;
        }
    const float mean = sum / n;
    dev[y * w + x] = sqrtf(fmaxf(sum_sq / n - mean * mean, 0.0f));
}

// One a-trous pass (atrous.wgsl:129). When `out` is non-null this is the last
// pass and it writes the remodulated Accum-shaped result instead.
__global__ __launch_bounds__(TX* TY) void k_dn_pass(const float4* __restrict__ src, float4* __restrict__ dst,
                                                    const float4* __restrict__ gnd, const float4* __restrict__ galb,
                                                    const float* __restrict__ dev, uint32_t w, uint32_t h,
                                                    int stride, lucida_denoise_params prm, GpuAccum* __restrict__ out) {
    const int x = blockIdx.x * TX + threadIdx.x, y = blockIdx.y * TY + threadIdx.y;
    if (x >= (int)w || y >= (int)h) return;
    const float K[5] = {1.0f / 16.0f, 1.0f / 4.0f, 3.0f / 8.0f, 1.0f / 4.0f, 1.0f / 16.0f};
    const uint32_t ci = y * w + x;
    const float4 cg = gnd[ci];
    const float4 cc = src[ci];
    const float ca = fmaxf(galb[ci].w, 1.0e-3f);
    const float lum_scale = prm.sigma_luminance * fmaxf(dev[ci], 1.0e-4f);
    const float sd = prm.sigma_depth * fmaxf(fabsf(cg.w), 1.0e-3f);
    const float sa = prm.sigma_albedo * ca;
    // The edge stops multiply to one exponential: nd^s * e^-(a+b+c) =
    // 2^(s*log2(nd) - (a+b+c)*log2(e)), with the stops' reciprocals taken
    // once per pixel and the hardware's exp2 / log2 (a relative error near
    // 1e-6 in a weight). nd^0 = 1 also at nd = 0.
    const float LOG2E = 1.4426950408889634f;
    const float rd = LOG2E / sd, rl = LOG2E / lum_scale, ra = LOG2E / sa;
    const float sn = prm.sigma_normal;
    float sx = 0.0f, sy = 0.0f, sz = 0.0f, ws = 0.0f;
    for (int ky = 0; ky < 5; ++ky) {
        const int py = y + (ky - 2) * stride;
        if (py < 0 || py >= (int)h) continue;
        for (int kx = 0; kx < 5; ++kx) {
            const int px = x + (kx - 2) * stride;
            if (px < 0 || px >= (int)w) continue;
            const uint32_t qi = py * w + px;
            const float4 qg = gnd[qi];
            const float4 qc = src[qi];
            const float qa = galb[qi].w;
            const float nd = fmaxf(cg.x * qg.x + cg.y * qg.y + cg.z * qg.z, 0.0f);
            const float ln = sn == 0.0f ? 0.0f : (nd > 0.0f ? sn * __builtin_amdgcn_logf(nd) : -INFINITY);
            const float e = ln - fabsf(cg.w - qg.w) * rd - fabsf(cc.w - qc.w) * rl - fabsf(ca - qa) * ra;
            const float wt = K[kx] * K[ky] * __builtin_amdgcn_exp2f(e);
            sx += qc.x * wt;
            sy += qc.y * wt;
            sz += qc.z * wt;
            ws += wt;
        }
    }
    float4 r = cc;
    if (ws > 0.0f) {
        r.x = sx / ws;
        r.y = sy / ws;
        r.z = sz / ws;
        r.w = lum3(r.x, r.y, r.z);
    }
    if (!out) {
        dst[ci] = r;
        return;
    }
    const float4 al = galb[ci];
    float4* o = reinterpret_cast<float4*>(out + ci);
    o[0] = make_float4(r.x * al.x, r.y * al.y, r.z * al.z, 1.0f);
    o[1] = make_float4(al.x, al.y, al.z, cg.w);
    o[2] = make_float4(cg.x, cg.y, cg.z, 0.0f);
    o[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Zero iterations: remodulate the demodulated colour unchanged.
__global__ __launch_bounds__(TX* TY) void k_dn_modulate(const float4* __restrict__ src, const float4* __restrict__ gnd,
                                                        const float4* __restrict__ galb, uint32_t w, uint32_t h,
                                                        GpuAccum* __restrict__ out) {
    const uint32_t x = blockIdx.x * TX + threadIdx.x, y = blockIdx.y * TY + threadIdx.y;
    if (x >= w || y >= h) return;
    const uint32_t ci = y * w + x;
    const float4 r = src[ci], al = galb[ci], cg = gnd[ci];
    float4* o = reinterpret_cast<float4*>(out + ci);
    o[0] = make_float4(r.x * al.x, r.y * al.y, r.z * al.z, 1.0f);
    o[1] = make_float4(al.x, al.y, al.z, cg.w);
    o[2] = make_float4(cg.x, cg.y, cg.z, 0.0f);
    o[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

}  // namespace

// Grows the context's a-trous planes to `pixels` (waits for queued work first).
lucida_status atrous_reserve(lucida_ctx* ctx, size_t pixels) {
    const size_t want = pixels * (16 * 4 + 4);
    if (want > ctx->dn_buf.bytes) {
        LUCIDA_HIP(hipStreamSynchronize(ctx->stream), LUCIDA_VALIDATION);
        LUCIDA_HIP(ctx->dn_buf.ensure(want), LUCIDA_OOM);
    }
    return LUCIDA_OK;
}

// Queues the filter of a w x h Accum film into `out` on the context's stream;
// the planes must have been reserved for w * h pixels.
void atrous_enqueue(lucida_ctx* ctx, const GpuAccum* in, uint32_t w, uint32_t h, const lucida_denoise_params& params,
                    GpuAccum* out) {
    const size_t n = (size_t)w * h;
    float4* a = ctx->dn_buf.as<float4>();
    float4* b = a + n;
    float4* gnd = a + 2 * n;
    float4* galb = a + 3 * n;
    float* dev = reinterpret_cast<float*>(a + 4 * n);
    const dim3 blk(TX, TY), grd((w + TX - 1) / TX, (h + TY - 1) / TY);
    hipStream_t s = ctx->stream;
    hipLaunchKernelGGL(k_dn_prepare, grd, blk, 0, s, in, w, h, a, gnd, galb);
    if (params.iterations == 0) {
        hipLaunchKernelGGL(k_dn_modulate, grd, blk, 0, s, a, gnd, galb, w, h, out);
    } else {
        hipLaunchKernelGGL(k_dn_deviation, grd, blk, 0, s, a, w, h, dev);
        for (uint32_t level = 0; level < params.iterations; ++level) {
            const bool last = level + 1 == params.iterations;
            hipLaunchKernelGGL(k_dn_pass, grd, blk, 0, s, a, b, gnd, galb, dev, w, h, 1 << level, params,
                               last ? out : nullptr);
            std::swap(a, b);
        }
    }
}

}  // namespace lucida

using namespace lucida;

extern "C" {

lucida_status lucida_denoise_atrous(lucida_ctx* ctx, const lucida_film* in, const lucida_denoise_params* params,
                                    lucida_film* out) {
    if (!ctx || !in || !params || !out) return set_error(LUCIDA_INVALID_ARG, "lucida_denoise_atrous: null argument");
    if (in == out) return set_error(LUCIDA_INVALID_ARG, "lucida_denoise_atrous: input and output must differ");
    if (in->width != out->width || in->height != out->height)
        return set_error(LUCIDA_INVALID_ARG, "lucida_denoise_atrous: film sizes differ");
    if (params->iterations > 16) return set_error(LUCIDA_INVALID_ARG, "lucida_denoise_atrous: at most 16 iterations");
    clear_error();
    lucida_status st = atrous_reserve(ctx, (size_t)in->width * in->height);
    if (st != LUCIDA_OK) return st;
    atrous_enqueue(ctx, in->accum, in->width, in->height, *params, out->accum);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    LUCIDA_HIP(hipStreamSynchronize(ctx->stream), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

}  // extern "C"
