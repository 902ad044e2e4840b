// Learned denoiser for gfx950: a 4-level encoder-decoder of 3x3 convolutions
// (channels 32/48/64/96, skip connections, ReLU), BF16 weights and activations
// with FP32 accumulation on the matrix cores, trained on the GPU by this file's
// own forward, backward and Adam kernels. There is no counterpart in the
// reference (its only denoiser is the a-trous filter, crates/core/src/denoise.rs);
// the inputs are the reference's Accum channels (gpu_layout.rs:345).
//
// Convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16 with the output
// channels as the GEMM's M side: a wave computes 16 output channels x 16
// consecutive pixels of one row per instruction, the K side walks (tap,
// 16-channel block) pairs. Activations are NHWC BF16 so every lane loads its
// 8 K-values with one 16-byte load; weights are pre-swizzled into per-lane
// fragments so every wave reads them with one coalesced 1 KB load per tile.
// The input-gradient pass is the same kernel over transposed, tap-flipped
// weights; the weight-gradient pass stages 32-pixel chunks through LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "context.hpp"
#include "net.hpp"

namespace lucida {
namespace {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline float b2f(bf16 v) { return (float)v; }
__device__ inline bf16 f2b(float v) { return (bf16)v; }

// ---------------------------------------------------------------------------
// Weight preparation: FP32 master weights W[tap][ci][co] -> BF16 fragments.
// Fragment (chunk c, m-tile mt, lane l, i) holds A[co][k] with co = mt*16 +
// l%16 and k = the i-th of 8 consecutive input channels of K-block
// kb = 2c + (l/16)/2, half (l/16)%2. transpose = 1 builds the input-gradient
// weights W'[tap][co][ci] = W[8 - tap][ci][co].
// ---------------------------------------------------------------------------
__global__ void k_wprep(const float* __restrict__ W, int cin, int cout, int transpose, bf16* __restrict__ frag) {
    const int in_c = transpose ? cout : cin;     // channels the fragment's K walks
    const int out_c = transpose ? cin : cout;    // channels the fragment's M walks
    const int kbs = 9 * in_c / 16;
    const int chunks = (kbs + 1) / 2;
    const int mts = out_c / 16;
    const int total = chunks * mts * 64 * 8;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const int i = e & 7, l = (e >> 3) & 63, rest = e >> 9;
        const int mt = rest % mts, c = rest / mts;
        const int g = l >> 4, r = l & 15;
        const int kb = 2 * c + (g >> 1);
        float v = 0.0f;
        if (kb < kbs) {
            const int tap = kb / (in_c / 16), cb = kb % (in_c / 16);
            const int k_ch = cb * 16 + 8 * (g & 1) + i;  // input channel of the GEMM
            const int m_ch = mt * 16 + r;                 // output channel of the GEMM
            v = transpose ? W[((8 - tap) * cin + m_ch) * cout + k_ch] : W[(tap * cin + k_ch) * cout + m_ch];
        }
        frag[e] = f2b(v);
    }
}

// ---------------------------------------------------------------------------
// Convolution (forward, or input gradient with prepared weights).
// in:  NHWC bf16 [n][h][w][cin]; out: [n][h][w][MT*16]. One wave: NT tiles
// of 16 consecutive pixels of a row (w % (16 * NT) == 0); each K chunk's
// weight fragments are loaded once and used for all NT tiles, and every
// output keeps the same sequence of MFMAs whatever NT is. bias (optional,
// fp32) and ReLU fused in the epilogue.
// ---------------------------------------------------------------------------
template <int MT, int NT>
__global__ __launch_bounds__(256) void k_conv(const bf16* __restrict__ in, int cin, const bf16x8* __restrict__ frag,
                                              const float* __restrict__ bias, int relu, bf16* __restrict__ out,
                                              int n, int h, int w) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * (blockDim.x >> 6)) + (threadIdx.x >> 6);
    const int row_tiles = w / (16 * NT);
    const int tiles = n * h * row_tiles;
    if (wave >= tiles) return;
    const int xt = wave % row_tiles, rowi = wave / row_tiles;
    const int y = rowi % h, img = rowi / h;
    const int x0 = xt * 16 * NT;
    const int g = lane >> 4, r = lane & 15;
    const int cbs = cin / 16;
    const int kbs = 9 * cbs;
    const int chunks = (kbs + 1) / 2;
    f32x4 acc[NT][MT];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int m = 0; m < MT; ++m) acc[t][m] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // A chunk's weight fragments and input columns; chunk c + 1's loads are
    // issued before chunk c's products, so a wave keeps its next loads in
    // flight while its MFMAs run.
    auto load_chunk = [&](int c, bf16x8* av, bf16x8* bv) {
        const int kb = 2 * c + (g >> 1);
        const bool kok = kb < kbs;
        int tap = 0, cb = 0;
        if (kok) {
            tap = kb / cbs;
            cb = kb - tap * cbs;
        }
        const int sy = y + tap / 3 - 1;
        const bool yok = kok && sy >= 0 && sy < h;
#pragma unroll
        for (int m = 0; m < MT; ++m) av[m] = frag[(c * MT + m) * 64 + lane];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int sx = x0 + 16 * t + r + tap % 3 - 1;
            if (yok && sx >= 0 && sx < w) {
                bv[t] = *reinterpret_cast<const bf16x8*>(in + (((size_t)img * h + sy) * w + sx) * cin + cb * 16 +
                                                         8 * (g & 1));
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) bv[t][i] = f2b(0.0f);
            }
        }
    };
    bf16x8 av[MT], bv[NT], av_next[MT], bv_next[NT];
    load_chunk(0, av_next, bv_next);
    for (int c = 0; c < chunks; ++c) {
#pragma unroll
        for (int m = 0; m < MT; ++m) av[m] = av_next[m];
#pragma unroll
        for (int t = 0; t < NT; ++t) bv[t] = bv_next[t];
        if (c + 1 < chunks) load_chunk(c + 1, av_next, bv_next);
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int m = 0; m < MT; ++m)
                acc[t][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bv[t], acc[t][m], 0, 0, 0);
    }
    const int cout = MT * 16;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        bf16* o = out + (((size_t)img * h + y) * w + x0 + 16 * t + r) * cout;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            bf16x4 v;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int co = m * 16 + 4 * g + j;
                float s = acc[t][m][j] + (bias ? bias[co] : 0.0f);
                if (relu) s = fmaxf(s, 0.0f);
                v[j] = f2b(s);
            }
            *reinterpret_cast<bf16x4*>(o + m * 16 + 4 * g) = v;
        }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: dW[tap][ci][co] += sum_p dpre[p][co] * in[p + off(tap)][ci],
// db[co] += sum_p dpre[p][co] (tap 4's blocks only). Grid: (tap, slice); each
// block walks its slice of 32-pixel chunks (32 consecutive pixels in NHWC
// order, the last chunk zero-padded), staging both operands in LDS, and its
// four waves split the (co-tile, ci-tile) pairs, at most WG_MAX_PAIRS each.
// Each block owns its slice's partial sums and stores them to its own row of
// `part` (row length 9 * cin * cout + cout, bias partials at the end from the
// tap-4 blocks); k_wgrad_sum then adds the rows up in slice order, so the
// gradients, and the training built on them, are the same on every run.
// ---------------------------------------------------------------------------
constexpr int WG_MAX_PAIRS = 12;
constexpr int WG_SLICES = 64;

__global__ __launch_bounds__(256) void k_wgrad(const bf16* __restrict__ in, int cin, const bf16* __restrict__ dpre,
                                               int cout, int n, int h, int w, int slices,
                                               float* __restrict__ part) {
    extern __shared__ unsigned char smem_raw[];
    bf16* s_in = reinterpret_cast<bf16*>(smem_raw);  // [32][cin]
    bf16* s_dp = s_in + 32 * cin;                    // [32][cout]
    const int tap = blockIdx.x / slices, slice = blockIdx.x % slices;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int g = lane >> 4, r = lane & 15;
    const int mts = cout / 16, nts = cin / 16;
    const int pairs = mts * nts;
    f32x4 acc[WG_MAX_PAIRS];
#pragma unroll
    for (int q = 0; q < WG_MAX_PAIRS; ++q) acc[q] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    float bsum = 0.0f;  // bias partial for channel threadIdx.x (tap 4 only)
    const size_t pixels = (size_t)n * h * w;
    const int chunks = (int)((pixels + 31) / 32);
    const int per = (chunks + slices - 1) / slices;
    const int c0 = slice * per, c1 = min(chunks, c0 + per);
    for (int ch = c0; ch < c1; ++ch) {
        const size_t p0 = (size_t)ch * 32;
        __syncthreads();
        // Stage: 32 pixels x cin of the shifted input, 32 x cout of dpre (16 B per load).
        for (int e = threadIdx.x; e < 32 * cin / 8; e += 256) {
            const int p = e / (cin / 8), part = e % (cin / 8);
            const size_t pf = p0 + p;
            const int x = (int)(pf % w), y = (int)((pf / w) % h);
            const size_t img = pf / ((size_t)w * h);
            const int sy = y + dy, sx = x + dx;
            bf16x8 v;
            if (pf < pixels && sy >= 0 && sy < h && sx >= 0 && sx < w) {
                v = *reinterpret_cast<const bf16x8*>(in + ((img * h + sy) * w + sx) * cin + part * 8);
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) v[i] = f2b(0.0f);
            }
            *reinterpret_cast<bf16x8*>(s_in + p * cin + part * 8) = v;
        }
        for (int e = threadIdx.x; e < 32 * cout / 8; e += 256) {
            const int p = e / (cout / 8), part = e % (cout / 8);
            bf16x8 v;
            if (p0 + p < pixels) {
                v = *reinterpret_cast<const bf16x8*>(dpre + (p0 + p) * cout + part * 8);
            } else {
#pragma unroll
                for (int i = 0; i < 8; ++i) v[i] = f2b(0.0f);
            }
            *reinterpret_cast<bf16x8*>(s_dp + p * cout + part * 8) = v;
        }
        __syncthreads();
        if (tap == 4 && (int)threadIdx.x < cout)
            for (int p = 0; p < 32; ++p) bsum += b2f(s_dp[p * cout + threadIdx.x]);
        // K = the chunk's 32 pixels; lane's k-range is pixels 8g .. 8g+7.
#pragma unroll
        for (int q = 0; q < WG_MAX_PAIRS; ++q) {
            const int pr = wv + 4 * q;
            if (pr < pairs) {
                const int mt = pr % mts, nt = pr / mts;
                bf16x8 av, bv;
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    av[i] = s_dp[(8 * g + i) * cout + mt * 16 + r];
                    bv[i] = s_in[(8 * g + i) * cin + nt * 16 + r];
                }
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[q], 0, 0, 0);
            }
        }
    }
    // D[m = co][n = ci]: lane holds co = mt*16 + 4g + j, ci = nt*16 + r.
    float* row = part + (size_t)slice * ((size_t)9 * cin * cout + cout);
#pragma unroll
    for (int q = 0; q < WG_MAX_PAIRS; ++q) {
        const int pr = wv + 4 * q;
        if (pr < pairs) {
            const int mt = pr % mts, nt = pr / mts;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int co = mt * 16 + 4 * g + j, ci = nt * 16 + r;
                row[(tap * cin + ci) * cout + co] = acc[q][j];
            }
        }
    }
    if (tap == 4 && (int)threadIdx.x < cout) row[(size_t)9 * cin * cout + threadIdx.x] = bsum;
}

// grad[e] = sum over slices k = 0, 1, ... of part[k][e], in that order.
__global__ __launch_bounds__(256) void k_wgrad_sum(const float* __restrict__ part, size_t row, int slices,
                                                   float* __restrict__ grad) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= row) return;
    float acc = 0.0f;
    for (int k = 0; k < slices; ++k) acc += part[(size_t)k * row + e];
    grad[e] = acc;
}

// ---------------------------------------------------------------------------
// Elementwise and resampling kernels (NHWC bf16, fp32 arithmetic).
// ---------------------------------------------------------------------------

// d <- d * (y > 0): the ReLU's backward, in place.
__global__ void k_relu_bwd(bf16* __restrict__ d, const bf16* __restrict__ y, size_t count) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
        if (!(b2f(y[e]) > 0.0f)) d[e] = f2b(0.0f);
}

// 2x2 average pool: out[n][h/2][w/2][c].
// 2x2 average pool and nearest-upsample + channel concat, eight channels
// (16 bytes) per thread; channel counts are multiples of 16. The pool sums
// the four taps in the same order for every channel.
__global__ void k_pool(const bf16* __restrict__ in, int n, int h, int w, int c, bf16* __restrict__ out) {
    const uint32_t c8 = (uint32_t)c / 8, w2 = (uint32_t)w / 2, h2 = (uint32_t)h / 2;
    const uint32_t total = (uint32_t)n * h2 * w2 * c8;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const uint32_t g = e % c8, p = e / c8;
        const uint32_t x = p % w2, q = p / w2, y = q % h2, img = q / h2;
        const bf16* b = in + (((size_t)img * h + 2 * y) * w + 2 * x) * c + 8 * g;
        const bf16x8 v00 = *reinterpret_cast<const bf16x8*>(b);
        const bf16x8 v01 = *reinterpret_cast<const bf16x8*>(b + c);
        const bf16x8 v10 = *reinterpret_cast<const bf16x8*>(b + (size_t)w * c);
        const bf16x8 v11 = *reinterpret_cast<const bf16x8*>(b + (size_t)w * c + c);
        bf16x8 r;
#pragma unroll
        for (int i = 0; i < 8; ++i) r[i] = f2b(0.25f * (b2f(v00[i]) + b2f(v01[i]) + b2f(v10[i]) + b2f(v11[i])));
        *reinterpret_cast<bf16x8*>(out + (size_t)p * c + 8 * g) = r;
    }
}
// Gradient through the pool plus the skip branch: out = up(dpool)/4 + dskip.
__global__ void k_pool_bwd_add(const bf16* __restrict__ dpool, const bf16* __restrict__ dskip, int n, int h, int w,
                               int c, bf16* __restrict__ out) {
    const size_t total = (size_t)n * h * w * c;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int ch = e % c;
        size_t p = e / c;
        const int x = p % w, y = (p / w) % h, img = p / ((size_t)w * h);
        const float a = b2f(dpool[(((size_t)img * (h / 2) + y / 2) * (w / 2) + x / 2) * c + ch]);
        out[e] = f2b(0.25f * a + (dskip ? b2f(dskip[e]) : 0.0f));
    }
}

// Nearest 2x upsample of `low` (c1 channels) concatenated with `skip` (c2).
__global__ void k_upcat(const bf16* __restrict__ low, int c1, const bf16* __restrict__ skip, int c2, int n, int h,
                        int w, bf16* __restrict__ out) {
    const uint32_t c8 = (uint32_t)(c1 + c2) / 8, l8 = (uint32_t)c1 / 8;
    const uint32_t total = (uint32_t)n * h * w * c8;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const uint32_t g = e % c8, p = e / c8;
        const uint32_t x = p % (uint32_t)w, q = p / (uint32_t)w, y = q % (uint32_t)h, img = q / (uint32_t)h;
        const bf16* src = g < l8 ? low + (((size_t)img * (h / 2) + y / 2) * (w / 2) + x / 2) * c1 + 8 * g
                                 : skip + (size_t)p * c2 + 8 * (g - l8);
        *reinterpret_cast<bf16x8*>(out + (size_t)e * 8) = *reinterpret_cast<const bf16x8*>(src);
    }
}
// Backward of k_upcat: dlow sums its 2x2 block, dskip is a slice.
__global__ void k_upcat_bwd(const bf16* __restrict__ dcat, int c1, int c2, int n, int h, int w, bf16* __restrict__ dlow,
                            bf16* __restrict__ dskip) {
    const int c = c1 + c2;
    const size_t low_total = (size_t)n * (h / 2) * (w / 2) * c1;
    const size_t skip_total = (size_t)n * h * w * c2;
    const size_t total = low_total + skip_total;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        if (e < low_total) {
            const int ch = e % c1;
            size_t p = e / c1;
            const int x = p % (w / 2), y = (p / (w / 2)) % (h / 2), img = p / ((size_t)(w / 2) * (h / 2));
            const bf16* b = dcat + (((size_t)img * h + 2 * y) * w + 2 * x) * c + ch;
            dlow[e] = f2b(b2f(b[0]) + b2f(b[c]) + b2f(b[(size_t)w * c]) + b2f(b[(size_t)w * c + c]));
        } else {
            const size_t f = e - low_total;
            const int ch = f % c2;
            const size_t p = f / c2;
            dskip[f] = dcat[p * c + c1 + ch];
        }
    }
}

// ---------------------------------------------------------------------------
// Features and loss.
// ---------------------------------------------------------------------------

// Input features from an Accum film (means): log1p(radiance / albedo) (3),
// albedo (3), normal (3), depth / (1 + depth) (1), log1p of the demodulated
// luminance's standard error (1), log2(samples) / 16 (1), the a-trous
// estimate log1p(filtered / albedo) (3), one zero channel; the a-trous
// estimate again, in FP32, into base_log (3 floats per pixel). With
// `target` non-null also writes the target log1p(ref radiance / albedo) (3
// floats per pixel) from the reference film, demodulated by the INPUT film's
// albedo so inference can remodulate with the same guide.
__global__ void k_features(const GpuAccum* __restrict__ film, const GpuAccum* __restrict__ base,
                           const GpuAccum* __restrict__ ref, int fw, int x0, int y0, int cw, int ch, int img,
                           bf16* __restrict__ feat, float* __restrict__ base_log, float* __restrict__ target) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cw || y >= ch) return;
    const size_t src = (size_t)(y0 + y) * fw + (x0 + x);
    const float4* a = reinterpret_cast<const float4*>(film + src);
    const float4 a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
    const float ns = fmaxf(a0.w, 1.0f);
    const float inv = 1.0f / ns;
    const float al[3] = {fmaxf(a1.x * inv, 1e-3f), fmaxf(a1.y * inv, 1e-3f), fmaxf(a1.z * inv, 1e-3f)};
    const float rd[3] = {a0.x * inv, a0.y * inv, a0.z * inv};
    const float dep = fmaxf(a1.w * inv, 0.0f);
    float f[16];
    for (int k = 0; k < 3; ++k) f[k] = log1pf(fmaxf(rd[k] / al[k], 0.0f));
    for (int k = 0; k < 3; ++k) f[3 + k] = al[k];
    f[6] = a2.x * inv;
    f[7] = a2.y * inv;
    f[8] = a2.z * inv;
    f[9] = dep / (1.0f + dep);
    // Variance of the per-sample radiance, luminance-weighted, over the mean's
    // standard error, demodulated by the albedo's luminance.
    const float var_l = 0.2126f * fmaxf(a3.x * inv - rd[0] * rd[0], 0.0f) +
                        0.7152f * fmaxf(a3.y * inv - rd[1] * rd[1], 0.0f) +
                        0.0722f * fmaxf(a3.z * inv - rd[2] * rd[2], 0.0f);
    const float al_l = 0.2126f * al[0] + 0.7152f * al[1] + 0.0722f * al[2];
    f[10] = log1pf(sqrtf(var_l * inv) / al_l);
    f[11] = log2f(ns) * (1.0f / 16.0f);
    // The a-trous estimate of the same film, demodulated like channels 0..2:
    // the network predicts a correction to it.
    const float4 b0 = reinterpret_cast<const float4*>(base + src)[0];
    const float binv = 1.0f / fmaxf(b0.w, 1.0f);
    f[12] = log1pf(fmaxf(b0.x * binv / al[0], 0.0f));
    f[13] = log1pf(fmaxf(b0.y * binv / al[1], 0.0f));
    f[14] = log1pf(fmaxf(b0.z * binv / al[2], 0.0f));
    f[15] = 0.0f;
    const size_t dst = ((size_t)img * ch + y) * cw + x;
    // The estimate the output corrects, kept in FP32: in BF16 (8 significant
    // bits) a zero correction could not even reproduce it.
    for (int k = 0; k < 3; ++k) base_log[dst * 3 + k] = f[12 + k];
    bf16x8 lo, hi;
    for (int k = 0; k < 8; ++k) {
        lo[k] = f2b(f[k]);
        hi[k] = f2b(f[8 + k]);
    }
    reinterpret_cast<bf16x8*>(feat + dst * 16)[0] = lo;
    reinterpret_cast<bf16x8*>(feat + dst * 16)[1] = hi;
    if (target) {
        const float4 r0 = reinterpret_cast<const float4*>(ref + src)[0];
        const float rinv = 1.0f / fmaxf(r0.w, 1.0f);
        const float rr[3] = {r0.x * rinv, r0.y * rinv, r0.z * rinv};
        for (int k = 0; k < 3; ++k) target[dst * 3 + k] = log1pf(fmaxf(rr[k] / al[k], 0.0f));
    }
}

// Mean squared error of (output + the FP32 a-trous estimate) against the
// target over the 3 used channels; writes dL/dout (bf16, zero on the
// padding channels) and a per-block partial loss.
__global__ void k_loss(const bf16* __restrict__ out, const float* __restrict__ base_log,
                       const float* __restrict__ target, size_t pixels, bf16* __restrict__ dout,
                       float* __restrict__ loss_partial) {
    __shared__ float red[256];
    float s = 0.0f;
    const float scale = 2.0f / (float)(pixels * 3);
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < pixels; p += (size_t)gridDim.x * blockDim.x) {
        bf16x8 d0, d1;
        for (int k = 0; k < 8; ++k) {
            d0[k] = f2b(0.0f);
            d1[k] = f2b(0.0f);
        }
        for (int k = 0; k < 3; ++k) {
            const float e = (b2f(out[p * 16 + k]) + base_log[p * 3 + k]) - target[p * 3 + k];
            s += e * e;
            d0[k] = f2b(scale * e);
        }
        reinterpret_cast<bf16x8*>(dout + p * 16)[0] = d0;
        reinterpret_cast<bf16x8*>(dout + p * 16)[1] = d1;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_partial[blockIdx.x] = red[0];
}

// Remodulated prediction into an Accum-shaped film: radiance =
// expm1(max(out + FP32 a-trous estimate, 0)) * albedo, samples = 1, guides
// copied as means.
__global__ void k_compose(const bf16* __restrict__ out, const float* __restrict__ base_log,
                          const GpuAccum* __restrict__ film, size_t pixels, GpuAccum* __restrict__ dst) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pixels) return;
    const float4* a = reinterpret_cast<const float4*>(film + p);
    const float4 a1 = a[1], a2 = a[2];
    const float inv = 1.0f / fmaxf(a[0].w, 1.0f);
    const float al[3] = {fmaxf(a1.x * inv, 1e-3f), fmaxf(a1.y * inv, 1e-3f), fmaxf(a1.z * inv, 1e-3f)};
    float rad[3];
    for (int k = 0; k < 3; ++k)
        rad[k] = expm1f(fmaxf(b2f(out[p * 16 + k]) + base_log[p * 3 + k], 0.0f)) * al[k];
    float4* o = reinterpret_cast<float4*>(dst + p);
    o[0] = make_float4(rad[0], rad[1], rad[2], 1.0f);
    o[1] = make_float4(al[0], al[1], al[2], a1.w * inv);
    o[2] = make_float4(a2.x * inv, a2.y * inv, a2.z * inv, 0.0f);
    o[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Adam (Kingma & Ba) on fp32 parameters; step t >= 1.
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, size_t count, float lr, float b1, float b2, float eps, float bc1,
                       float bc2) {
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
        const float gr = grad[e];
        const float mm = b1 * m[e] + (1.0f - b1) * gr;
        const float vv = b2 * v[e] + (1.0f - b2) * gr * gr;
        m[e] = mm;
        v[e] = vv;
        p[e] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    }
}

// The a-trous filter the network corrects: DenoiseParams' defaults
// (crates/core/src/denoise.rs:62-94).
const lucida_denoise_params kAtrousDefaults = {4u, 64.0f, 0.08f, 4.0f, 0.1f};

inline unsigned grid1(size_t n, int block = 256) {
    size_t g = (n + block - 1) / block;
    return (unsigned)(g > 65535 * 16 ? 65535 * 16 : (g ? g : 1));
}

}  // namespace

// --- host side --------------------------------------------------------------------

const NetLayer kNetLayers[kNetNumLayers] = {
    {16, 32, 1},  {32, 32, 1},  {32, 48, 1}, {48, 48, 1}, {48, 64, 1}, {64, 64, 1}, {64, 96, 1}, {96, 96, 1},
    {160, 64, 1}, {64, 64, 1},  {112, 48, 1}, {48, 48, 1}, {80, 32, 1}, {32, 32, 1}, {32, 16, 0},
};

size_t net_param_count() {
    size_t c = 0;
    for (const auto& l : kNetLayers) c += (size_t)9 * l.cin * l.cout + l.cout;
    return c;
}

template <int MT, int NT>
static void launch_conv_t(hipStream_t s, const bf16* in, int cin, const bf16x8* f, const float* bias, int relu,
                          bf16* out, int n, int h, int w) {
    const int tiles = n * h * (w / (16 * NT));
    hipLaunchKernelGGL((k_conv<MT, NT>), dim3((tiles + 3) / 4), dim3(256), 0, s, in, cin, f, bias, relu, out, n, h, w);
}

// Tiles per wave: 4 for 16 output channels, 2 above (more tiles cost
// occupancy the weight reuse does not repay), fewer when the row width does
// not divide.
template <int MT>
static void launch_conv_m(hipStream_t s, const bf16* in, int cin, const bf16x8* f, const float* bias, int relu,
                          bf16* out, int n, int h, int w) {
    constexpr int NT_MAX = MT == 1 ? 4 : 2;
    if (NT_MAX == 4 && w % 64 == 0)
        launch_conv_t<MT, NT_MAX>(s, in, cin, f, bias, relu, out, n, h, w);
    else if (w % 32 == 0)
        launch_conv_t<MT, 2>(s, in, cin, f, bias, relu, out, n, h, w);
    else
        launch_conv_t<MT, 1>(s, in, cin, f, bias, relu, out, n, h, w);
}

static lucida_status launch_conv(hipStream_t s, const bf16* in, int cin, int cout, const bf16* frag, const float* bias,
                                 int relu, bf16* out, int n, int h, int w) {
    const bf16x8* f = reinterpret_cast<const bf16x8*>(frag);
    switch (cout / 16) {
        case 1: launch_conv_m<1>(s, in, cin, f, bias, relu, out, n, h, w); break;
        case 2: launch_conv_m<2>(s, in, cin, f, bias, relu, out, n, h, w); break;
        case 3: launch_conv_m<3>(s, in, cin, f, bias, relu, out, n, h, w); break;
        case 4: launch_conv_m<4>(s, in, cin, f, bias, relu, out, n, h, w); break;
        case 5: launch_conv_m<5>(s, in, cin, f, bias, relu, out, n, h, w); break;
        case 6: launch_conv_m<6>(s, in, cin, f, bias, relu, out, n, h, w); break;
        case 7: launch_conv_m<7>(s, in, cin, f, bias, relu, out, n, h, w); break;
        case 10: launch_conv_m<10>(s, in, cin, f, bias, relu, out, n, h, w); break;
        default: return set_error(LUCIDA_INVALID_ARG, "conv: unsupported output channel count");
    }
    return LUCIDA_OK;
}

static size_t frag_elems(int in_c, int out_c) {
    const int kbs = 9 * in_c / 16;
    return (size_t)((kbs + 1) / 2) * (out_c / 16) * 64 * 8;
}

}  // namespace lucida

using namespace lucida;

// The network object: fp32 master parameters, Adam moments, bf16 fragments,
// and activation / gradient buffers sized for the largest batch seen.
struct lucida_net {
    lucida_ctx* ctx = nullptr;
    DevBuf params, grads, adam_m, adam_v;  // fp32, layout: per layer W[9][cin][cout] then b[cout]
    DevBuf wg_part;                        // weight-gradient partials: WG_SLICES rows of one layer's W and b
    DevBuf frag_fwd, frag_bwd;             // bf16 fragments
    size_t frag_fwd_off[kNetNumLayers], frag_bwd_off[kNetNumLayers], param_off[kNetNumLayers];
    DevBuf acts, dacts, target, loss_partial;
    DevBuf base;  // a-trous estimate of the film being loaded (Accum-shaped)
    DevBuf base_log;  // its log1p(estimate / albedo) per pixel of the batch, FP32 (3 floats)
    size_t act_capacity = 0;
    uint32_t step = 0;
    bool frags_valid = false;
    int n = 0, h = 0, w = 0;
};

namespace lucida {

// Activation tensors of one forward pass, as offsets into net->acts.
struct ActPlan {
    size_t x;          // features [n][h][w][16]
    size_t l[15];      // outputs of layers 0..14
    size_t p0, p1, p2; // pooled skip0, skip1, skip2
    size_t c2, c1, c0; // concatenations before layers 8, 10, 12
    size_t total;
};

static ActPlan plan_acts(int n, int h, int w) {
    ActPlan a{};
    size_t off = 0;
    auto take = [&](int hh, int ww, int c) {
        size_t o = off;
        off += (size_t)n * hh * ww * c;
        off = (off + 63) & ~size_t(63);
        return o;
    };
    const int H[4] = {h, h / 2, h / 4, h / 8}, Wd[4] = {w, w / 2, w / 4, w / 8};
    a.x = take(H[0], Wd[0], 16);
    const int lvl[15] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    for (int i = 0; i < 15; ++i) a.l[i] = take(H[lvl[i]], Wd[lvl[i]], kNetLayers[i].cout);
    a.p0 = take(H[1], Wd[1], 32);
    a.p1 = take(H[2], Wd[2], 48);
    a.p2 = take(H[3], Wd[3], 64);
    a.c2 = take(H[2], Wd[2], 160);
    a.c1 = take(H[1], Wd[1], 112);
    a.c0 = take(H[0], Wd[0], 80);
    a.total = off;
    return a;
}

static lucida_status net_prep_frags(lucida_net* net) {
    hipStream_t s = net->ctx->stream;
    const float* P = net->params.as<float>();
    bf16* ff = net->frag_fwd.as<bf16>();
    bf16* fb = net->frag_bwd.as<bf16>();
    for (int i = 0; i < kNetNumLayers; ++i) {
        const NetLayer& L = kNetLayers[i];
        hipLaunchKernelGGL(k_wprep, dim3(256), dim3(256), 0, s, P + net->param_off[i], L.cin, L.cout, 0,
                           ff + net->frag_fwd_off[i]);
        hipLaunchKernelGGL(k_wprep, dim3(256), dim3(256), 0, s, P + net->param_off[i], L.cin, L.cout, 1,
                           fb + net->frag_bwd_off[i]);
    }
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    net->frags_valid = true;
    return LUCIDA_OK;
}

static lucida_status net_ensure_acts(lucida_net* net, int n, int h, int w) {
    if (w % 128 != 0 || h % 8 != 0 || n < 1)
        return set_error(LUCIDA_INVALID_ARG, "denoiser: width must be a multiple of 128 and height of 8");
    ActPlan a = plan_acts(n, h, w);
    if (a.total > net->act_capacity) {
        LUCIDA_HIP(hipStreamSynchronize(net->ctx->stream), LUCIDA_VALIDATION);
        LUCIDA_HIP(net->acts.ensure(a.total * 2), LUCIDA_OOM);
        LUCIDA_HIP(net->dacts.ensure(a.total * 2), LUCIDA_OOM);
        net->act_capacity = a.total;
    }
    LUCIDA_HIP(net->target.ensure((size_t)n * h * w * 3 * 4), LUCIDA_OOM);
    LUCIDA_HIP(net->base_log.ensure((size_t)n * h * w * 3 * 4), LUCIDA_OOM);
    LUCIDA_HIP(net->loss_partial.ensure(4096 * 4), LUCIDA_OOM);
    net->n = n;
    net->h = h;
    net->w = w;
    return LUCIDA_OK;
}

// Forward over the features already in acts[x].
static lucida_status net_forward(lucida_net* net) {
    if (!net->frags_valid) {
        lucida_status st = net_prep_frags(net);
        if (st != LUCIDA_OK) return st;
    }
    hipStream_t s = net->ctx->stream;
    const int n = net->n, H0 = net->h, W0 = net->w;
    const ActPlan a = plan_acts(n, H0, W0);
    bf16* A = net->acts.as<bf16>();
    const float* P = net->params.as<float>();
    const bf16* F = net->frag_fwd.as<bf16>();
    const int H[4] = {H0, H0 / 2, H0 / 4, H0 / 8}, Wd[4] = {W0, W0 / 2, W0 / 4, W0 / 8};
    auto conv = [&](int li, size_t in_off, int lvl) {
        const NetLayer& L = kNetLayers[li];
        return launch_conv(s, A + in_off, L.cin, L.cout, F + net->frag_fwd_off[li],
                           P + net->param_off[li] + (size_t)9 * L.cin * L.cout, L.relu, A + a.l[li], n, H[lvl],
                           Wd[lvl]);
    };
    auto pool = [&](size_t in_off, int lvl, int c, size_t out_off) {
        size_t total = (size_t)n * H[lvl + 1] * Wd[lvl + 1] * c / 8;
        hipLaunchKernelGGL(k_pool, dim3(grid1(total)), dim3(256), 0, s, A + in_off, n, H[lvl], Wd[lvl], c,
                           A + out_off);
    };
    auto upcat = [&](size_t low, int c1, size_t skip, int c2, int lvl, size_t out_off) {
        size_t total = (size_t)n * H[lvl] * Wd[lvl] * (c1 + c2) / 8;
        hipLaunchKernelGGL(k_upcat, dim3(grid1(total)), dim3(256), 0, s, A + low, c1, A + skip, c2, n, H[lvl],
                           Wd[lvl], A + out_off);
    };
    lucida_status st;
#define TRY(x)               \
    if ((st = (x)) != LUCIDA_OK) return st;
    TRY(conv(0, a.x, 0));
    TRY(conv(1, a.l[0], 0));
    pool(a.l[1], 0, 32, a.p0);
    TRY(conv(2, a.p0, 1));
    TRY(conv(3, a.l[2], 1));
    pool(a.l[3], 1, 48, a.p1);
    TRY(conv(4, a.p1, 2));
    TRY(conv(5, a.l[4], 2));
    pool(a.l[5], 2, 64, a.p2);
    TRY(conv(6, a.p2, 3));
    TRY(conv(7, a.l[6], 3));
    upcat(a.l[7], 96, a.l[5], 64, 2, a.c2);
    TRY(conv(8, a.c2, 2));
    TRY(conv(9, a.l[8], 2));
    upcat(a.l[9], 64, a.l[3], 48, 1, a.c1);
    TRY(conv(10, a.c1, 1));
    TRY(conv(11, a.l[10], 1));
    upcat(a.l[11], 48, a.l[1], 32, 0, a.c0);
    TRY(conv(12, a.c0, 0));
    TRY(conv(13, a.l[12], 0));
    TRY(conv(14, a.l[13], 0));
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    return LUCIDA_OK;
}

// Backward from dL/dout in dacts[l[14]]; accumulates parameter gradients.
static lucida_status net_backward(lucida_net* net) {
    hipStream_t s = net->ctx->stream;
    const int n = net->n, H0 = net->h, W0 = net->w;
    const ActPlan a = plan_acts(n, H0, W0);
    bf16* A = net->acts.as<bf16>();
    bf16* D = net->dacts.as<bf16>();
    float* G = net->grads.as<float>();
    const bf16* FB = net->frag_bwd.as<bf16>();
    const int H[4] = {H0, H0 / 2, H0 / 4, H0 / 8}, Wd[4] = {W0, W0 / 2, W0 / 4, W0 / 8};
    // Every layer's W and b gradients are written whole by its k_wgrad_sum;
    // the partials buffer is sized once, for the largest layer.
    size_t max_row = 0;
    for (int li = 0; li < kNetNumLayers; ++li)
        max_row = std::max(max_row, (size_t)9 * kNetLayers[li].cin * kNetLayers[li].cout + kNetLayers[li].cout);
    LUCIDA_HIP(net->wg_part.ensure((size_t)WG_SLICES * max_row * 4), LUCIDA_OOM);
    lucida_status st;
    // d(layer li's pre-activation) -> weight grads, and d(its input) into d_in.
    auto layer_bwd = [&](int li, size_t in_off, int lvl, size_t d_in_off) -> lucida_status {
        const NetLayer& L = kNetLayers[li];
        const size_t count = (size_t)n * H[lvl] * Wd[lvl] * L.cout;
        if (L.relu)
            hipLaunchKernelGGL(k_relu_bwd, dim3(grid1(count)), dim3(256), 0, s, D + a.l[li], A + a.l[li], count);
        if ((L.cin / 16) * (L.cout / 16) > 4 * WG_MAX_PAIRS)
            return set_error(LUCIDA_INVALID_ARG, "weight gradient: too many channel tiles for one block");
        const size_t lds = (size_t)32 * (L.cin + L.cout) * 2;
        const size_t row = (size_t)9 * L.cin * L.cout + L.cout;  // the layer's W then b, as in `grads`
        hipLaunchKernelGGL(k_wgrad, dim3(9 * WG_SLICES), dim3(256), lds, s, A + in_off, L.cin, D + a.l[li], L.cout,
                           n, H[lvl], Wd[lvl], WG_SLICES, net->wg_part.as<float>());
        hipLaunchKernelGGL(k_wgrad_sum, dim3((unsigned)((row + 255) / 256)), dim3(256), 0, s,
                           net->wg_part.as<float>(), row, WG_SLICES, G + net->param_off[li]);
        if (d_in_off == (size_t)-1) return LUCIDA_OK;
        return launch_conv(s, D + a.l[li], L.cout, L.cin, FB + net->frag_bwd_off[li], nullptr, 0, D + d_in_off, n,
                           H[lvl], Wd[lvl]);
    };
    auto upcat_bwd = [&](size_t dcat, int c1, int c2, int lvl, size_t dlow, size_t dskip) {
        size_t total = (size_t)n * H[lvl] * Wd[lvl] * c2 + (size_t)n * H[lvl + 1] * Wd[lvl + 1] * c1;
        hipLaunchKernelGGL(k_upcat_bwd, dim3(grid1(total)), dim3(256), 0, s, D + dcat, c1, c2, n, H[lvl], Wd[lvl],
                           D + dlow, D + dskip);
    };
    auto pool_bwd_add = [&](size_t dpool, size_t dskip, int lvl, int c, size_t out) {
        size_t total = (size_t)n * H[lvl] * Wd[lvl] * c;
        hipLaunchKernelGGL(k_pool_bwd_add, dim3(grid1(total)), dim3(256), 0, s, D + dpool, D + dskip, n, H[lvl],
                           Wd[lvl], c, D + out);
    };
    // Scratch gradient slots: the d of a concatenation input is computed into the
    // cat buffers' d slots; d(skip) into the pooled buffers' d slots is not
    // needed, the skip gradients go to a.l[*] d slots directly.
    TRY(layer_bwd(14, a.l[13], 0, a.l[13]));
    TRY(layer_bwd(13, a.l[12], 0, a.l[12]));
    TRY(layer_bwd(12, a.c0, 0, a.c0));
    // dcat0 -> d l[11] (low, 48 ch at level 1) and dskip0 (32 ch at level 0) parked in d l[0]'s slot
    upcat_bwd(a.c0, 48, 32, 0, a.l[11], a.l[0]);
    TRY(layer_bwd(11, a.l[10], 1, a.l[10]));
    TRY(layer_bwd(10, a.c1, 1, a.c1));
    upcat_bwd(a.c1, 64, 48, 1, a.l[9], a.l[2]);  // dskip1 parked in d l[2]'s slot
    TRY(layer_bwd(9, a.l[8], 2, a.l[8]));
    TRY(layer_bwd(8, a.c2, 2, a.c2));
    upcat_bwd(a.c2, 96, 64, 2, a.l[7], a.l[4]);  // dskip2 parked in d l[4]'s slot
    TRY(layer_bwd(7, a.l[6], 3, a.l[6]));
    TRY(layer_bwd(6, a.p2, 3, a.p2));
    pool_bwd_add(a.p2, a.l[4], 2, 64, a.l[5]);   // d l[5] = pool_bwd(d p2) + dskip2
    TRY(layer_bwd(5, a.l[4], 2, a.l[4]));
    TRY(layer_bwd(4, a.p1, 2, a.p1));
    pool_bwd_add(a.p1, a.l[2], 1, 48, a.l[3]);   // d l[3] = pool_bwd(d p1) + dskip1
    TRY(layer_bwd(3, a.l[2], 1, a.l[2]));
    TRY(layer_bwd(2, a.p0, 1, a.p0));
    pool_bwd_add(a.p0, a.l[0], 0, 32, a.l[1]);   // d l[1] = pool_bwd(d p0) + dskip0
    TRY(layer_bwd(1, a.l[0], 0, a.l[0]));
    TRY(layer_bwd(0, a.x, 0, (size_t)-1));
#undef TRY
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    return LUCIDA_OK;
}

}  // namespace lucida

extern "C" {

lucida_status lucida_net_create(lucida_ctx* ctx, uint32_t seed, lucida_net** out) {
    if (!ctx || !out) return set_error(LUCIDA_INVALID_ARG, "lucida_net_create: null argument");
    auto* net = new lucida_net();
    net->ctx = ctx;
    const size_t np = net_param_count();
    size_t off = 0, ff = 0, fb = 0;
    for (int i = 0; i < kNetNumLayers; ++i) {
        const NetLayer& L = kNetLayers[i];
        net->param_off[i] = off;
        off += (size_t)9 * L.cin * L.cout + L.cout;
        net->frag_fwd_off[i] = ff;
        ff += frag_elems(L.cin, L.cout);
        net->frag_bwd_off[i] = fb;
        fb += frag_elems(L.cout, L.cin);
    }
    // He-normal initialisation from a counter-based stream (Box-Muller over
    // PCG draws); the input's six padding channels and the output's thirteen
    // padding channels start (and, receiving no gradient, stay) at zero. The
    // output layer starts at zero, so an untrained network returns the a-trous
    // estimate it corrects.
    std::vector<float> host(np, 0.0f);
    uint32_t state = seed * 747796405u + 1u;
    auto uni = [&]() {
        state = state * 747796405u + 2891336453u;
        uint32_t word = ((state >> ((state >> 28) + 4u)) ^ state) * 277803737u;
        return ((float)(((word >> 22) ^ word) >> 8) + 0.5f) * (1.0f / 16777216.0f);
    };
    for (int i = 0; i < kNetNumLayers; ++i) {
        const NetLayer& L = kNetLayers[i];
        const int cin_used = i == 0 ? kNetInputsUsed : L.cin;
        const int cout_used = i == kNetNumLayers - 1 ? 3 : L.cout;
        const float stdv = i == kNetNumLayers - 1 ? 0.0f : std::sqrt(2.0f / (9.0f * cin_used));
        float* W = host.data() + net->param_off[i];
        for (int t = 0; t < 9; ++t)
            for (int ci = 0; ci < L.cin; ++ci)
                for (int co = 0; co < L.cout; ++co) {
                    float u1 = uni(), u2 = uni();
                    float z = std::sqrt(-2.0f * std::log(u1)) * std::cos(6.28318530718f * u2);
                    W[(t * L.cin + ci) * L.cout + co] = (ci < cin_used && co < cout_used) ? z * stdv : 0.0f;
                }
    }
    lucida_status st = LUCIDA_OK;
    if (net->params.ensure(np * 4) != hipSuccess || net->grads.ensure(np * 4) != hipSuccess ||
        net->adam_m.ensure(np * 4) != hipSuccess || net->adam_v.ensure(np * 4) != hipSuccess ||
        net->frag_fwd.ensure(ff * 2) != hipSuccess || net->frag_bwd.ensure(fb * 2) != hipSuccess)
        st = set_error(LUCIDA_OOM, "denoiser parameter allocation failed");
    if (st == LUCIDA_OK &&
        (hipMemcpyAsync(net->params.ptr, host.data(), np * 4, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
         hipMemsetAsync(net->adam_m.ptr, 0, np * 4, ctx->stream) != hipSuccess ||
         hipMemsetAsync(net->adam_v.ptr, 0, np * 4, ctx->stream) != hipSuccess ||
         hipMemsetAsync(net->grads.ptr, 0, np * 4, ctx->stream) != hipSuccess ||
         hipStreamSynchronize(ctx->stream) != hipSuccess))
        st = set_error(LUCIDA_VALIDATION, "denoiser parameter upload failed");
    if (st != LUCIDA_OK) {
        delete net;
        return st;
    }
    *out = net;
    return LUCIDA_OK;
}

void lucida_net_destroy(lucida_net* net) {
    if (!net) return;
    (void)hipStreamSynchronize(net->ctx->stream);
    DevBuf* b[] = {&net->params, &net->grads, &net->adam_m, &net->adam_v, &net->frag_fwd, &net->frag_bwd,
                   &net->acts, &net->dacts, &net->target, &net->loss_partial, &net->base, &net->base_log,
                   &net->wg_part};
    for (DevBuf* x : b) x->release();
    delete net;
}

size_t lucida_net_param_count(void) { return net_param_count(); }

lucida_status lucida_net_get_params(lucida_net* net, float* out, int which) {
    if (!net || !out || which < 0 || which > 1) return set_error(LUCIDA_INVALID_ARG, "lucida_net_get_params: bad argument");
    const DevBuf& b = which == 0 ? net->params : net->grads;
    LUCIDA_HIP(hipMemcpyAsync(out, b.ptr, net_param_count() * 4, hipMemcpyDeviceToHost, net->ctx->stream),
               LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(net->ctx->stream), LUCIDA_READBACK);
    return LUCIDA_OK;
}

lucida_status lucida_net_set_params(lucida_net* net, const float* in) {
    if (!net || !in) return set_error(LUCIDA_INVALID_ARG, "lucida_net_set_params: null argument");
    LUCIDA_HIP(hipMemcpyAsync(net->params.ptr, in, net_param_count() * 4, hipMemcpyHostToDevice, net->ctx->stream),
               LUCIDA_VALIDATION);
    LUCIDA_HIP(hipStreamSynchronize(net->ctx->stream), LUCIDA_VALIDATION);
    net->frags_valid = false;
    return LUCIDA_OK;
}

// Features of a batch of crops: film k contributes the crop (x0[k], y0[k], w, h).
static lucida_status load_batch(lucida_net* net, const lucida_film* const* inputs, const lucida_film* const* refs,
                                uint32_t count, const uint32_t* x0, const uint32_t* y0, uint32_t w, uint32_t h) {
    if (count > 4096 || w > 16384 || h > 16384 || (uint64_t)count * w * h > (1ull << 26))
        return set_error(LUCIDA_INVALID_ARG, "denoiser: batch larger than 2^26 pixels");
    for (uint32_t k = 0; k < count; ++k) {
        const lucida_film* f = inputs[k];
        if (!f || (refs && !refs[k])) return set_error(LUCIDA_INVALID_ARG, "denoiser: null film");
        const uint64_t cx = x0 ? x0[k] : 0, cy = y0 ? y0[k] : 0;
        if (cx + w > f->width || cy + h > f->height) return set_error(LUCIDA_INVALID_ARG, "crop outside the film");
        if (refs && (refs[k]->width != f->width || refs[k]->height != f->height))
            return set_error(LUCIDA_INVALID_ARG, "reference film size differs");
    }
    lucida_status st = net_ensure_acts(net, (int)count, (int)h, (int)w);
    if (st != LUCIDA_OK) return st;
    size_t max_pixels = 0;
    for (uint32_t k = 0; k < count; ++k)
        max_pixels = std::max<size_t>(max_pixels, (size_t)inputs[k]->width * inputs[k]->height);
    if (max_pixels * sizeof(GpuAccum) > net->base.bytes) {
        LUCIDA_HIP(hipStreamSynchronize(net->ctx->stream), LUCIDA_VALIDATION);
        LUCIDA_HIP(net->base.ensure(max_pixels * sizeof(GpuAccum)), LUCIDA_OOM);
    }
    st = atrous_reserve(net->ctx, max_pixels);
    if (st != LUCIDA_OK) return st;
    const ActPlan a = plan_acts(net->n, net->h, net->w);
    bf16* X = net->acts.as<bf16>() + a.x;
    for (uint32_t k = 0; k < count; ++k) {
        const lucida_film* f = inputs[k];
        const uint32_t cx = x0 ? x0[k] : 0, cy = y0 ? y0[k] : 0;
        // The whole film is filtered (crops then see the filter's real
        // neighbourhood); stream order keeps one scratch film enough.
        atrous_enqueue(net->ctx, f->accum, f->width, f->height, kAtrousDefaults, net->base.as<GpuAccum>());
        hipLaunchKernelGGL(k_features, dim3((w + 127) / 128, h), dim3(128), 0, net->ctx->stream, f->accum,
                           net->base.as<GpuAccum>(), refs ? refs[k]->accum : nullptr, (int)f->width, (int)cx, (int)cy,
                           (int)w, (int)h, (int)k, X, net->base_log.as<float>(),
                           refs ? net->target.as<float>() : nullptr);
    }
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    return LUCIDA_OK;
}

lucida_status lucida_net_forward_backward(lucida_net* net, const lucida_film* const* inputs,
                                          const lucida_film* const* refs, uint32_t count, const uint32_t* x0,
                                          const uint32_t* y0, uint32_t w, uint32_t h, float* loss) {
    if (!net || !inputs || !refs || count == 0) return set_error(LUCIDA_INVALID_ARG, "lucida_net_forward_backward: bad argument");
    clear_error();
    lucida_status st = load_batch(net, inputs, refs, count, x0, y0, w, h);
    if (st != LUCIDA_OK) return st;
    st = net_forward(net);
    if (st != LUCIDA_OK) return st;
    const ActPlan a = plan_acts(net->n, net->h, net->w);
    const size_t pixels = (size_t)net->n * net->h * net->w;
    const unsigned blocks = 1024;
    hipLaunchKernelGGL(k_loss, dim3(blocks), dim3(256), 0, net->ctx->stream, net->acts.as<bf16>() + a.l[14],
                       net->base_log.as<float>(), net->target.as<float>(), pixels, net->dacts.as<bf16>() + a.l[14],
                       net->loss_partial.as<float>());
    st = net_backward(net);
    if (st != LUCIDA_OK) return st;
    if (loss) {
        std::vector<float> part(blocks);
        LUCIDA_HIP(hipMemcpyAsync(part.data(), net->loss_partial.ptr, blocks * 4, hipMemcpyDeviceToHost,
                                  net->ctx->stream),
                   LUCIDA_READBACK);
        LUCIDA_HIP(hipStreamSynchronize(net->ctx->stream), LUCIDA_READBACK);
        double s = 0;
        for (float v : part) s += v;
        *loss = (float)(s / (double)(pixels * 3));
    }
    return LUCIDA_OK;
}

lucida_status lucida_net_adam(lucida_net* net, float lr) {
    if (!net) return set_error(LUCIDA_INVALID_ARG, "lucida_net_adam: null net");
    net->step += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - std::pow(b1, (float)net->step), bc2 = 1.0f - std::pow(b2, (float)net->step);
    const size_t np = net_param_count();
    hipLaunchKernelGGL(k_adam, dim3(grid1(np)), dim3(256), 0, net->ctx->stream, net->params.as<float>(),
                       net->grads.as<float>(), net->adam_m.as<float>(), net->adam_v.as<float>(), np, lr, b1, b2, eps,
                       bc1, bc2);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    return net_prep_frags(net);
}

lucida_status lucida_net_train_step(lucida_net* net, const lucida_film* const* inputs, const lucida_film* const* refs,
                                    uint32_t count, const uint32_t* x0, const uint32_t* y0, uint32_t w, uint32_t h,
                                    float lr, float* loss) {
    lucida_status st = lucida_net_forward_backward(net, inputs, refs, count, x0, y0, w, h, loss);
    if (st != LUCIDA_OK) return st;
    return lucida_net_adam(net, lr);
}

lucida_status lucida_net_denoise(lucida_net* net, const lucida_film* in, lucida_film* out) {
    if (!net || !in || !out) return set_error(LUCIDA_INVALID_ARG, "lucida_net_denoise: null argument");
    if (in == out || in->width != out->width || in->height != out->height)
        return set_error(LUCIDA_INVALID_ARG, "lucida_net_denoise: films must differ and share a size");
    clear_error();
    const lucida_film* ins[1] = {in};
    lucida_status st = load_batch(net, ins, nullptr, 1, nullptr, nullptr, in->width, in->height);
    if (st != LUCIDA_OK) return st;
    st = net_forward(net);
    if (st != LUCIDA_OK) return st;
    const ActPlan a = plan_acts(net->n, net->h, net->w);
    const size_t pixels = (size_t)in->width * in->height;
    hipLaunchKernelGGL(k_compose, dim3(grid1(pixels)), dim3(256), 0, net->ctx->stream, net->acts.as<bf16>() + a.l[14],
                       net->base_log.as<float>(), in->accum, pixels, out->accum);
    LUCIDA_HIP(hipGetLastError(), LUCIDA_SHADER);
    LUCIDA_HIP(hipStreamSynchronize(net->ctx->stream), LUCIDA_VALIDATION);
    return LUCIDA_OK;
}

// Copies one activation tensor of the last forward pass (or its gradient) to
// the host as floats, for parity tests: index 0 = features, 1..15 = layers.
lucida_status lucida_net_read_tensor(lucida_net* net, int index, int gradient, float* out, size_t count) {
    if (!net || !out || index < 0 || index > 15) return set_error(LUCIDA_INVALID_ARG, "lucida_net_read_tensor: bad argument");
    if (net->n == 0) return set_error(LUCIDA_INVALID_ARG, "lucida_net_read_tensor: no forward pass yet");
    const ActPlan a = plan_acts(net->n, net->h, net->w);
    const size_t off = index == 0 ? a.x : a.l[index - 1];
    const int lvl_of[16] = {0, 0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    const int c = index == 0 ? 16 : kNetLayers[index - 1].cout;
    const int div = 1 << lvl_of[index];
    const size_t want = (size_t)net->n * (net->h / div) * (net->w / div) * c;
    if (count != want) return set_error(LUCIDA_INVALID_ARG, "lucida_net_read_tensor: wrong element count");
    std::vector<uint16_t> tmp(want);
    const DevBuf& b = gradient ? net->dacts : net->acts;
    LUCIDA_HIP(hipMemcpyAsync(tmp.data(), b.as<uint16_t>() + off, want * 2, hipMemcpyDeviceToHost, net->ctx->stream),
               LUCIDA_READBACK);
    LUCIDA_HIP(hipStreamSynchronize(net->ctx->stream), LUCIDA_READBACK);
    for (size_t i = 0; i < want; ++i) {
        uint32_t bits = (uint32_t)tmp[i] << 16;
        std::memcpy(&out[i], &bits, 4);
    }
    return LUCIDA_OK;
}

}  // extern "C"
