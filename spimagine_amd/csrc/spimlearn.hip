// spimlearn.hip -- libspimlearn.so (include/spimlearn.h): a small convolutional network that
// reconstructs a full-resolution max projection from a half-resolution, quarter-step one,
// with inference and training on gfx950.
//
// The 32->32 3x3 convolutions (forward and input-gradient) are implicit GEMMs on the matrix
// cores: v_mfma_f32_32x32x16_bf16, A = weights (32 output channels x 16 input channels of
// one tap), B = 32 consecutive pixels of one row x the same 16 input channels, FP32
// accumulation over 9 taps x 2 channel halves.  A workgroup computes a 32 x 16 pixel tile;
// its (16+2) x (32+2) x 32-channel bf16 input halo sits in LDS with an 80-byte pixel stride
// (16 lanes' 16-byte reads then touch 16 distinct 4-bank groups).  Weights live in
// registers as pre-packed fragments (18 x 16 bytes per lane).  The thin layers (2->32 input,
// 32->4 head fused with the pixel shuffle and the bilinear base), pooling, upsampling, the
// loss and the weight gradients (FP32 VALU, deterministic per-block partials + reduction)
// are plain kernels.  Activations: bf16, channels fastest (n, h, w, 32).  Gradients: FP32.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "spimlearn.h"

namespace {

constexpr int C = 32;
// 32->32 layers in parameter order
enum Layer { kE0, kE0b, kE1, kE1b, kE2a, kE2, kU1, kU0, kU0b, kLayers };
constexpr int kInW = 9 * 2 * C, kInParams = kInW + C;
constexpr int kL = 9 * C * C, kLParams = kL + C;
constexpr int kHeadW = 9 * C * 4, kHeadParams = kHeadW + 4;
constexpr int kOffIn = 0;
constexpr int kOffL0 = kInParams;
constexpr int kOffHead = kOffL0 + kLayers * kLParams;
constexpr int kParams = kOffHead + kHeadParams;  // 84,996
constexpr int kFragPerLayer = 18 * 64;           // uint4 fragments per layer
constexpr int kMaxWgradBlocks = 512;

typedef __bf16 bf16x8 __attribute__((__vector_size__(16)));
typedef float f32x16 __attribute__((__vector_size__(64)));

__device__ __forceinline__ uint32_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ float bf2f(uint32_t b) { return __uint_as_float(b << 16); }
__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xFFFF0000u); }
__device__ __forceinline__ uint32_t pack2(float a, float b) { return f2bf(a) | (f2bf(b) << 16); }
// ReLU derivative from the stored bf16 activation: strictly positive
__device__ __forceinline__ bool bf_pos(uint32_t b) { return (b & 0x8000u) == 0u && (b & 0x7FFFu) != 0u; }

// ------------------------------------------------------------------ 32->32 conv (MFMA)
constexpr int kTW = 32, kTH = 16, kLW = kTW + 2, kLH = kTH + 2, kPixU4 = 5;

// The (kTH + 2) x (kTW + 2) halo tile of bf16 activations (C channels, 80-byte pixel
// stride) at (x0 - 1, y0 - 1) of image img into LDS, zero outside the image: every chunk's
// load issued before the first store (one round trip to memory instead of one per chunk).
__device__ __forceinline__ void load_halo_tile(const uint16_t* __restrict__ in, uint4* tile,
                                               int x0, int y0, int img, int h, int w) {
    constexpr int kChunks = kLH * kLW * 4, kIter = (kChunks + 255) / 256;
    const int tid = threadIdx.x;
    const size_t plane = (size_t)h * w;
    uint4 v[kIter];
#pragma unroll
    for (int k = 0; k < kIter; ++k) {
        const int c = tid + 256 * k;
        const int pix = c >> 2, q = c & 3;
        const int ty = pix / kLW, tx = pix - ty * kLW;
        const int gy = y0 - 1 + ty, gx = x0 - 1 + tx;
        v[k] = make_uint4(0u, 0u, 0u, 0u);
        if (c < kChunks && gy >= 0 && gy < h && gx >= 0 && gx < w)
            v[k] = *(const uint4*)(in + ((size_t)img * plane + (size_t)gy * w + gx) * C + q * 8);
    }
#pragma unroll
    for (int k = 0; k < kIter; ++k) {
        const int c = tid + 256 * k;
        if (c < kChunks) tile[(c >> 2) * kPixU4 + (c & 3)] = v[k];
    }
}

// DGRAD = false: out(bf16) = relu(conv(in bf16) + bias).
// DGRAD = true : out(f32)  = conv(in f32 * [mask > 0]) with the input-gradient fragments.
template <bool DGRAD>
__global__ __launch_bounds__(256) void k_conv32(const void* __restrict__ in,
                                                const uint16_t* __restrict__ mask,
                                                const uint4* __restrict__ frag,
                                                const float* __restrict__ bias,
                                                void* __restrict__ out, int h, int w) {
    __shared__ uint4 tile[kLH * kLW * kPixU4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int x0 = blockIdx.x * kTW, y0 = blockIdx.y * kTH, img = blockIdx.z;
    const size_t plane = (size_t)h * w;
    constexpr int kChunks = kLH * kLW * 4;  // 16-byte chunks of the halo tile
    if constexpr (!DGRAD) load_halo_tile((const uint16_t*)in, tile, x0, y0, img, h, w);
    for (int c = DGRAD ? tid : kChunks; c < kChunks; c += 256) {
        const int pix = c >> 2, q = c & 3;
        const int ty = pix / kLW, tx = pix - ty * kLW;
        const int gy = y0 - 1 + ty, gx = x0 - 1 + tx;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (gy >= 0 && gy < h && gx >= 0 && gx < w) {
            const size_t p = (size_t)img * plane + (size_t)gy * w + gx;
            if constexpr (DGRAD) {
                const float4* src = (const float4*)((const float*)in + p * C + q * 8);
                const float4 a = src[0], b = src[1];
                float f[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
                const uint4 m = *(const uint4*)(mask + p * C + q * 8);
                const uint32_t mw[4] = {m.x, m.y, m.z, m.w};
#pragma unroll
                for (int j = 0; j < 8; ++j)
                    if (!bf_pos((mw[j >> 1] >> (16 * (j & 1))) & 0xFFFFu)) f[j] = 0.0f;
                v = make_uint4(pack2(f[0], f[1]), pack2(f[2], f[3]), pack2(f[4], f[5]),
                               pack2(f[6], f[7]));
            } else {
                v = *(const uint4*)((const uint16_t*)in + p * C + q * 8);
            }
        }
        tile[pix * kPixU4 + q] = v;
    }
    uint4 wf[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) wf[k] = frag[k * 64 + lane];
    __syncthreads();
    const int px = lane & 31, kh = lane >> 5;
    const int ox = x0 + px;
    for (int rr = 0; rr < kTH / 4; ++rr) {
        const int r = wv * (kTH / 4) + rr, oy = y0 + r;
        if (oy >= h) break;  // uniform over the wave: the MFMAs below run with all lanes
        f32x16 acc = {};
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int base = ((r + t / 3) * kLW + px + t % 3) * kPixU4 + kh;
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const uint4 b = tile[base + 2 * hh];
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    __builtin_bit_cast(bf16x8, wf[2 * t + hh]), __builtin_bit_cast(bf16x8, b),
                    acc, 0, 0, 0);
            }
        }
        if (ox < w) {
            // accumulator register i of lane l: output channel 8 (i/4) + 4 (l/32) + i%4,
            // pixel l%32
            const size_t p = (size_t)img * plane + (size_t)oy * w + ox;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int co0 = 8 * g + 4 * kh;
                if constexpr (DGRAD) {
                    *(float4*)((float*)out + p * C + co0) =
                        make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
                } else {
                    const float v0 = fmaxf(acc[4 * g] + bias[co0], 0.0f);
                    const float v1 = fmaxf(acc[4 * g + 1] + bias[co0 + 1], 0.0f);
                    const float v2 = fmaxf(acc[4 * g + 2] + bias[co0 + 2], 0.0f);
                    const float v3 = fmaxf(acc[4 * g + 3] + bias[co0 + 3], 0.0f);
                    *(uint2*)((uint16_t*)out + p * C + co0) = make_uint2(pack2(v0, v1), pack2(v2, v3));
                }
            }
        }
    }
}

// fragments of the 6 layers: forward A[t][hh](lane, j) = W[t][16 hh + 8 (lane/32) + j][lane%32];
// input gradient (taps mirrored, channels swapped) A'[t][hh](lane, j) =
// W[8 - t][lane%32][16 hh + 8 (lane/32) + j]
// Head fragments (fwd + kLayers kFragPerLayer): the 32->4 head as a 32->32 tile whose
// output rows 0-3 hold the bf16-rounded fp32 weights W[t][ci][co] and rows 8-11 the bf16
// remainders W - hi (hi + lo carries 16 bits of each weight's mantissa); other rows 0.
__global__ void k_pack(const float* __restrict__ params, uint4* __restrict__ fwd,
                       uint4* __restrict__ dgrad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kLayers * kFragPerLayer) {
        const int hi_ = i - kLayers * kFragPerLayer;
        if (hi_ >= kFragPerLayer) return;
        const int k = hi_ / 64, lane = hi_ % 64, t = k >> 1, hh = k & 1, row = lane & 31;
        const float* W = params + kOffHead;
        uint32_t f[4];
        for (int j = 0; j < 8; j += 2) {
            uint32_t b[2];
            for (int e = 0; e < 2; ++e) {
                const int ci = 16 * hh + 8 * (lane >> 5) + j + e;
                b[e] = 0u;
                if (row < 4) {
                    b[e] = f2bf(W[(t * C + ci) * 4 + row]);
                } else if (row >= 8 && row < 12) {
                    const float wv = W[(t * C + ci) * 4 + row - 8];
                    b[e] = f2bf(wv - bf2f(f2bf(wv)));
                }
            }
            f[j / 2] = b[0] | (b[1] << 16);
        }
        fwd[i] = make_uint4(f[0], f[1], f[2], f[3]);
        return;
    }
    const int L = i / kFragPerLayer, k = (i / 64) % 18, lane = i % 64;
    const int t = k >> 1, hh = k & 1;
    const float* W = params + kOffL0 + L * kLParams;
    uint32_t f[4], d[4];
    for (int j = 0; j < 8; j += 2) {
        const int kk = 16 * hh + 8 * (lane >> 5) + j;
        f[j / 2] = f2bf(W[(t * C + kk) * C + (lane & 31)]) |
                   (f2bf(W[(t * C + kk + 1) * C + (lane & 31)]) << 16);
        d[j / 2] = f2bf(W[((8 - t) * C + (lane & 31)) * C + kk]) |
                   (f2bf(W[((8 - t) * C + (lane & 31)) * C + kk + 1]) << 16);
    }
    fwd[i] = make_uint4(f[0], f[1], f[2], f[3]);
    dgrad[i] = make_uint4(d[0], d[1], d[2], d[3]);
}

// ------------------------------------------------------------------ thin layers (VALU)
// a0 = relu(conv_in(x)), x = (n, 2, h, w) float32 planes; one thread per pixel
__global__ __launch_bounds__(256) void k_conv_in(const float* __restrict__ x,
                                                 const float* __restrict__ prm,
                                                 uint16_t* __restrict__ out, int h, int w) {
    __shared__ float sw[kInParams];
    for (int i = threadIdx.x; i < kInParams; i += 256) sw[i] = prm[kOffIn + i];
    __syncthreads();
    const int xx = blockIdx.x * 16 + (threadIdx.x & 15), yy = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int img = blockIdx.z;
    if (xx >= w || yy >= h) return;
    float acc[C];
#pragma unroll
    for (int co = 0; co < C; ++co) acc[co] = sw[kInW + co];
    for (int t = 0; t < 9; ++t) {
        const int sy = yy + t / 3 - 1, sx = xx + t % 3 - 1;
        if (sy < 0 || sy >= h || sx < 0 || sx >= w) continue;
#pragma unroll
        for (int ci = 0; ci < 2; ++ci) {
            const float v = x[((size_t)(img * 2 + ci) * h + sy) * w + sx];
#pragma unroll
            for (int co = 0; co < C; ++co) acc[co] = fmaf(v, sw[(t * 2 + ci) * C + co], acc[co]);
        }
    }
    uint4* o = (uint4*)(out + (((size_t)img * h + yy) * w + xx) * C);
#pragma unroll
    for (int q = 0; q < 4; ++q)
        o[q] = make_uint4(pack2(fmaxf(acc[8 * q], 0.f), fmaxf(acc[8 * q + 1], 0.f)),
                          pack2(fmaxf(acc[8 * q + 2], 0.f), fmaxf(acc[8 * q + 3], 0.f)),
                          pack2(fmaxf(acc[8 * q + 4], 0.f), fmaxf(acc[8 * q + 5], 0.f)),
                          pack2(fmaxf(acc[8 * q + 6], 0.f), fmaxf(acc[8 * q + 7], 0.f)));
}

// bilinear x2 upsampling of plane 0 (half-pixel centres, edge-clamped: the
// align_corners=False convention), sampled at output pixel (Y, X)
__device__ __forceinline__ float bilinear_up2(const float* p, int h, int w, int Y, int X) {
    const float sy = fmaxf((Y + 0.5f) * 0.5f - 0.5f, 0.0f), sx = fmaxf((X + 0.5f) * 0.5f - 0.5f, 0.0f);
    const int y0 = (int)sy, x0 = (int)sx;
    const int y1 = y0 + (y0 < h - 1 ? 1 : 0), x1 = x0 + (x0 < w - 1 ? 1 : 0);
    const float ly = sy - y0, lx = sx - x0;
    const float a = p[(size_t)y0 * w + x0], b = p[(size_t)y0 * w + x1];
    const float c = p[(size_t)y1 * w + x0], d = p[(size_t)y1 * w + x1];
    return (1.0f - ly) * ((1.0f - lx) * a + lx * b) + ly * ((1.0f - lx) * c + lx * d);
}

// y = pixel_shuffle2(head(u0)) + bilinear_up2(x plane 0) on the matrix cores: the 32->4
// head as k_conv32's tiles (32 x 16 low-res pixels a workgroup, 18 MFMAs a row) with the
// head fragments of k_pack; accumulator registers 0-3 and 4-7 of lanes 0-31 hold the
// high and low weight parts' sums of channels 0-3 for pixel lane%32.
__global__ __launch_bounds__(256) void k_head(const uint16_t* __restrict__ u0,
                                              const uint4* __restrict__ frag,
                                              const float* __restrict__ x,
                                              const float* __restrict__ prm,
                                              float* __restrict__ y, int h, int w) {
    __shared__ uint4 tile[kLH * kLW * kPixU4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int x0 = blockIdx.x * kTW, y0 = blockIdx.y * kTH, img = blockIdx.z;
    load_halo_tile(u0, tile, x0, y0, img, h, w);
    uint4 wf[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) wf[k] = frag[k * 64 + lane];
    const float* bias = prm + kOffHead + kHeadW;
    const float* x0p = x + (size_t)img * 2 * h * w;
    float* yo = y + (size_t)img * 4 * h * w;
    __syncthreads();
    const int px = lane & 31, kh = lane >> 5;
    const int ox = x0 + px;
    for (int rr = 0; rr < kTH / 4; ++rr) {
        const int r = wv * (kTH / 4) + rr, oy = y0 + r;
        if (oy >= h) break;  // uniform over the wave: the MFMAs below run with all lanes
        f32x16 acc = {};
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int base = ((r + t / 3) * kLW + px + t % 3) * kPixU4 + kh;
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const uint4 b = tile[base + 2 * hh];
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                    __builtin_bit_cast(bf16x8, wf[2 * t + hh]), __builtin_bit_cast(bf16x8, b),
                    acc, 0, 0, 0);
            }
        }
        if (kh == 0 && ox < w) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int Y = 2 * oy + (c >> 1), X = 2 * ox + (c & 1);
                yo[(size_t)Y * 2 * w + X] =
                    (acc[c] + acc[4 + c] + bias[c]) + bilinear_up2(x0p, h, w, Y, X);
            }
        }
    }
}

// du0(p, ci) = sum_t sum_co dr(p - off(t), co) W[t][ci][co]
__global__ __launch_bounds__(256) void k_head_dgrad(const float* __restrict__ dr,
                                                    const float* __restrict__ prm,
                                                    float* __restrict__ du0, int h, int w) {
    __shared__ float sw[kHeadW];
    for (int i = threadIdx.x; i < kHeadW; i += 256) sw[i] = prm[kOffHead + i];
    __syncthreads();
    const int xx = blockIdx.x * 16 + (threadIdx.x & 15), yy = blockIdx.y * 16 + (threadIdx.x >> 4);
    const int img = blockIdx.z;
    if (xx >= w || yy >= h) return;
    float acc[C];
#pragma unroll
    for (int ci = 0; ci < C; ++ci) acc[ci] = 0.0f;
    for (int t = 0; t < 9; ++t) {
        const int sy = yy - (t / 3 - 1), sx = xx - (t % 3 - 1);
        if (sy < 0 || sy >= h || sx < 0 || sx >= w) continue;
        const float4 d = *(const float4*)(dr + (((size_t)img * h + sy) * w + sx) * 4);
#pragma unroll
        for (int ci = 0; ci < C; ++ci) {
            const float* wr = sw + (t * C + ci) * 4;
            acc[ci] = fmaf(d.x, wr[0], fmaf(d.y, wr[1], fmaf(d.z, wr[2], fmaf(d.w, wr[3], acc[ci]))));
        }
    }
    float4* o = (float4*)(du0 + (((size_t)img * h + yy) * w + xx) * C);
#pragma unroll
    for (int q = 0; q < 8; ++q) o[q] = make_float4(acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]);
}

// ------------------------------------------------------------------ resampling
// out (n, h/2, w/2) = 2x2 mean of in (n, h, w); one thread per (output pixel, 8 channels)
__global__ void k_pool2(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int n, int h, int w) {
    const int h2 = h / 2, w2 = w / 2;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * h2 * w2 * 4) return;
    const int q = i & 3;
    const size_t p = i >> 2;
    const int xx = p % w2, yy = (p / w2) % h2, img = p / ((size_t)w2 * h2);
    float s[8] = {};
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const uint4 u = *(const uint4*)(in + (((size_t)img * h + 2 * yy + a) * w + 2 * xx + b) * C + 8 * q);
            const uint32_t ww[4] = {u.x, u.y, u.z, u.w};
            for (int j = 0; j < 4; ++j) {
                s[2 * j] += bf_lo(ww[j]);
                s[2 * j + 1] += bf_hi(ww[j]);
            }
        }
    *(uint4*)(out + p * C + 8 * q) = make_uint4(pack2(.25f * s[0], .25f * s[1]), pack2(.25f * s[2], .25f * s[3]),
                                               pack2(.25f * s[4], .25f * s[5]), pack2(.25f * s[6], .25f * s[7]));
}

// out (n, h, w) = nearest_up2(low (n, h/2, w/2)) + skip (n, h, w)
__global__ void k_up2_add(const uint16_t* __restrict__ low, const uint16_t* __restrict__ skip,
                          uint16_t* __restrict__ out, int n, int h, int w) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * h * w * 4) return;
    const int q = i & 3;
    const size_t p = i >> 2;
    const int xx = p % w, yy = (p / w) % h, img = p / ((size_t)w * h);
    const uint4 a = *(const uint4*)(low + (((size_t)img * (h / 2) + yy / 2) * (w / 2) + xx / 2) * C + 8 * q);
    const uint4 b = *(const uint4*)(skip + p * C + 8 * q);
    const uint32_t aw[4] = {a.x, a.y, a.z, a.w}, bw[4] = {b.x, b.y, b.z, b.w};
    uint32_t o[4];
    for (int j = 0; j < 4; ++j)
        o[j] = pack2(bf_lo(aw[j]) + bf_lo(bw[j]), bf_hi(aw[j]) + bf_hi(bw[j]));
    *(uint4*)(out + p * C + 8 * q) = make_uint4(o[0], o[1], o[2], o[3]);
}

// din (n, h, w) += dout (n, h/2, w/2) / 4 at the 2x2 children; one thread per (pixel, 4 ch)
__global__ void k_pool2_bwd(const float* __restrict__ dout, float* __restrict__ din, int n, int h, int w) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * h * w * 8) return;
    const int q = i & 7;
    const size_t p = i >> 3;
    const int xx = p % w, yy = (p / w) % h, img = p / ((size_t)w * h);
    const float4 d = *(const float4*)(dout + (((size_t)img * (h / 2) + yy / 2) * (w / 2) + xx / 2) * C + 4 * q);
    float4* o = (float4*)(din + p * C + 4 * q);
    float4 v = *o;
    v.x += .25f * d.x;
    v.y += .25f * d.y;
    v.z += .25f * d.z;
    v.w += .25f * d.w;
    *o = v;
}

// dlow (n, h/2, w/2) = sum of ds (n, h, w) over the 2x2 children
__global__ void k_up2_bwd(const float* __restrict__ ds, float* __restrict__ dlow, int n, int h, int w) {
    const int h2 = h / 2, w2 = w / 2;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)n * h2 * w2 * 8) return;
    const int q = i & 7;
    const size_t p = i >> 3;
    const int xx = p % w2, yy = (p / w2) % h2, img = p / ((size_t)w2 * h2);
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
            const float4 d = *(const float4*)(ds + (((size_t)img * h + 2 * yy + a) * w + 2 * xx + b) * C + 4 * q);
            s.x += d.x;
            s.y += d.y;
            s.z += d.z;
            s.w += d.w;
        }
    *(float4*)(dlow + p * C + 4 * q) = s;
}

// ------------------------------------------------------------------ loss
// dr (n, h, w, 4) = 2 (y - target) / count at the pixel-shuffle positions; per-block sums
// of squared error
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ y, const float* __restrict__ tgt,
                                              float* __restrict__ dr, float* __restrict__ partial,
                                              int n, int h, int w, float scale) {
    __shared__ float red[256];
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    float s = 0.0f;
    if (i < (size_t)n * h * w) {
        const int xx = i % w, yy = (i / w) % h, img = i / ((size_t)w * h);
        float d4[4];
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) {
                const size_t o = ((size_t)img * 2 * h + 2 * yy + a) * 2 * w + 2 * xx + b;
                const float d = y[o] - tgt[o];
                s = fmaf(d, d, s);
                d4[2 * a + b] = scale * d;
            }
        *(float4*)(dr + i * 4) = make_float4(d4[0], d4[1], d4[2], d4[3]);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ------------------------------------------------------------------ weight gradients
// dW[t][ci][co] = sum_p x(p + off(t), ci) g(p, co) [mask(p, co) > 0];  db[co] = sum_p g(p, co).
// Blocks stride over 16x16 pixel tiles; thread (ci, 4 output channels) keeps 9 x 4 sums;
// each block writes its partial sums (9 CIN COUT + COUT floats).
// XPLANAR: x is (n, CIN, h, w) float32 planes, else (n, h, w, CIN) bf16.
template <int CIN, int COUT, bool XPLANAR, bool MASK>
__global__ __launch_bounds__(256) void k_wgrad(const void* __restrict__ x, const float* __restrict__ g,
                                               const uint16_t* __restrict__ mask,
                                               float* __restrict__ partial, int n, int h, int w) {
    constexpr int NW = 9 * CIN * COUT + COUT;
    constexpr int ACTIVE = CIN * COUT / 4;
    static_assert(ACTIVE <= 256, "one thread per (input channel, 4 output channels)");
    __shared__ __attribute__((aligned(16))) float xs[18 * 18 * CIN];
    __shared__ __attribute__((aligned(16))) float gs[256 * COUT];
    const int tid = threadIdx.x;
    const int ci = tid % CIN, cb = tid / CIN;
    float acc[9][4] = {};
    float bacc[4] = {};
    const int ntx = (w + 15) / 16, nty = (h + 15) / 16;
    const int ntiles = n * ntx * nty;
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int img = tile / (ntx * nty), tr = tile % (ntx * nty);
        const int ty0 = (tr / ntx) * 16, tx0 = (tr % ntx) * 16;
        __syncthreads();
        for (int e = tid; e < 18 * 18 * CIN; e += 256) {
            const int pix = e / CIN, c = e % CIN;
            const int yy = ty0 - 1 + pix / 18, xx = tx0 - 1 + pix % 18;
            float v = 0.0f;
            if (yy >= 0 && yy < h && xx >= 0 && xx < w) {
                if constexpr (XPLANAR)
                    v = ((const float*)x)[((size_t)(img * CIN + c) * h + yy) * w + xx];
                else
                    v = bf2f(((const uint16_t*)x)[(((size_t)img * h + yy) * w + xx) * CIN + c]);
            }
            xs[e] = v;
        }
        for (int e = tid; e < 256 * COUT; e += 256) {
            const int pix = e / COUT, c = e % COUT;
            const int yy = ty0 + pix / 16, xx = tx0 + pix % 16;
            float v = 0.0f;
            if (yy < h && xx < w) {
                const size_t o = (((size_t)img * h + yy) * w + xx) * COUT + c;
                v = g[o];
                if (MASK && !bf_pos(mask[o])) v = 0.0f;
            }
            gs[e] = v;
        }
        __syncthreads();
        if (tid < ACTIVE) {
            for (int p = 0; p < 256; ++p) {
                const float4 g4 = *(const float4*)(gs + p * COUT + 4 * cb);
                const int py = p >> 4, px = p & 15;
#pragma unroll
                for (int t = 0; t < 9; ++t) {
                    const float xv = xs[((py + t / 3) * 18 + px + t % 3) * CIN + ci];
                    acc[t][0] = fmaf(xv, g4.x, acc[t][0]);
                    acc[t][1] = fmaf(xv, g4.y, acc[t][1]);
                    acc[t][2] = fmaf(xv, g4.z, acc[t][2]);
                    acc[t][3] = fmaf(xv, g4.w, acc[t][3]);
                }
                if (ci == 0) {
                    bacc[0] += g4.x;
                    bacc[1] += g4.y;
                    bacc[2] += g4.z;
                    bacc[3] += g4.w;
                }
            }
        }
    }
    if (tid < ACTIVE) {
        float* out = partial + (size_t)blockIdx.x * NW;
#pragma unroll
        for (int t = 0; t < 9; ++t)
#pragma unroll
            for (int k = 0; k < 4; ++k) out[(t * CIN + ci) * COUT + 4 * cb + k] = acc[t][k];
        if (ci == 0)
            for (int k = 0; k < 4; ++k) out[9 * CIN * COUT + 4 * cb + k] = bacc[k];
    }
}

__global__ void k_reduce(const float* __restrict__ partial, int nblocks, int nw, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nw) return;
    float s = 0.0f;
    for (int b = 0; b < nblocks; ++b) s += partial[(size_t)b * nw + i];
    out[i] = s;
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int n, float lr, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + 1e-8f);
}

// ------------------------------------------------------------------ judge helpers
// Squared error of a frame (h x w) against gx x gy exact patches of c x c pixels: patch
// (ty, tx) covers the frame pixels from row ty (h / gy) + oy and column tx (w / gx) + ox
// on; the patches are stored one after another (tx fastest), rows of c.  One workgroup
// sums everything in a fixed order (deterministic): *sse = (accumulate ? *sse : 0) + sum.
__global__ __launch_bounds__(256) void k_judge_patches(const float* __restrict__ frame, int w,
                                                       int h, const float* __restrict__ patches,
                                                       int c, int gx, int gy, int ox, int oy,
                                                       int accumulate, float* __restrict__ sse) {
    __shared__ float red[256];
    const int cc = c * c, n = gx * gy * cc;
    const int cw = w / gx, ch = h / gy;
    float v = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) {
        const int k = i / cc, r = i - k * cc;
        const int ty = k / gx, tx = k - ty * gx;
        const int py = r / c, px = r - py * c;
        const int y = ty * ch + oy + py, x = tx * cw + ox + px;
        const float d = frame[(size_t)y * w + x] - patches[i];
        v = fmaf(d, d, v);
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) sse[0] = (accumulate ? sse[0] : 0.0f) + red[0];
}

// The same sum for up to kJudgeFrames frames at once against their sparse exact frames
// (patch 1: the gx x gy grid of pixels (tx w/gx + ox, ty h/gy + oy), sparse frame k at
// sparse + k gx gy, row-major), with the sum of fourth powers beside it (the spread of the
// squared errors, for the judge's confidence bound): block b of frame k (grid
// kJudgeBlocks x n) sums its contiguous share of the grid in a fixed order into
// part[k kJudgeBlocks + b] and part[kJudgeFrames kJudgeBlocks + k kJudgeBlocks + b];
// k_judge_reduce adds a frame's partial sums in order (deterministic).
constexpr int kJudgeFrames = 16;
constexpr int kJudgeBlocks = 32;
struct JudgeFrames {
    const float* f[kJudgeFrames];
};

__global__ __launch_bounds__(256) void k_judge_grid(JudgeFrames frames, int w, int h,
                                                    const float* __restrict__ sparse, int gx,
                                                    int gy, int ox, int oy,
                                                    float* __restrict__ part) {
    __shared__ float red[2][256];
    const int k = blockIdx.y, b = blockIdx.x;
    const int n = gx * gy, per = (n + kJudgeBlocks - 1) / kJudgeBlocks;
    const int i0 = b * per, i1 = min(n, i0 + per);
    const int cw = w / gx, ch = h / gy;
    const float* frame = frames.f[k];
    const float* sp = sparse + (size_t)k * n;
    float v2 = 0.0f, v4 = 0.0f;
    for (int i = i0 + (int)threadIdx.x; i < i1; i += 256) {
        const int ty = i / gx, tx = i - ty * gx;
        const float d = frame[(size_t)(ty * ch + oy) * w + tx * cw + ox] - sp[i];
        const float d2 = d * d;
        v2 += d2;
        v4 = fmaf(d2, d2, v4);
    }
    red[0][threadIdx.x] = v2;
    red[1][threadIdx.x] = v4;
    __syncthreads();
    for (int r = 128; r > 0; r >>= 1) {
        if (threadIdx.x < r) {
            red[0][threadIdx.x] += red[0][threadIdx.x + r];
            red[1][threadIdx.x] += red[1][threadIdx.x + r];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[k * kJudgeBlocks + b] = red[0][0];
        part[kJudgeFrames * kJudgeBlocks + k * kJudgeBlocks + b] = red[1][0];
    }
}

__global__ void k_judge_reduce(const float* __restrict__ part, int n, float* __restrict__ sse,
                               float* __restrict__ sse4) {
    const int k = threadIdx.x;
    if (k >= n) return;
    float v = 0.0f, v4 = 0.0f;
    for (int b = 0; b < kJudgeBlocks; ++b) {
        v += part[k * kJudgeBlocks + b];
        v4 += part[kJudgeFrames * kJudgeBlocks + k * kJudgeBlocks + b];
    }
    sse[k] = v;
    if (sse4) sse4[k] = v4;
}

// The judge's decision with a confidence bound: with m2 = sse / npix and m4 = sse4 / npix
// the mean squared and fourth-power errors of the judge's pixels, the squared error's
// variance is m4 - m2^2 and the standard error of m2 is sqrt((m4 - m2^2) / npix);
// flags[k] = 1 (fall back to the exact frame) unless m2 + z standard errors <= mse_max.
__global__ void k_judge_flags_bound(const float* __restrict__ sse, const float* __restrict__ sse4,
                                    int n, float inv_npix, float mse_max, float z,
                                    uint32_t* __restrict__ flags) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const float m2 = sse[k] * inv_npix, m4 = sse4[k] * inv_npix;
    const float ucb = m2 + z * sqrtf(fmaxf(m4 - m2 * m2, 0.0f) * inv_npix);
    flags[k] = ucb <= mse_max ? 0u : 1u;
}

// flags[k] = 1 when frame k's judged mean squared error sse[k] / npix exceeds mse_max (or
// is not a number): the frame falls back to the exact render
__global__ void k_judge_flags(const float* __restrict__ sse, int n, float inv_npix, float mse_max,
                              uint32_t* __restrict__ flags) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) flags[k] = sse[k] * inv_npix <= mse_max ? 0u : 1u;
}

__global__ void k_fill(float* __restrict__ p, int n, float v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// dst (2h, 2w) = nearest-neighbour x2 of src (h, w)
__global__ void k_upsample2(const float* __restrict__ src, int w, int h, float* __restrict__ dst) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)4 * w * h) return;
    const int X = i % (2 * w), Y = i / (2 * w);
    dst[i] = src[(size_t)(Y >> 1) * w + (X >> 1)];
}

}  // namespace

// ---------------------------------------------------------------------------------- host
struct spl_net {
    int device = 0;
    hipStream_t stream = nullptr;
    float *params = nullptr, *grad = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    uint4 *frag_fwd = nullptr, *frag_dgrad = nullptr;
    int adam_t = 0;
    // workspace for (n, h, w)
    size_t cap = 0;  // n h w the buffers hold
    uint16_t *a0 = nullptr, *e0 = nullptr, *e0b = nullptr, *s0 = nullptr, *u0 = nullptr,
             *u0b = nullptr;                                                // h x w
    uint16_t *p1 = nullptr, *e1 = nullptr, *e1b = nullptr, *s1 = nullptr, *u1 = nullptr;  // h/2
    uint16_t *p2 = nullptr, *e2a = nullptr, *e2 = nullptr;                // h/4
    float *gH0 = nullptr, *gH1 = nullptr, *gM0 = nullptr, *gM1 = nullptr, *gQ0 = nullptr,
          *gQ1 = nullptr, *dr = nullptr, *y = nullptr;
    float *wpart = nullptr, *lpart = nullptr;
    size_t lpart_n = 0;
    float* judge_part = nullptr;  // spl_judge_sse_frames partial sums (2 x kJudgeFrames blocks)
};

namespace {

thread_local std::string g_err;
thread_local const spl_net* g_err_net = nullptr;

int fail(const spl_net* net, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    g_err_net = net;
    return code;
}

#define HIPCALL(net, expr)                                                                    \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((net), e_ == hipErrorOutOfMemory ? SPL_ERR_OUT_OF_MEMORY : SPL_ERR_HIP, \
                        "%s failed: %s", #expr, hipGetErrorString(e_));                     \
    } while (0)

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

void free_work(spl_net* t) {
    void* bufs[] = {t->a0,  t->e0,  t->e0b, t->s0,  t->u0,  t->u0b, t->p1,  t->e1,
                    t->e1b, t->s1,  t->u1,  t->p2,  t->e2a, t->e2,  t->gH0, t->gH1,
                    t->gM0, t->gM1, t->gQ0, t->gQ1, t->dr,  t->y,   t->lpart};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    t->a0 = t->e0 = t->e0b = t->s0 = t->u0 = t->u0b = nullptr;
    t->p1 = t->e1 = t->e1b = t->s1 = t->u1 = t->p2 = t->e2a = t->e2 = nullptr;
    t->gH0 = t->gH1 = t->gM0 = t->gM1 = t->gQ0 = t->gQ1 = t->dr = t->y = t->lpart = nullptr;
    t->cap = 0;
    t->lpart_n = 0;
}

int check_shape(const spl_net* t, int n, int h, int w) {
    if (n < 1 || h < 4 || w < 4 || h % 4 || w % 4 || h > 16384 || w > 16384)
        return fail(t, SPL_ERR_INVALID_ARG, "invalid batch shape n=%d h=%d w=%d (h, w multiples of 4)", n, h, w);
    if ((size_t)n * h * w > ((size_t)1 << 31) / C)
        return fail(t, SPL_ERR_INVALID_ARG, "batch too large");
    return SPL_OK;
}

// Work buffers for n images of h x w (grown, never shrunk: calls of different batch sizes
// reuse the largest allocation).
int ensure_work(spl_net* t, int n, int h, int w) {
    const size_t px = (size_t)n * h * w;
    if (t->cap >= px && t->a0) {
        t->lpart_n = (px + 255) / 256;  // the loss's blocks for this call's pixels
        return SPL_OK;
    }
    free_work(t);
    const size_t b16 = px * C * 2, f32 = px * C * 4;
    struct {
        void** p;
        size_t bytes;
    } plan[] = {
        {(void**)&t->a0, b16},      {(void**)&t->e0, b16},      {(void**)&t->s0, b16},
        {(void**)&t->u0, b16},      {(void**)&t->e0b, b16},     {(void**)&t->u0b, b16},
        {(void**)&t->e1b, b16 / 4},
        {(void**)&t->p1, b16 / 4},  {(void**)&t->e1, b16 / 4},
        {(void**)&t->s1, b16 / 4},  {(void**)&t->u1, b16 / 4},  {(void**)&t->p2, b16 / 16},
        {(void**)&t->e2a, b16 / 16}, {(void**)&t->e2, b16 / 16}, {(void**)&t->gH0, f32},
        {(void**)&t->gH1, f32},     {(void**)&t->gM0, f32 / 4}, {(void**)&t->gM1, f32 / 4},
        {(void**)&t->gQ0, f32 / 16}, {(void**)&t->gQ1, f32 / 16}, {(void**)&t->dr, px * 16},
        {(void**)&t->y, px * 16},
    };
    for (auto& e : plan) HIPCALL(t, hipMalloc(e.p, e.bytes));
    t->lpart_n = (px + 255) / 256;
    HIPCALL(t, hipMalloc(&t->lpart, t->lpart_n * sizeof(float)));
    t->cap = px;
    return SPL_OK;
}

inline dim3 grid16(int n, int h, int w) { return dim3((w + 15) / 16, (h + 15) / 16, n); }
inline unsigned blocks_for(size_t items) { return (unsigned)((items + 255) / 256); }

void conv32_fwd(spl_net* t, int L, const uint16_t* in, uint16_t* out, int n, int h, int w, hipStream_t s) {
    dim3 grid((w + kTW - 1) / kTW, (h + kTH - 1) / kTH, n);
    hipLaunchKernelGGL(k_conv32<false>, grid, dim3(256), 0, s, in, nullptr,
                       t->frag_fwd + (size_t)L * kFragPerLayer, t->params + kOffL0 + L * kLParams + kL,
                       out, h, w);
}

void conv32_dgrad(spl_net* t, int L, const float* g, const uint16_t* mask, float* out, int n, int h, int w,
                  hipStream_t s) {
    dim3 grid((w + kTW - 1) / kTW, (h + kTH - 1) / kTH, n);
    hipLaunchKernelGGL(k_conv32<true>, grid, dim3(256), 0, s, g, mask,
                       t->frag_dgrad + (size_t)L * kFragPerLayer, nullptr, out, h, w);
}

template <int CIN, int COUT, bool XPLANAR, bool MASK>
void wgrad(spl_net* t, const void* x, const float* g, const uint16_t* mask, float* dst, int n, int h, int w,
           hipStream_t s) {
    constexpr int NW = 9 * CIN * COUT + COUT;
    const int ntiles = n * ((w + 15) / 16) * ((h + 15) / 16);
    const int blocks = ntiles < kMaxWgradBlocks ? ntiles : kMaxWgradBlocks;
    hipLaunchKernelGGL((k_wgrad<CIN, COUT, XPLANAR, MASK>), dim3(blocks), dim3(256), 0, s, x, g, mask,
                       t->wpart, n, h, w);
    hipLaunchKernelGGL(k_reduce, dim3((NW + 255) / 256), dim3(256), 0, s, t->wpart, blocks, NW, dst);
}

void forward(spl_net* t, const float* x, int n, int h, int w, float* y, hipStream_t s) {
    const int h2 = h / 2, w2 = w / 2, h4 = h / 4, w4 = w / 4;
    hipLaunchKernelGGL(k_conv_in, grid16(n, h, w), dim3(256), 0, s, x, t->params, t->a0, h, w);
    conv32_fwd(t, kE0, t->a0, t->e0, n, h, w, s);
    conv32_fwd(t, kE0b, t->e0, t->e0b, n, h, w, s);
    hipLaunchKernelGGL(k_pool2, dim3(blocks_for((size_t)n * h2 * w2 * 4)), dim3(256), 0, s, t->e0b, t->p1, n, h, w);
    conv32_fwd(t, kE1, t->p1, t->e1, n, h2, w2, s);
    conv32_fwd(t, kE1b, t->e1, t->e1b, n, h2, w2, s);
    hipLaunchKernelGGL(k_pool2, dim3(blocks_for((size_t)n * h4 * w4 * 4)), dim3(256), 0, s, t->e1b, t->p2, n, h2, w2);
    conv32_fwd(t, kE2a, t->p2, t->e2a, n, h4, w4, s);
    conv32_fwd(t, kE2, t->e2a, t->e2, n, h4, w4, s);
    hipLaunchKernelGGL(k_up2_add, dim3(blocks_for((size_t)n * h2 * w2 * 4)), dim3(256), 0, s, t->e2, t->e1b, t->s1, n, h2, w2);
    conv32_fwd(t, kU1, t->s1, t->u1, n, h2, w2, s);
    hipLaunchKernelGGL(k_up2_add, dim3(blocks_for((size_t)n * h * w * 4)), dim3(256), 0, s, t->u1, t->e0b, t->s0, n, h, w);
    conv32_fwd(t, kU0, t->s0, t->u0, n, h, w, s);
    conv32_fwd(t, kU0b, t->u0, t->u0b, n, h, w, s);
    hipLaunchKernelGGL(k_head, dim3((w + kTW - 1) / kTW, (h + kTH - 1) / kTH, n), dim3(256), 0, s,
                       t->u0b, t->frag_fwd + (size_t)kLayers * kFragPerLayer, x, t->params, y, h,
                       w);
}

float* layer_grad(spl_net* t, int L) { return t->grad + kOffL0 + L * kLParams; }

// loss + backward into t->grad; leaves the per-block squared-error sums in t->lpart.
// Gradient buffers: gH* at h x w, gM* at h/2, gQ* at h/4 (float32, 32 channels).
void backward(spl_net* t, const float* x, const float* tgt, int n, int h, int w, hipStream_t s) {
    const int h2 = h / 2, w2 = w / 2, h4 = h / 4, w4 = w / 4;
    const float scale = 2.0f / (float)((size_t)n * 4 * h * w);
    hipLaunchKernelGGL(k_loss, dim3(t->lpart_n), dim3(256), 0, s, t->y, tgt, t->dr, t->lpart, n, h, w, scale);
    // head(u0b)
    wgrad<C, 4, false, false>(t, t->u0b, t->dr, nullptr, t->grad + kOffHead, n, h, w, s);
    hipLaunchKernelGGL(k_head_dgrad, grid16(n, h, w), dim3(256), 0, s, t->dr, t->params, t->gH0, h, w);
    // u0b = relu(conv(u0))
    wgrad<C, C, false, true>(t, t->u0, t->gH0, t->u0b, layer_grad(t, kU0b), n, h, w, s);
    conv32_dgrad(t, kU0b, t->gH0, t->u0b, t->gH1, n, h, w, s);  // d u0
    // u0 = relu(conv(s0)), s0 = up2(u1) + e0b
    wgrad<C, C, false, true>(t, t->s0, t->gH1, t->u0, layer_grad(t, kU0), n, h, w, s);
    conv32_dgrad(t, kU0, t->gH1, t->u0, t->gH0, n, h, w, s);  // d s0 (= the e0b share)
    hipLaunchKernelGGL(k_up2_bwd, dim3(blocks_for((size_t)n * h2 * w2 * 8)), dim3(256), 0, s, t->gH0, t->gM0, n, h, w);
    // u1 = relu(conv(s1)), s1 = up2(e2) + e1b
    wgrad<C, C, false, true>(t, t->s1, t->gM0, t->u1, layer_grad(t, kU1), n, h2, w2, s);
    conv32_dgrad(t, kU1, t->gM0, t->u1, t->gM1, n, h2, w2, s);  // d s1 (= the e1b share)
    hipLaunchKernelGGL(k_up2_bwd, dim3(blocks_for((size_t)n * h4 * w4 * 8)), dim3(256), 0, s, t->gM1, t->gQ0, n, h2, w2);
    // e2 = relu(conv(e2a))
    wgrad<C, C, false, true>(t, t->e2a, t->gQ0, t->e2, layer_grad(t, kE2), n, h4, w4, s);
    conv32_dgrad(t, kE2, t->gQ0, t->e2, t->gQ1, n, h4, w4, s);
    // e2a = relu(conv(p2)), p2 = pool(e1b)
    wgrad<C, C, false, true>(t, t->p2, t->gQ1, t->e2a, layer_grad(t, kE2a), n, h4, w4, s);
    conv32_dgrad(t, kE2a, t->gQ1, t->e2a, t->gQ0, n, h4, w4, s);  // d p2
    hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks_for((size_t)n * h2 * w2 * 8)), dim3(256), 0, s, t->gQ0, t->gM1, n, h2, w2);
    // e1b = relu(conv(e1))
    wgrad<C, C, false, true>(t, t->e1, t->gM1, t->e1b, layer_grad(t, kE1b), n, h2, w2, s);
    conv32_dgrad(t, kE1b, t->gM1, t->e1b, t->gM0, n, h2, w2, s);  // d e1
    // e1 = relu(conv(p1)), p1 = pool(e0b)
    wgrad<C, C, false, true>(t, t->p1, t->gM0, t->e1, layer_grad(t, kE1), n, h2, w2, s);
    conv32_dgrad(t, kE1, t->gM0, t->e1, t->gM1, n, h2, w2, s);  // d p1
    hipLaunchKernelGGL(k_pool2_bwd, dim3(blocks_for((size_t)n * h * w * 8)), dim3(256), 0, s, t->gM1, t->gH0, n, h, w);
    // e0b = relu(conv(e0))
    wgrad<C, C, false, true>(t, t->e0, t->gH0, t->e0b, layer_grad(t, kE0b), n, h, w, s);
    conv32_dgrad(t, kE0b, t->gH0, t->e0b, t->gH1, n, h, w, s);  // d e0
    // e0 = relu(conv(a0))
    wgrad<C, C, false, true>(t, t->a0, t->gH1, t->e0, layer_grad(t, kE0), n, h, w, s);
    conv32_dgrad(t, kE0, t->gH1, t->e0, t->gH0, n, h, w, s);  // d a0
    // a0 = relu(conv_in(x))
    wgrad<2, C, true, true>(t, x, t->gH0, t->a0, t->grad + kOffIn, n, h, w, s);
}

int read_sse(spl_net* t, double* sse, hipStream_t s) {
    std::vector<float> part(t->lpart_n);
    HIPCALL(t, hipMemcpyAsync(part.data(), t->lpart, part.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    HIPCALL(t, hipStreamSynchronize(s));
    double acc = 0.0;
    for (float v : part) acc += v;
    *sse = acc;
    return SPL_OK;
}

int repack(spl_net* t, hipStream_t s) {
    hipLaunchKernelGGL(k_pack, dim3(((kLayers + 1) * kFragPerLayer + 255) / 256), dim3(256), 0, s,
                       t->params, t->frag_fwd, t->frag_dgrad);
    HIPCALL(t, hipGetLastError());
    return SPL_OK;
}

}  // namespace

extern "C" {

int spl_abi_version(void) { return SPL_ABI_VERSION; }
int spl_param_count(void) { return kParams; }

const char* spl_last_error(const spl_net* t) { return g_err_net == t ? g_err.c_str() : ""; }

int spl_create(int device, spl_net** out) {
    if (!out) return fail(nullptr, SPL_ERR_INVALID_ARG, "NULL output pointer");
    *out = nullptr;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd == 0)
        return fail(nullptr, SPL_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= nd)
        return fail(nullptr, SPL_ERR_INVALID_ARG, "device %d out of range [0, %d)", device, nd);
    spl_net* t = new (std::nothrow) spl_net();
    if (!t) return fail(nullptr, SPL_ERR_OUT_OF_MEMORY, "out of host memory");
    t->device = device;
    DeviceGuard g(device);
    hipError_t e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
    const size_t pb = kParams * sizeof(float);
    if (e == hipSuccess) e = hipMalloc(&t->params, pb);
    if (e == hipSuccess) e = hipMalloc(&t->grad, pb);
    if (e == hipSuccess) e = hipMalloc(&t->adam_m, pb);
    if (e == hipSuccess) e = hipMalloc(&t->adam_v, pb);
    if (e == hipSuccess) e = hipMalloc(&t->frag_fwd, (kLayers + 1) * kFragPerLayer * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc(&t->frag_dgrad, kLayers * kFragPerLayer * sizeof(uint4));
    if (e == hipSuccess) e = hipMalloc(&t->wpart, (size_t)kMaxWgradBlocks * kLParams * sizeof(float));
    if (e == hipSuccess) e = hipMemsetAsync(t->params, 0, pb, t->stream);
    if (e == hipSuccess) e = hipMemsetAsync(t->adam_m, 0, pb, t->stream);
    if (e == hipSuccess) e = hipMemsetAsync(t->adam_v, 0, pb, t->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
    if (e != hipSuccess) {
        int rc = fail(nullptr, SPL_ERR_HIP, "network creation failed: %s", hipGetErrorString(e));
        spl_destroy(t);
        return rc;
    }
    if (repack(t, t->stream) != SPL_OK || hipStreamSynchronize(t->stream) != hipSuccess) {
        spl_destroy(t);
        return fail(nullptr, SPL_ERR_HIP, "weight packing failed");
    }
    *out = t;
    return SPL_OK;
}

int spl_destroy(spl_net* t) {
    if (!t) return SPL_OK;
    {
        DeviceGuard g(t->device);
        if (t->stream) (void)hipStreamSynchronize(t->stream);
        free_work(t);
        void* bufs[] = {t->params,   t->grad,       t->adam_m, t->adam_v,
                        t->frag_fwd, t->frag_dgrad, t->wpart,  t->judge_part};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (t->stream) (void)hipStreamDestroy(t->stream);
    }
    delete t;
    return SPL_OK;
}

int spl_set_params(spl_net* t, const float* host, int n) {
    if (!t) return fail(nullptr, SPL_ERR_INVALID_ARG, "NULL network");
    if (!host || n != kParams) return fail(t, SPL_ERR_INVALID_ARG, "expected %d parameters, got %d", kParams, n);
    DeviceGuard g(t->device);
    const size_t pb = kParams * sizeof(float);
    HIPCALL(t, hipMemcpyAsync(t->params, host, pb, hipMemcpyHostToDevice, t->stream));
    HIPCALL(t, hipMemsetAsync(t->adam_m, 0, pb, t->stream));
    HIPCALL(t, hipMemsetAsync(t->adam_v, 0, pb, t->stream));
    t->adam_t = 0;
    int rc = repack(t, t->stream);
    if (rc) return rc;
    HIPCALL(t, hipStreamSynchronize(t->stream));
    return SPL_OK;
}

int spl_get_params(spl_net* t, float* host, int n) {
    if (!t) return fail(nullptr, SPL_ERR_INVALID_ARG, "NULL network");
    if (!host || n != kParams) return fail(t, SPL_ERR_INVALID_ARG, "expected %d parameters, got %d", kParams, n);
    DeviceGuard g(t->device);
    HIPCALL(t, hipMemcpyAsync(host, t->params, kParams * sizeof(float), hipMemcpyDeviceToHost, t->stream));
    HIPCALL(t, hipStreamSynchronize(t->stream));
    return SPL_OK;
}

int spl_forward(spl_net* t, const float* d_in, int n, int h, int w, float* d_out, void* stream) {
    if (!t) return fail(nullptr, SPL_ERR_INVALID_ARG, "NULL network");
    int rc = check_shape(t, n, h, w);
    if (rc) return rc;
    if (!d_in || !d_out) return fail(t, SPL_ERR_INVALID_ARG, "NULL buffer");
    DeviceGuard g(t->device);
    hipStream_t s = stream ? (hipStream_t)stream : t->stream;
    if (t->cap < (size_t)n * h * w) HIPCALL(t, hipStreamSynchronize(s));  // buffers regrown
    rc = ensure_work(t, n, h, w);
    if (rc) return rc;
    forward(t, d_in, n, h, w, d_out, s);
    HIPCALL(t, hipGetLastError());
    return SPL_OK;
}

int spl_gradients(spl_net* t, const float* d_in, const float* d_target, int n, int h, int w, float* host_grad,
                  double* sse, void* stream) {
    if (!t) return fail(nullptr, SPL_ERR_INVALID_ARG, "NULL network");
    int rc = check_shape(t, n, h, w);
    if (rc) return rc;
    if (!d_in || !d_target || !host_grad) return fail(t, SPL_ERR_INVALID_ARG, "NULL buffer");
    DeviceGuard g(t->device);
    hipStream_t s = stream ? (hipStream_t)stream : t->stream;
    HIPCALL(t, hipStreamSynchronize(s));
    rc = ensure_work(t, n, h, w);
    if (rc) return rc;
    forward(t, d_in, n, h, w, t->y, s);
    backward(t, d_in, d_target, n, h, w, s);
    HIPCALL(t, hipGetLastError());
    HIPCALL(t, hipMemcpyAsync(host_grad, t->grad, kParams * sizeof(float), hipMemcpyDeviceToHost, s));
    double v = 0.0;
    rc = read_sse(t, &v, s);
    if (rc) return rc;
    if (sse) *sse = v;
    return SPL_OK;
}

int spl_train_step(spl_net* t, const float* d_in, const float* d_target, int n, int h, int w, float lr,
                   double* sse, void* stream) {
    if (!t) return fail(nullptr, SPL_ERR_INVALID_ARG, "NULL network");
    int rc = check_shape(t, n, h, w);
    if (rc) return rc;
    if (!d_in || !d_target) return fail(t, SPL_ERR_INVALID_ARG, "NULL buffer");
    if (!(lr > 0.0f)) return fail(t, SPL_ERR_INVALID_ARG, "learning rate must be positive");
    DeviceGuard g(t->device);
    hipStream_t s = stream ? (hipStream_t)stream : t->stream;
    if (t->cap < (size_t)n * h * w) HIPCALL(t, hipStreamSynchronize(s));  // buffers regrown
    rc = ensure_work(t, n, h, w);
    if (rc) return rc;
    forward(t, d_in, n, h, w, t->y, s);
    backward(t, d_in, d_target, n, h, w, s);
    ++t->adam_t;
    const float bc1 = 1.0f - powf(0.9f, (float)t->adam_t), bc2 = 1.0f - powf(0.999f, (float)t->adam_t);
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, s, t->params, t->grad, t->adam_m,
                       t->adam_v, kParams, lr, bc1, bc2);
    rc = repack(t, s);
    if (rc) return rc;
    HIPCALL(t, hipGetLastError());
    if (sse) return read_sse(t, sse, s);
    return SPL_OK;
}

int spl_judge_sse(spl_net* t, const float* d_frame, int w, int h, const float* d_patches,
                  int patch, int gx, int gy, int ox, int oy, float* d_sse, int accumulate,
                  void* stream) {
    if (!t) return fail(nullptr, SPL_ERR_INVALID_ARG, "NULL network");
    if (!d_frame || !d_patches || !d_sse || w <= 0 || h <= 0 || patch < 1 || gx < 1 || gy < 1 ||
        ox < 0 || oy < 0 || ox + patch > w / gx || oy + patch > h / gy)
        return fail(t, SPL_ERR_INVALID_ARG, "invalid judge arguments");
    DeviceGuard g(t->device);
    hipStream_t s = stream ? (hipStream_t)stream : t->stream;
    hipLaunchKernelGGL(k_judge_patches, dim3(1), dim3(256), 0, s, d_frame, w, h, d_patches, patch,
                       gx, gy, ox, oy, accumulate, d_sse);
    HIPCALL(t, hipGetLastError());
    return SPL_OK;
}

int spl_judge_sse_frames(spl_net* t, const float* const* frames, int n, int w, int h,
                         const float* d_sparse, int gx, int gy, int ox, int oy, float* d_sse,
                         float* d_sse4, void* stream) {
    if (!t) return fail(nullptr, SPL_ERR_INVALID_ARG, "NULL network");
    if (!frames || !d_sparse || !d_sse || n <= 0 || w <= 0 || h <= 0 || gx < 1 || gy < 1 ||
        ox < 0 || oy < 0 || ox >= w / gx || oy >= h / gy)
        return fail(t, SPL_ERR_INVALID_ARG, "invalid judge arguments");
    for (int k = 0; k < n; ++k)
        if (!frames[k]) return fail(t, SPL_ERR_INVALID_ARG, "NULL frame %d", k);
    DeviceGuard g(t->device);
    hipStream_t s = stream ? (hipStream_t)stream : t->stream;
    if (!t->judge_part)
        HIPCALL(t, hipMalloc(&t->judge_part, sizeof(float) * 2 * kJudgeFrames * kJudgeBlocks));
    const size_t per = (size_t)gx * gy;
    for (int k0 = 0; k0 < n; k0 += kJudgeFrames) {
        const int m = std::min(kJudgeFrames, n - k0);
        JudgeFrames jf{};
        for (int k = 0; k < m; ++k) jf.f[k] = frames[k0 + k];
        hipLaunchKernelGGL(k_judge_grid, dim3(kJudgeBlocks, m), dim3(256), 0, s, jf, w, h,
                           d_sparse + k0 * per, gx, gy, ox, oy, t->judge_part);
        hipLaunchKernelGGL(k_judge_reduce, dim3(1), dim3(64), 0, s, (const float*)t->judge_part, m,
                           d_sse + k0, d_sse4 ? d_sse4 + k0 : nullptr);
    }
    HIPCALL(t, hipGetLastError());
    return SPL_OK;
}

int spl_judge_flags_bound(spl_net* t, const float* d_sse, const float* d_sse4, int n, int npix,
                          float mse_max, float z, unsigned int* d_flags, void* stream) {
    if (!t) return fail(nullptr, SPL_ERR_INVALID_ARG, "NULL network");
    if (!d_sse || !d_sse4 || !d_flags || n <= 0 || npix <= 0 || !(z >= 0.0f))
        return fail(t, SPL_ERR_INVALID_ARG, "invalid judge arguments");
    DeviceGuard g(t->device);
    hipStream_t s = stream ? (hipStream_t)stream : t->stream;
    hipLaunchKernelGGL(k_judge_flags_bound, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       d_sse, d_sse4, n, 1.0f / (float)npix, mse_max, z, (uint32_t*)d_flags);
    HIPCALL(t, hipGetLastError());
    return SPL_OK;
}

int spl_judge_flags(spl_net* t, const float* d_sse, int n, int npix, float mse_max,
                    unsigned int* d_flags, void* stream) {
    if (!t) return fail(nullptr, SPL_ERR_INVALID_ARG, "NULL network");
    if (!d_sse || !d_flags || n <= 0 || npix <= 0)
        return fail(t, SPL_ERR_INVALID_ARG, "invalid judge-flag arguments");
    DeviceGuard g(t->device);
    hipStream_t s = stream ? (hipStream_t)stream : t->stream;
    hipLaunchKernelGGL(k_judge_flags, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_sse, n,
                       1.0f / (float)npix, mse_max, (uint32_t*)d_flags);
    HIPCALL(t, hipGetLastError());
    return SPL_OK;
}

int spl_upsample2(spl_net* t, const float* d_src, int w, int h, float* d_dst, void* stream) {
    if (!t) return fail(nullptr, SPL_ERR_INVALID_ARG, "NULL network");
    if (!d_src || !d_dst || w <= 0 || h <= 0) return fail(t, SPL_ERR_INVALID_ARG, "invalid upsample arguments");
    DeviceGuard g(t->device);
    hipStream_t s = stream ? (hipStream_t)stream : t->stream;
    const size_t n = (size_t)4 * w * h;
    hipLaunchKernelGGL(k_upsample2, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_src, w, h, d_dst);
    HIPCALL(t, hipGetLastError());
    return SPL_OK;
}

}  // extern "C"
