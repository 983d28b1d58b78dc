// render_kernels.h -- the ray-march and 2D post-pass kernels (gfx950).
//
// One work-item per pixel/ray, as in the reference's 2D NDRange (volumerender.py:366-386).
// Workgroups of 256 cover 16x16 pixel tiles, each wave64 an 8x8 sub-tile so the 64 rays
// of a wave sample neighbouring voxels.  Tiles are handed out XCD-aware (see tile_pixel).
#pragma once
#include <type_traits>

#include "ray_math.h"
#include "volume_sampler.h"

namespace spr {

constexpr int kLoopUnroll = 16;  // LOOPUNROLL, volume_kernel.cl:16

struct RayParams {
    float invP[16];
    float invM[16];
    float box[6];
    float min_val, max_val, gamma, alpha_pow;
    float iso_val;
    int max_steps;     // iso: S
    int nouter;        // max_project: (S / numParts) / 16
    int current_part;  // max_project: multipass part
    uint32_t W, H;
    uint32_t ntx, nty, ntiles;  // 16x16 tiles per row, tile rows, total tiles
    uint32_t tile_map;          // kTileMap* below
    uint32_t row_pad;           // staged kernels: pad LDS slab rows to an odd multiple of 16 B
    uint32_t box16;             // staged kernels: 16-sample groups per staged box (1, 2, 4)
    uint32_t skip16;            // staged iso march with skipping: 16-sample chunks per
                                // coarse skip test (1: none)
    const uint32_t* order;      // kTileMapCostXcd: workgroup -> work item (k_tile_order)
    const uint32_t* gate;       // max projections: render only if *gate != 0 (nullptr: always)
};

// a max projection whose device-side gate is closed leaves its outputs untouched (uniform
// over the launch: every thread reads the same word)
__device__ __forceinline__ bool gate_closed(const RayParams& p) {
    return p.gate != nullptr && *p.gate == 0u;
}

// Workgroup -> 16x16 tile assignment.  The dispatcher hands consecutive workgroups to
// the 8 XCDs round-robin (workgroup b runs on XCD b % 8), each XCD with its own L2.
//   kTileMapInterleaved  XCD x renders tile rows x, x+8, x+16, ...: whole rows of tiles
//                        (neighbouring beams through the volume) share one L2, and every
//                        XCD gets a similar share of the frame's hit rays.  The grid is
//                        8 * ceil(nty/8) * ntx workgroups; the surplus ones exit at once.
//   kTileMapBands        XCD x renders one contiguous horizontal band of the frame.
//   kTileMapLinear       tile = workgroup id (row-major).
//   kTileMapCostXcd      as kTileMapInterleaved, but each XCD takes its tile rows' work
//                        in descending estimated cost (p.order, from k_tile_cost and
//                        k_tile_order), row-major among equal costs: the dispatcher hands
//                        the slowest work out first, so the frame does not end on a few
//                        long tiles started late.  The order changes only which
//                        workgroup renders a tile and when: every pixel is the same.
constexpr uint32_t kTileMapInterleaved = 0, kTileMapBands = 1, kTileMapLinear = 2,
                   kTileMapCostXcd = 3;

__host__ __device__ inline uint32_t tile_grid(uint32_t tile_map, uint32_t ntx, uint32_t nty) {
    return tile_map == kTileMapInterleaved || tile_map == kTileMapCostXcd
               ? 8u * ((nty + 7u) / 8u) * ntx
               : ntx * nty;
}

// pixel of this thread inside 16x16 tile (tx, ty): each wave64 takes an 8x8 quarter
__device__ __forceinline__ void tile_thread_pixel(uint32_t tx, uint32_t ty, uint32_t* x, uint32_t* y) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    *x = tx * 16u + (wave & 1u) * 8u + (lane & 7u);
    *y = ty * 16u + (wave >> 1) * 8u + (lane >> 3);
}

__device__ __forceinline__ bool tile_pixel(const RayParams& p, uint32_t* x, uint32_t* y) {
    const uint32_t b = blockIdx.x;
    uint32_t tx, ty;
    if (p.tile_map == kTileMapInterleaved) {
        const uint32_t xcd = b & 7u, k = b >> 3;
        ty = xcd + 8u * (k / p.ntx);
        tx = k % p.ntx;
    } else if (p.tile_map == kTileMapCostXcd) {
        const uint32_t t = p.order[b];  // >= ntiles: a surplus workgroup (ty >= nty)
        tx = t % p.ntx;
        ty = t / p.ntx;
    } else if (p.tile_map == kTileMapBands) {
        const uint32_t q = p.ntiles >> 3, r = p.ntiles & 7u;
        const uint32_t xcd = b & 7u, k = b >> 3;
        const uint32_t t = xcd * q + min(xcd, r) + k;
        tx = t % p.ntx;
        ty = t / p.ntx;
    } else {
        tx = b % p.ntx;
        ty = b / p.ntx;
    }
    tile_thread_pixel(tx, ty, x, y);
    return ty < p.nty && *x < p.W && *y < p.H;
}

__device__ __forceinline__ float apply_gamma(float c, float gamma) {
    // clamp(pow(colVal, gamma), 0, 1) (volume_kernel.cl:162); pow(x, 1) == x
    float g = (gamma == 1.0f) ? c : powf(c, gamma);
    return fminf(fmaxf(g, 0.0f), 1.0f);
}

// volume_kernel.cl:20-185 (float) and :190-345 (short, also used for uint8)
// one ray per thread, every sample read from global memory (all modes and filters)
template <class Vol, int FILTER>
__device__ __forceinline__ void mip_rays(const Vol& vol, const RayParams& p, uint32_t x, uint32_t y,
                                         bool inside, float* __restrict__ out,
                                         float* __restrict__ alpha) {
    constexpr bool kShort = !std::is_same<typename Vol::value_type, float>::value;
    if (!inside || gate_closed(p)) return;
    const uint32_t o = x + p.W * y;
    F4 orig, direc;
    ray_setup(p.invP, p.invM, x, y, p.W, p.H, &orig, &direc);
    float tnear, tfar;
    if (!intersect_box(orig, direc, p.box, &tnear, &tfar)) {
        out[o] = 0.0f;
        alpha[o] = kShort ? 0.0f : -1.0f;
        return;
    }
    if (tnear < 0.0f) tnear = 0.0f;
    const float dt = fabsf(tfar - tnear) / (float)(p.nouter * kLoopUnroll);
    orig = add4(orig, scl4((float)p.current_part * dt, direc));
    const F4 delta = scl4(0.5f * dt, direc);
    F4 pos = scl4(0.5f, add4(mk4(1.0f + orig.x, 1.0f + orig.y, 1.0f + orig.z, 1.0f + orig.w),
                             scl4(tnear, direc)));
    float px = pos.x, py = pos.y, pz = pos.z;
    float col = 0.0f;
    if (p.alpha_pow == 0.0f) {
        for (int i = 0; i <= p.nouter; ++i) {
#pragma unroll
            for (int j = 0; j < kLoopUnroll; ++j) {
                float v = vol.template sample<FILTER>(px, py, pz);
                col = fmaxf(col, v);
                px = px + delta.x;
                py = py + delta.y;
                pz = pz + delta.z;
            }
        }
        col = (p.max_val == 0.0f) ? col : (col - p.min_val) / (p.max_val - p.min_val);
    } else {
        float cumsum = 1.0f;
        const float a2 = p.alpha_pow * p.alpha_pow;
        const float a2s = 0.1f * p.alpha_pow * p.alpha_pow;
        for (int i = 0; i <= p.nouter; ++i) {
            for (int j = 0; j < kLoopUnroll; ++j) {
                float v = vol.template sample<FILTER>(px, py, pz);
                v = (p.max_val == 0.0f) ? v : (v - p.min_val) / (p.max_val - p.min_val);
                col = fmaxf(col, cumsum * v);
                if (kShort)
                    cumsum *= (1.0f - a2s * v);
                else
                    cumsum *= (1.0f - a2 * fminf(fmaxf(v, 0.0f), 1.0f));
                px = px + delta.x;
                py = py + delta.y;
                pz = pz + delta.z;
                if (cumsum <= 0.01f) break;  // inner loop only (volume_kernel.cl:148)
            }
        }
    }
    col = apply_gamma(col, p.gamma);
    const float a = kShort ? tnear : 1.0f;
    if (p.current_part == 0) {
        out[o] = col;
        alpha[o] = a;
    } else {
        out[o] = fmaxf(col, out[o]);
        alpha[o] = fmaxf(a, alpha[o]);
    }
}

template <class Vol, int FILTER>
__global__ __launch_bounds__(256) void k_max_project(Vol vol, RayParams p,
                                                     float* __restrict__ out,
                                                     float* __restrict__ alpha) {
    uint32_t x, y;
    const bool inside = tile_pixel(p, &x, &y);
    mip_rays<Vol, FILTER>(vol, p, x, y, inside, out, alpha);
}

// outputs of a ray without an iso crossing (iso_kernel.cl:133-141)
__device__ __forceinline__ void iso_miss(uint32_t o, float* __restrict__ out,
                                         float* __restrict__ alpha, float* __restrict__ depth,
                                         float* __restrict__ normals) {
    out[o] = 0.0f;
    alpha[o] = 0.0f;
    depth[o] = INFINITY;
    normals[3 * o] = 0.0f;
    normals[3 * o + 1] = 0.0f;
    normals[3 * o + 2] = 0.0f;
}

// A ray whose march crossed the iso value between samples i-1 and i (iso_kernel.cl:
// 143-225): 10-substep refinement from sample i-1, the 4-tap gradient at the refined
// point, Phong shading; t_hit enters as tnear + i dt.
template <int FILTER, class Vol>
__device__ __forceinline__ void iso_finish(const Vol& vol, const RayParams& p, uint32_t o,
                                           float tnear, float dt, F4 delta, F4 pos0, int i,
                                           int isGreater, F4 direc, float* __restrict__ out,
                                           float* __restrict__ alpha, float* __restrict__ depth,
                                           float* __restrict__ normals) {
    const float iso = p.iso_val;
    float t_hit = tnear + (float)i * dt;
    // 10-substep refinement from step i-1 (iso_kernel.cl:145-162)
    const float fb = 10.0f;
    F4 delta2 = mk4(delta.x / fb, delta.y / fb, delta.z / fb, delta.w / fb);
    float dt2 = dt / fb;
    F4 pos = add4(pos0, scl4((float)(i - 1), delta));
    // the ten substep samples are loaded together (their positions do not depend on the
    // values), then scanned for the first one across: the same samples and the same
    // running sums as a loop that loads one substep at a time and breaks
    float vs[10];
    {
        F4 q = pos;
#pragma unroll
        for (int j = 0; j < 10; j++) {
            vs[j] = vol.template sample<FILTER>(q.x, q.y, q.z);
            q = add4(q, delta2);
        }
    }
    for (int j = 1; j <= 10; j++) {
        pos = add4(pos, delta2);
        t_hit += dt2;
        if ((int)(vs[j - 1] > iso) != isGreater) break;
    }
    // Phong at the refined point, 4-tap gradient: weights (2,1) in x,y and (1,1) in z
    F4 light = normalize4(mult(p.invM, mk4(2.0f, -1.0f, -2.0f, 0.0f)));
    float h = dt;
    h *= p.gamma * p.gamma;
    float h2 = 2.0f * h;
    F4 n;
    n.x = 2.0f * vol.template sample<FILTER>(pos.x + h, pos.y, pos.z) -
          2.0f * vol.template sample<FILTER>(pos.x + (-h), pos.y, pos.z);
    n.x = n.x + vol.template sample<FILTER>(pos.x + h2, pos.y, pos.z);
    n.x = n.x - vol.template sample<FILTER>(pos.x + (-h2), pos.y, pos.z);
    n.y = 2.0f * vol.template sample<FILTER>(pos.x, pos.y + h, pos.z) -
          2.0f * vol.template sample<FILTER>(pos.x, pos.y + (-h), pos.z);
    n.y = n.y + vol.template sample<FILTER>(pos.x, pos.y + h2, pos.z);
    n.y = n.y - vol.template sample<FILTER>(pos.x, pos.y + (-h2), pos.z);
    n.z = vol.template sample<FILTER>(pos.x, pos.y, pos.z + h) -
          vol.template sample<FILTER>(pos.x, pos.y, pos.z + (-h));
    n.z = n.z + vol.template sample<FILTER>(pos.x, pos.y, pos.z + h2);
    n.z = n.z - vol.template sample<FILTER>(pos.x, pos.y, pos.z + (-h2));
    n.w = 0.0f;
    n = scl4(1.0f - (float)(2 * isGreater), normalize4(n));
    F4 refl = sub4(scl4(2.0f * dot4(light, n), n), light);
    float diffuse = fmaxf(0.0f, dot4(light, n));
    float specular = powf(fmaxf(0.0f, dot4(normalize4(refl), normalize4(direc))), 10.0f);
    float col = 0.3f + 0.4f * diffuse;
    col = col + ((float)(diffuse > 0.0f) * 0.3f) * specular;
    out[o] = col;
    alpha[o] = tnear;
    depth[o] = t_hit;
    normals[3 * o] = n.x;
    normals[3 * o + 1] = n.y;
    normals[3 * o + 2] = n.z;
}

// iso_kernel.cl:17-225
template <class Vol, int FILTER>
__global__ __launch_bounds__(256) void k_iso_surface(Vol vol, RayParams p,
                                                     float* __restrict__ out,
                                                     float* __restrict__ alpha,
                                                     float* __restrict__ depth,
                                                     float* __restrict__ normals) {
    uint32_t x, y;
    if (!tile_pixel(p, &x, &y)) return;
    const uint32_t o = x + p.W * y;
    F4 orig, direc;
    ray_setup(p.invP, p.invM, x, y, p.W, p.H, &orig, &direc);
    float tnear, tfar;
    bool hit = intersect_box(orig, direc, p.box, &tnear, &tfar);
    int isGreater = 0;
    int i = 1;
    float dt = 0.0f;
    F4 pos0 = mk4(0, 0, 0, 0), delta = mk4(0, 0, 0, 0);
    const float iso = p.iso_val;
    if (hit) {
        if (tnear < 0.0f) tnear = 0.0f;
        dt = 1.0f * (tfar - tnear) / ((float)p.max_steps - 1.0f);
        delta = scl4(0.5f * dt, direc);
        pos0 = scl4(0.5f, add4(mk4(1.0f + orig.x, 1.0f + orig.y, 1.0f + orig.z, 1.0f + orig.w),
                               scl4(tnear, direc)));
        float px = pos0.x, py = pos0.y, pz = pos0.z;
        float v = vol.template sample<FILTER>(px, py, pz);
        isGreater = v > iso;
        hit = false;
        for (i = 1; i < p.max_steps; i++) {
            px = px + delta.x;
            py = py + delta.y;
            pz = pz + delta.z;
            v = vol.template sample<FILTER>(px, py, pz);
            if ((int)(v > iso) != isGreater) {
                hit = true;
                break;
            }
        }
    }
    if (!hit) {
        iso_miss(o, out, alpha, depth, normals);
        return;
    }
    iso_finish<FILTER>(vol, p, o, tnear, dt, delta, pos0, i, isGreater, direc, out, alpha, depth,
                       normals);
}

struct ConvWeights {
    float w[16];
    int nh;
};

// convolve_2d.cl: normalised separable blur over the in-range taps.  AXIS 0 = x, 1 = y;
// C = 1 (scalar, conv_x/conv_y) or 3 (normals, conv_vec_x/conv_vec_y).
template <int AXIS, int C>
__global__ __launch_bounds__(256) void k_conv(const float* __restrict__ in, float* __restrict__ out,
                                              int W, int H, ConvWeights cw) {
    int i = blockIdx.x * 16 + (threadIdx.x & 15);
    int j = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (i >= W || j >= H) return;
    const int Nh = cw.nh;
    const int c = AXIS == 0 ? i : j;
    const int N = AXIS == 0 ? W : H;
    const int hs = ((c - Nh / 2) < 0) ? Nh / 2 - c : 0;
    const int he = ((c + Nh / 2) >= N) ? Nh - (c + Nh / 2 - N + 1) : Nh;
    const int start = c - Nh / 2;
    float r[C];
#pragma unroll
    for (int k = 0; k < C; ++k) r[k] = 0.0f;
    float s = 0.0f;
    for (int ht = hs; ht < he; ++ht) {
        const float w = cw.w[ht];
        const size_t q = AXIS == 0 ? (size_t)(start + ht) + (size_t)j * W
                                   : (size_t)i + (size_t)(start + ht) * W;
        s += w;
#pragma unroll
        for (int k = 0; k < C; ++k) r[k] += w * in[C * q + k];
    }
    const size_t o = (size_t)i + (size_t)j * W;
#pragma unroll
    for (int k = 0; k < C; ++k) out[C * o + k] = r[k] / s;
}

// occlusion.cl:41-82.  Point i of every pixel draws its radius and angle from
// lcg_random(x + a_i, y + b_i) and lcg_random(x + c_i, y + d_i), whose offsets a_i .. d_i
// (rand_int of i and i*i) do not depend on the pixel: the workgroup computes them once
// into an LDS table for the first kOccTable points (the same expressions, so the same
// values), and the pixels read them from there.
constexpr int kOccTable = 256;

__device__ __forceinline__ void occ_offsets(uint32_t i, float& ra, float& rb, float& pa,
                                            float& pb) {
    const uint32_t ii = i * i;
    ra = lcg_rand_int(i, ii, 0, 1000);
    rb = lcg_rand_int(ii, i, 294, 97701);
    pa = lcg_rand_int(ii, i, 0, 1997);
    pb = lcg_rand_int(i, ii, 569, 17633);
}

static __global__ __launch_bounds__(256) void k_occlusion(float* __restrict__ out, int W, int H,
                                                          int radius, int n_points,
                                                          const float* __restrict__ depth) {
    __shared__ float4 tab[kOccTable];
    const int nt = min(n_points, kOccTable);
    for (int i = (int)threadIdx.x; i < nt; i += 256) {
        float4 t;
        occ_offsets((uint32_t)i, t.x, t.y, t.z, t.w);
        tab[i] = t;
    }
    __syncthreads();
    int x = blockIdx.x * 16 + (threadIdx.x & 15);
    int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const float MPI_2 = 6.2831853071795f;
    const float depth0 = depth[x + y * W];
    float occ = 0.0f;
    for (int i = 0; i < n_points; ++i) {
        float oa, ob, oc, od;
        if (i < kOccTable) {
            const float4 t = tab[i];
            oa = t.x;
            ob = t.y;
            oc = t.z;
            od = t.w;
        } else {
            occ_offsets((uint32_t)i, oa, ob, oc, od);
        }
        float ra = (float)x + oa;
        float rb = (float)y + ob;
        float r = (float)(uint32_t)radius * lcg_random((uint32_t)ra, (uint32_t)rb);
        float pa = (float)x + oc;
        float pb = (float)y + od;
        float phi = MPI_2 * lcg_random((uint32_t)pa, (uint32_t)pb);
        float sn, cs;
        spr_sincosf(phi, &sn, &cs);
        int x2 = clampi((int)((float)x + r * cs), 0, W - 1);
        int y2 = clampi((int)((float)y + r * sn), 0, H - 1);
        float d = depth[x2 + y2 * W];
        occ += (d < depth0) ? 1.0f : 0.0f;
    }
    out[x + W * y] = occ / (float)(uint32_t)n_points;
}

// iso_kernel.cl:505-588 -- the final iso output (overwrites `out`)
static __global__ __launch_bounds__(256) void k_shading(float* __restrict__ out, RayParams p,
                                                 float occ_strength,
                                                 const float* __restrict__ normals,
                                                 const float* __restrict__ depth,
                                                 const float* __restrict__ occl) {
    uint32_t x = blockIdx.x * 16 + (threadIdx.x & 15);
    uint32_t y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= p.W || y >= p.H) return;
    F4 orig, direc;
    ray_setup(p.invP, p.invM, x, y, p.W, p.H, &orig, &direc);
    F4 light = normalize4(mult(p.invM, mk4(2.0f, -1.0f, -2.0f, 0.0f)));
    const uint32_t o = x + p.W * y;
    F4 n = mk4(normals[3 * o], normals[3 * o + 1], normals[3 * o + 2], 0.0f);
    float occ = occl[o];
    float d = depth[o];
    n = normalize4(n);
    F4 refl = sub4(scl4(2.0f * dot4(light, n), n), light);
    float diffuse = fmaxf(0.0f, dot4(light, n));
    float specular = powf(fmaxf(0.0f, dot4(normalize4(refl), normalize4(direc))), 10.0f);
    float col = 0.5f + 0.3f * diffuse;
    col = col + ((float)(diffuse > 0.0f) * 0.2f) * specular;
    col = (1.0f - occ_strength) * col + (occ_strength * col) * (1.0f - occ);
    // iso_kernel.cl:579 multiplies by the double literal 1.0
    col = (float)((double)((1.0f - occ_strength) * col) + (1.0 * (double)occ_strength) * (double)col);
    col *= (d < INFINITY) ? 1.0f : 0.0f;
    out[o] = col;
}

// rays of a w x h frame whose box test hits (for the reference sample count)
static __global__ __launch_bounds__(256) void k_count_hits(RayParams p, unsigned long long* __restrict__ count) {
    uint32_t x, y;
    bool inside = tile_pixel(p, &x, &y);
    int hit = 0;
    if (inside) {
        F4 orig, direc;
        ray_setup(p.invP, p.invM, x, y, p.W, p.H, &orig, &direc);
        float tnear, tfar;
        hit = intersect_box(orig, direc, p.box, &tnear, &tfar) ? 1 : 0;
    }
    int n = __syncthreads_count(hit);
    if (threadIdx.x == 0 && n > 0) atomicAdd(count, (unsigned long long)n);
}

// per-workgroup min/max partials of a volume (GLWidget._get_min_max replacement)
template <typename T>
__global__ __launch_bounds__(256) void k_minmax(const T* __restrict__ data, size_t n,
                                                float* __restrict__ partial) {
    __shared__ float smin[256], smax[256];
    float mn = INFINITY, mx = -INFINITY;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        float v = (float)data[i];
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = smin[0];
        partial[2 * blockIdx.x + 1] = smax[0];
    }
}

}  // namespace spr
