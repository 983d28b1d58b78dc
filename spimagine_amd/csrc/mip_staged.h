// mip_staged.h -- max projection with LDS staging of the volume (per wave or per workgroup).
//
// Same result, bit for bit, as k_max_project on the padded layout (volume_kernel.cl:20-185,
// :190-345 with alpha_pow == 0): identical sample positions, weights and lerp order; only
// where the eight corners are read from changes.
//
// Why: the 64 rays of a wave sit ~1 voxel apart on a plane oblique to the volume axes, so
// a wave-wide global gather touches ~55 distinct 64-byte granules and the vector-L1 tag
// path, not HBM or VALU, limits the plain kernel.  Here each wave walks its rays in chunks
// of the reference's 16-sample inner loop: it bounds the padded-voxel box those 16 samples
// of its 64 rays touch (DPP reductions), copies the box into its own LDS slab with aligned
// 16-byte row loads (one granule per 4 lanes), then gathers the corners from LDS.  A box
// larger than the slab (rays of one wave with very different dt, near the box silhouette)
// is sampled from global memory for that chunk instead.  In the per-wave form there is no
// workgroup barrier: each wave owns its slab and LDS operations of one wave complete in
// order.  In the shared form the four waves of a 16x16 tile stage one union box (fewer
// bytes than four overlapping boxes, and boxes of oblique views fit more often) with two
// workgroup barriers per box; every ray of a workgroup takes the same number of samples
// (the reference's fixed count), so the waves advance in step.
#pragma once
#include "render_kernels.h"

namespace spr {

typedef unsigned short spr_us2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_min_u16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(spr_us2, a),
                                                                  __builtin_bit_cast(spr_us2, b)));
}

// wave64 minima of three packed u16 pairs at once (row_shr 1/2/4/8, row_bcast 15/31):
// the three chains interleave, so no chain waits on its own DPP read; uniform results
__device__ __forceinline__ void wave_min3_u16x2(uint32_t& a, uint32_t& b, uint32_t& c) {
#define SPR_MIN3_STEP(CTRL, RM)                                                                \
    {                                                                                          \
        const uint32_t da = __builtin_amdgcn_update_dpp(a, a, CTRL, RM, 0xf, false);           \
        const uint32_t db = __builtin_amdgcn_update_dpp(b, b, CTRL, RM, 0xf, false);           \
        const uint32_t dc = __builtin_amdgcn_update_dpp(c, c, CTRL, RM, 0xf, false);           \
        a = pk_min_u16(a, da);                                                                 \
        b = pk_min_u16(b, db);                                                                 \
        c = pk_min_u16(c, dc);                                                                 \
    }
    SPR_MIN3_STEP(0x111, 0xf)
    SPR_MIN3_STEP(0x112, 0xf)
    SPR_MIN3_STEP(0x114, 0xf)
    SPR_MIN3_STEP(0x118, 0xf)
    SPR_MIN3_STEP(0x142, 0xa)
    SPR_MIN3_STEP(0x143, 0xc)
#undef SPR_MIN3_STEP
    a = __builtin_amdgcn_readlane(a, 63);
    b = __builtin_amdgcn_readlane(b, 63);
    c = __builtin_amdgcn_readlane(c, 63);
}

// The same minima over each half of the wave: (a, b, c) of lanes 0-31 and (a2, b2, c2) of
// lanes 32-63 (the reduction without its last step), uniform results
__device__ __forceinline__ void wave_min3_u16x2_halves(uint32_t& a, uint32_t& b, uint32_t& c,
                                                       uint32_t& a2, uint32_t& b2, uint32_t& c2) {
#define SPR_MIN3_STEP(CTRL, RM)                                                                \
    {                                                                                          \
        const uint32_t da = __builtin_amdgcn_update_dpp(a, a, CTRL, RM, 0xf, false);           \
        const uint32_t db = __builtin_amdgcn_update_dpp(b, b, CTRL, RM, 0xf, false);           \
        const uint32_t dc = __builtin_amdgcn_update_dpp(c, c, CTRL, RM, 0xf, false);           \
        a = pk_min_u16(a, da);                                                                 \
        b = pk_min_u16(b, db);                                                                 \
        c = pk_min_u16(c, dc);                                                                 \
    }
    SPR_MIN3_STEP(0x111, 0xf)
    SPR_MIN3_STEP(0x112, 0xf)
    SPR_MIN3_STEP(0x114, 0xf)
    SPR_MIN3_STEP(0x118, 0xf)
    SPR_MIN3_STEP(0x142, 0xa)
#undef SPR_MIN3_STEP
    a2 = __builtin_amdgcn_readlane(a, 63);
    b2 = __builtin_amdgcn_readlane(b, 63);
    c2 = __builtin_amdgcn_readlane(c, 63);
    a = __builtin_amdgcn_readlane(a, 31);
    b = __builtin_amdgcn_readlane(b, 31);
    c = __builtin_amdgcn_readlane(c, 31);
}

// q = n / d and n - q d for n, d < 2^16 (d > 0): a float quotient off by at most one,
// corrected (no integer division)
__device__ __forceinline__ void divmod16(uint32_t n, uint32_t d, float rd, uint32_t& q,
                                         uint32_t& m) {
    q = (uint32_t)((float)n * rd);
    int32_t r = (int32_t)n - (int32_t)__umul24(q, d);
    if (r < 0) {
        q -= 1u;
        r += (int32_t)d;
    } else if (r >= (int32_t)d) {
        q += 1u;
        r -= (int32_t)d;
    }
    m = (uint32_t)r;
}

// wave64 minima of three unsigned 32-bit values (the same DPP chain as wave_min3_u16x2);
// uniform results
__device__ __forceinline__ void wave_min3_u32(uint32_t& a, uint32_t& b, uint32_t& c) {
#define SPR_MIN3_STEP(CTRL, RM)                                                                \
    {                                                                                          \
        const uint32_t da = __builtin_amdgcn_update_dpp(a, a, CTRL, RM, 0xf, false);           \
        const uint32_t db = __builtin_amdgcn_update_dpp(b, b, CTRL, RM, 0xf, false);           \
        const uint32_t dc = __builtin_amdgcn_update_dpp(c, c, CTRL, RM, 0xf, false);           \
        a = min(a, da);                                                                        \
        b = min(b, db);                                                                        \
        c = min(c, dc);                                                                        \
    }
    SPR_MIN3_STEP(0x111, 0xf)
    SPR_MIN3_STEP(0x112, 0xf)
    SPR_MIN3_STEP(0x114, 0xf)
    SPR_MIN3_STEP(0x118, 0xf)
    SPR_MIN3_STEP(0x142, 0xa)
    SPR_MIN3_STEP(0x143, 0xc)
#undef SPR_MIN3_STEP
    a = __builtin_amdgcn_readlane(a, 63);
    b = __builtin_amdgcn_readlane(b, 63);
    c = __builtin_amdgcn_readlane(c, 63);
}

// The bricks (brick_bounds.h) covering the padded-voxel box [LX, HX] x [LY, HY] x [LZ, HZ]
// (inclusive; every corner a sample of the box reads lies inside): each lane loads up to
// two of them (uniform box; lanes past the count load nothing).  Returns false, loading
// nothing, when the box covers more than kBrickMaxQuery bricks.  keys: this lane's
// (min key, ~max key) so far (all ones: none).
template <typename V>
__device__ __forceinline__ bool brick_load(const V& vol, uint32_t LX, uint32_t LY, uint32_t LZ,
                                           uint32_t HX, uint32_t HY, uint32_t HZ, uint32_t lane,
                                           BrickKey& k) {
    k.kmin = k.nkmax = 0xFFFFFFFFu;
    const uint32_t bx0 = LX >> kBrickShift, by0 = LY >> kBrickShift, bz0 = LZ >> kBrickShift;
    const uint32_t sx = (HX >> kBrickShift) - bx0 + 1u, sy = (HY >> kBrickShift) - by0 + 1u,
                   sz = (HZ >> kBrickShift) - bz0 + 1u;
    const uint32_t n = sx * sy * sz;
    if (n > kBrickMaxQuery || vol.bricks.keys == nullptr) return false;
    const float rsx = __builtin_amdgcn_rcpf((float)sx), rsy = __builtin_amdgcn_rcpf((float)sy);
#pragma unroll
    for (uint32_t r = 0; r < kBrickMaxQuery / 64u; ++r) {
        const uint32_t j = lane + 64u * r;
        if (j < n) {
            uint32_t q, dx, dy, dz;
            divmod16(j, sx, rsx, q, dx);
            divmod16(q, sy, rsy, dz, dy);
            const uint32_t idx =
                (__umul24(bz0 + dz, vol.bricks.nby) + by0 + dy) * vol.bricks.nbx + bx0 + dx;
            const BrickKey b = vol.bricks.keys[idx];
            k.kmin = min(k.kmin, b.kmin);
            k.nkmax = min(k.nkmax, b.nkmax);
        }
    }
    return true;
}

// The voxel pair (x, x+1) from a dword-aligned slab address a and the byte shift sh of the
// pair within that dword, read as two aligned dwords (one ds_read2_b32) and shifted into
// place with v_alignbyte: LDS accesses that are not dword aligned stall the LDS pipeline.
// The four corner rows of a sample share the shift (row and plane strides are multiples
// of 16).  Reads up to 6 bytes past the pair (the slab arrays carry 16 bytes of slack).
__device__ __forceinline__ void lds_pair_at(const uint8_t* a, uint32_t sh, float& v0, float& v1,
                                            uint16_t) {
    const uint32_t* d = reinterpret_cast<const uint32_t*>(a);
    const uint32_t w = __builtin_amdgcn_alignbyte(d[1], d[0], sh);
    v0 = (float)(w & 0xFFFFu);
    v1 = (float)(w >> 16);
}
__device__ __forceinline__ void lds_pair_at(const uint8_t* a, uint32_t sh, float& v0, float& v1,
                                            uint8_t) {
    const uint32_t* d = reinterpret_cast<const uint32_t*>(a);
    const uint32_t w = __builtin_amdgcn_alignbyte(d[1], d[0], sh);
    v0 = (float)(w & 0xFFu);
    v1 = (float)((w >> 8) & 0xFFu);
}
__device__ __forceinline__ void lds_pair_at(const uint8_t* a, uint32_t, float& v0, float& v1, float) {
    const float* d = reinterpret_cast<const float*>(a);
    v0 = d[0];
    v1 = d[1];
}

// 16-byte aligned load / store units
struct alignas(16) Unit16 {
    uint32_t w[4];
};

// Ray set-up of the staged kernels (volume_kernel.cl:45-106) without outputs: whether the
// ray hits the box, and its sample 0 position and step in texture units.
__device__ __forceinline__ bool staged_ray_pos(const RayParams& p, uint32_t x, uint32_t y,
                                               float& px, float& py, float& pz, float& dx,
                                               float& dy, float& dz, float& tnear) {
    F4 orig, direc;
    ray_setup(p.invP, p.invM, x, y, p.W, p.H, &orig, &direc);
    float tfar;
    if (!intersect_box(orig, direc, p.box, &tnear, &tfar)) return false;
    if (tnear < 0.0f) tnear = 0.0f;
    const float dt = fabsf(tfar - tnear) / (float)(p.nouter * kLoopUnroll);
    orig = add4(orig, scl4((float)p.current_part * dt, direc));
    const F4 delta = scl4(0.5f * dt, direc);
    const F4 pos = scl4(0.5f, add4(mk4(1.0f + orig.x, 1.0f + orig.y, 1.0f + orig.z, 1.0f + orig.w),
                                   scl4(tnear, direc)));
    px = pos.x;
    py = pos.y;
    pz = pos.z;
    dx = delta.x;
    dy = delta.y;
    dz = delta.z;
    return true;
}

// staged_ray_pos, writing a missed ray's outputs (volume_kernel.cl:85-91; the short kernel's
// alpha 0, volume_kernel.cl:258-264)
template <bool kShort>
__device__ __forceinline__ bool staged_ray_begin(const RayParams& p, uint32_t x, uint32_t y,
                                                 bool inside, float* __restrict__ out,
                                                 float* __restrict__ alpha, float& px, float& py,
                                                 float& pz, float& dx, float& dy, float& dz,
                                                 float& tnear) {
    if (!inside) return false;
    if (staged_ray_pos(p, x, y, px, py, pz, dx, dy, dz, tnear)) return true;
    const uint32_t o = x + p.W * y;
    out[o] = 0.0f;
    alpha[o] = kShort ? 0.0f : -1.0f;
    return false;
}

// A staged box's first voxel along x: aligned to 4 bytes (its 16-byte row loads need only
// dword alignment), so a row carries at most 3 bytes before the box.
template <typename T>
constexpr uint32_t kBoxAlign = sizeof(T) >= 4 ? 1u : 4u / (uint32_t)sizeof(T);
// The max projection's nearest filter keeps 16-byte aligned boxes: a nearest sample is one
// LDS read, the copy is most of a box's time, and with 4-byte aligned rows the frames took
// ~6% longer (1024^2 frames of 1024x1024x512 uint16, random poses: 1.50 against 1.41 ms),
// while the linear filter's took ~5% less.
template <typename T, int FILT>
constexpr uint32_t kBoxAlignMip = FILT == 1 ? 16u / (uint32_t)sizeof(T) : kBoxAlign<T>;

// Padded-voxel box of samples first..last of a ray whose sample 0 is at (px, py, pz),
// packed for a min-reduction: b0 = lo.x | lo.y << 16, b1 = lo.z | (0xFFFF - hi.x) << 16,
// b2 = (0xFFFF - hi.y) | (0xFFFF - hi.z) << 16 (hi exclusive).  Inactive lanes give all
// ones (no effect).  Positions other than sample 0 are estimated with one fma; mg* covers
// their distance from the sequentially added positions (at most 63 additions, each off by
// half an ulp).  EST = false: first is 0 and sample 0's corner is exact.
template <typename V, bool EST = false>
__device__ __forceinline__ void staged_box_n(const V& vol, bool active, float first, float last,
                                             float px, float py, float pz, float dx, float dy,
                                             float dz, float mgx, float mgy, float mgz,
                                             uint32_t& b0, uint32_t& b1, uint32_t& b2) {
    b0 = b1 = b2 = 0xFFFFFFFFu;
    if (!active) return;
    const float ex = fmaf(last, dx, px), ey = fmaf(last, dy, py), ez = fmaf(last, dz, pz);
    constexpr bool P2 = V::kPow2;
    const float sx = EST ? fmaf(first, dx, px) : px, sy = EST ? fmaf(first, dy, py) : py,
                sz = EST ? fmaf(first, dz, pz) : pz;
    const float u0x = texel<P2>(sx, vol.fnx), u0y = texel<P2>(sy, vol.fny),
                u0z = texel<P2>(sz, vol.fnz);
    const float u1x = texel<P2>(ex, vol.fnx), u1y = texel<P2>(ey, vol.fny),
                u1z = texel<P2>(ez, vol.fnz);
    uint32_t l0x, l0y, l0z, h0x, h0y, h0z;
    if constexpr (EST) {
        l0x = vol.corner1(floorf(u0x - mgx), vol.mx);
        l0y = vol.corner1(floorf(u0y - mgy), vol.my);
        l0z = vol.corner1(floorf(u0z - mgz), vol.mz);
        h0x = vol.corner1(floorf(u0x + mgx), vol.mx);
        h0y = vol.corner1(floorf(u0y + mgy), vol.my);
        h0z = vol.corner1(floorf(u0z + mgz), vol.mz);
    } else {
        l0x = h0x = vol.corner1(floorf(u0x), vol.mx);
        l0y = h0y = vol.corner1(floorf(u0y), vol.my);
        l0z = h0z = vol.corner1(floorf(u0z), vol.mz);
    }
    const uint32_t lx = min(l0x, vol.corner1(floorf(u1x - mgx), vol.mx));
    const uint32_t ly = min(l0y, vol.corner1(floorf(u1y - mgy), vol.my));
    const uint32_t lz = min(l0z, vol.corner1(floorf(u1z - mgz), vol.mz));
    const uint32_t hx = max(h0x, vol.corner1(floorf(u1x + mgx), vol.mx)) + 1u;
    const uint32_t hy = max(h0y, vol.corner1(floorf(u1y + mgy), vol.my)) + 1u;
    const uint32_t hz = max(h0z, vol.corner1(floorf(u1z + mgz), vol.mz)) + 1u;
    b0 = lx | (ly << 16);
    b1 = lz | ((0xFFFFu - hx) << 16);
    b2 = (0xFFFFu - hy) | ((0xFFFFu - hz) << 16);
}

template <typename V, int CH>
__device__ __forceinline__ void staged_chunk_box(const V& vol, bool active,
                                                 float px, float py, float pz, float dx, float dy,
                                                 float dz, float mgx, float mgy, float mgz,
                                                 uint32_t& b0, uint32_t& b1, uint32_t& b2) {
    staged_box_n(vol, active, 0.0f, (float)(CH - 1), px, py, pz, dx, dy, dz, mgx, mgy, mgz, b0, b1,
                 b2);
}

// Copies rows r = z SY + y of the box into slab[r SX TS ...] with NT threads (t = this
// thread's index among them): 16-byte buffer loads, nu units per row, NT / nu rows a pass.
using U32x4 = __attribute__((ext_vector_type(4))) uint32_t;

// Copies rows r = z SY + y of the box into slab[r rowb ...] with NT threads (t = this
// thread's index among them): 16-byte buffer loads, nu units per row, R = NT / nu rows a
// pass.  Thread t copies unit t mod nu of rows t / nu + k R; its row's (z, y) and byte
// offsets advance by fixed steps (no division or wide multiply in the loop), and four rows
// are loaded before the first is stored, so four loads are in flight per thread.
template <uint32_t NT>
__device__ __forceinline__ void staged_copy_rows(__amdgpu_buffer_rsrc_t rsrc, uint8_t* slab,
                                                 uint32_t t, uint32_t LY, uint32_t LZ,
                                                 uint32_t gx0, uint32_t nu, uint32_t rowb,
                                                 uint32_t SY, uint32_t SZ, uint32_t PY,
                                                 uint32_t pitch) {
    const uint32_t R = NT / nu;                 // uniform
    uint32_t lr, lu;
    divmod16(t, nu, __builtin_amdgcn_rcpf((float)nu), lr, lu);
    if (lr >= R) return;
    const uint32_t nrows = SY * SZ;
    const uint32_t dz = R / SY, dy = R - dz * SY;  // uniform row step in (z, y)
    uint32_t zr, yr;
    divmod16(lr, SY, __builtin_amdgcn_rcpf((float)SY), zr, yr);
    uint32_t row = __umul24(LZ + zr, PY) + LY + yr;  // padded row index of row lr
    const uint32_t drow = dz * PY + dy, wrap = PY - SY;
    const uint32_t gx = gx0 + lu * 16u;
    uint32_t dst = __umul24(lr, rowb) + lu * 16u;   // slab byte offset of row lr's unit
    const uint32_t ddst = R * rowb;
    for (uint32_t r = lr; r < nrows; r += 4u * R) {
        uint32_t off[4], at[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            off[k] = __umul24(row, pitch) + gx;
            at[k] = dst;
            dst += ddst;
            row += drow;
            yr += dy;
            if (yr >= SY) {  // next plane of the box
                yr -= SY;
                row += wrap;
            }
        }
        U32x4 v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            v[k] = r + (uint32_t)k * R < nrows ? __builtin_amdgcn_raw_buffer_load_b128(rsrc, off[k], 0, 0)
                                               : U32x4{0u, 0u, 0u, 0u};
        // all four loads issue before any store (keeps the compiler from sinking the
        // conditional rows' loads into their branches, one round trip each)
        __asm__ volatile("" ::"v"(v[0]), "v"(v[1]), "v"(v[2]), "v"(v[3]));
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (r + (uint32_t)k * R < nrows) {
                Unit16 u;
                u.w[0] = v[k][0];
                u.w[1] = v[k][1];
                u.w[2] = v[k][2];
                u.w[3] = v[k][3];
                *reinterpret_cast<Unit16*>(slab + at[k]) = u;
            }
        }
    }
}

// The next box's rows in flight while the current box is sampled: staged_prefetch_rows
// issues the loads of staged_copy_rows (one wave, NT = 64) into registers, at most
// kPrefetchRows rows per lane (npass = ceil(rows / R), uniform); staged_commit_rows stores
// them into the slab once the wave is done with the previous box.
constexpr int kPrefetchRows = 8;

__device__ __forceinline__ void staged_prefetch_rows(__amdgpu_buffer_rsrc_t rsrc, uint32_t t,
                                                     uint32_t LY, uint32_t gx0, uint32_t nu,
                                                     uint32_t SY, uint32_t PY, uint32_t pitch,
                                                     uint32_t nrows, int npass,
                                                     U32x4 (&v)[kPrefetchRows]) {
    const uint32_t R = 64u / nu;
    uint32_t lr, lu;
    divmod16(t, nu, __builtin_amdgcn_rcpf((float)nu), lr, lu);
    if (lr >= R) return;
    const uint32_t dz = R / SY, dy = R - dz * SY;
    uint32_t zr, yr;
    divmod16(lr, SY, __builtin_amdgcn_rcpf((float)SY), zr, yr);
    uint32_t row = __umul24(zr, PY) + LY + yr;
    const uint32_t drow = dz * PY + dy, wrap = PY - SY;
    const uint32_t gx = gx0 + lu * 16u;
#pragma unroll
    for (int k = 0; k < kPrefetchRows; ++k) {
        if (k < npass && lr + (uint32_t)k * R < nrows)
            v[k] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, __umul24(row, pitch) + gx, 0, 0);
        row += drow;
        yr += dy;
        if (yr >= SY) {
            yr -= SY;
            row += wrap;
        }
    }
}

__device__ __forceinline__ void staged_commit_rows(uint8_t* slab, uint32_t t, uint32_t nu,
                                                   uint32_t rowb, uint32_t nrows, int npass,
                                                   const U32x4 (&v)[kPrefetchRows]) {
    const uint32_t R = 64u / nu;
    uint32_t lr, lu;
    divmod16(t, nu, __builtin_amdgcn_rcpf((float)nu), lr, lu);
    if (lr >= R) return;
    uint32_t dst = __umul24(lr, rowb) + lu * 16u;
    const uint32_t ddst = R * rowb;
#pragma unroll
    for (int k = 0; k < kPrefetchRows; ++k) {
        if (k < npass && lr + (uint32_t)k * R < nrows) {
            Unit16 u;
            u.w[0] = v[k][0];
            u.w[1] = v[k][1];
            u.w[2] = v[k][2];
            u.w[3] = v[k][3];
            *reinterpret_cast<Unit16*>(slab + dst) = u;
        }
        dst += ddst;
    }
}

// Opacity mode (alpha_pow > 0; volume_kernel.cl:120-150, the plain kernel's mip_rays): each
// sample v is normalised by the transfer range, col = max(col, cumsum v), cumsum *= 1 - a v
// (v clamped to [0, 1] for float volumes), and the reference's inner 16-sample loop breaks
// once cumsum <= 0.01 — after the position advanced past the breaking sample; the next
// 16-sample group starts over, so a spent ray takes one sample per group.  A lane whose
// group has broken keeps its col, cumsum and position (brk, reset at each group start).
struct Opacity {
    float cumsum;
    bool brk;
    bool norm;       // max_val != 0
    float mn, den;   // min_val, max_val - min_val
    float a;         // 0.1 alpha_pow^2 (integer voxels) or alpha_pow^2 (float voxels)
    float rcp;       // 1 / den, correctly rounded
    static constexpr bool kFast = false;
};
// The same state for launches whose normalisation (v - mn) / den runs without a division
// (norm_quotient): integer voxels, max_val != 0, |den| in [2^-32, 2^32], |mn| in
// [2^-20, 2^39] (fast_norm_ok).
struct OpacityFast : Opacity {
    static constexpr bool kFast = true;
};

// x / den correctly rounded without a division: q0 = x r with r = RN(1/den), then two fma
// residual corrections q' = q + (x - den q) r.  After the first q is a faithful quotient,
// and with r the correctly rounded reciprocal the second gives RN(x / den) (Markstein's
// theorem, P. Markstein, IBM J. Res. Dev. 34(1), 1990) while nothing overflows or
// underflows: |den| in [2^-32, 2^32] and x = 0 or |x| in [2^-60, 2^40].  With integer
// voxels v in [0, 65535] and |mn| in [2^-20, 2^39], x = v - mn is 0 (v == mn) or at
// least 2^-44 in magnitude (Sterbenz: v - mn is exact when v and mn are within a factor
// of two, and at least mn / 2 otherwise), and at most 2^40.  Checked against IEEE
// division on the CPU over 5.4e9 pairs (random exponents, integer and lerp-like x,
// quotients next to rounding midpoints, divisors with significands near 1 and 2).
__device__ __forceinline__ float norm_quotient(float x, float den, float rcp) {
    const float q0 = x * rcp;
    const float q1 = fmaf(fmaf(-den, q0, x), rcp, q0);
    return fmaf(fmaf(-den, q1, x), rcp, q1);
}
__host__ inline bool fast_norm_ok(float min_val, float max_val) {
    const float den = max_val - min_val, amn = std::fabs(min_val), aden = std::fabs(den);
    return max_val != 0.0f && aden >= 0x1p-32f && aden <= 0x1p32f && amn >= 0x1p-20f &&
           amn <= 0x1p39f;
}

// Whether a ray of the opacity mode may stop before its last sample (uniform).  With every
// sample of the volume in [vmin, vmax] (brick bounds widened for float voxels), normalised
// as the samples are to [lo, hi]: when every step multiplies cumsum by a factor in [0, 1]
// -- integer voxels: lo >= 0 and fl(a hi) <= 1; float voxels (v clamped to [0, 1] in the
// transparency update): a <= 1 -- cumsum never grows and stays >= 0, so a later sample
// adds fl(cumsum' v) <= fl(cumsum hi) to col (rounding is monotone).  *hi_out = max(hi, 0).
template <typename T>
__device__ __forceinline__ bool opacity_can_stop(const Opacity& o, float vmin, float vmax,
                                                 float& hi_out) {
    constexpr bool kShort = !std::is_same<T, float>::value;
    hi_out = 0.0f;
    if (!(vmin <= vmax)) return false;  // a NaN voxel: never stop
    const float m = bound_margin<T>(vmin, vmax);
    float lo = vmin - m, hi = vmax + m;
    if (o.norm) {
        if (!(o.den > 0.0f)) return false;
        lo = (lo - o.mn) / o.den;
        hi = (hi - o.mn) / o.den;
    }
    hi_out = fmaxf(hi, 0.0f);
    if (kShort) return lo >= 0.0f && o.a * hi_out <= 1.0f;
    return o.a <= 1.0f;
}

// a ray whose later samples can no longer raise col (opacity_can_stop held)
__device__ __forceinline__ bool opacity_spent(const Opacity& o, float col, float hi) {
    return o.cumsum * hi <= col;
}

template <bool kShort, typename OP>
__device__ __forceinline__ void opacity_step(OP& o, float v, float& col, float& px,
                                             float& py, float& pz, float dx, float dy,
                                             float dz) {
    if constexpr (OP::kFast)
        v = norm_quotient(v - o.mn, o.den, o.rcp);
    else
        v = o.norm ? (v - o.mn) / o.den : v;
    const float c = fmaxf(col, o.cumsum * v);
    const float cs = kShort ? o.cumsum * (1.0f - o.a * v)
                            : o.cumsum * (1.0f - o.a * fminf(fmaxf(v, 0.0f), 1.0f));
    if (!o.brk) {
        col = c;
        o.cumsum = cs;
        px = px + dx;
        py = py + dy;
        pz = pz + dz;
        o.brk = cs <= 0.01f;
    }
}

// CH samples from the slab holding the box (lox, LY, LZ, row bytes rowb, SY rows a plane);
// slab offset of corner (c+1) relative to the box computed exactly in float (every term is
// an integer below 2^24).  FILT 1: the nearest filter, one voxel per sample (nearest's
// voxel is one of the linear sample's 8 corners, so it lies in the box staged for them).
template <typename V, int CH, bool OPAC = false, int FILT = 0, typename OP = Opacity>
__device__ __forceinline__ void staged_sample_slab(const V& vol,
                                                   const uint8_t* slab, uint32_t lox, uint32_t LY,
                                                   uint32_t LZ, uint32_t rowb, uint32_t SY,
                                                   float& px, float& py, float& pz, float dx,
                                                   float dy, float dz, float& col, OP& op) {
    using T = typename V::value_type;
    constexpr uint32_t TS = sizeof(T);
    const uint32_t sy = rowb, sz = rowb * SY;
    const float fsy = (float)sy, fsz = (float)sz;
    const float oy = 1.0f - (float)LY, oz = 1.0f - (float)LZ;
    const float ox = (1.0f - (float)lox) * (float)TS;
    if constexpr (FILT == 1) {
        const float ny = -(float)LY, nz = -(float)LZ, nx = -(float)(lox * TS);
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            float ix, iy, iz;
            vol.nearest_f(px, py, pz, &ix, &iy, &iz);
            const uint32_t q =
                (uint32_t)fmaf(iz + nz, fsz, fmaf(iy + ny, fsy, fmaf(ix, (float)TS, nx)));
            const float v = (float)*reinterpret_cast<const T*>(slab + q);
            if constexpr (OPAC) {
                opacity_step<!std::is_same<T, float>::value>(op, v, col, px, py, pz, dx, dy, dz);
            } else {
                col = fmaxf(col, v);
                px = px + dx;
                py = py + dy;
                pz = pz + dz;
            }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        float cx, cy, cz, ax, ay, az;
        vol.locate_f(px, py, pz, &cx, &cy, &cz, &ax, &ay, &az);
        const float bf = fmaf(cz + oz, fsz, fmaf(cy + oy, fsy, fmaf(cx, (float)TS, ox)));
        const uint32_t q = (uint32_t)bf;
        const uint8_t* a = slab + (q & ~3u);  // slab is 16-byte aligned, sy and sz are
        const uint32_t sh = q & 3u;            // multiples of 16
        float v000, v100, v010, v110, v001, v101, v011, v111;
        lds_pair_at(a, sh, v000, v100, T());
        lds_pair_at(a + sy, sh, v010, v110, T());
        lds_pair_at(a + sz, sh, v001, v101, T());
        lds_pair_at(a + sy + sz, sh, v011, v111, T());
        const float v = trilerp(v000, v100, v010, v110, v001, v101, v011, v111, ax, ay, az);
        if constexpr (OPAC) {
            opacity_step<!std::is_same<T, float>::value>(op, v, col, px, py, pz, dx, dy, dz);
        } else {
            col = fmaxf(col, v);
            px = px + dx;
            py = py + dy;
            pz = pz + dz;
        }
    }
}

template <typename V, int CH, bool OPAC = false, int FILT = 0, typename OP = Opacity>
__device__ __forceinline__ void staged_sample_global(const V& vol, float& px,
                                                     float& py, float& pz, float dx, float dy,
                                                     float dz, float& col, OP& op) {
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        const float v = FILT == 1 ? vol.nearest(px, py, pz) : vol.linear(px, py, pz);
        if constexpr (OPAC) {
            opacity_step<!std::is_same<typename V::value_type, float>::value>(op, v, col, px, py,
                                                                             pz, dx, dy, dz);
        } else {
            col = fmaxf(col, v);
            px = px + dx;
            py = py + dy;
            pz = pz + dz;
        }
    }
}

// Two-channel (U16Pair) forms of the two samplers: one locate and one address per sample,
// both channels interpolated with the single-channel operation order, so each plane is
// bit-identical to a single-channel render of that channel.
template <typename V, int CH>
__device__ __forceinline__ void staged_sample_slab_pair(const V& vol, const uint8_t* slab,
                                                        uint32_t lox, uint32_t LY, uint32_t LZ,
                                                        uint32_t rowb, uint32_t SY, float& px,
                                                        float& py, float& pz, float dx, float dy,
                                                        float dz, float& col, float& col2) {
    const uint32_t sy = rowb, sz = rowb * SY;
    const float fsy = (float)sy, fsz = (float)sz;
    const float oy = 1.0f - (float)LY, oz = 1.0f - (float)LZ;
    const float ox = (1.0f - (float)lox) * 4.0f;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        float cx, cy, cz, ax, ay, az;
        vol.locate_f(px, py, pz, &cx, &cy, &cz, &ax, &ay, &az);
        // a multiple of 4: voxels are 4-byte aligned in the slab
        const uint32_t q = (uint32_t)fmaf(cz + oz, fsz, fmaf(cy + oy, fsy, fmaf(cx, 4.0f, ox)));
        const uint32_t* a = reinterpret_cast<const uint32_t*>(slab + q);
        const uint32_t* b = reinterpret_cast<const uint32_t*>(slab + q + sy);
        const uint32_t* c = reinterpret_cast<const uint32_t*>(slab + q + sz);
        const uint32_t* d = reinterpret_cast<const uint32_t*>(slab + q + sy + sz);
        const uint32_t a0 = a[0], a1 = a[1], b0 = b[0], b1 = b[1];
        const uint32_t c0 = c[0], c1 = c[1], d0 = d[0], d1 = d[1];
        col = fmaxf(col, trilerp((float)(a0 & 0xFFFFu), (float)(a1 & 0xFFFFu), (float)(b0 & 0xFFFFu),
                                 (float)(b1 & 0xFFFFu), (float)(c0 & 0xFFFFu), (float)(c1 & 0xFFFFu),
                                 (float)(d0 & 0xFFFFu), (float)(d1 & 0xFFFFu), ax, ay, az));
        col2 = fmaxf(col2, trilerp((float)(a0 >> 16), (float)(a1 >> 16), (float)(b0 >> 16),
                                   (float)(b1 >> 16), (float)(c0 >> 16), (float)(c1 >> 16),
                                   (float)(d0 >> 16), (float)(d1 >> 16), ax, ay, az));
        px = px + dx;
        py = py + dy;
        pz = pz + dz;
    }
}

template <typename V, int CH>
__device__ __forceinline__ void staged_sample_global_pair(const V& vol, float& px, float& py,
                                                          float& pz, float dx, float dy, float dz,
                                                          float& col, float& col2) {
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        float cx, cy, cz, ax, ay, az;
        vol.locate_f(px, py, pz, &cx, &cy, &cz, &ax, &ay, &az);
        const uint32_t off = vol.byte_offset_f(cx, cy, cz);
        const __amdgpu_buffer_rsrc_t r = vol.rsrc();
        float l00[2], h00[2], l10[2], h10[2], l01[2], h01[2], l11[2], h11[2];
        PairBits<U16Pair>::load2(r, off, 0, l00, h00);
        PairBits<U16Pair>::load2(r, off, vol.soff_y, l10, h10);
        PairBits<U16Pair>::load2(r, off, vol.soff_z, l01, h01);
        PairBits<U16Pair>::load2(r, off, vol.soff_yz, l11, h11);
        col = fmaxf(col, trilerp(l00[0], h00[0], l10[0], h10[0], l01[0], h01[0], l11[0], h11[0],
                                 ax, ay, az));
        col2 = fmaxf(col2, trilerp(l00[1], h00[1], l10[1], h10[1], l01[1], h01[1], l11[1], h11[1],
                                   ax, ay, az));
        px = px + dx;
        py = py + dy;
        pz = pz + dz;
    }
}

// the sampler matching the voxel type (col2: the second channel of a U16Pair volume)
template <typename V, int N, bool OPAC, int FILT = 0, typename OP = Opacity>
__device__ __forceinline__ void sample_slab_any(const V& vol, const uint8_t* slab, uint32_t lox,
                                                uint32_t LY, uint32_t LZ, uint32_t rowb,
                                                uint32_t SY, float& px, float& py, float& pz,
                                                float dx, float dy, float dz, float& col,
                                                float& col2, OP& op) {
    if constexpr (std::is_same<typename V::value_type, U16Pair>::value)
        staged_sample_slab_pair<V, N>(vol, slab, lox, LY, LZ, rowb, SY, px, py, pz, dx, dy, dz, col,
                                      col2);
    else
        staged_sample_slab<V, N, OPAC, FILT>(vol, slab, lox, LY, LZ, rowb, SY, px, py, pz, dx, dy,
                                             dz, col, op);
}

template <typename V, int N, bool OPAC, int FILT = 0, typename OP = Opacity>
__device__ __forceinline__ void sample_global_any(const V& vol, float& px, float& py, float& pz,
                                                  float dx, float dy, float dz, float& col,
                                                  float& col2, OP& op) {
    if constexpr (std::is_same<typename V::value_type, U16Pair>::value)
        staged_sample_global_pair<V, N>(vol, px, py, pz, dx, dy, dz, col, col2);
    else
        staged_sample_global<V, N, OPAC, FILT>(vol, px, py, pz, dx, dy, dz, col, op);
}

// the transfer and the outputs (volume_kernel.cl:159-183); OPAC: col is already normalised
template <bool kShort, bool OPAC = false>
__device__ __forceinline__ void staged_ray_end(const RayParams& p, uint32_t o, float col,
                                               float tnear, float* __restrict__ out,
                                               float* __restrict__ alpha) {
    if constexpr (!OPAC) col = (p.max_val == 0.0f) ? col : (col - p.min_val) / (p.max_val - p.min_val);
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

// the output value alone (the second plane of a pair render; alpha is the first plane's)
__device__ __forceinline__ void staged_plane_end(const RayParams& p, uint32_t o, float col,
                                                 float* __restrict__ out) {
    col = (p.max_val == 0.0f) ? col : (col - p.min_val) / (p.max_val - p.min_val);
    col = apply_gamma(col, p.gamma);
    out[o] = p.current_part == 0 ? col : fmaxf(col, out[o]);
}

// Lanes without a ray take the first active lane's ray (act: the wave's ballot of active,
// nonzero), so every lane's LDS address is valid; the shuffles run with every lane of the
// wave enabled.
__device__ __forceinline__ void shadow_first_active(uint64_t act, bool active, float& px,
                                                    float& py, float& pz, float& dx, float& dy,
                                                    float& dz) {
    const int src = __builtin_ctzll(act);
    const float spx = __shfl(px, src), spy = __shfl(py, src), spz = __shfl(pz, src);
    const float sdx = __shfl(dx, src), sdy = __shfl(dy, src), sdz = __shfl(dz, src);
    px = active ? px : spx;
    py = active ? py : spy;
    pz = active ? pz : spz;
    dx = active ? dx : sdx;
    dy = active ? dy : sdy;
    dz = active ? dz : sdz;
}

// The sampling loop of mip_staged_rays over groups g_begin .. g_end-1 of CH samples of
// every ray of the wave (workgroup when SHARED), from positions (px, py, pz) on: stages
// the boxes, samples them and folds the samples into col (col2: a pair's second channel).
// active: this lane has a ray; live: the wave has one (lanes without a ray follow an
// active lane's positions).  HALVES: a box too large for the slab is first tried as two
// half-wave boxes.  SKIP (per-wave slabs, volumes with brick bounds): the plain max
// projection skips a box whose bricks' maximum cannot exceed any active ray's running
// maximum (its samples cannot change col; only the positions advance, by the same
// additions) and stops once every ray holds the volume's maximum; the opacity mode stops
// once no later sample can raise any ray's colour (opacity_spent).
template <typename T, bool POW2, int SLAB_BYTES, int CH, bool SHARED, bool WIDE, bool OPAC,
          bool HALVES = true, bool SKIP = false, int FILT = 0, bool FASTDIV = false>
__device__ __forceinline__ void mip_staged_march(const PaddedVolume<T, POW2, WIDE>& vol,
                                                 const RayParams& p, bool active, bool live,
                                                 uint8_t* slab, uint32_t* s_box, float& px,
                                                 float& py, float& pz, float dx, float dy,
                                                 float dz, int g_begin, int g_end, float& col,
                                                 float& col2) {
    constexpr bool kShort = !std::is_same<T, float>::value;
    constexpr bool kPair = std::is_same<T, U16Pair>::value;
    constexpr uint32_t E = 16u / sizeof(T);  // voxels per 16-byte unit
    constexpr uint32_t TS = sizeof(T);
    constexpr uint32_t NT = SHARED ? 256u : 64u;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t PY = vol.py, pitch = vol.px_bytes;
    const float mgx = 1e-5f * vol.fnx + 0.01f, mgy = 1e-5f * vol.fny + 0.01f,
                mgz = 1e-5f * vol.fnz + 0.01f;

    using PV = PaddedVolume<T, POW2, WIDE>;
    // The wave's (workgroup's when SHARED) box of samples first..last of every ray, from the
    // reduced per-lane boxes; EST: sample `first` is estimated (a box ahead of the rays).
    struct Box {
        uint32_t lox, LY, LZ, rowb, SY, SZ, nu;
        uint32_t hx;  // last padded x the box's samples read
        uint64_t bytes;
    };
    // a reduced box (packed as staged_box_n) in slab terms; true when it fits the slab
    auto decode = [&](uint32_t r0, uint32_t r1, uint32_t r2, Box& bx) -> bool {
        const uint32_t LX = r0 & 0xFFFFu;
        bx.LY = r0 >> 16;
        bx.LZ = r1 & 0xFFFFu;
        const uint32_t HX = 0xFFFFu - (r1 >> 16), HY = 0xFFFFu - (r2 & 0xFFFFu),
                       HZ = 0xFFFFu - (r2 >> 16);
        bx.lox = LX & ~(kBoxAlignMip<T, FILT> - 1u);
        bx.hx = HX;
        const uint32_t SX = (HX + 1u - bx.lox + E - 1u) & ~(E - 1u);  // voxels, multiple of E
        bx.SY = HY - bx.LY + 1u;
        bx.SZ = HZ - bx.LZ + 1u;
        bx.nu = SX / E;                                  // 16-byte units per row
        // slab row stride: SX voxels, plus 16 bytes when row_pad asks for an odd number of
        // 16-byte units (rows a few apart then fall on different LDS banks)
        bx.rowb = (bx.nu + (p.row_pad & ~bx.nu & 1u)) * 16u;
        bx.bytes = bx.nu <= NT ? (uint64_t)bx.rowb * bx.SY * bx.SZ : ~0ull;
        return bx.bytes <= (uint64_t)SLAB_BYTES;
    };
    auto box_of = [&](auto est, float first, float last, Box& bx) -> bool {
        uint32_t b0, b1, b2;
        staged_box_n<PV, decltype(est)::value>(vol, active, first, last, px, py, pz, dx, dy, dz,
                                               mgx, mgy, mgz, b0, b1, b2);
        uint32_t r0 = b0, r1 = b1, r2 = b2;
        wave_min3_u16x2(r0, r1, r2);
        if constexpr (SHARED) {
            const uint32_t w = threadIdx.x >> 6;
            if (lane == 0) {
                s_box[3 * w] = r0;
                s_box[3 * w + 1] = r1;
                s_box[3 * w + 2] = r2;
            }
            __syncthreads();
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                r0 = pk_min_u16(r0, s_box[3 * k]);
                r1 = pk_min_u16(r1, s_box[3 * k + 1]);
                r2 = pk_min_u16(r2, s_box[3 * k + 2]);
            }
        }
        return decode(r0, r1, r2, bx);
    };
    // the slab's contents visible to the wave (workgroup when SHARED)
    auto visible = [&]() {
        if constexpr (SHARED) {
            __syncthreads();
        } else {
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
    };
    // copies a box that fits into the slab (rows addressed from plane LZ on: 32-bit offsets
    // whatever the volume's size) and makes it visible
    auto copy_box = [&](const Box& bx) {
        staged_copy_rows<NT>(vol.rsrc_plane(bx.LZ), slab, SHARED ? threadIdx.x : lane, bx.LY, 0u,
                             bx.lox * TS, bx.nu, bx.rowb, bx.SY, bx.SZ, PY, pitch);
        visible();
    };
    // Stages the box of the next n samples of every ray into the slab.  Returns false, with
    // nothing staged, when the box does not fit.
    auto stage = [&](int n, Box& bx) -> bool {
        const bool fits = box_of(std::false_type(), 0.0f, (float)(n - 1), bx);
        if (fits) {
            copy_box(bx);
        } else {
            visible();
        }
        return fits;
    };
    // this lane's share of the brick bounds of box bx (false: too many bricks to ask)
    auto box_keys = [&](const Box& bx, BrickKey& k) -> bool {
        return brick_load(vol, bx.lox, bx.LY, bx.LZ, bx.hx, bx.LY + bx.SY - 1u, bx.LZ + bx.SZ - 1u,
                          lane, k);
    };
    // SKIP, plain max projection: 0 = sample the box, 1 = skip it (its bricks' maximum is at
    // most every active ray's running maximum), 2 = stop (every active ray holds the
    // volume's maximum).  Uniform.
    auto skip_test = [&](BrickKey k) -> int {
        uint32_t c = active ? order_key(col) : 0xFFFFFFFFu;  // col >= 0, never NaN
        wave_min3_u32(k.kmin, k.nkmax, c);
        const float bmin = key_value<T>(k.kmin), bmax = key_value<T>(~k.nkmax);
        const float cmin = key_value<float>(c);
        if (vol.vmax + bound_margin<T>(vol.vmin, vol.vmax) <= cmin) return 2;
        return bmax + bound_margin<T>(bmin, bmax) <= cmin ? 1 : 0;
    };
    // the positions of n samples, as the sampling loops advance them
    auto advance = [&](int n) {
#pragma unroll 1
        for (int j = 0; j < n; ++j) {
            px = px + dx;
            py = py + dy;
            pz = pz + dz;
        }
    };
    // the wave's reads of the slab complete before the next box overwrites it
    auto release = [&]() {
        if constexpr (!SHARED) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        }
    };
    // stages the boxes of the next CH samples of each half of the wave side by side in the
    // slab (ha at its start, hb after it) when both fit; false, with nothing staged, if not
    auto stage_halves = [&](Box& ha, Box& hb) -> bool {
        uint32_t a0, a1, a2, b0, b1, b2;
        staged_box_n<PV, false>(vol, active, 0.0f, (float)(CH - 1), px, py, pz, dx, dy, dz,
                                mgx, mgy, mgz, a0, a1, a2);
        wave_min3_u16x2_halves(a0, a1, a2, b0, b1, b2);
        // a half without an active lane has no box (all ones): then the other half's box is
        // the wave's, which did not fit
        if (!decode(a0, a1, a2, ha) || !decode(b0, b1, b2, hb) ||
            ha.bytes + hb.bytes > (uint64_t)SLAB_BYTES)
            return false;
        staged_copy_rows<NT>(vol.rsrc_plane(ha.LZ), slab, lane, ha.LY, 0u, ha.lox * TS, ha.nu,
                             ha.rowb, ha.SY, ha.SZ, PY, pitch);
        staged_copy_rows<NT>(vol.rsrc_plane(hb.LZ), slab + ha.bytes, lane, hb.LY, 0u,
                             hb.lox * TS, hb.nu, hb.rowb, hb.SY, hb.SZ, PY, pitch);
        visible();
        return true;
    };

    static_assert(kLoopUnroll % CH == 0, "a chunk is a whole number of boxes");
    static_assert(!(OPAC && kPair), "pair renders are plain max projections");
    static_assert(FILT == 0 || (!kPair && !WIDE), "nearest: one-channel volumes below 4 GiB");
    const int G = CH == 16 ? (int)p.box16 : 1;  // groups per box, uniform
    static_assert(!FASTDIV || (OPAC && kShort), "division-free normalisation: integer voxels");
    std::conditional_t<FASTDIV, OpacityFast, Opacity> op;
    op.cumsum = 1.0f;
    op.brk = false;
    op.norm = p.max_val != 0.0f;
    op.mn = p.min_val;
    op.den = p.max_val - p.min_val;
    op.rcp = 1.0f / op.den;
    op.a = kShort ? 0.1f * p.alpha_pow * p.alpha_pow : p.alpha_pow * p.alpha_pow;
    int done = 0;  // samples of the current ray position sequence taken so far (uniform)
    // before n samples: a new reference 16-sample group starts the break state afresh
    auto group = [&](int n) {
        if constexpr (OPAC) {
            if ((done & 15) == 0) op.brk = false;
        }
        done += n;
    };
    static_assert(!SKIP || !SHARED, "skipping runs with per-wave slabs");
    // SKIP, opacity mode: whether the volume's bounds let a ray stop early, and the largest
    // normalised sample (uniform)
    float vmax_n = 0.0f;
    bool can_stop = false;
    if constexpr (SKIP && OPAC) {
        can_stop = opacity_can_stop<T>(op, vol.vmin, vol.vmax, vmax_n);
    }
    // per-wave slabs: the next box's rows are loaded into registers while this box is
    // sampled (when the next box fits and needs at most kPrefetchRows rows per lane); SKIP:
    // with its brick bounds
    Box nb{};
    bool pf = false, pf_keys = false;
    int pf_n = 0;
    U32x4 pv[kPrefetchRows];
    BrickKey nk{};
    for (int i = g_begin; i < g_end; i += G) {
        const int ng = min(G, g_end - i);
        if constexpr (SKIP && OPAC) {
            if (can_stop && __ballot(active && !opacity_spent(op, col, vmax_n)) == 0) break;
        }
        Box b;
        bool fits;
        if constexpr (SKIP && !OPAC) {
            // the box's bricks: loaded with the prefetched rows, or now
            BrickKey k = nk;
            bool have = pf && pf_keys;
            if (!pf) {
                fits = box_of(std::false_type(), 0.0f, (float)(ng * CH - 1), b);
                have = box_keys(b, k);
            }
            const int t = have ? skip_test(k) : 0;
            if (t == 2) break;
            if (t == 1) {
                pf = false;  // prefetched rows (if any) are dropped
                advance(ng * CH);
                continue;
            }
            if (pf) {
                staged_commit_rows(slab, lane, nb.nu, nb.rowb, nb.SY * nb.SZ, pf_n, pv);
                visible();
                b = nb;
                fits = true;
            } else if (fits) {
                copy_box(b);
            } else {
                visible();
            }
        } else {
            if (!SHARED && pf) {
                staged_commit_rows(slab, lane, nb.nu, nb.rowb, nb.SY * nb.SZ, pf_n, pv);
                visible();
                b = nb;
                fits = true;
            } else {
                fits = stage(ng * CH, b);
            }
        }
        pf = false;
        if constexpr (!SHARED && !OPAC) {
            const int i2 = i + G;
            if (i2 < g_end) {
                const int ng2 = min(G, g_end - i2);
                if (box_of(std::true_type(), (float)(ng * CH), (float)((ng + ng2) * CH - 1), nb)) {
                    const uint32_t R = 64u / nb.nu;
                    pf_n = (int)((nb.SY * nb.SZ + R - 1u) / R);
                    if (pf_n <= kPrefetchRows) {
                        staged_prefetch_rows(vol.rsrc_plane(nb.LZ), lane, nb.LY, nb.lox * TS, nb.nu,
                                             nb.SY, PY, pitch, nb.SY * nb.SZ, pf_n, pv);
                        pf = true;
                        if constexpr (SKIP) pf_keys = box_keys(nb, nk);
                    }
                }
            }
        }
        if (fits) {
            if (live) {
#pragma unroll 1
                for (int k = 0; k < ng; ++k) {
                    group(CH);
                    sample_slab_any<PV, CH, OPAC, FILT>(vol, slab, b.lox, b.LY, b.LZ, b.rowb, b.SY, px, py,
                                                  pz, dx, dy, dz, col, col2, op);
                }
            }
            release();
            continue;
        }
        if constexpr (FILT == 1) {
            // the nearest filter reads one voxel per sample: a box that does not fit is
            // cheaper gathered from global memory than split and staged piecewise
            if (live) {
#pragma unroll 1
                for (int k = 0; k < ng; ++k) {
                    group(CH);
                    sample_global_any<PV, CH, OPAC, FILT>(vol, px, py, pz, dx, dy, dz, col, col2,
                                                          op);
                }
            }
            release();
            continue;
        }
        // a box that does not fit the slab: the ngf groups of box bf staged per half-wave,
        // split into 8- and 4-sample boxes, or gathered from global memory
        auto fallback = [&](const Box& bf, int ngf) {
            if constexpr (HALVES && CH == 16 && !SHARED && !kPair) {
                // a box too large for the slab because the wave's two halves (rows 0-3 and 4-7
                // of its 8x8 rays) run apart, e.g. along the volume's silhouette where the rays'
                // lengths change from row to row: one box per half, side by side in the slab.
                // Lanes without a ray follow the wave's first active lane, in the first half.
                Box ha, hb;
                if (ngf == 1 && stage_halves(ha, hb)) {
                    if (live) {
                        const bool second = active && lane >= 32u;
                        group(CH);
                        sample_slab_any<PV, CH, OPAC, FILT>(
                            vol, second ? slab + ha.bytes : slab,
                            second ? hb.lox : ha.lox, second ? hb.LY : ha.LY, second ? hb.LZ : ha.LZ,
                            second ? hb.rowb : ha.rowb, second ? hb.SY : ha.SY, px, py, pz, dx, dy, dz,
                            col, col2, op);
                    }
                    release();
                    return;
                }
            }
            if constexpr (CH == 16 && !SHARED) {
                if (bf.bytes <= (uint64_t)(4 * ngf * SLAB_BYTES)) {
                    // a box up to 4x the slab: halves of 8 samples with a box each, then quarters
                    // of 4; a quarter that still does not fit samples global memory.  (Rays far
                    // apart, e.g. a sparse frame, give boxes that no split brings under the
                    // slab: those go to global memory directly.)
#pragma unroll 1
                    for (int h = 0; h < 2 * ngf; ++h) {
                        Box hb;
                        if (stage(8, hb)) {
                            group(8);
                            sample_slab_any<PV, 8, OPAC, FILT>(vol, slab, hb.lox, hb.LY, hb.LZ, hb.rowb, hb.SY,
                                                         px, py, pz, dx, dy, dz, col, col2, op);
                            release();
                            continue;
                        }
#pragma unroll 1
                        for (int q = 0; q < 2; ++q) {
                            Box qb;
                            const bool qfits = stage(4, qb);
                            group(4);
                            if (qfits)
                                sample_slab_any<PV, 4, OPAC, FILT>(vol, slab, qb.lox, qb.LY, qb.LZ, qb.rowb,
                                                             qb.SY, px, py, pz, dx, dy, dz, col, col2, op);
                            else
                                sample_global_any<PV, 4, OPAC, FILT>(vol, px, py, pz, dx, dy, dz, col, col2, op);
                            release();
                        }
                    }
                } else {
#pragma unroll 1
                    for (int k = 0; k < ngf; ++k) {
                        group(CH);
                        sample_global_any<PV, CH, OPAC, FILT>(vol, px, py, pz, dx, dy, dz, col, col2, op);
                    }
                    release();
                }
            } else {
                if (live) {
#pragma unroll 1
                    for (int k = 0; k < ngf; ++k) {
                        group(CH);
                        sample_global_any<PV, CH, OPAC, FILT>(vol, px, py, pz, dx, dy, dz, col, col2, op);
                    }
                }
                release();
            }
        };
        if constexpr (CH == 16 && !SHARED && !kPair) {
            if (ng > 1) {
                // a box of several 16-sample groups that does not fit: each group on its own,
                // as a box of one group (oblique rays: 32-sample boxes often do not fit where
                // 16-sample boxes do)
#pragma unroll 1
                for (int k = 0; k < ng; ++k) {
                    Box bk;
                    if (stage(CH, bk)) {
                        if (live) {
                            group(CH);
                            sample_slab_any<PV, CH, OPAC, FILT>(vol, slab, bk.lox, bk.LY, bk.LZ,
                                                                bk.rowb, bk.SY, px, py, pz, dx, dy,
                                                                dz, col, col2, op);
                        }
                        release();
                    } else {
                        fallback(bk, 1);
                    }
                }
                continue;
            }
        }
        fallback(b, ng);
    }
}

// A wave's (workgroup's when SHARED) rays of a max projection, from set-up to outputs.
// CH: samples per unrolled group (16, 8 or 4); a box holds p.box16 groups of 16 samples
// when CH is 16 (one group otherwise).  The samples, their order and their positions are
// those of the reference's 16-sample inner loop whatever the grouping is (a max does not
// depend on how the sequence is cut); a smaller CH keeps a box small when the ray step is
// several voxels long, a larger box spreads the box reduction and copy over more samples
// when it is short.  With per-wave slabs and CH 16, a box larger than the slab is split
// into halves of 8 samples and those into quarters of 4, each staged on its own (when it
// is at most 4x the slab); only what still does not fit is sampled from global memory.
// SHARED = false: each wave stages the box of its own 64 rays into its own slab
// (SLAB_BYTES, wave-level ordering only).  SHARED = true: the workgroup's four waves
// stage the union box of its 256 rays into one slab of SLAB_BYTES with all 256 threads,
// two workgroup barriers per box; s_box holds the four waves' partial boxes.
// OPAC: the opacity mode (alpha_pow > 0, Opacity); boxes are not loaded ahead (a spent
// lane's position does not advance with the group).
template <typename T, bool POW2, int SLAB_BYTES, int CH, bool SHARED, bool WIDE = false,
          bool OPAC = false, bool HALVES = true, bool SKIP = false, int FILT = 0,
          bool FASTDIV = false>
__device__ __forceinline__ void mip_staged_rays(const PaddedVolume<T, POW2, WIDE>& vol, const RayParams& p,
                                                uint32_t x, uint32_t y, bool inside, uint8_t* slab,
                                                uint32_t* s_box, float* __restrict__ out,
                                                float* __restrict__ alpha) {
    constexpr bool kShort = !std::is_same<T, float>::value;
    constexpr bool kPair = std::is_same<T, U16Pair>::value;  // out: two planes of W x H
    if (gate_closed(p)) return;  // uniform over the launch
    float px = 0.5f, py = 0.5f, pz = 0.5f, dx = 0.0f, dy = 0.0f, dz = 0.0f, tnear = 0.0f;
    const bool active = staged_ray_begin<kShort>(p, x, y, inside, out, alpha, px, py, pz, dx, dy,
                                                 dz, tnear);
    if constexpr (kPair) {
        if (inside && !active) out[x + p.W * y + p.W * p.H] = 0.0f;  // the miss, second plane
    }
    const uint64_t act = __ballot(active);
    if constexpr (SHARED) {
        if (__syncthreads_or(active ? 1 : 0) == 0) return;  // uniform over the workgroup
    } else {
        if (act == 0) return;
    }
    const bool live = act != 0;  // uniform over the wave
    if (live) shadow_first_active(act, active, px, py, pz, dx, dy, dz);
    float col = 0.0f, col2 = 0.0f;
    mip_staged_march<T, POW2, SLAB_BYTES, CH, SHARED, WIDE, OPAC, HALVES, SKIP, FILT, FASTDIV>(
        vol, p, active, live, slab, s_box, px, py, pz, dx, dy, dz, 0,
        (p.nouter + 1) * (kLoopUnroll / CH), col, col2);
    if (active) {
        staged_ray_end<kShort, OPAC>(p, x + p.W * y, col, tnear, out, alpha);
        if constexpr (kPair) staged_plane_end(p, x + p.W * y + p.W * p.H, col2, out);
    }
}

// one frame: one workgroup per 16x16 tile (tile map as k_max_project)
// SLAB_BYTES: per wave (SHARED = false) or per workgroup (SHARED = true)
template <typename T, bool POW2, int SLAB_BYTES, int CH, bool SHARED, bool WIDE = false,
          bool OPAC = false, bool SKIP = false>
__global__ __launch_bounds__(256) void k_max_project_staged(PaddedVolume<T, POW2, WIDE> vol, RayParams p,
                                                            float* __restrict__ out,
                                                            float* __restrict__ alpha) {
    constexpr int kLds = SHARED ? SLAB_BYTES : 4 * SLAB_BYTES;
    __shared__ Unit16 lds[kLds / 16 + 1];  // +16 B: lds_pair reads past a pair
    __shared__ uint32_t s_box[12];
    uint8_t* slab = reinterpret_cast<uint8_t*>(lds) + (SHARED ? 0u : (threadIdx.x >> 6) * SLAB_BYTES);
    uint32_t x = 0, y = 0;
    const bool inside = tile_pixel(p, &x, &y);
    mip_staged_rays<T, POW2, SLAB_BYTES, CH, SHARED, WIDE, OPAC, true, SKIP>(vol, p, x, y, inside,
                                                                             slab, s_box, out, alpha);
}

// Estimated cost of each wave of a frame of the per-wave staged kernels, for the
// cost-ordered tile map.  A wave's rays take their 16-sample boxes from a slab when the
// box fits (the rays run close together) and otherwise split the box or gather from
// global memory, at ~3.4x the time per sample (measured).  The estimate: 0 for a wave
// without a ray, else 1 + the number of kCostProbes 16-sample groups, spread along the
// rays, whose box does not fit slab_bytes (0 .. kCostProbes + 1).  One workgroup per
// 16x16 tile, row-major (grid p.ntiles), 256 threads; no outputs but cost[4 tile + wave]
// (wave: the 8x8 quarter of tile_thread_pixel).
constexpr int kCostProbes = 6;

template <typename T, bool POW2, bool WIDE = false>
__global__ __launch_bounds__(256) void k_tile_cost(PaddedVolume<T, POW2, WIDE> vol, RayParams p,
                                                   uint32_t slab_bytes,
                                                   uint32_t* __restrict__ cost) {
    using PV = PaddedVolume<T, POW2, WIDE>;
    constexpr uint32_t E = 16u / sizeof(T);
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6, t = blockIdx.x;
    uint32_t x, y;
    tile_thread_pixel(t % p.ntx, t / p.ntx, &x, &y);
    float px = 0.5f, py = 0.5f, pz = 0.5f, dx = 0.0f, dy = 0.0f, dz = 0.0f, tnear = 0.0f;
    const bool active = x < p.W && y < p.H &&
                        staged_ray_pos(p, x, y, px, py, pz, dx, dy, dz, tnear);
    const uint64_t act = __ballot(active);
    uint32_t c = 0;
    if (act != 0) {
        c = 1;
        // lanes without a ray take the first active lane's ray, as in mip_staged_rays
        const int src = __builtin_ctzll(act);
        const float spx = __shfl(px, src), spy = __shfl(py, src), spz = __shfl(pz, src);
        const float sdx = __shfl(dx, src), sdy = __shfl(dy, src), sdz = __shfl(dz, src);
        px = active ? px : spx;
        py = active ? py : spy;
        pz = active ? pz : spz;
        dx = active ? dx : sdx;
        dy = active ? dy : sdy;
        dz = active ? dz : sdz;
        const float mgx = 1e-5f * vol.fnx + 0.01f, mgy = 1e-5f * vol.fny + 0.01f,
                    mgz = 1e-5f * vol.fnz + 0.01f;
        const int last_group = p.nouter;  // groups 0 .. nouter of 16 samples
#pragma unroll 1
        for (int m = 0; m < kCostProbes; ++m) {
            const float first = (float)(kLoopUnroll * ((m * last_group) / (kCostProbes - 1)));
            uint32_t b0, b1, b2;
            staged_box_n<PV, true>(vol, active, first, first + (float)(kLoopUnroll - 1), px, py,
                                   pz, dx, dy, dz, mgx, mgy, mgz, b0, b1, b2);
            wave_min3_u16x2(b0, b1, b2);
            // the box's slab bytes as mip_staged_rays' box_of (rows of whole 16-byte units)
            const uint32_t LX = b0 & 0xFFFFu, LY = b0 >> 16, LZ = b1 & 0xFFFFu;
            const uint32_t HX = 0xFFFFu - (b1 >> 16), HY = 0xFFFFu - (b2 & 0xFFFFu),
                           HZ = 0xFFFFu - (b2 >> 16);
            const uint32_t lox = LX & ~(kBoxAlign<T> - 1u);
            const uint32_t nu = ((HX + 1u - lox + E - 1u) & ~(E - 1u)) / E;
            const uint64_t bytes = (uint64_t)nu * 16u * (HY - LY + 1u) * (HZ - LZ + 1u);
            c += nu <= 64u && bytes <= (uint64_t)slab_bytes ? 0u : 1u;
        }
    }
    if (lane == 0) cost[4u * t + wave] = c;
}

// Workgroup -> work item order for kTileMapCostXcd: XCD x takes the items of tile rows
// x, x+8, x+16, ... in descending cost, row-major among equal costs (neighbouring items
// run close in time and share the XCD's L2).  Items: 16x16 tiles (per_wave = 0; a tile
// costs its slowest wave, as a workgroup holds its LDS until its last wave ends) or the
// tiles' 8x8 quarters, item 4 tile + quarter (per_wave = 1, one wave per workgroup).
// order[8 k + x] = XCD x's k-th item; slots left over (nty not a multiple of 8) get the
// item past the last (below the frame).  One workgroup of 8 waves, wave x orders XCD x's
// items: a count per cost value, then a stable scatter by ballots over 64 items at a time.
constexpr uint32_t kCostValues = kCostProbes + 2;

static __global__ __launch_bounds__(512) void k_tile_order(const uint32_t* __restrict__ cost,
                                                    RayParams p, uint32_t per_wave,
                                                    uint32_t* __restrict__ order) {
    const uint32_t lane = threadIdx.x & 63u, x = threadIdx.x >> 6;
    const uint32_t Q = per_wave ? 4u : 1u;                // items per tile
    const uint32_t row = p.ntx * Q;                       // items per tile row
    const uint32_t rows = x < p.nty ? (p.nty - x + 7u) / 8u : 0u;  // tile rows of XCD x
    const uint32_t cnt = rows * row;
    const uint32_t slots = (p.nty + 7u) / 8u * row;       // per XCD
    const uint64_t below = (1ull << lane) - 1ull;
    auto item = [&](uint32_t j) { return (x + 8u * (j / row)) * row + j % row; };
    auto value = [&](uint32_t e) -> uint32_t {
        const uint32_t c = per_wave ? cost[e]
                                    : max(max(cost[4 * e], cost[4 * e + 1]),
                                          max(cost[4 * e + 2], cost[4 * e + 3]));
        return min(c, kCostValues - 1u);
    };
    uint32_t start[kCostValues];  // wave-uniform
#pragma unroll
    for (uint32_t v = 0; v < kCostValues; ++v) start[v] = 0;
    for (uint32_t j0 = 0; j0 < cnt; j0 += 64u) {
        const uint32_t j = j0 + lane;
        const uint32_t v = j < cnt ? value(item(j)) : kCostValues;
#pragma unroll
        for (uint32_t k = 0; k < kCostValues; ++k) start[k] += (uint32_t)__popcll(__ballot(v == k));
    }
    // descending: the costliest value first
    uint32_t run = 0;
#pragma unroll
    for (int k = (int)kCostValues - 1; k >= 0; --k) {
        const uint32_t h = start[k];
        start[k] = run;
        run += h;
    }
    for (uint32_t j0 = 0; j0 < cnt; j0 += 64u) {
        const uint32_t j = j0 + lane;
        const uint32_t e = j < cnt ? item(j) : 0u;
        const uint32_t v = j < cnt ? value(e) : kCostValues;
        uint32_t pos = 0;
#pragma unroll
        for (uint32_t k = 0; k < kCostValues; ++k) {
            const uint64_t m = __ballot(v == k);
            if (v == k) pos = start[k] + (uint32_t)__popcll(m & below);
            start[k] += (uint32_t)__popcll(m);
        }
        if (j < cnt) order[8u * pos + x] = e;
    }
    for (uint32_t k = cnt + lane; k < slots; k += 64u) order[8u * k + x] = p.ntiles * Q;
}

// Pixel of this thread in a one-wave workgroup: an 8x8 quarter of a 16x16 tile.  Tile
// rows are dealt to the XCDs as kTileMapInterleaved, the four quarters of a tile to
// consecutive workgroups of its XCD; kTileMapCostXcd: item p.order[b] = 4 tile + quarter.
__device__ __forceinline__ bool wave_tile_pixel(const RayParams& p, uint32_t* x, uint32_t* y) {
    const uint32_t b = blockIdx.x, lane = threadIdx.x & 63u;
    uint32_t tx, ty, q;
    if (p.tile_map == kTileMapCostXcd) {
        const uint32_t e = p.order[b];  // >= 4 ntiles: a surplus workgroup (ty >= nty)
        q = e & 3u;
        tx = (e >> 2) % p.ntx;
        ty = (e >> 2) / p.ntx;
    } else {
        const uint32_t xcd = b & 7u, k = b >> 3;
        q = k & 3u;
        ty = xcd + 8u * ((k >> 2) / p.ntx);
        tx = (k >> 2) % p.ntx;
    }
    *x = tx * 16u + (q & 1u) * 8u + (lane & 7u);
    *y = ty * 16u + (q >> 1) * 8u + (lane >> 3);
    return ty < p.nty && *x < p.W && *y < p.H;
}

// one frame with one wave per workgroup (64 threads, one slab of SLAB_BYTES): a CU takes a
// new wave's work as soon as any of its waves ends, instead of when a workgroup's slowest
// of four waves does.  Grid: 4 x tile_grid (interleaved or cost-ordered maps).
template <typename T, bool POW2, int SLAB_BYTES, int CH, bool WIDE = false, bool OPAC = false,
          bool SKIP = false, int FILT = 0, bool FASTDIV = false>
__global__ __launch_bounds__(64) void k_max_project_wave(PaddedVolume<T, POW2, WIDE> vol,
                                                         RayParams p, float* __restrict__ out,
                                                         float* __restrict__ alpha) {
    __shared__ Unit16 lds[SLAB_BYTES / 16 + 1];  // +16 B: lds_pair reads past a pair
    uint32_t x = 0, y = 0;
    const bool inside = wave_tile_pixel(p, &x, &y);
    mip_staged_rays<T, POW2, SLAB_BYTES, CH, false, WIDE, OPAC, true, SKIP, FILT, FASTDIV>(
        vol, p, x, y, inside, reinterpret_cast<uint8_t*>(lds), nullptr, out, alpha);
}

// Iso-surface march through the per-wave LDS slabs (iso_kernel.cl:83-131, the plain
// kernel's k_iso_surface): sample k of a ray at pos0 + the running sum of k deltas, k = 0
// .. max_steps-1; sample 0 sets isGreater, the first later sample on the other side of
// the iso value is the crossing.  A wave stages the box of its still-marching rays' next
// CH samples; rays that crossed (and lanes without a ray) follow a marching lane's
// positions, so every LDS read stays in the box, and their samples are ignored.  The
// wave stops marching when no lane marches; the crossing's refinement, gradient and
// shading then run per lane from global memory (iso_finish), as in k_iso_surface.
// FILT 1: the nearest filter (one voxel per sample, inside the box staged for the linear
// sample's corners).
template <typename V, int CH, bool SLAB, int FILT = 0>
__device__ __forceinline__ void iso_march_chunk(const V& vol, const uint8_t* slab, uint32_t lox,
                                                uint32_t LY, uint32_t LZ, uint32_t rowb,
                                                uint32_t SY, float& px, float& py, float& pz,
                                                float dx, float dy, float dz, int k0, int nsteps,
                                                float iso, bool march, int& gt, bool& found,
                                                int& ihit) {
    using T = typename V::value_type;
    constexpr uint32_t TS = sizeof(T);
    const uint32_t sy = rowb, sz = rowb * SY;
    const float fsy = (float)sy, fsz = (float)sz;
    const float oy = 1.0f - (float)LY, oz = 1.0f - (float)LZ;
    const float ox = (1.0f - (float)lox) * (float)TS;
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        float v;
        if constexpr (SLAB && FILT == 1) {
            float ix, iy, iz;
            vol.nearest_f(px, py, pz, &ix, &iy, &iz);
            const uint32_t q = (uint32_t)fmaf(iz - (float)LZ, fsz,
                                              fmaf(iy - (float)LY, fsy, (ix - (float)lox) * (float)TS));
            v = (float)*reinterpret_cast<const T*>(slab + q);
        } else if constexpr (FILT == 1) {
            v = vol.nearest(px, py, pz);
        } else if constexpr (SLAB) {
            float cx, cy, cz, ax, ay, az;
            vol.locate_f(px, py, pz, &cx, &cy, &cz, &ax, &ay, &az);
            const float bf = fmaf(cz + oz, fsz, fmaf(cy + oy, fsy, fmaf(cx, (float)TS, ox)));
            const uint32_t q = (uint32_t)bf;
            const uint8_t* a = slab + (q & ~3u);
            const uint32_t sh = q & 3u;
            float v000, v100, v010, v110, v001, v101, v011, v111;
            lds_pair_at(a, sh, v000, v100, T());
            lds_pair_at(a + sy, sh, v010, v110, T());
            lds_pair_at(a + sz, sh, v001, v101, T());
            lds_pair_at(a + sy + sz, sh, v011, v111, T());
            v = trilerp(v000, v100, v010, v110, v001, v101, v011, v111, ax, ay, az);
        } else {
            v = vol.linear(px, py, pz);
        }
        const int k = k0 + j;
        const int g = (int)(v > iso);
        if (k == 0) {
            gt = g;
        } else if (march && !found && k < nsteps && g != gt) {
            found = true;
            ihit = k;
        }
        px = px + dx;
        py = py + dy;
        pz = pz + dz;
    }
}

// SKIP (a volume with brick bounds): a chunk after the first is skipped -- only the
// positions advance, by the same additions -- when its box's bricks show that no marching
// ray can change side there: bricks' maximum <= iso for rays that start below or at the
// iso value, bricks' minimum > iso for rays that start above it.
// WAVE: one wave per workgroup (grid 4 x the interleaved tile grid, as k_max_project_wave),
// else four waves, one 16x16 tile, per workgroup.  COARSE (SKIP, CH 16): the coarse skip
// tests of p.skip16 chunks (a kernel of its own: compiled into the kernel and switched off,
// they made 200-step frames 7% slower).
template <typename T, bool POW2, int SLAB_BYTES, int CH, bool WIDE = false, bool SKIP = false,
          int FILT = 0, bool WAVE = false, bool COARSE = false>
__global__ __launch_bounds__(WAVE ? 64 : 256) void k_iso_staged(PaddedVolume<T, POW2, WIDE> vol,
                                                                RayParams p,
                                                                float* __restrict__ out,
                                                                float* __restrict__ alpha,
                                                                float* __restrict__ depth,
                                                                float* __restrict__ normals) {
    using PV = PaddedVolume<T, POW2, WIDE>;
    static_assert(FILT == 0 || !WIDE, "nearest: volumes below 4 GiB");
    constexpr uint32_t E = 16u / sizeof(T);
    constexpr uint32_t TS = sizeof(T);
    constexpr uint32_t NW = WAVE ? 1u : 4u;
    __shared__ Unit16 lds[NW * SLAB_BYTES / 16 + 1];  // +16 B: lds_pair reads past a pair
    const uint32_t lane = threadIdx.x & 63u;
    uint8_t* slab = reinterpret_cast<uint8_t*>(lds) + (WAVE ? 0u : (threadIdx.x >> 6) * SLAB_BYTES);
    uint32_t x = 0, y = 0;
    const bool inside = WAVE ? wave_tile_pixel(p, &x, &y) : tile_pixel(p, &x, &y);
    const uint32_t o = x + p.W * y;
    F4 orig = mk4(0, 0, 0, 0), direc = mk4(0, 0, 0, 0), pos0 = mk4(0.5f, 0.5f, 0.5f, 0.0f),
       delta = mk4(0, 0, 0, 0);
    float tnear = 0.0f, tfar = 0.0f, dt = 0.0f;
    bool active = false;
    if (inside) {
        ray_setup(p.invP, p.invM, x, y, p.W, p.H, &orig, &direc);
        active = intersect_box(orig, direc, p.box, &tnear, &tfar);
        if (active) {
            if (tnear < 0.0f) tnear = 0.0f;
            dt = 1.0f * (tfar - tnear) / ((float)p.max_steps - 1.0f);
            delta = scl4(0.5f * dt, direc);
            pos0 = scl4(0.5f, add4(mk4(1.0f + orig.x, 1.0f + orig.y, 1.0f + orig.z, 1.0f + orig.w),
                                   scl4(tnear, direc)));
        }
    }
    float px = pos0.x, py = pos0.y, pz = pos0.z;
    const float dx = delta.x, dy = delta.y, dz = delta.z;
    const float mgx = 1e-5f * vol.fnx + 0.01f, mgy = 1e-5f * vol.fny + 0.01f,
                mgz = 1e-5f * vol.fnz + 0.01f;
    const uint32_t PY = vol.py, pitch = vol.px_bytes;
    int gt = 0, ihit = 0;
    bool found = false;
    const int nsteps = p.max_steps;
    // SKIP: after a skipped chunk the next p.skip16 chunks are first tested as one box (one
    // box reduction and one brick query for all of them); after a failed test the chunks
    // it covered are tested one by one, and a box over more bricks than one query takes
    // ends the coarse tests of the wave
    int coarse_from = 0;
    bool last_skipped = false;
    for (int k0 = 0; k0 < nsteps; k0 += CH) {
        const bool march = active && !found;
        const uint64_t live = __ballot(march);
        if (live == 0) break;  // uniform
        // lanes not marching take a marching lane's position and step for this chunk
        const int src = __builtin_ctzll(live);
        float qx = __shfl(px, src), qy = __shfl(py, src), qz = __shfl(pz, src);
        const float ex = __shfl(dx, src), ey = __shfl(dy, src), ez = __shfl(dz, src);
        float sx = march ? px : qx, sy_ = march ? py : qy, sz_ = march ? pz : qz;
        const float tx = march ? dx : ex, ty = march ? dy : ey, tz = march ? dz : ez;
        if constexpr (SKIP && CH == 16 && COARSE) {
            if (last_skipped && k0 >= coarse_from) {
                const int NC = (int)p.skip16 * CH;
                uint32_t c0, c1, c2;
                staged_box_n<PV>(vol, march, 0.0f, (float)(NC - 1), sx, sy_, sz_, tx, ty, tz,
                                 mgx, mgy, mgz, c0, c1, c2);
                wave_min3_u16x2(c0, c1, c2);
                BrickKey k;
                bool cskip = false;
                const bool queried =
                    brick_load(vol, c0 & 0xFFFFu, c0 >> 16, c1 & 0xFFFFu, 0xFFFFu - (c1 >> 16),
                               0xFFFFu - (c2 & 0xFFFFu), 0xFFFFu - (c2 >> 16), lane, k);
                if (queried) {
                    uint32_t unused = 0;
                    wave_min3_u32(k.kmin, k.nkmax, unused);
                    const float bmin = key_value<T>(k.kmin), bmax = key_value<T>(~k.nkmax);
                    const float m = bound_margin<T>(bmin, bmax);
                    const bool same_side = gt ? bmin - m > p.iso_val : bmax + m <= p.iso_val;
                    cskip = __ballot(march && !same_side) == 0;
                }
                if (cskip) {
#pragma unroll 4
                    for (int j = 0; j < NC; ++j) {
                        sx = sx + tx;
                        sy_ = sy_ + ty;
                        sz_ = sz_ + tz;
                    }
                    if (march) {
                        px = sx;
                        py = sy_;
                        pz = sz_;
                    }
                    k0 += NC - CH;
                    continue;
                }
                coarse_from = queried ? k0 + NC : nsteps;
            }
        }
        uint32_t b0, b1, b2;
        staged_box_n<PV>(vol, march, 0.0f, (float)(CH - 1), sx, sy_, sz_, tx, ty, tz, mgx, mgy,
                         mgz, b0, b1, b2);
        uint32_t r0 = b0, r1 = b1, r2 = b2;
        wave_min3_u16x2(r0, r1, r2);
        const uint32_t LX = r0 & 0xFFFFu, LY = r0 >> 16, LZ = r1 & 0xFFFFu;
        const uint32_t HX = 0xFFFFu - (r1 >> 16), HY = 0xFFFFu - (r2 & 0xFFFFu),
                       HZ = 0xFFFFu - (r2 >> 16);
        const uint32_t lox = LX & ~(kBoxAlign<T> - 1u);
        const uint32_t SX = (HX + 1u - lox + E - 1u) & ~(E - 1u);
        const uint32_t SY = HY - LY + 1u, SZ = HZ - LZ + 1u;
        const uint32_t nu = SX / E;
        const uint32_t rowb = nu * 16u;
        const uint64_t bytes = nu <= 64u ? (uint64_t)rowb * SY * SZ : ~0ull;
        const bool fits = bytes <= (uint64_t)SLAB_BYTES;
        bool skip = false;
        if constexpr (SKIP) {
            BrickKey k;
            if (k0 > 0 && brick_load(vol, LX, LY, LZ, HX, HY, HZ, lane, k)) {
                uint32_t unused = 0;
                wave_min3_u32(k.kmin, k.nkmax, unused);
                const float bmin = key_value<T>(k.kmin), bmax = key_value<T>(~k.nkmax);
                const float m = bound_margin<T>(bmin, bmax);
                const bool same_side = gt ? bmin - m > p.iso_val : bmax + m <= p.iso_val;
                skip = __ballot(march && !same_side) == 0;
            }
        }
        if (skip) {
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                sx = sx + tx;
                sy_ = sy_ + ty;
                sz_ = sz_ + tz;
            }
        } else {
            if (fits)
                staged_copy_rows<64>(vol.rsrc_plane(LZ), slab, lane, LY, 0u, lox * TS, nu, rowb, SY,
                                     SZ, PY, pitch);
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (fits)
                iso_march_chunk<PV, CH, true, FILT>(vol, slab, lox, LY, LZ, rowb, SY, sx, sy_, sz_,
                                                    tx, ty, tz, k0, nsteps, p.iso_val, march, gt,
                                                    found, ihit);
            else
                iso_march_chunk<PV, CH, false, FILT>(vol, slab, lox, LY, LZ, rowb, SY, sx, sy_, sz_,
                                                     tx, ty, tz, k0, nsteps, p.iso_val, march, gt,
                                                     found, ihit);
        }
        last_skipped = skip;
        if (march) {  // a marching lane's own position moves on
            px = sx;
            py = sy_;
            pz = sz_;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    if (!inside) return;
    if (!found) {
        iso_miss(o, out, alpha, depth, normals);
        return;
    }
    iso_finish<FILT>(vol, p, o, tnear, dt, delta, pos0, ihit, gt, direc, out, alpha, depth,
                     normals);
}

// Samples per staged box for a frame: the ray advances at most N_a / (16 nouter) voxels
// per sample along axis a (its whole path inside the box spans at most the volume).
// 16-sample boxes while 15 steps span at most 40 voxels -- a box that does not fit the
// slab is staged per half-wave, split into 8- and 4-sample boxes or gathered, so rays
// along an axis keep long boxes -- else 8 (measured on 1024^2 frames of a 512^3 volume,
// 4 poses: at 200 steps, 2.7 voxels a step, 16-sample boxes 28% faster than the 4-sample
// boxes a 12-voxel bound chose; at 50 and 100 steps 8-sample boxes beat 4 by 33% and 26%
// and 16 by 2-6%; at 400, 16 beat 8 by 18%).
__host__ inline int staged_chunk(int nouter, int nmax) {
    const float step = (float)nmax / (16.0f * (float)(nouter > 0 ? nouter : 1));
    return 15.0f * step <= 40.0f ? 16 : 8;
}

// The plain max projection (not the opacity mode, whose samples cost more, so that staging
// weighs less): 16-sample groups while 15 steps span at most 36 voxels, and boxes of two
// groups only while 31 steps span at most 40.  1024^2 frames of a 512^3 uint16 volume, 8
// poses, ms (two-group boxes / one group / 8-sample groups): 200 steps 0.595 / 0.581 /
// 0.562, 240 steps 0.642 / 0.625 / 0.640, 320 steps 0.704 / 0.691 / 0.782, 400 steps 0.764
// / 0.766 / 0.932; the opacity mode is fastest with two-group boxes at each of them.
__host__ inline int staged_chunk_mip(int nouter, int nmax) {
    const float step = (float)nmax / (16.0f * (float)(nouter > 0 ? nouter : 1));
    return 15.0f * step <= 36.0f ? 16 : 8;
}
__host__ inline uint32_t staged_groups_mip(int nouter, int nmax) {
    const float step = (float)nmax / (16.0f * (float)(nouter > 0 ? nouter : 1));
    return 31.0f * step <= 40.0f ? 2u : 1u;
}

// 16-sample chunks per coarse skip test of the staged iso march (RayParams::skip16): four
// while 63 steps span at most 40 voxels, else none (1024^2 frames of a 512^3 volume: at
// 896 steps four chunks a test made the iso frames 5% faster, 0.843 against 0.890 ms, and
// eight 3%; at 200 steps, whose 64-sample boxes mostly cover more bricks than one query
// takes, the tests cost 6%).
__host__ inline uint32_t iso_skip_chunks(int ch, int max_steps, int nmax) {
    const float step = (float)nmax / (float)(max_steps > 0 ? max_steps : 1);
    return ch == 16 && 63.0f * step <= 40.0f ? 4u : 1u;
}

// ---------------------------------------------------------------------------------
// Batched frames (many volumes x poses in one launch).  A job is one frame: a resident
// padded volume, the pose's inverse matrices and the two output planes.  In a uniform
// batch every other setting (frame size, box, min/max/gamma, max_steps, numParts) is the
// batch's; in a mixed batch (BatchQueues::mixed) each job carries its own frame size, box,
// transfer range, gamma and step count, and the voxel type, filter and alpha_pow are shared.

struct BatchJob {
    const void* base;  // padded volume of the job (same voxel type across the batch)
    uint32_t nbytes, px_bytes, py;
    int nx, ny, nz;
    float invP[16];
    float invM[16];
    float* out;
    float* alpha;
    // mixed batches: the frame's own settings, and its first item in each queue
    float box[6];
    float min_val, max_val, gamma;
    int nouter;
    uint32_t W, H;
    uint32_t first[8];
};

// Work queue of a persistent batch launch: the (job, tile) items whose tile row is
// congruent to q mod 8 form queue q, drained first by the workgroups with
// blockIdx % 8 == q (one XCD under round-robin dispatch, keeping a tile row's neighbours
// in one L2), then shared by every workgroup once its own queue is empty.
struct BatchQueues {
    uint32_t count[8];  // items per queue
    uint32_t tiles[8];  // uniform batch: tiles per frame in queue q
    uint32_t mixed;     // 1: jobs of their own sizes and settings (BatchJob::first)
    uint32_t njobs;
};

template <typename T, bool POW2>
__device__ __forceinline__ PaddedVolume<T, POW2> job_volume(const BatchJob& j) {
    PaddedVolume<T, POW2> v;
    v.base = j.base;
    v.nbytes = j.nbytes;
    v.total_bytes = j.nbytes;
    v.px_bytes = j.px_bytes;
    v.py = j.py;
    v.soff_y = j.px_bytes;
    v.soff_z = j.px_bytes * j.py;
    v.soff_yz = v.soff_y + v.soff_z;
    v.nx = j.nx;
    v.ny = j.ny;
    v.nz = j.nz;
    v.fnx = (float)j.nx;
    v.fny = (float)j.ny;
    v.fnz = (float)j.nz;
    v.mx = (float)(j.nx - 1);
    v.my = (float)(j.ny - 1);
    v.mz = (float)(j.nz - 1);
    v.bricks.keys = nullptr;  // batched frames sample every box
    v.bricks.nbx = v.bricks.nby = 0;
    v.vmin = v.vmax = 0.0f;
    return v;
}

// zero the queue counters (a kernel, so the batch can be captured in a HIP graph)
static __global__ void k_reset_counters(uint32_t* __restrict__ counters) {
    if (threadIdx.x < 8) counters[threadIdx.x] = 0u;
}

// Persistent launch: a fixed grid of workgroups claims (job, tile) items from the eight
// queues with one atomic per item until every queue is drained; every workgroup leaves
// after one failed claim on each queue.
// MODE 0: staged plain max projection (linear filter, alpha_pow == 0; CH samples per
// staged box); MODE 3: staged opacity mode (linear filter, alpha_pow > 0, per-wave slabs);
// MODE 1 / 2: every sample from global memory with the linear / nearest filter (any
// alpha_pow; CH and SHARED unused)
template <typename T, bool POW2, int SLAB_BYTES, int MODE, int CH, bool SHARED>
__global__ __launch_bounds__(256) void k_max_project_batch(const BatchJob* __restrict__ jobs,
                                                           RayParams shared, BatchQueues qs,
                                                           uint32_t* __restrict__ counters) {
    constexpr int kLds = SHARED ? SLAB_BYTES : 4 * SLAB_BYTES;
    __shared__ Unit16 lds[kLds / 16 + 1];  // +16 B: lds_pair reads past a pair
    __shared__ uint32_t s_box[12];
    __shared__ uint32_t s_item;
    uint8_t* slab = reinterpret_cast<uint8_t*>(lds) + (SHARED ? 0u : (threadIdx.x >> 6) * SLAB_BYTES);
    uint32_t q = blockIdx.x & 7u;
    for (uint32_t drained = 0; drained < 8u;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&counters[q], 1u);
        __syncthreads();
        const uint32_t item = s_item;
        __syncthreads();
        if (item >= qs.count[q]) {
            q = (q + 1u) & 7u;
            ++drained;
            continue;
        }
        uint32_t jid, t;
        if (qs.mixed) {
            // the job whose items of queue q hold `item`: the last one whose first item is
            // at most `item` (jobs without items in queue q share their successor's first)
            uint32_t lo = 0, hi = qs.njobs;
            while (hi - lo > 1u) {
                const uint32_t mid = (lo + hi) >> 1;
                if (jobs[mid].first[q] <= item)
                    lo = mid;
                else
                    hi = mid;
            }
            jid = lo;
            t = item - jobs[jid].first[q];
        } else {
            const uint32_t per = qs.tiles[q];
            jid = item / per;
            t = item - jid * per;
        }
        const BatchJob& j = jobs[jid];
        RayParams p = shared;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            p.invP[k] = j.invP[k];
            p.invM[k] = j.invM[k];
        }
        if (qs.mixed) {
            p.W = j.W;
            p.H = j.H;
            p.ntx = (j.W + 15u) / 16u;
            p.nty = (j.H + 15u) / 16u;
            p.ntiles = p.ntx * p.nty;
#pragma unroll
            for (int k = 0; k < 6; ++k) p.box[k] = j.box[k];
            p.min_val = j.min_val;
            p.max_val = j.max_val;
            p.gamma = j.gamma;
            p.nouter = j.nouter;
        }
        const uint32_t ty = q + 8u * (t / p.ntx), tx = t % p.ntx;
        uint32_t x, y;
        tile_thread_pixel(tx, ty, &x, &y);
        const bool inside = x < p.W && y < p.H;
        const PaddedVolume<T, POW2> vol = job_volume<T, POW2>(j);
        if constexpr (MODE == 0 || MODE == 3)
            // (no half-wave boxes: 1.4% slower here, where the queues leave no frame tail)
            mip_staged_rays<T, POW2, SLAB_BYTES, CH, SHARED, false, MODE == 3, false>(
                vol, p, x, y, inside, slab, s_box, j.out, j.alpha);
        else
            mip_rays<PaddedVolume<T, POW2>, MODE == 2 ? 1 : 0>(vol, p, x, y, inside, j.out,
                                                                j.alpha);
    }
}

}  // namespace spr

namespace spr {

// ---------------------------------------------------------------------------------
// Coarse max projection of two volumes (the learned reconstruction's input planes).
// Not the reference's sample sequence: sample k of a ray sits at p0 + k d (one fma, no
// running sum), so the samples of a ray can be split between waves.  A workgroup serves
// one 8x8 pixel tile of one volume (blockIdx.y); wave w marches samples
// [w L, (w+1) L) of all 64 rays (L = a quarter of the (nouter+1)*16 samples) with the
// per-wave LDS staging of the exact kernels, and the four partial maxima meet in LDS.
// Output: out[v][y][x] and alpha[y][x] (alpha from volume 0), transfer as
// volume_kernel.cl:159-170.
template <typename T, bool POW2, int SLAB_BYTES, int CH>
__global__ __launch_bounds__(256) void k_coarse_mip_pair(PaddedVolume<T, POW2> va,
                                                         PaddedVolume<T, POW2> vb, RayParams p,
                                                         float* __restrict__ out,
                                                         float* __restrict__ alpha) {
    constexpr bool kShort = !std::is_same<T, float>::value;
    constexpr uint32_t E = 16u / sizeof(T);
    constexpr uint32_t TS = sizeof(T);
    __shared__ Unit16 lds[4 * SLAB_BYTES / 16 + 1];  // +16 B: lds_pair reads past a pair
    __shared__ float red[4][64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint8_t* slab = reinterpret_cast<uint8_t*>(lds) + wave * SLAB_BYTES;
    const PaddedVolume<T, POW2> vol = blockIdx.y == 0 ? va : vb;
    float* outv = out + (size_t)blockIdx.y * p.W * p.H;
    const uint32_t ntx8 = (p.W + 7u) / 8u;
    const uint32_t x = (blockIdx.x % ntx8) * 8u + (lane & 7u), y = (blockIdx.x / ntx8) * 8u + (lane >> 3);
    const bool inside = x < p.W && y < p.H;
    float p0x = 0.5f, p0y = 0.5f, p0z = 0.5f, dx = 0.0f, dy = 0.0f, dz = 0.0f, tnear = 0.0f;
    bool active = false;
    if (inside) {
        F4 orig, direc;
        ray_setup(p.invP, p.invM, x, y, p.W, p.H, &orig, &direc);
        float tfar;
        if (intersect_box(orig, direc, p.box, &tnear, &tfar)) {
            if (tnear < 0.0f) tnear = 0.0f;
            const float dt = fabsf(tfar - tnear) / (float)(p.nouter * kLoopUnroll);
            const F4 delta = scl4(0.5f * dt, direc);
            const F4 pos = scl4(0.5f, add4(mk4(1.0f + orig.x, 1.0f + orig.y, 1.0f + orig.z,
                                               1.0f + orig.w),
                                           scl4(tnear, direc)));
            p0x = pos.x;
            p0y = pos.y;
            p0z = pos.z;
            dx = delta.x;
            dy = delta.y;
            dz = delta.z;
            active = true;
        }
    }
    const uint64_t act = __ballot(active);
    float col = 0.0f;
    if (act != 0) {  // uniform over the wave (every wave of the block sees the same rays)
        const int src = __builtin_ctzll(act);
        const float sx = __shfl(p0x, src), sy = __shfl(p0y, src), sz = __shfl(p0z, src);
        const float sdx = __shfl(dx, src), sdy = __shfl(dy, src), sdz = __shfl(dz, src);
        p0x = active ? p0x : sx;
        p0y = active ? p0y : sy;
        p0z = active ? p0z : sz;
        dx = active ? dx : sdx;
        dy = active ? dy : sdy;
        dz = active ? dz : sdz;
        const uint32_t PY = vol.py, pitch = vol.px_bytes;
        const float mgx = 1e-5f * vol.fnx + 0.01f, mgy = 1e-5f * vol.fny + 0.01f,
                    mgz = 1e-5f * vol.fnz + 0.01f;
        static_assert(CH == 4, "a quarter of (nouter+1)*16 samples is a multiple of 4 only");
        const int L = (p.nouter + 1) * kLoopUnroll / 4;
        const int k0 = (int)wave * L;
        for (int c = 0; c < L; c += CH) {
            const float kc = (float)(k0 + c);
            const float px = fmaf(kc, dx, p0x), py = fmaf(kc, dy, p0y), pz = fmaf(kc, dz, p0z);
            uint32_t b0, b1, b2;
            // the margins cover fma(k + CH - 1, d, p0) against fma(CH - 1, d, fma(k, d, p0))
            staged_chunk_box<PaddedVolume<T, POW2>, CH>(vol, active, px, py, pz, dx, dy, dz, mgx,
                                                        mgy, mgz, b0, b1, b2);
            uint32_t r0 = b0, r1 = b1, r2 = b2;
            wave_min3_u16x2(r0, r1, r2);
            const uint32_t LX = r0 & 0xFFFFu, LY = r0 >> 16, LZ = r1 & 0xFFFFu;
            const uint32_t HX = 0xFFFFu - (r1 >> 16), HY = 0xFFFFu - (r2 & 0xFFFFu),
                           HZ = 0xFFFFu - (r2 >> 16);
            const uint32_t lox = LX & ~(kBoxAlign<T> - 1u);
            const uint32_t SX = (HX + 1u - lox + E - 1u) & ~(E - 1u);
            const uint32_t SY = HY - LY + 1u, SZ = HZ - LZ + 1u;
            const uint32_t nu = SX / E;
            const bool fits = (uint64_t)SX * SY * SZ * TS <= (uint64_t)SLAB_BYTES && nu <= 64u;
            if (fits) {
                staged_copy_rows<64>(vol.rsrc_plane(LZ), slab, lane, LY, 0u, lox * TS, nu, SX * TS, SY,
                                     SZ, PY, pitch);
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                const uint32_t rowb = SX * TS, szb = rowb * SY;
                const float fsy = (float)rowb, fsz = (float)szb;
                const float oy = 1.0f - (float)LY, oz = 1.0f - (float)LZ;
                const float ox = (1.0f - (float)lox) * (float)TS;
#pragma unroll
                for (int j = 0; j < CH; ++j) {
                    const float kj = kc + (float)j;
                    float cx, cy, cz, ax, ay, az;
                    vol.locate_f(fmaf(kj, dx, p0x), fmaf(kj, dy, p0y), fmaf(kj, dz, p0z), &cx, &cy,
                                 &cz, &ax, &ay, &az);
                    const uint32_t q = (uint32_t)fmaf(cz + oz, fsz, fmaf(cy + oy, fsy, fmaf(cx, (float)TS, ox)));
                    float v000, v100, v010, v110, v001, v101, v011, v111;
                    const uint8_t* aq = slab + (q & ~3u);
                    const uint32_t sh = q & 3u;
                    lds_pair_at(aq, sh, v000, v100, T());
                    lds_pair_at(aq + rowb, sh, v010, v110, T());
                    lds_pair_at(aq + szb, sh, v001, v101, T());
                    lds_pair_at(aq + rowb + szb, sh, v011, v111, T());
                    col = fmaxf(col, trilerp(v000, v100, v010, v110, v001, v101, v011, v111, ax, ay, az));
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
            } else {
#pragma unroll
                for (int j = 0; j < CH; ++j) {
                    const float kj = kc + (float)j;
                    col = fmaxf(col, vol.linear(fmaf(kj, dx, p0x), fmaf(kj, dy, p0y), fmaf(kj, dz, p0z)));
                }
            }
        }
    }
    red[wave][lane] = col;
    __syncthreads();
    if (wave != 0 || !inside) return;
    const uint32_t o = x + p.W * y;
    if (!active) {
        outv[o] = 0.0f;
        if (blockIdx.y == 0) alpha[o] = kShort ? 0.0f : -1.0f;
        return;
    }
    col = fmaxf(fmaxf(red[0][lane], red[1][lane]), fmaxf(red[2][lane], red[3][lane]));
    col = (p.max_val == 0.0f) ? col : (col - p.min_val) / (p.max_val - p.min_val);
    outv[o] = apply_gamma(col, p.gamma);
    if (blockIdx.y == 0) alpha[o] = kShort ? tnear : 1.0f;
}

}  // namespace spr
