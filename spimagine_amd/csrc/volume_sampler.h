// volume_sampler.h -- software trilinear / nearest sampling of a resident volume.
//
// Semantics of the reference's image sampler (volume_kernel.cl:42-43, utils.cl:75):
// CLK_NORMALIZED_COORDS_TRUE | CLK_ADDRESS_CLAMP_TO_EDGE | CLK_FILTER_{LINEAR,NEAREST},
// OpenCL 1.2 section 8.2.  LINEAR: u = s*N, u' = u - 0.5 (two float32 roundings),
// i0 = floor(u'), a = u' - i0, corners clamped to [0, N-1], then lerp along x, y, z.
// Integer voxels (u16/u8) are interpolated in float32 and not rounded back.
// Hardware texture filtering is not used: its reduced-precision weights would not
// reproduce these float32 weights.
#pragma once
#include "brick_bounds.h"
#include "ray_math.h"

namespace spr {

template <typename T>
__device__ __forceinline__ float voxel_f(const T* __restrict__ p, uint32_t i) {
    return (float)p[i];
}

// texel coordinate u' = s*N - 0.5.  When N is a power of two s*N is exact, so one fused
// multiply-add gives the same single rounding as the two-step form.
template <bool POW2>
__device__ __forceinline__ float texel(float s, float fn) {
    if (POW2) return fmaf(s, fn, -0.5f);
    float u = s * fn;
    return u - 0.5f;
}

// floor of a texel coordinate kept in [-1, N] before the integer conversion; every
// value outside that range selects the same clamped corner pair
__device__ __forceinline__ int floor_idx(float f, float fn) { return (int)fminf(fmaxf(f, -1.0f), fn); }
__device__ __forceinline__ int clampi(int a, int lo, int hi) { return min(max(a, lo), hi); }

// x, then y, then z (the lerp order every sampler and the oracle use)
__device__ __forceinline__ float trilerp(float v000, float v100, float v010, float v110, float v001,
                                         float v101, float v011, float v111, float ax, float ay,
                                         float az) {
    float c00 = lerpf(v000, v100, ax);
    float c10 = lerpf(v010, v110, ax);
    float c01 = lerpf(v001, v101, ax);
    float c11 = lerpf(v011, v111, ax);
    float c0 = lerpf(c00, c10, ay);
    float c1 = lerpf(c01, c11, ay);
    return lerpf(c0, c1, az);
}

template <typename T>
struct LinearVolume {
    using value_type = T;
    const T* __restrict__ data;
    int nx, ny, nz;
    float fnx, fny, fnz;

    __device__ __forceinline__ float linear(float px, float py, float pz) const {
        float ux = px * fnx, uy = py * fny, uz = pz * fnz;
        ux = ux - 0.5f;
        uy = uy - 0.5f;
        uz = uz - 0.5f;
        float fx = floorf(ux), fy = floorf(uy), fz = floorf(uz);
        float ax = ux - fx, ay = uy - fy, az = uz - fz;
        int ix = floor_idx(fx, fnx), iy = floor_idx(fy, fny), iz = floor_idx(fz, fnz);
        uint32_t x0 = clampi(ix, 0, nx - 1), x1 = clampi(ix + 1, 0, nx - 1);
        uint32_t y0 = clampi(iy, 0, ny - 1), y1 = clampi(iy + 1, 0, ny - 1);
        uint32_t z0 = clampi(iz, 0, nz - 1), z1 = clampi(iz + 1, 0, nz - 1);
        uint32_t r00 = (z0 * (uint32_t)ny + y0) * (uint32_t)nx;
        uint32_t r10 = (z0 * (uint32_t)ny + y1) * (uint32_t)nx;
        uint32_t r01 = (z1 * (uint32_t)ny + y0) * (uint32_t)nx;
        uint32_t r11 = (z1 * (uint32_t)ny + y1) * (uint32_t)nx;
        float c00 = lerpf(voxel_f(data, r00 + x0), voxel_f(data, r00 + x1), ax);
        float c10 = lerpf(voxel_f(data, r10 + x0), voxel_f(data, r10 + x1), ax);
        float c01 = lerpf(voxel_f(data, r01 + x0), voxel_f(data, r01 + x1), ax);
        float c11 = lerpf(voxel_f(data, r11 + x0), voxel_f(data, r11 + x1), ax);
        float c0 = lerpf(c00, c10, ay);
        float c1 = lerpf(c01, c11, ay);
        return lerpf(c0, c1, az);
    }

    __device__ __forceinline__ float nearest(float px, float py, float pz) const {
        float ux = px * fnx, uy = py * fny, uz = pz * fnz;
        uint32_t ix = clampi(floor_idx(floorf(ux), fnx), 0, nx - 1);
        uint32_t iy = clampi(floor_idx(floorf(uy), fny), 0, ny - 1);
        uint32_t iz = clampi(floor_idx(floorf(uz), fnz), 0, nz - 1);
        return voxel_f(data, (iz * (uint32_t)ny + iy) * (uint32_t)nx + ix);
    }

    template <int FILTER>
    __device__ __forceinline__ float sample(float px, float py, float pz) const {
        if (FILTER == 1) return nearest(px, py, pz);
        return linear(px, py, pz);
    }
};

// ---------------------------------------------------------------------------------
// Padded ("apron") layout: a (nz+2, ny+2, nx+2) copy whose outer shell replicates the
// nearest edge voxel, padded[p] = vol[clamp(p - 1, 0, N - 1)] per axis.  A linear sample
// with i0 = floor(u') clamped to [-1, N-1] then reads the corner pair (i0+1, i0+2) of the
// padded grid, which holds exactly the clamp-to-edge pair (vol[clamp(i0)],
// vol[clamp(i0+1)]): no per-corner clamping, and the x-neighbour pair is one (unaligned)
// load.  The four pair loads of a sample share one byte offset; the y and z neighbours
// are added through the buffer instruction's scalar offset.

template <typename T>
struct PairBits;
template <>
struct PairBits<uint16_t> {
    __device__ __forceinline__ static void load(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff,
                                                float* lo, float* hi) {
        uint32_t v = __builtin_amdgcn_raw_buffer_load_b32(r, off, soff, 0);
        *lo = (float)(v & 0xFFFFu);
        *hi = (float)(v >> 16);
    }
    __device__ __forceinline__ static float one(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        return (float)__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0);
    }
};
template <>
struct PairBits<uint8_t> {
    __device__ __forceinline__ static void load(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff,
                                                float* lo, float* hi) {
        uint32_t v = __builtin_amdgcn_raw_buffer_load_b16(r, off, soff, 0);
        *lo = (float)(v & 0xFFu);
        *hi = (float)((v >> 8) & 0xFFu);
    }
    __device__ __forceinline__ static float one(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        return (float)__builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
    }
};
template <>
struct PairBits<float> {
    __device__ __forceinline__ static void load(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff,
                                                float* lo, float* hi) {
        auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0);
        *lo = __builtin_bit_cast(float, (uint32_t)v[0]);
        *hi = __builtin_bit_cast(float, (uint32_t)v[1]);
    }
    __device__ __forceinline__ static float one(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    }
};

// Two uint16 channels per voxel, channel 0 in the low half: the 2x mean- and max-pooled
// copies of a volume interleaved, so that one box, one address and one read per corner
// pair serve both planes of a coarse pair render.
struct U16Pair {
    uint32_t bits;
};
template <>
struct PairBits<U16Pair> {
    // the corner pair (x, x+1) of both channels: lo[c] = channel c at x, hi[c] at x+1
    __device__ __forceinline__ static void load2(__amdgpu_buffer_rsrc_t r, uint32_t off,
                                                 uint32_t soff, float* lo, float* hi) {
        auto v = __builtin_amdgcn_raw_buffer_load_b64(r, off, soff, 0);
        const uint32_t a = (uint32_t)v[0], b = (uint32_t)v[1];
        lo[0] = (float)(a & 0xFFFFu);
        lo[1] = (float)(a >> 16);
        hi[0] = (float)(b & 0xFFFFu);
        hi[1] = (float)(b >> 16);
    }
};

// word 3 of a gfx9-family buffer resource (raw buffer, 32-bit data format)
constexpr int kBufferWord3 = 0x00020000;

// WIDE: the padded volume may exceed 4 GiB, beyond a buffer resource's 32-bit offsets.
// Staging then addresses rows relative to a resource rebased at the box's first z-plane
// (rsrc_plane), and linear() gathers with 64-bit flat loads.
template <typename T, bool POW2, bool WIDE = false>
struct PaddedVolume {
    using value_type = T;
    static constexpr bool kPow2 = POW2;
    static constexpr bool kWide = WIDE;
    const void* base;             // padded volume
    uint32_t nbytes;              // its size in bytes (the buffer's num_records; !WIDE)
    uint32_t px_bytes;            // row pitch, padded_pitch<T>(nx)*sizeof(T)
    uint32_t py;                  // padded rows per plane, ny+2
    uint32_t soff_y, soff_z, soff_yz;  // byte strides of the y, z and y+z neighbours
    uint64_t total_bytes;         // size in bytes (any size)
    int nx, ny, nz;
    float fnx, fny, fnz;           // N
    float mx, my, mz;              // N - 1
    BrickGrid bricks;              // per-brick bounds (bricks.keys == nullptr: none)
    float vmin, vmax;              // the volume's minimum and maximum voxel

    // raw buffer resource over the padded volume (uniform: lives in SGPRs)
    __device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc() const {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)nbytes,
                                                 kBufferWord3);
    }

    // raw buffer resource starting at padded plane z (uniform z): offsets of rows from
    // plane z on, up to 4 GiB past it
    __device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_plane(uint32_t z) const {
        const uint64_t skip = (uint64_t)z * soff_z;
        const uint64_t rest = total_bytes - skip;
        const uint32_t n = rest > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)rest;
        return __builtin_amdgcn_make_buffer_rsrc(
            const_cast<uint8_t*>(static_cast<const uint8_t*>(base) + skip), (short)0, (int)n,
            kBufferWord3);
    }

    __device__ __forceinline__ uint32_t byte_offset(uint32_t ix1, uint32_t iy1, uint32_t iz1) const {
        uint32_t row = __umul24(iz1, py) + iy1;
        return __umul24(row, px_bytes) + ix1 * (uint32_t)sizeof(T);
    }

    // padded index of one axis' lower corner: floor(u') clamped to [-1, N-1], plus one
    __device__ __forceinline__ static uint32_t corner1(float f, float m) {
        return (uint32_t)((int)__builtin_amdgcn_fmed3f(f, -1.0f, m) + 1);
    }

    // lower-corner padded indices and trilinear weights of a normalised position
    __device__ __forceinline__ void locate(float px, float py_, float pz, uint32_t* ix1,
                                           uint32_t* iy1, uint32_t* iz1, float* ax, float* ay,
                                           float* az) const {
        float ux = texel<POW2>(px, fnx), uy = texel<POW2>(py_, fny), uz = texel<POW2>(pz, fnz);
        float fx = floorf(ux), fy = floorf(uy), fz = floorf(uz);
        *ax = ux - fx;
        *ay = uy - fy;
        *az = uz - fz;
        *ix1 = corner1(fx, mx);
        *iy1 = corner1(fy, my);
        *iz1 = corner1(fz, mz);
    }

    // Lower-corner floors clamped to [-1, N-1] (as floats) and trilinear weights.  Index
    // arithmetic on these stays exact in float32 while it stays below 2^24, which the
    // padded layout guarantees for (nz+2)*(ny+2) and the row pitch.
    __device__ __forceinline__ void locate_f(float px, float py_, float pz, float* cx, float* cy,
                                             float* cz, float* ax, float* ay, float* az) const {
        float ux = texel<POW2>(px, fnx), uy = texel<POW2>(py_, fny), uz = texel<POW2>(pz, fnz);
        float fx = floorf(ux), fy = floorf(uy), fz = floorf(uz);
        *ax = ux - fx;
        *ay = uy - fy;
        *az = uz - fz;
        *cx = __builtin_amdgcn_fmed3f(fx, -1.0f, mx);
        *cy = __builtin_amdgcn_fmed3f(fy, -1.0f, my);
        *cz = __builtin_amdgcn_fmed3f(fz, -1.0f, mz);
    }

    // byte offset of the padded corner (cx+1, cy+1, cz+1): row = (cz+1)*PY + (cy+1) exact
    // in float, then one 24-bit multiply-add with the row pitch
    __device__ __forceinline__ uint32_t byte_offset_f(float cx, float cy, float cz) const {
        const float row = fmaf(cz + 1.0f, (float)py, cy + 1.0f);
        const uint32_t xb = (uint32_t)fmaf(cx, (float)sizeof(T), (float)sizeof(T));
        return __umul24((uint32_t)row, px_bytes) + xb;
    }

    __device__ __forceinline__ float linear(float px, float py_, float pz) const {
        float cx, cy, cz, ax, ay, az;
        locate_f(px, py_, pz, &cx, &cy, &cz, &ax, &ay, &az);
        float v000, v100, v010, v110, v001, v101, v011, v111;
        if constexpr (WIDE) {
            const uint32_t row = (uint32_t)fmaf(cz + 1.0f, (float)py, cy + 1.0f);
            const uint32_t xb = (uint32_t)fmaf(cx, (float)sizeof(T), (float)sizeof(T));
            const T* q = reinterpret_cast<const T*>(static_cast<const uint8_t*>(base) +
                                                    (uint64_t)row * px_bytes + xb);
            const T* qy = reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(q) + soff_y);
            const T* qz = reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(q) + soff_z);
            const T* qyz = reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(q) + soff_yz);
            v000 = (float)q[0];
            v100 = (float)q[1];
            v010 = (float)qy[0];
            v110 = (float)qy[1];
            v001 = (float)qz[0];
            v101 = (float)qz[1];
            v011 = (float)qyz[0];
            v111 = (float)qyz[1];
        } else {
            const uint32_t off = byte_offset_f(cx, cy, cz);
            const __amdgpu_buffer_rsrc_t r = rsrc();
            PairBits<T>::load(r, off, 0, &v000, &v100);
            PairBits<T>::load(r, off, soff_y, &v010, &v110);
            PairBits<T>::load(r, off, soff_z, &v001, &v101);
            PairBits<T>::load(r, off, soff_yz, &v011, &v111);
        }
        return trilerp(v000, v100, v010, v110, v001, v101, v011, v111, ax, ay, az);
    }

    // padded indices of nearest's voxel as floats (exact integers): the staged kernels'
    // slab offsets of a nearest sample
    __device__ __forceinline__ void nearest_f(float px, float py_, float pz, float* ix1,
                                              float* iy1, float* iz1) const {
        *ix1 = fminf(fmaxf(floorf(px * fnx), 0.0f), mx) + 1.0f;
        *iy1 = fminf(fmaxf(floorf(py_ * fny), 0.0f), my) + 1.0f;
        *iz1 = fminf(fmaxf(floorf(pz * fnz), 0.0f), mz) + 1.0f;
    }

    __device__ __forceinline__ float nearest(float px, float py_, float pz) const {
        float ux = px * fnx, uy = py_ * fny, uz = pz * fnz;
        uint32_t ix1 = (uint32_t)((int)fminf(fmaxf(floorf(ux), 0.0f), mx) + 1);
        uint32_t iy1 = (uint32_t)((int)fminf(fmaxf(floorf(uy), 0.0f), my) + 1);
        uint32_t iz1 = (uint32_t)((int)fminf(fmaxf(floorf(uz), 0.0f), mz) + 1);
        return PairBits<T>::one(rsrc(), byte_offset(ix1, iy1, iz1));
    }

    template <int FILTER>
    __device__ __forceinline__ float sample(float px, float py_, float pz) const {
        if (FILTER == 1) return nearest(px, py_, pz);
        return linear(px, py_, pz);
    }
};

// Row pitch of the padded layout in elements: nx+2 rounded up to a 16-byte multiple, so
// every padded row starts 16-byte aligned (the staged kernel moves rows in 16-byte units).
template <typename T>
__host__ __device__ inline uint32_t padded_pitch(int nx) {
    const uint32_t e = 16u / (uint32_t)sizeof(T);
    return ((uint32_t)nx + 2u + e - 1u) / e * e;
}

template <class V>
struct is_padded {
    static constexpr bool value = false;
};
template <typename T, bool POW2, bool WIDE>
struct is_padded<PaddedVolume<T, POW2, WIDE>> {
    static constexpr bool value = true;
};

// padded[z][y][x] = vol[clamp(z-1)][clamp(y-1)][clamp(x-1)] for x < pitch (the columns
// past nx+1 also replicate the last voxel)
template <typename T>
__global__ __launch_bounds__(256) void k_pad_volume(const T* __restrict__ src, T* __restrict__ dst,
                                                    int nx, int ny, int nz) {
    const int PXP = (int)padded_pitch<T>(nx), PY = ny + 2;
    const int x = blockIdx.x * 256 + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= PXP) return;
    const int sx = min(max(x - 1, 0), nx - 1);
    const int sy = min(max(y - 1, 0), ny - 1);
    const int sz = min(max(z - 1, 0), nz - 1);
    dst[((size_t)z * PY + y) * PXP + x] = src[((size_t)sz * ny + sy) * nx + sx];
}

}  // namespace spr
