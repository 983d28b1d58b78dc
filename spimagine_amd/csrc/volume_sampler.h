// volume_sampler.h -- software trilinear / nearest sampling of a resident volume.
//
// Semantics of the reference's image sampler (volume_kernel.cl:42-43, utils.cl:104):
// CLK_NORMALIZED_COORDS_TRUE | CLK_ADDRESS_CLAMP_TO_EDGE | CLK_FILTER_{LINEAR,NEAREST},
// OpenCL 1.2 section 8.2.  LINEAR: u = s*N, u' = u - 0.5 (two float32 roundings),
// i0 = floor(u'), a = u' - i0, corners clamped to [0, N-1], then lerp along x, y, z.
// Integer voxels (u16/u8) are interpolated in float32 and not rounded back.
// Hardware texture filtering is not used: its reduced-precision weights would not
// reproduce these float32 weights.
#pragma once
#include "ray_math.h"

namespace spr {

template <typename T>
__device__ __forceinline__ float voxel_f(const T* __restrict__ p, uint32_t i) {
    return (float)p[i];
}

// floor of a texel coordinate kept in [-1, N] before the integer conversion; every
// value outside that range selects the same clamped corner pair
__device__ __forceinline__ int floor_idx(float f, float fn) { return (int)fminf(fmaxf(f, -1.0f), fn); }
__device__ __forceinline__ int clampi(int a, int lo, int hi) { return min(max(a, lo), hi); }

template <typename T>
struct LinearVolume {
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

}  // namespace spr
