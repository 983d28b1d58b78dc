// brick_bounds.h -- per-brick value bounds of a padded volume, for exact empty-space
// skipping and early ray termination.
//
// The padded copy (volume_sampler.h) is cut into 8x8x8 bricks of padded voxels; for each
// brick the library keeps the minimum and maximum voxel as order-preserving 32-bit keys
// (uint16/uint8: the value itself; float: the IEEE bits mapped so that unsigned order is
// numeric order), stored as (min key, ~max key) so that one unsigned min-reduction bounds a
// box from both sides.  A brick holding a NaN gets the widest keys, which no skip test
// accepts.  Built on upload (k_brick_bounds in spimrender.hip).
//
// Why the bounds are exact for the renderer: every trilinear sample is a convex
// combination of the 8 corners it reads, evaluated as three levels of fma lerps.  For
// integer voxels the lerp operands and differences are integers below 2^24, so each lerp
// lies between its two endpoints exactly and the sample lies in [min, max] of its corners.
// For float voxels each lerp can leave that interval by a few ulps of the corners'
// magnitude; float bounds are widened by 2^-16 of the magnitude (plus 1e-30 for
// denormals), far above the at most ~9 ulps three lerps can add.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace spr {

constexpr uint32_t kBrickShift = 3;  // 8x8x8 padded voxels per brick
constexpr uint32_t kBrick = 1u << kBrickShift;
constexpr uint32_t kBrickMaxQuery = 128;  // bricks one wave query may cover (2 per lane)

struct BrickKey {
    uint32_t kmin;   // order key of the brick's minimum
    uint32_t nkmax;  // ~ order key of its maximum
};

// order-preserving key of a voxel value
__host__ __device__ inline uint32_t order_key(uint16_t v) { return v; }
__host__ __device__ inline uint32_t order_key(uint8_t v) { return v; }
__host__ __device__ inline uint32_t order_key_bits(uint32_t b) {
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ inline uint32_t order_key(float v) { return order_key_bits(__float_as_uint(v)); }

// the value of a key (float: the inverse map; the NaN keys of a NaN brick stay NaN)
template <typename T>
__host__ __device__ inline float key_value(uint32_t k) {
    if constexpr (std::is_same<T, float>::value) {
        const uint32_t b = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
#ifdef __HIP_DEVICE_COMPILE__
        return __uint_as_float(b);
#else
        float f;
        __builtin_memcpy(&f, &b, 4);
        return f;
#endif
    } else {
        return (float)k;
    }
}

// widening of a float bound (see the file comment); integer voxels need none
template <typename T>
__device__ __forceinline__ float bound_margin(float lo, float hi) {
    if constexpr (std::is_same<T, float>::value)
        return fmaxf(fabsf(lo), fabsf(hi)) * 0x1p-16f + 1e-30f;
    else
        return 0.0f;
}

// Per-brick keys of a padded volume: keys[(bz nby + by) nbx + bx] covers padded voxels
// [8 bx, 8 bx + 8) x [8 by, ...) x [8 bz, ...).  keys == nullptr: no bounds.
struct BrickGrid {
    const BrickKey* keys;
    uint32_t nbx, nby;
};

}  // namespace spr
