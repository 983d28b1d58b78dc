// split_rays.h -- max projection of frames whose rays lie far apart (a sparse grid of
// pixels, e.g. one in 8 x 8 of a frame): each ray's samples are split into segments
// marched by different waves.
//
// A frame of scattered rays has no sample locality between neighbouring lanes, so the
// LDS-staged kernels fall back to global gathers, and with one ray per lane a small frame
// fills few waves: 8 sparse frames of 64 x 64 rays are 512 waves of ~900 dependent
// samples each on a 256-CU device.  Here a wave serves 8 x 8 rays of one frame over one
// segment of SEG samples; its lanes first advance their ray's position over the samples
// before the segment with the march's own additions (p = p + d, one rounding each, so
// every sample position is bit-identical to the sequential march of volume_kernel.cl:
// 121-131 and 293-298), then sample the segment through the staged kernels' LDS slabs
// (k_split_staged) or from global memory (k_split_march).  A ray's segments meet in its
// output word with an atomic max on the float's bits (every partial maximum is >= +0;
// only values > 0 are posted), exact because a maximum does not depend on how its
// sequence is cut.  A finishing pass
// applies the transfer (volume_kernel.cl:159-170) and writes alpha, as the one-ray-per-
// lane kernels do.  Plain max projection only (alpha_pow 0), linear filter, one pass.
#pragma once

#include "mip_staged.h"

namespace spr {

// Samples per segment: 64 -> a 1024^2 frame's 1/16-stride grid of 896-step rays is 15
// segments x 64 waves.
constexpr int kSplitSegDefault = 64;

// the job holding item `item` of a job table whose job k starts at jobs[k].first[which]
__device__ __forceinline__ uint32_t split_job_of(const BatchJob* __restrict__ jobs,
                                                 uint32_t njobs, uint32_t which, uint32_t item) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].first[which] <= item)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// frame jid is left untouched when its gate is closed (gates == nullptr: every frame runs)
__device__ __forceinline__ bool split_gated(const uint32_t* __restrict__ gates, uint32_t jid) {
    return gates != nullptr && gates[jid] == 0u;
}

// a job's ray parameters: the launch's (staging settings) with the job's matrices, frame
// size, box, transfer and step count; one pass
__device__ __forceinline__ RayParams split_params(const RayParams& base, const BatchJob& j) {
    RayParams p = base;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        p.invP[k] = j.invP[k];
        p.invM[k] = j.invM[k];
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) p.box[k] = j.box[k];
    p.W = j.W;
    p.H = j.H;
    p.min_val = j.min_val;
    p.max_val = j.max_val;
    p.gamma = j.gamma;
    p.nouter = j.nouter;
    p.current_part = 0;
    p.gate = nullptr;
    return p;
}

// Output words of every job's pixels to +0 (the identity of the segments' maximum).
// Pixels: job k's pixels start at jobs[k].first[1].
static __global__ __launch_bounds__(256) void k_split_clear(const BatchJob* __restrict__ jobs,
                                                            uint32_t njobs, uint32_t npix,
                                                            const uint32_t* __restrict__ gates) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= npix) return;
    const uint32_t jid = split_job_of(jobs, njobs, 1, i);
    if (split_gated(gates, jid)) return;
    const BatchJob& j = jobs[jid];
    j.out[i - j.first[1]] = 0.0f;
}

// One wave = 8 x 8 rays x one segment.  Items: job k's start at jobs[k].first[0]; within a
// job, item = group * nseg + segment (a group's segments side by side).
template <typename T, bool POW2, int SEG>
__global__ __launch_bounds__(256) void k_split_march(const BatchJob* __restrict__ jobs,
                                                     uint32_t njobs, uint32_t nitems,
                                                     RayParams base,
                                                     const uint32_t* __restrict__ gates) {
    const uint32_t item = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (item >= nitems) return;  // uniform per wave
    const uint32_t jid = split_job_of(jobs, njobs, 0, item);
    if (split_gated(gates, jid)) return;
    const BatchJob& j = jobs[jid];
    const uint32_t total = (j.nouter + 1u) * (uint32_t)kLoopUnroll;
    const uint32_t nseg = (total + SEG - 1u) / SEG;
    const uint32_t local = item - j.first[0];
    const uint32_t g = local / nseg, seg = local - g * nseg;
    const uint32_t ngx = (j.W + 7u) / 8u, lane = threadIdx.x & 63u;
    const uint32_t x = (g % ngx) * 8u + (lane & 7u), y = (g / ngx) * 8u + (lane >> 3);
    if (x >= j.W || y >= j.H) return;
    const RayParams p = split_params(base, j);
    float px, py, pz, dx, dy, dz, tnear;
    if (!staged_ray_pos(p, x, y, px, py, pz, dx, dy, dz, tnear)) return;
    const uint32_t s0 = seg * (uint32_t)SEG, s1 = min(total, s0 + (uint32_t)SEG);
#pragma unroll 4
    for (uint32_t k = 0; k < s0; ++k) {
        px = px + dx;
        py = py + dy;
        pz = pz + dz;
    }
    const PaddedVolume<T, POW2> vol = job_volume<T, POW2>(j);
    float col = 0.0f;
#pragma unroll 4
    for (uint32_t k = s0; k < s1; ++k) {
        col = fmaxf(col, vol.linear(px, py, pz));
        px = px + dx;
        py = py + dy;
        pz = pz + dz;
    }
    if (col > 0.0f)
        atomicMax(reinterpret_cast<uint32_t*>(j.out) + x + j.W * y, __float_as_uint(col));
}

// The same work items marched through the per-wave LDS slab of the staged kernels (one
// wave per workgroup, 20,464-byte slab: eight workgroups per CU): the segment's 16-sample
// boxes are staged whole, per half-wave, or split into 8- and 4-sample boxes, as in the
// single-frame kernel (mip_staged_march), and sampled from global memory only when none
// fits.  Rays a few voxels apart (one pixel in 8 x 8 of a 1024^2 frame of a 512^3
// volume) fit as half-wave boxes.  Lanes without a ray follow the first active lane.
constexpr int kSplitSlab = 20464;

template <typename T, bool POW2, int SEG>
__global__ __launch_bounds__(64) void k_split_staged(const BatchJob* __restrict__ jobs,
                                                     uint32_t njobs, uint32_t nitems,
                                                     RayParams base,
                                                     const uint32_t* __restrict__ gates) {
    static_assert(SEG % kLoopUnroll == 0, "segments of whole 16-sample groups");
    __shared__ Unit16 lds[kSplitSlab / 16 + 1];  // +16 B: lds_pair reads past a pair
    const uint32_t item = blockIdx.x;
    if (item >= nitems) return;
    const uint32_t jid = split_job_of(jobs, njobs, 0, item);
    if (split_gated(gates, jid)) return;  // uniform
    const BatchJob& j = jobs[jid];
    const uint32_t total = (j.nouter + 1u) * (uint32_t)kLoopUnroll;
    const uint32_t nseg = (total + SEG - 1u) / SEG;
    const uint32_t local = item - j.first[0];
    const uint32_t g = local / nseg, seg = local - g * nseg;
    const uint32_t ngx = (j.W + 7u) / 8u, lane = threadIdx.x & 63u;
    const uint32_t x = (g % ngx) * 8u + (lane & 7u), y = (g / ngx) * 8u + (lane >> 3);
    const RayParams p = split_params(base, j);
    float px = 0.5f, py = 0.5f, pz = 0.5f, dx = 0.0f, dy = 0.0f, dz = 0.0f, tnear = 0.0f;
    const bool active =
        x < j.W && y < j.H && staged_ray_pos(p, x, y, px, py, pz, dx, dy, dz, tnear);
    const uint64_t act = __ballot(active);
    if (act == 0) return;  // uniform
    const uint32_t s0 = seg * (uint32_t)SEG, s1 = min(total, s0 + (uint32_t)SEG);
    if (active) {
#pragma unroll 4
        for (uint32_t k = 0; k < s0; ++k) {
            px = px + dx;
            py = py + dy;
            pz = pz + dz;
        }
    }
    shadow_first_active(act, active, px, py, pz, dx, dy, dz);
    const PaddedVolume<T, POW2> vol = job_volume<T, POW2>(j);
    float col = 0.0f, col2 = 0.0f;
    mip_staged_march<T, POW2, kSplitSlab, 16, false, false, false, true, false>(
        vol, p, active, true, reinterpret_cast<uint8_t*>(lds), nullptr, px, py, pz, dx, dy, dz,
        (int)(s0 / (uint32_t)kLoopUnroll), (int)(s1 / (uint32_t)kLoopUnroll), col, col2);
    if (active && col > 0.0f)
        atomicMax(reinterpret_cast<uint32_t*>(j.out) + x + j.W * y, __float_as_uint(col));
}

// The transfer and alpha of every ray (volume_kernel.cl:159-183); a ray that misses the box
// gets the missed ray's outputs (volume_kernel.cl:85-91, :258-264).
template <bool kShort>
__global__ __launch_bounds__(256) void k_split_finish(const BatchJob* __restrict__ jobs,
                                                      uint32_t njobs, uint32_t npix,
                                                      RayParams base,
                                                      const uint32_t* __restrict__ gates) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= npix) return;
    const uint32_t jid = split_job_of(jobs, njobs, 1, i);
    if (split_gated(gates, jid)) return;
    const BatchJob& j = jobs[jid];
    const uint32_t o = i - j.first[1];
    const uint32_t y = o / j.W, x = o - y * j.W;
    const RayParams p = split_params(base, j);
    float px, py, pz, dx, dy, dz, tnear;
    if (!staged_ray_pos(p, x, y, px, py, pz, dx, dy, dz, tnear)) {
        j.out[o] = 0.0f;
        j.alpha[o] = kShort ? 0.0f : -1.0f;
        return;
    }
    staged_ray_end<kShort>(p, o, j.out[o], tnear, j.out, j.alpha);
}

}  // namespace spr
