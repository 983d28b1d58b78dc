// launch_typed.hip -- kernel launchers of one voxel type and one power-of-two rule.
//
// Compiled once per (SPR_VOXEL, SPR_POW2) pair (spimagine_amd/build.py): SPR_VOXEL 0 =
// float, 1 = uint16_t, 2 = uint8_t; SPR_POW2 1 = every volume dimension a power of two
// (PaddedVolume's one-rounding texel rule).  Each compilation instantiates
// TypedLaunch<T, P2> (spr_internal.h) and, for uint16_t, the channel-pair launcher.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "spr_internal.h"

#ifndef SPR_VOXEL
#error "compile with -DSPR_VOXEL=0|1|2 -DSPR_POW2=0|1"
#endif

// SPR_PART 0: the staged launchers (built with the max-ILP scheduler); 1: the persistent
// batch launch and the plain gather kernels (default scheduler); undefined: both.
#ifndef SPR_PART
#define SPR_PART 2
#endif

namespace spr {
// The per-wave slab of the four-wave staged kernels for slab_kb 19: two workgroups per CU
// of four slabs, their 16 bytes of slack and up to 52 bytes of box and queue words, within
// 81,920 B each (LDS is allocated in 512-byte granules): 5% more than 19 KiB, 1.4% faster
// on the batch launch.
constexpr uint32_t kQuadSlab19 = 20448;
namespace {

template <typename T, bool P2, int CH, bool SH, typename... Args>
void launch_staged_sh(int slab_kb, dim3 grid, hipStream_t s, Args... args) {
    constexpr int F = SH ? 4 : 1;  // a workgroup slab holds the four waves' share
    if (slab_kb == 8)
        hipLaunchKernelGGL((k_max_project_staged<T, P2, 8192 * F, CH, SH>), grid, dim3(256), 0, s, args...);
    else if (slab_kb == 16 || (SH && slab_kb == 19))
        hipLaunchKernelGGL((k_max_project_staged<T, P2, 16384 * F, CH, SH>), grid, dim3(256), 0, s, args...);
    else if (slab_kb == 19)  // per-wave only: two workgroups of 4 slabs fill the 160 KiB
        hipLaunchKernelGGL((k_max_project_staged<T, P2, kQuadSlab19, CH, false>), grid, dim3(256), 0, s, args...);
    else
        hipLaunchKernelGGL((k_max_project_staged<T, P2, 12288 * F, CH, SH>), grid, dim3(256), 0, s, args...);
}

template <typename T, bool P2, int CH, typename... Args>
void launch_staged_ch(int slab_kb, bool shared, dim3 grid, hipStream_t s, Args... args) {
    if (shared)
        launch_staged_sh<T, P2, CH, true>(slab_kb, grid, s, args...);
    else
        launch_staged_sh<T, P2, CH, false>(slab_kb, grid, s, args...);
}

// The cost-ordered tile map (kTileMapCostXcd) for the per-wave staged kernels: the waves'
// estimated costs (k_tile_cost) and the workgroups' order (k_tile_order), computed on
// stream s into *buf from the stream-ordered allocator (the caller frees it after the
// frame's launch).  per_wave: one wave per workgroup (k_max_project_wave) instead of a
// 16x16 tile.  Sets p.tile_map and p.order and the frame's grid; other tile maps, and
// frames of more than kMaxOrderedTiles tiles, are left as they are.
constexpr uint32_t kMaxOrderedTiles = 65536;

template <typename T, bool P2, bool WIDE>
hipError_t order_tiles(int tile_map, bool per_wave, const PaddedVolume<T, P2, WIDE>& v,
                       uint32_t slab_bytes, RayParams& p, hipStream_t s, void** buf, dim3* grid) {
    if (tile_map != (int)kTileMapCostXcd || p.ntiles > kMaxOrderedTiles) return hipSuccess;
    const uint32_t slots = tile_grid((uint32_t)tile_map, p.ntx, p.nty) << (per_wave ? 2 : 0);
    hipError_t e = hipMallocAsync(buf, (size_t)(4 * p.ntiles + slots) * sizeof(uint32_t), s);
    if (e != hipSuccess) return e;
    uint32_t* cost = (uint32_t*)*buf;
    uint32_t* order = cost + 4 * p.ntiles;
    hipLaunchKernelGGL((k_tile_cost<T, P2, WIDE>), dim3(p.ntiles), dim3(256), 0, s, v, p,
                       slab_bytes, cost);
    p.tile_map = (uint32_t)tile_map;
    hipLaunchKernelGGL(k_tile_order, dim3(1), dim3(512), 0, s, (const uint32_t*)cost, p,
                       per_wave ? 1u : 0u, order);
    p.order = order;
    *grid = dim3(slots);
    return hipGetLastError();
}

// The slab of a one-wave workgroup for slab_kb 19: LDS is allocated in 512-byte granules,
// so eight one-wave workgroups per CU fit 20,480 B each (the slab plus its 16 bytes of
// slack: all 160 KiB), 5% more than 19 KiB (19,952 B was ~0.5-1% faster than 19 KiB, and
// 20,464 B 1.3% faster again on the 1024^2 max projection).
constexpr uint32_t kWaveSlab19 = 20464;

// one wave per workgroup: the plain max projection (slab 16 or 19 KiB) or the opacity mode;
// skip: the variant that skips boxes by their brick bounds (19 KiB slabs)
template <typename T, bool P2, bool OPAC, int CH, typename... Args>
void launch_wave_ch(int slab_kb, bool skip, bool fastdiv, dim3 grid, hipStream_t s, Args... args) {
    // the opacity mode of integer voxels with its normalisation free of divisions
    if constexpr (OPAC && !std::is_same<T, float>::value) {
        if (fastdiv) {
            if (skip)
                hipLaunchKernelGGL((k_max_project_wave<T, P2, kWaveSlab19, CH, false, true, true, 0, true>), grid, dim3(64), 0, s, args...);
            else
                hipLaunchKernelGGL((k_max_project_wave<T, P2, kWaveSlab19, CH, false, true, false, 0, true>), grid, dim3(64), 0, s, args...);
            return;
        }
    }
    if (slab_kb == 16 && !OPAC)
        hipLaunchKernelGGL((k_max_project_wave<T, P2, 16384, CH, false, OPAC>), grid, dim3(64), 0, s, args...);
    else if (skip)
        hipLaunchKernelGGL((k_max_project_wave<T, P2, kWaveSlab19, CH, false, OPAC, true>), grid, dim3(64), 0, s, args...);
    else
        hipLaunchKernelGGL((k_max_project_wave<T, P2, kWaveSlab19, CH, false, OPAC>), grid, dim3(64), 0, s, args...);
}

template <typename T, bool P2, bool OPAC, typename... Args>
void launch_wave(int slab_kb, bool skip, bool fastdiv, int ch, dim3 grid, hipStream_t s,
                 Args... args) {
    if (ch == 16)
        launch_wave_ch<T, P2, OPAC, 16>(slab_kb, skip, fastdiv, grid, s, args...);
    else if (ch == 8)
        launch_wave_ch<T, P2, OPAC, 8>(slab_kb, skip, fastdiv, grid, s, args...);
    else
        launch_wave_ch<T, P2, OPAC, 4>(slab_kb, skip, fastdiv, grid, s, args...);
}

// the nearest filter staged: one wave per workgroup, 19 KiB slabs; OPAC: the opacity mode;
// SKIP: boxes skipped (max projection) or waves stopped (opacity mode) by brick bounds
template <typename T, bool P2, bool OPAC, bool SKIP, typename... Args>
void launch_wave_nearest_sk(int ch, dim3 grid, hipStream_t s, Args... args) {
    if (ch == 16)
        hipLaunchKernelGGL((k_max_project_wave<T, P2, kWaveSlab19, 16, false, OPAC, SKIP, 1>), grid, dim3(64), 0, s, args...);
    else if (ch == 8)
        hipLaunchKernelGGL((k_max_project_wave<T, P2, kWaveSlab19, 8, false, OPAC, SKIP, 1>), grid, dim3(64), 0, s, args...);
    else
        hipLaunchKernelGGL((k_max_project_wave<T, P2, kWaveSlab19, 4, false, OPAC, SKIP, 1>), grid, dim3(64), 0, s, args...);
}

template <typename T, bool P2, typename... Args>
void launch_wave_nearest(bool opac, bool skip, int ch, dim3 grid, hipStream_t s, Args... args) {
    if (opac)
        skip ? launch_wave_nearest_sk<T, P2, true, true>(ch, grid, s, args...)
             : launch_wave_nearest_sk<T, P2, true, false>(ch, grid, s, args...);
    else
        skip ? launch_wave_nearest_sk<T, P2, false, true>(ch, grid, s, args...)
             : launch_wave_nearest_sk<T, P2, false, false>(ch, grid, s, args...);
}

// One wave per workgroup for the per-wave staged kernels?  wg_waves 0 (default): yes.  A
// CU takes a new wave's work as soon as any of its waves ends, instead of when a
// workgroup's slowest of four waves does: 1024^2 frames 7% faster; 2048^2 frames of the
// 1024x1024x512 volume 10.6% (max projection), 7% (opacity mode) and 8.5% (iso march)
// faster, of the 2048x2048x512 one 9.7% and 3.5%.  (Before the max-ILP build and the
// 20,464-byte slabs four waves were 10% faster on 2048^2 frames.)
bool one_wave_for(const LaunchTuning& t, const RayParams&) {
    if (t.wg_waves != 0) return t.wg_waves == 1;
    return true;
}

// a per-wave slab's bytes for the slab_kb tuning value
uint32_t slab_bytes_of(int slab_kb) { return slab_kb == 19 ? kQuadSlab19 : (uint32_t)slab_kb * 1024u; }

// volumes above 4 GiB: the same per-wave slabs as below 4 GiB (on 2048^2 frames of a
// 2048x2048x512 volume 2.3% faster for the max projection and 6.5% for the opacity mode
// than 16 KiB slabs); OPAC: the opacity mode; SKIP: boxes skipped by their brick bounds
template <typename T, bool P2, bool OPAC, bool SKIP, int CH>
void launch_staged_wide_ch(const PaddedVolume<T, P2, true>& v, bool one_wave, dim3 grid,
                           hipStream_t s, const RayParams& p, float* out, float* alpha) {
    if (one_wave)  // grid: one workgroup per 8x8 quarter of the interleaved tile grid
        hipLaunchKernelGGL((k_max_project_wave<T, P2, kWaveSlab19, CH, true, OPAC, SKIP>),
                           dim3(4 * tile_grid(kTileMapInterleaved, p.ntx, p.nty)), dim3(64), 0, s,
                           v, p, out, alpha);
    else
        hipLaunchKernelGGL((k_max_project_staged<T, P2, kQuadSlab19, CH, false, true, OPAC, SKIP>), grid,
                           dim3(256), 0, s, v, p, out, alpha);
}

template <typename T, bool P2, bool OPAC, bool SKIP>
void launch_staged_wide(const VolSlot& vs, int ch, bool one_wave, dim3 grid, hipStream_t s,
                        const RayParams& p, float* out, float* alpha) {
    const PaddedVolume<T, P2, true> v = padded_volume<T, P2, true>(vs);
    if (ch == 16)
        launch_staged_wide_ch<T, P2, OPAC, SKIP, 16>(v, one_wave, grid, s, p, out, alpha);
    else if (ch == 8)
        launch_staged_wide_ch<T, P2, OPAC, SKIP, 8>(v, one_wave, grid, s, p, out, alpha);
    else
        launch_staged_wide_ch<T, P2, OPAC, SKIP, 4>(v, one_wave, grid, s, p, out, alpha);
}

// the iso-surface march of a volume above 4 GiB staged: one wave per workgroup in
// 20,464-byte slabs, or four waves in 16 KiB slabs (20,448 B measured 0.7% slower on
// 2048^2 frames)
template <typename T, bool P2, bool SKIP>
void launch_iso_wide(const VolSlot& vs, int ch, bool one_wave, dim3 grid, hipStream_t s,
                     const RayParams& p, float* out, float* alpha, float* depth, float* normals) {
    const PaddedVolume<T, P2, true> v = padded_volume<T, P2, true>(vs);
    const dim3 g1(4 * tile_grid(kTileMapInterleaved, p.ntx, p.nty));
    if (one_wave && ch == 16 && SKIP && p.skip16 > 1)
        hipLaunchKernelGGL((k_iso_staged<T, P2, kWaveSlab19, 16, true, SKIP, 0, true, true>), g1, dim3(64), 0, s, v, p, out, alpha, depth, normals);
    else if (one_wave && ch == 16)
        hipLaunchKernelGGL((k_iso_staged<T, P2, kWaveSlab19, 16, true, SKIP, 0, true>), g1, dim3(64), 0, s, v, p, out, alpha, depth, normals);
    else if (one_wave && ch == 8)
        hipLaunchKernelGGL((k_iso_staged<T, P2, kWaveSlab19, 8, true, SKIP, 0, true>), g1, dim3(64), 0, s, v, p, out, alpha, depth, normals);
    else if (one_wave)
        hipLaunchKernelGGL((k_iso_staged<T, P2, kWaveSlab19, 4, true, SKIP, 0, true>), g1, dim3(64), 0, s, v, p, out, alpha, depth, normals);
    else if (ch == 16)
        hipLaunchKernelGGL((k_iso_staged<T, P2, 16384, 16, true, SKIP>), grid, dim3(256), 0, s, v, p, out, alpha, depth, normals);
    else if (ch == 8)
        hipLaunchKernelGGL((k_iso_staged<T, P2, 16384, 8, true, SKIP>), grid, dim3(256), 0, s, v, p, out, alpha, depth, normals);
    else
        hipLaunchKernelGGL((k_iso_staged<T, P2, 16384, 4, true, SKIP>), grid, dim3(256), 0, s, v, p, out, alpha, depth, normals);
}

// four waves per workgroup, per-wave or workgroup slabs; skip: per-wave 19 KiB slabs that
// skip boxes by their brick bounds
template <typename T, bool P2, typename... Args>
void launch_staged(int slab_kb, bool shared, bool skip, int ch, dim3 grid, hipStream_t s,
                   Args... args) {
    if (skip) {
        if (ch == 16)
            hipLaunchKernelGGL((k_max_project_staged<T, P2, kQuadSlab19, 16, false, false, false, true>), grid, dim3(256), 0, s, args...);
        else if (ch == 8)
            hipLaunchKernelGGL((k_max_project_staged<T, P2, kQuadSlab19, 8, false, false, false, true>), grid, dim3(256), 0, s, args...);
        else
            hipLaunchKernelGGL((k_max_project_staged<T, P2, kQuadSlab19, 4, false, false, false, true>), grid, dim3(256), 0, s, args...);
        return;
    }
    if (ch == 16)
        launch_staged_ch<T, P2, 16>(slab_kb, shared, grid, s, args...);
    else if (ch == 8)
        launch_staged_ch<T, P2, 8>(slab_kb, shared, grid, s, args...);
    else
        launch_staged_ch<T, P2, 4>(slab_kb, shared, grid, s, args...);
}

// the opacity mode (alpha_pow > 0) staged: per-wave 19 KiB slabs; SKIP: rays stop once no
// later sample can change them
template <typename T, bool P2, bool SKIP, typename... Args>
void launch_staged_opac(int ch, dim3 grid, hipStream_t s, Args... args) {
    if (ch == 16)
        hipLaunchKernelGGL((k_max_project_staged<T, P2, kQuadSlab19, 16, false, false, true, SKIP>), grid, dim3(256), 0, s, args...);
    else if (ch == 8)
        hipLaunchKernelGGL((k_max_project_staged<T, P2, kQuadSlab19, 8, false, false, true, SKIP>), grid, dim3(256), 0, s, args...);
    else
        hipLaunchKernelGGL((k_max_project_staged<T, P2, kQuadSlab19, 4, false, false, true, SKIP>), grid, dim3(256), 0, s, args...);
}

// the iso-surface march staged: per-wave slabs; one_wave: one wave per workgroup, grid 4 x
// the interleaved tile grid (with coarse skip tests when p.skip16 asks for them)
template <typename T, bool P2, bool SKIP, int FILT = 0, typename... Args>
void launch_iso_staged(int ch, bool one_wave, dim3 grid, hipStream_t s, const RayParams& p,
                       Args... args) {
    const dim3 g1(4 * tile_grid(kTileMapInterleaved, p.ntx, p.nty));
    if (one_wave && ch == 16 && SKIP && p.skip16 > 1)
        hipLaunchKernelGGL((k_iso_staged<T, P2, kWaveSlab19, 16, false, SKIP, FILT, true, true>), g1, dim3(64), 0, s, args...);
    else if (one_wave && ch == 16)
        hipLaunchKernelGGL((k_iso_staged<T, P2, kWaveSlab19, 16, false, SKIP, FILT, true>), g1, dim3(64), 0, s, args...);
    else if (one_wave && ch == 8)
        hipLaunchKernelGGL((k_iso_staged<T, P2, kWaveSlab19, 8, false, SKIP, FILT, true>), g1, dim3(64), 0, s, args...);
    else if (one_wave)
        hipLaunchKernelGGL((k_iso_staged<T, P2, kWaveSlab19, 4, false, SKIP, FILT, true>), g1, dim3(64), 0, s, args...);
    else if (ch == 16)
        hipLaunchKernelGGL((k_iso_staged<T, P2, kQuadSlab19, 16, false, SKIP, FILT>), grid, dim3(256), 0, s, args...);
    else if (ch == 8)
        hipLaunchKernelGGL((k_iso_staged<T, P2, kQuadSlab19, 8, false, SKIP, FILT>), grid, dim3(256), 0, s, args...);
    else
        hipLaunchKernelGGL((k_iso_staged<T, P2, kQuadSlab19, 4, false, SKIP, FILT>), grid, dim3(256), 0, s, args...);
}

// the plain (unstaged) kernels on the slot's layout: linear layout (a volume above 4 GiB
// with the nearest filter included), or staging switched off
template <typename T, bool P2, typename F>
hipError_t visit_layout(const VolSlot& vs, F&& f) {
    if (vs.layout == SPR_LAYOUT_LINEAR) return f(linear_volume<T>(vs));
    return f(padded_volume<T, P2>(vs));
}

// one persistent batch launch: counter reset + the kernel, grid = CUs x resident workgroups
template <typename T, bool P2, int SLAB, int MODE, int CH, bool SH>
hipError_t launch_batch_kernel(int num_cus, const BatchJob* d_jobs, const RayParams& p,
                               const BatchQueues& q, uint32_t* counters, hipStream_t s) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, k_max_project_batch<T, P2, SLAB, MODE, CH, SH>, 256, 0);
    if (e != hipSuccess) return e;
    if (per_cu < 1) per_cu = 1;
    const uint32_t grid = (uint32_t)(num_cus * per_cu + 7) / 8u * 8u;
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, counters);
    hipLaunchKernelGGL((k_max_project_batch<T, P2, SLAB, MODE, CH, SH>), dim3(grid), dim3(256), 0, s,
                       d_jobs, p, q, counters);
    return hipGetLastError();
}

template <typename T, bool P2, int SLAB, bool SH>
hipError_t launch_batch_staged(int ch, int num_cus, const BatchJob* d_jobs, const RayParams& p,
                               const BatchQueues& q, uint32_t* counters, hipStream_t s) {
    constexpr int B = SLAB * (SH ? 4 : 1);
    if (ch == 16) return launch_batch_kernel<T, P2, B, 0, 16, SH>(num_cus, d_jobs, p, q, counters, s);
    if (ch == 8) return launch_batch_kernel<T, P2, B, 0, 8, SH>(num_cus, d_jobs, p, q, counters, s);
    return launch_batch_kernel<T, P2, B, 0, 4, SH>(num_cus, d_jobs, p, q, counters, s);
}

template <typename T, bool P2, int SLAB>
hipError_t launch_batch_slab(bool shared, int ch, int num_cus, const BatchJob* d_jobs,
                             const RayParams& p, const BatchQueues& q, uint32_t* counters,
                             hipStream_t s) {
    return shared ? launch_batch_staged<T, P2, SLAB, true>(ch, num_cus, d_jobs, p, q, counters, s)
                  : launch_batch_staged<T, P2, SLAB, false>(ch, num_cus, d_jobs, p, q, counters, s);
}

// both channels of a pair slot: one wave per workgroup (20,464-byte slabs) or four
template <bool P2, int CH>
void launch_pair_ch(const VolSlot& vs, bool one_wave, dim3 grid, hipStream_t s, const RayParams& p,
                    float* out, float* alpha) {
    if (one_wave)
        hipLaunchKernelGGL((k_max_project_wave<U16Pair, P2, kWaveSlab19, CH>),
                           dim3(4 * tile_grid(kTileMapInterleaved, p.ntx, p.nty)), dim3(64), 0, s,
                           padded_volume<U16Pair, P2>(vs), p, out, alpha);
    else
        hipLaunchKernelGGL((k_max_project_staged<U16Pair, P2, kQuadSlab19, CH, false>), grid, dim3(256), 0,
                           s, padded_volume<U16Pair, P2>(vs), p, out, alpha);
}

}  // namespace

#if SPR_PART != 1
template <typename T, bool P2>
hipError_t TypedLaunch<T, P2>::max_project(const LaunchTuning& t, const VolSlot& vs, RayParams p,
                                           bool nearest, float* out, float* alpha, hipStream_t s) {
    const dim3 grid(tile_grid(p.tile_map, p.ntx, p.nty)), block(256);
    const bool opac = p.alpha_pow != 0.0f;
    const bool staged = !nearest && !opac && t.slab_kb > 0;
    // the opacity mode: per-wave staging (the slab_kb 19 slabs, volumes above 4 GiB too)
    const bool staged_opac = !nearest && opac && t.slab_kb > 0 && t.staging == 0;
    const int nmax = std::max(vs.nx, std::max(vs.ny, vs.nz));
    // skipping by brick bounds: per-wave slabs of the default size, a volume with bounds
    const bool skip = t.skip != 0 && vs.d_bricks != nullptr && t.staging == 0 &&
                      (t.slab_kb == 19 || vs.pad_wide);
    // the plain max projection's box lengths (chunk 0: from the ray step)
    if (t.chunk == 0 && !opac) p.box16 = staged_groups_mip(p.nouter, nmax);
    // the nearest filter (plain or opacity mode) staged: padded slots below 4 GiB, per-wave
    // 19 KiB slabs, one wave per workgroup
    if (nearest && t.slab_kb > 0 && t.staging == 0 && vs.layout == SPR_LAYOUT_PADDED) {
        const PaddedVolume<T, P2> v = padded_volume<T, P2>(vs);
        const int ch = chunk_for(t, p.nouter, nmax, !opac);
        dim3 g(4 * tile_grid(kTileMapInterleaved, p.ntx, p.nty));
        void* obuf = nullptr;
        hipError_t e = order_tiles(t.tile_map, true, v, kWaveSlab19, p, s, &obuf, &g);
        if (e == hipSuccess) {
            launch_wave_nearest<T, P2>(opac, t.skip != 0 && vs.d_bricks != nullptr, ch, g, s, v,
                                       p, out, alpha);
            e = hipGetLastError();
        }
        if (obuf) (void)hipFreeAsync(obuf, s);
        return e;
    }
    if ((staged || staged_opac) && vs.pad_wide) {
        const bool one_wave = one_wave_for(t, p);
        const int ch = chunk_for(t, p.nouter, nmax, !opac);
        if (staged)
            skip ? launch_staged_wide<T, P2, false, true>(vs, ch, one_wave, grid, s, p, out, alpha)
                 : launch_staged_wide<T, P2, false, false>(vs, ch, one_wave, grid, s, p, out, alpha);
        else
            skip ? launch_staged_wide<T, P2, true, true>(vs, ch, one_wave, grid, s, p, out, alpha)
                 : launch_staged_wide<T, P2, true, false>(vs, ch, one_wave, grid, s, p, out, alpha);
        return hipGetLastError();
    }
    if ((staged || staged_opac) && vs.layout == SPR_LAYOUT_PADDED) {
        const PaddedVolume<T, P2> v = padded_volume<T, P2>(vs);
        const int ch = chunk_for(t, p.nouter, nmax, !opac);
        // one wave per workgroup: per-wave slabs of 16 or 19 KiB
        const bool one_wave = one_wave_for(t, p) && (staged_opac || (t.staging == 0 && t.slab_kb >= 16));
        // (one wave: tile rows interleaved over the XCDs unless cost-ordered)
        dim3 g = one_wave ? dim3(4 * tile_grid(kTileMapInterleaved, p.ntx, p.nty)) : grid;
        void* obuf = nullptr;
        hipError_t e = hipSuccess;
        if (staged_opac || t.staging == 0)  // per-wave slabs
            e = order_tiles(t.tile_map, one_wave, v,
                            one_wave ? (staged_opac || t.slab_kb != 16 ? kWaveSlab19 : 16384u)
                                     : staged_opac ? kQuadSlab19 : slab_bytes_of(t.slab_kb),
                            p, s, &obuf, &g);
        if (e == hipSuccess) {
            if (one_wave && staged_opac)
                launch_wave<T, P2, true>(19, skip, fast_norm_ok(p.min_val, p.max_val), ch, g, s,
                                         v, p, out, alpha);
            else if (one_wave)
                launch_wave<T, P2, false>(t.slab_kb, skip, false, ch, g, s, v, p, out, alpha);
            else if (staged_opac)
                skip ? launch_staged_opac<T, P2, true>(ch, g, s, v, p, out, alpha)
                     : launch_staged_opac<T, P2, false>(ch, g, s, v, p, out, alpha);
            else
                launch_staged<T, P2>(t.slab_kb, t.staging == 1, skip, ch, g, s, v, p, out, alpha);
            e = hipGetLastError();
        }
        if (obuf) (void)hipFreeAsync(obuf, s);
        return e;
    }
    return plain_max_project(vs, p, nearest, out, alpha, s);
}

template <typename T, bool P2>
hipError_t TypedLaunch<T, P2>::iso_march(const LaunchTuning& t, const VolSlot& vs, RayParams p,
                                         bool nearest, float* out, float* alpha, float* depth,
                                         float* normals, hipStream_t s) {
    const dim3 grid(tile_grid(p.tile_map, p.ntx, p.nty)), block(256);
    // the march staged in LDS: linear filter, per-wave staging, padded or wide
    const bool staged_iso = !nearest && t.slab_kb > 0 && t.staging == 0;
    const int nmax = std::max(vs.nx, std::max(vs.ny, vs.nz));
    const int ch = chunk_for(t, p.max_steps / kLoopUnroll, nmax);
    const bool skip = t.skip != 0 && vs.d_bricks != nullptr;
    p.skip16 = iso_skip_chunks(ch, p.max_steps, nmax);
    // one wave per workgroup as for the max projection (interleaved tile maps)
    const bool one_wave = one_wave_for(t, p) && p.tile_map == kTileMapInterleaved;
    if (staged_iso && vs.pad_wide) {
        skip ? launch_iso_wide<T, P2, true>(vs, ch, one_wave, grid, s, p, out, alpha, depth, normals)
             : launch_iso_wide<T, P2, false>(vs, ch, one_wave, grid, s, p, out, alpha, depth, normals);
        return hipGetLastError();
    }
    if (staged_iso && vs.layout == SPR_LAYOUT_PADDED) {
        const PaddedVolume<T, P2> v = padded_volume<T, P2>(vs);
        skip ? launch_iso_staged<T, P2, true>(ch, one_wave, grid, s, p, v, p, out, alpha, depth, normals)
             : launch_iso_staged<T, P2, false>(ch, one_wave, grid, s, p, v, p, out, alpha, depth, normals);
        return hipGetLastError();
    }
    // the nearest filter's march staged the same way (padded slots below 4 GiB)
    if (nearest && t.slab_kb > 0 && t.staging == 0 && vs.layout == SPR_LAYOUT_PADDED) {
        const PaddedVolume<T, P2> v = padded_volume<T, P2>(vs);
        skip ? launch_iso_staged<T, P2, true, 1>(ch, one_wave, grid, s, p, v, p, out, alpha, depth, normals)
             : launch_iso_staged<T, P2, false, 1>(ch, one_wave, grid, s, p, v, p, out, alpha, depth, normals);
        return hipGetLastError();
    }
    return plain_iso_march(vs, p, nearest, out, alpha, depth, normals, s);
}

#endif  // SPR_PART != 1

#if SPR_PART != 0
// the plain (one ray per thread, global gathers) kernels: latency-bound gathers, which
// the default scheduler serves better (the max-ILP build made the nearest one 23% slower)
template <typename T, bool P2>
hipError_t TypedLaunch<T, P2>::plain_max_project(const VolSlot& vs, const RayParams& p,
                                                 bool nearest, float* out, float* alpha,
                                                 hipStream_t s) {
    const dim3 grid(tile_grid(p.tile_map, p.ntx, p.nty)), block(256);
    return visit_layout<T, P2>(vs, [&](auto v) {
        using V = decltype(v);
        if (nearest)
            hipLaunchKernelGGL((k_max_project<V, 1>), grid, block, 0, s, v, p, out, alpha);
        else
            hipLaunchKernelGGL((k_max_project<V, 0>), grid, block, 0, s, v, p, out, alpha);
        return hipGetLastError();
    });
}

template <typename T, bool P2>
hipError_t TypedLaunch<T, P2>::plain_iso_march(const VolSlot& vs, const RayParams& p,
                                               bool nearest, float* out, float* alpha,
                                               float* depth, float* normals, hipStream_t s) {
    const dim3 grid(tile_grid(p.tile_map, p.ntx, p.nty)), block(256);
    return visit_layout<T, P2>(vs, [&](auto v) {
        using V = decltype(v);
        if (nearest)
            hipLaunchKernelGGL((k_iso_surface<V, 1>), grid, block, 0, s, v, p, out, alpha, depth,
                               normals);
        else
            hipLaunchKernelGGL((k_iso_surface<V, 0>), grid, block, 0, s, v, p, out, alpha, depth,
                               normals);
        return hipGetLastError();
    });
}

template <typename T, bool P2>
hipError_t TypedLaunch<T, P2>::batch(const LaunchTuning& t, int mode, int ch, const BatchJob* d_jobs,
                                     const RayParams& p, const BatchQueues& q, uint32_t* counters,
                                     hipStream_t s) {
    const int cus = t.num_cus;
    if (mode == 1) return launch_batch_kernel<T, P2, 16, 1, 16, false>(cus, d_jobs, p, q, counters, s);
    if (mode == 2) return launch_batch_kernel<T, P2, 16, 2, 16, false>(cus, d_jobs, p, q, counters, s);
    if (mode == 3) {
        if (ch == 16) return launch_batch_kernel<T, P2, kQuadSlab19, 3, 16, false>(cus, d_jobs, p, q, counters, s);
        if (ch == 8) return launch_batch_kernel<T, P2, kQuadSlab19, 3, 8, false>(cus, d_jobs, p, q, counters, s);
        return launch_batch_kernel<T, P2, kQuadSlab19, 3, 4, false>(cus, d_jobs, p, q, counters, s);
    }
    const bool sh = t.staging == 1;
    if (t.slab_kb == 8) return launch_batch_slab<T, P2, 8192>(sh, ch, cus, d_jobs, p, q, counters, s);
    if (t.slab_kb == 16 || (sh && t.slab_kb == 19))
        return launch_batch_slab<T, P2, 16384>(sh, ch, cus, d_jobs, p, q, counters, s);
    if (t.slab_kb == 19)
        return launch_batch_staged<T, P2, kQuadSlab19, false>(ch, cus, d_jobs, p, q, counters, s);
    return launch_batch_slab<T, P2, 12288>(sh, ch, cus, d_jobs, p, q, counters, s);
}

#endif  // SPR_PART != 0

#if SPR_PART != 1
template <typename T, bool P2>
hipError_t TypedLaunch<T, P2>::split(int seg, bool staged, const BatchJob* d_jobs, uint32_t njobs,
                                     uint32_t nitems, uint32_t npix, const RayParams& base,
                                     const uint32_t* gates, hipStream_t s) {
    constexpr bool kShort = !std::is_same<T, float>::value;
    const dim3 pg((npix + 255u) / 256u), block(256);
    hipLaunchKernelGGL(k_split_clear, pg, block, 0, s, d_jobs, njobs, npix, gates);
    if (staged) {  // one wave (work item) per workgroup
        const dim3 ig(nitems), wave(64);
        if (seg == 32)
            hipLaunchKernelGGL((k_split_staged<T, P2, 32>), ig, wave, 0, s, d_jobs, njobs, nitems, base, gates);
        else if (seg == 128)
            hipLaunchKernelGGL((k_split_staged<T, P2, 128>), ig, wave, 0, s, d_jobs, njobs, nitems, base, gates);
        else
            hipLaunchKernelGGL((k_split_staged<T, P2, 64>), ig, wave, 0, s, d_jobs, njobs, nitems, base, gates);
    } else {  // four work items per workgroup
        const dim3 ig((nitems + 3u) / 4u);
        if (seg == 32)
            hipLaunchKernelGGL((k_split_march<T, P2, 32>), ig, block, 0, s, d_jobs, njobs, nitems, base, gates);
        else if (seg == 128)
            hipLaunchKernelGGL((k_split_march<T, P2, 128>), ig, block, 0, s, d_jobs, njobs, nitems, base, gates);
        else
            hipLaunchKernelGGL((k_split_march<T, P2, 64>), ig, block, 0, s, d_jobs, njobs, nitems, base, gates);
    }
    hipLaunchKernelGGL(k_split_finish<kShort>, pg, block, 0, s, d_jobs, njobs, npix, base, gates);
    return hipGetLastError();
}

template <typename T, bool P2>
hipError_t TypedLaunch<T, P2>::coarse_pair(const VolSlot& a, const VolSlot& b, const RayParams& p,
                                           float* out, float* alpha, hipStream_t s) {
    const dim3 grid((unsigned)(((p.W + 7) / 8) * ((p.H + 7) / 8)), 2), block(256);
    hipLaunchKernelGGL((k_coarse_mip_pair<T, P2, 12288, 4>), grid, block, 0, s,
                       padded_volume<T, P2>(a), padded_volume<T, P2>(b), p, out, alpha);
    return hipGetLastError();
}

#endif  // SPR_PART != 1

#if SPR_VOXEL == 0
template struct TypedLaunch<float, (SPR_POW2 != 0)>;
#elif SPR_VOXEL == 1
template struct TypedLaunch<uint16_t, (SPR_POW2 != 0)>;

#if SPR_PART != 1
template <bool P2>
hipError_t launch_pair(const LaunchTuning& t, const VolSlot& vs, RayParams p, float* out,
                       float* alpha, hipStream_t s) {
    const dim3 grid(tile_grid(p.tile_map, p.ntx, p.nty));
    const int nmax = std::max(vs.nx, std::max(vs.ny, vs.nz));
    const int ch = chunk_for(t, p.nouter, nmax);
    // (the one-wave kernel deals tiles interleaved; a cost-ordered map falls back to four)
    const bool one_wave = one_wave_for(t, p) && p.tile_map != kTileMapCostXcd;
    if (ch == 16)
        launch_pair_ch<P2, 16>(vs, one_wave, grid, s, p, out, alpha);
    else if (ch == 8)
        launch_pair_ch<P2, 8>(vs, one_wave, grid, s, p, out, alpha);
    else
        launch_pair_ch<P2, 4>(vs, one_wave, grid, s, p, out, alpha);
    return hipGetLastError();
}
template hipError_t launch_pair<(SPR_POW2 != 0)>(const LaunchTuning&, const VolSlot&, RayParams,
                                                 float*, float*, hipStream_t);
#endif  // SPR_PART != 1
#elif SPR_VOXEL == 2
template struct TypedLaunch<uint8_t, (SPR_POW2 != 0)>;
#else
#error "SPR_VOXEL must be 0, 1 or 2"
#endif

}  // namespace spr
