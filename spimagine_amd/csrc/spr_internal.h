// spr_internal.h -- what the translation units of libspimrender.so share: the resident
// volume record, the launch tuning state, the volume accessors built from a slot, and the
// per-voxel-type launchers.
//
// The kernels are templates over the voxel type and the padded layout's power-of-two rule;
// their launchers (launch_typed.hip) are compiled once per (voxel type, power-of-two) pair
// so the six device compilations run in parallel, and spimrender.hip (the C ABI) calls
// them through TypedLaunch<T, P2>.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "mip_staged.h"
#include "render_kernels.h"
#include "split_rays.h"
#include "spimrender.h"

namespace spr {

// one resident volume
struct VolSlot {
    bool used = false;
    int layout = SPR_LAYOUT_AUTO;  // SPR_LAYOUT_LINEAR or SPR_LAYOUT_PADDED once filled
    void* d_vol = nullptr;         // LINEAR: the volume as uploaded
    size_t vol_bytes = 0;
    void* d_pad = nullptr;         // PADDED: (nz+2, ny+2, pitch) apron copy
    size_t pad_bytes = 0;
    bool pad_wide = false;         // LINEAR volume above 4 GiB with a padded copy (d_pad)
    int dtype = -1, nx = 0, ny = 0, nz = 0;
    uint32_t generation = 0;       // bumped whenever the slot is filled or freed
    // per-brick bounds of the padded copy (brick_bounds.h; nullptr: none) and the
    // volume's minimum and maximum voxel
    BrickKey* d_bricks = nullptr;
    uint32_t nbx = 0, nby = 0, nbz = 0;
    float vmin = 0.0f, vmax = 0.0f;
};

// spr_set_tuning state (include/spimrender.h SPR_TUNE_*) plus the device's CU count
struct LaunchTuning {
    int tile_map = (int)kTileMapInterleaved;
    int slab_kb = 19;  // per-wave LDS slab of the staged max projection, 0 = off
    int staging = 0;   // 0 per-wave boxes, 1 one box per workgroup
    int chunk = 0;     // samples per staged box, 0 = from the ray step (32 when it allows 16)
    int row_pad = 0;   // pad staged slab rows to an odd number of 16-byte units
    int wg_waves = 0;  // per-wave staged kernels: waves per workgroup (1, 4; 0 = default, 1)
    int skip = 1;      // skip boxes / stop rays by the volume's brick bounds (SPR_TUNE_SKIP)
    int num_cus = 0;
};

inline bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }
inline bool slot_pow2(const VolSlot& v) { return is_pow2(v.nx) && is_pow2(v.ny) && is_pow2(v.nz); }

template <typename T>
LinearVolume<T> linear_volume(const VolSlot& s) {
    LinearVolume<T> v;
    v.data = (const T*)s.d_vol;
    v.nx = s.nx;
    v.ny = s.ny;
    v.nz = s.nz;
    v.fnx = (float)s.nx;
    v.fny = (float)s.ny;
    v.fnz = (float)s.nz;
    return v;
}

template <typename T, bool POW2, bool WIDE = false>
PaddedVolume<T, POW2, WIDE> padded_volume(const VolSlot& s) {
    PaddedVolume<T, POW2, WIDE> v;
    v.base = s.d_pad;
    v.nbytes = WIDE ? 0xFFFFFFFFu : (uint32_t)s.pad_bytes;
    v.total_bytes = s.pad_bytes;
    const uint32_t PY = (uint32_t)s.ny + 2;
    v.px_bytes = padded_pitch<T>(s.nx) * (uint32_t)sizeof(T);
    v.py = PY;
    v.soff_y = v.px_bytes;
    v.soff_z = v.px_bytes * PY;
    v.soff_yz = v.soff_y + v.soff_z;
    v.nx = s.nx;
    v.ny = s.ny;
    v.nz = s.nz;
    v.fnx = (float)s.nx;
    v.fny = (float)s.ny;
    v.fnz = (float)s.nz;
    v.mx = (float)(s.nx - 1);
    v.my = (float)(s.ny - 1);
    v.mz = (float)(s.nz - 1);
    v.bricks.keys = s.d_bricks;
    v.bricks.nbx = s.nbx;
    v.bricks.nby = s.nby;
    v.vmin = s.vmin;
    v.vmax = s.vmax;
    return v;
}

// samples per unrolled group of the staged kernels: the tuning value (32 and 64 are boxes
// of 2 and 4 groups of 16), else from the ray step (mip: the plain max projection's rule)
inline int chunk_for(const LaunchTuning& t, int nouter, int nmax, bool mip = false) {
    return t.chunk ? std::min(t.chunk, 16)
                   : mip ? staged_chunk_mip(nouter, nmax) : staged_chunk(nouter, nmax);
}

// the staged kernels' launch options carried in RayParams
inline void staged_tuning(const LaunchTuning& t, RayParams& p) {
    p.row_pad = (uint32_t)t.row_pad;
    // chunk 0 (from the ray step): where the step allows 16-sample groups, boxes of two
    // (a box that does not fit the slab falls back to one group at a time)
    p.box16 = t.chunk > 16 ? (uint32_t)(t.chunk / 16) : t.chunk == 0 ? 2u : 1u;
}

// Launchers of one voxel type T (float, uint16_t, uint8_t) and one power-of-two rule P2
// (every dimension of the volume a power of two: the texel rule of PaddedVolume); defined
// in launch_typed.hip, one compilation per (T, P2).  Linear-layout slots use P2 = false.
template <typename T, bool P2>
struct TypedLaunch {
    // max_project (plain or opacity mode) of one frame; p.nouter / p.current_part set
    static hipError_t max_project(const LaunchTuning& t, const VolSlot& vs, RayParams p,
                                  bool nearest, float* out, float* alpha, hipStream_t s);
    // the iso march (outputs before the 2D post-passes)
    static hipError_t iso_march(const LaunchTuning& t, const VolSlot& vs, RayParams p, bool nearest,
                                float* out, float* alpha, float* depth, float* normals,
                                hipStream_t s);
    // the plain kernels (one ray per thread, every sample gathered from global memory) of
    // the slot's layout: linear-layout slots, staging switched off (slab_kb 0)
    static hipError_t plain_max_project(const VolSlot& vs, const RayParams& p, bool nearest,
                                        float* out, float* alpha, hipStream_t s);
    static hipError_t plain_iso_march(const VolSlot& vs, const RayParams& p, bool nearest,
                                      float* out, float* alpha, float* depth, float* normals,
                                      hipStream_t s);
    // the persistent batch kernel (mode: see dispatch in spimrender.hip)
    static hipError_t batch(const LaunchTuning& t, int mode, int ch, const BatchJob* d_jobs,
                            const RayParams& p, const BatchQueues& q, uint32_t* counters,
                            hipStream_t s);
    // max projections of frames of scattered rays, each ray's samples split into segments
    // of seg (32, 64 or 128) samples (split_rays.h), staged in LDS slabs or gathered from
    // global memory; job k's first item and first pixel in first[0] and first[1]; base: the
    // launch's staging settings; gates: nullptr or one device flag per job (0: skip the job)
    static hipError_t split(int seg, bool staged, const BatchJob* d_jobs, uint32_t njobs,
                            uint32_t nitems, uint32_t npix, const RayParams& base,
                            const uint32_t* gates, hipStream_t s);
    // coarse max projections of two padded slots of this type and shape
    static hipError_t coarse_pair(const VolSlot& a, const VolSlot& b, const RayParams& p,
                                  float* out, float* alpha, hipStream_t s);
};

// both channels of a padded channel-pair (SPR_DTYPE_U16X2) slot (launch_typed.hip, the
// uint16 compilations)
template <bool P2>
hipError_t launch_pair(const LaunchTuning& t, const VolSlot& vs, RayParams p, float* out,
                       float* alpha, hipStream_t s);

}  // namespace spr
