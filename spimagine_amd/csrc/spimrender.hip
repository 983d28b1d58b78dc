// spimrender.hip -- libspimrender.so: C ABI (include/spimrender.h) over the gfx950
// ray-casting kernels.  Host side of spimagine's VolumeRenderer
// (spimagine/volumerender/volumerender.py): volume upload and layout, parameter state,
// kernel dispatch in the reference's launch order, the device-to-host output copies, and
// the batched (many volumes x poses, HIP graph) launch.  The ray-march kernels are
// launched through TypedLaunch (launch_typed.hip, one compilation per voxel type and
// power-of-two rule).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "spr_internal.h"

using namespace spr;

constexpr size_t kMaxSlots = 4096;

// Device scratch of spr_render_batch_noalloc: the persistent launch's eight queue counters
// (the first 256 bytes) and a job table of up to SPR_NOALLOC_MAX_JOBS jobs.  Each device
// holds its own copy until the process ends.  Host code never passes this name to a kernel
// or a copy: it uses the device address hipGetSymbolAddress returns for each device
// (noalloc_scratch_address, called at spr_create with that device current).
constexpr size_t kScratchCounterBytes = 256;
constexpr size_t kScratchBytes =
    kScratchCounterBytes + (size_t)SPR_NOALLOC_MAX_JOBS * sizeof(BatchJob);
static_assert(sizeof(BatchJob) == 256, "the header states the scratch size per job");
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// scratch's device address per device, and the host staging table of the no-allocation
// batch (one call at a time, across every context and device)
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_noalloc_mu;
static BatchJob g_noalloc_jobs[SPR_NOALLOC_MAX_JOBS];

struct spr_batch_graph {
    int n = 0, w = 0, h = 0;
    BatchJob* d_jobs = nullptr;
    uint32_t* d_counters = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t last_stream = nullptr;
    int dtype = -1;
    bool pow2 = false;  // the captured kernel's texel rule (every volume a power of two)
    // (slot, generation) of every job: a launch after any of these slots was refilled or
    // released is refused
    std::vector<std::pair<int, uint32_t>> refs;
};

struct spr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex table_mu;  // serialises spr_volume_add / spr_volume_release
    int layout_req = SPR_LAYOUT_AUTO;  // layout for the next upload
    // slot 0: the renderer's volume.  Capacity is reserved once (kMaxSlots) so adding a
    // slot never moves the others while renders on other threads read them.
    std::vector<VolSlot> vols;
    std::atomic<uint32_t> next_generation{1};
    spr_params params{};
    bool has_params = false;
    // w x h output buffers kept between spr_render calls (VolumeRenderer.buf*, reset_buffer)
    int w = 0, h = 0;
    float *d_out = nullptr, *d_alpha = nullptr, *d_depth = nullptr, *d_normals = nullptr,
          *d_occ = nullptr;
    // iso post-pass scratch (buf_tmp, buf_tmp_vec), grown on demand
    float *d_tmp = nullptr, *d_tmp_vec = nullptr;
    size_t scratch_px = 0;
    unsigned long long* d_count = nullptr;
    float* d_partial = nullptr;
    LaunchTuning tune;  // spr_set_tuning state and the device's CU count
    // pinned staging of batch job tables: a ring of host buffers, each reused once the
    // copy that last read it has run (stage_jobs)
    std::mutex stage_mu;
    struct Stage {
        void* host = nullptr;
        size_t bytes = 0;
        hipEvent_t done = nullptr;
    } stage[8];
    int stage_next = 0;
    void* d_scratch = nullptr;  // the device's scratch (spr_render_batch_noalloc)
};

namespace {

// last failure of the calling thread and the context it concerned (NULL: spr_create or a
// NULL-context call); thread-local so concurrent callers of one context never share it
thread_local std::string g_err;
thread_local const spr_ctx* g_err_ctx = nullptr;

int fail(spr_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    g_err_ctx = ctx;
    return code;
}

#define HIPCALL(ctx, expr)                                                                    \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            return fail((ctx), e_ == hipErrorOutOfMemory ? SPR_ERR_OUT_OF_MEMORY : SPR_ERR_HIP, \
                        "%s failed: %s", #expr, hipGetErrorString(e_));                     \
        }                                                                                     \
    } while (0)

size_t dtype_size(int dtype) {
    switch (dtype) {
        case SPR_DTYPE_F32: return 4;
        case SPR_DTYPE_U16: return 2;
        case SPR_DTYPE_U8: return 1;
        case SPR_DTYPE_U16X2: return 4;
        default: return 0;
    }
}

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

void free_outputs(spr_ctx* c) {
    float** bufs[] = {&c->d_out, &c->d_alpha, &c->d_depth, &c->d_normals, &c->d_occ};
    for (float** b : bufs) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
    }
    c->w = c->h = 0;
}

void free_scratch(spr_ctx* c) {
    if (c->d_tmp) (void)hipFree(c->d_tmp);
    if (c->d_tmp_vec) (void)hipFree(c->d_tmp_vec);
    c->d_tmp = c->d_tmp_vec = nullptr;
    c->scratch_px = 0;
}

void free_slot(VolSlot& v) {
    if (v.d_vol) (void)hipFree(v.d_vol);
    if (v.d_pad) (void)hipFree(v.d_pad);
    if (v.d_bricks) (void)hipFree(v.d_bricks);
    v = VolSlot();
}

// any filled slot
const VolSlot* any_slot_of(const spr_ctx* c, int slot) {
    if (slot < 0 || slot >= (int)c->vols.size() || !c->vols[slot].used) return nullptr;
    return &c->vols[slot];
}

// a filled one-channel slot (what every entry point but spr_render_pair samples)
const VolSlot* slot_of(const spr_ctx* c, int slot) {
    const VolSlot* v = any_slot_of(c, slot);
    return v && v->dtype != SPR_DTYPE_U16X2 ? v : nullptr;
}

int no_volume(spr_ctx* c, int slot) {
    const VolSlot* v = any_slot_of(c, slot);
    if (v) return fail(c, SPR_ERR_INVALID_ARG, "slot %d holds a channel pair (spr_render_pair only)", slot);
    return fail(c, SPR_ERR_NO_VOLUME, "no volume in slot %d", slot);
}

int ensure_outputs(spr_ctx* c, int w, int h) {
    if (c->w == w && c->h == h && c->d_out) return SPR_OK;
    free_outputs(c);
    const size_t px = (size_t)w * (size_t)h;
    HIPCALL(c, hipMalloc(&c->d_out, px * 4));
    HIPCALL(c, hipMalloc(&c->d_alpha, px * 4));
    HIPCALL(c, hipMalloc(&c->d_depth, px * 4));
    HIPCALL(c, hipMalloc(&c->d_normals, px * 12));
    HIPCALL(c, hipMalloc(&c->d_occ, px * 4));
    HIPCALL(c, hipMemsetAsync(c->d_out, 0, px * 4, c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_alpha, 0, px * 4, c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_depth, 0, px * 4, c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_normals, 0, px * 12, c->stream));
    HIPCALL(c, hipMemsetAsync(c->d_occ, 0, px * 4, c->stream));
    c->w = w;
    c->h = h;
    return SPR_OK;
}

int ensure_scratch(spr_ctx* c, size_t px) {
    if (c->scratch_px >= px) return SPR_OK;
    free_scratch(c);
    HIPCALL(c, hipMalloc(&c->d_tmp, px * 4));
    HIPCALL(c, hipMalloc(&c->d_tmp_vec, px * 12));
    c->scratch_px = px;
    return SPR_OK;
}

RayParams make_params(const spr_params& prm, int tile_map, int w, int h) {
    RayParams p;
    memcpy(p.invP, prm.inv_p, sizeof(p.invP));
    memcpy(p.invM, prm.inv_m, sizeof(p.invM));
    memcpy(p.box, prm.box, sizeof(p.box));
    p.min_val = prm.min_val;
    p.max_val = prm.max_val;
    p.gamma = prm.gamma;
    p.alpha_pow = prm.alpha_pow;
    p.iso_val = prm.max_val / 2.0f;
    p.max_steps = prm.max_steps;
    p.nouter = 0;
    p.current_part = 0;
    p.W = (uint32_t)w;
    p.H = (uint32_t)h;
    p.ntx = (uint32_t)((w + 15) / 16);
    p.nty = (uint32_t)((h + 15) / 16);
    p.ntiles = p.ntx * p.nty;
    // the cost-ordered map needs an order computed for the frame (order_tiles); every
    // other launch deals tiles interleaved
    p.tile_map = tile_map == (int)kTileMapCostXcd ? kTileMapInterleaved : (uint32_t)tile_map;
    p.row_pad = 0;
    p.box16 = 1;
    p.skip16 = 1;
    p.order = nullptr;
    p.gate = nullptr;
    return p;
}

RayParams make_params(const spr_ctx* c, int w, int h) { return make_params(c->params, c->tune.tile_map, w, h); }

int check_params(spr_ctx* c, const spr_params& p) {
    if (p.filter != SPR_FILTER_LINEAR && p.filter != SPR_FILTER_NEAREST)
        return fail(c, SPR_ERR_INVALID_ARG, "unknown filter %d", p.filter);
    if (p.max_steps < 1)
        return fail(c, SPR_ERR_INVALID_ARG, "max_steps must be positive (got %d)", p.max_steps);
    return SPR_OK;
}

// convolve_2d.cl weight exp((float)(-k*(ht-Nh/2.f)*(ht-Nh/2.f)/Nh/Nh)), float32 argument
ConvWeights conv_weights(int nh, float k) {
    ConvWeights cw;
    cw.nh = nh;
    for (int ht = 0; ht < 16; ++ht) {
        if (ht >= nh) {
            cw.w[ht] = 0.0f;
            continue;
        }
        float d = (float)ht - (float)nh / 2.0f;
        float a = -k * d;
        a = a * d;
        a = a / (float)nh;
        a = a / (float)nh;
        cw.w[ht] = expf(a);
    }
    return cw;
}


int check_frame(spr_ctx* c, int w, int h) {
    if (!c->has_params) return fail(c, SPR_ERR_INVALID_ARG, "spr_set_params was not called");
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768)
        return fail(c, SPR_ERR_INVALID_ARG, "invalid output size %d x %d", w, h);
    return SPR_OK;
}

int check_ready(spr_ctx* c, int w, int h) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!c->vols[0].used) return fail(c, SPR_ERR_NO_VOLUME, "no data provided, set_data(data) before");
    return check_frame(c, w, h);
}

int check_parts(spr_ctx* c, const spr_params& prm, int num_parts) {
    if (num_parts < 1 || (prm.max_steps / num_parts) / kLoopUnroll < 1)
        return fail(c, SPR_ERR_INVALID_ARG,
                    "numParts=%d leaves no 16-sample block of max_steps=%d "
                    "(the reference divides by zero here)",
                    num_parts, prm.max_steps);
    return SPR_OK;
}

// Calls f(TypedLaunch<T, P2>()) for the slot's voxel type; P2: the padded copy's texel rule
// (every dimension a power of two; linear-layout slots without a padded copy use false).
template <typename F>
hipError_t visit_launch(const VolSlot& vs, F&& f) {
    const bool p2 = (vs.layout == SPR_LAYOUT_PADDED || vs.pad_wide) && slot_pow2(vs);
    switch (vs.dtype) {
        case SPR_DTYPE_F32:
            return p2 ? f(TypedLaunch<float, true>()) : f(TypedLaunch<float, false>());
        case SPR_DTYPE_U16:
            return p2 ? f(TypedLaunch<uint16_t, true>()) : f(TypedLaunch<uint16_t, false>());
        default:
            return p2 ? f(TypedLaunch<uint8_t, true>()) : f(TypedLaunch<uint8_t, false>());
    }
}

// render(method, numParts, currentPart) into device buffers; the reference's launch
// order (volumerender.py:366-386 for max_project, :451-497 for iso_surface)
// tmp (w*h floats) and tmp_vec (3*w*h floats) are the iso post-pass scratch
// gate: max projection only, rendered only if *gate != 0 when the kernels run
int render_impl(spr_ctx* c, const VolSlot& vs, const spr_params& prm, int method, int num_parts,
                int current_part, int w, int h, float* out, float* alpha, float* depth,
                float* normals, float* occ, float* tmp, float* tmp_vec, hipStream_t s,
                const uint32_t* gate = nullptr) {
    RayParams p = make_params(prm, c->tune.tile_map, w, h);
    staged_tuning(c->tune, p);
    p.gate = gate;
    const bool nearest = prm.filter == SPR_FILTER_NEAREST;
    int rc;
    if (method == SPR_METHOD_MAX_PROJECT) {
        rc = check_parts(c, prm, num_parts);
        if (rc) return rc;
        if (!out || !alpha) return fail(c, SPR_ERR_INVALID_ARG, "out/alpha buffers are required");
        p.nouter = (prm.max_steps / num_parts) / kLoopUnroll;
        p.current_part = current_part;
        HIPCALL(c, visit_launch(vs, [&](auto L) {
            return L.max_project(c->tune, vs, p, nearest, out, alpha, s);
        }));
        return SPR_OK;
    }
    if (method == SPR_METHOD_ISO_SURFACE) {
        if (prm.max_steps < 2)
            return fail(c, SPR_ERR_INVALID_ARG, "iso_surface needs max_steps >= 2");
        if (!out || !alpha || !depth || !normals || !occ)
            return fail(c, SPR_ERR_INVALID_ARG, "iso_surface needs all five output buffers");
        if (prm.occ_n_points < 1 || prm.occ_radius < 0)
            return fail(c, SPR_ERR_INVALID_ARG, "invalid occlusion parameters");
        HIPCALL(c, visit_launch(vs, [&](auto L) {
            return L.iso_march(c->tune, vs, p, nearest, out, alpha, depth, normals, s);
        }));
        dim3 g2((w + 15) / 16, (h + 15) / 16), b2(256);
        const ConvWeights wv = conv_weights(7, 5.0f), ws = conv_weights(5, 10.0f);
        // _convolve_vec(buf_normals, 7) (volumerender.py:405-416, 470)
        hipLaunchKernelGGL((k_conv<0, 3>), g2, b2, 0, s, normals, tmp_vec, w, h, wv);
        hipLaunchKernelGGL((k_conv<1, 3>), g2, b2, 0, s, tmp_vec, normals, w, h, wv);
        // occlusion (:472-481)
        hipLaunchKernelGGL(k_occlusion, g2, b2, 0, s, occ, w, h, prm.occ_radius, prm.occ_n_points,
                           depth);
        // _convolve_scalar(buf_occlusion, 5) (:392-403, 483)
        hipLaunchKernelGGL((k_conv<0, 1>), g2, b2, 0, s, occ, tmp, w, h, ws);
        hipLaunchKernelGGL((k_conv<1, 1>), g2, b2, 0, s, tmp, occ, w, h, ws);
        // shading overwrites the output (:485-497)
        hipLaunchKernelGGL(k_shading, g2, b2, 0, s, out, p, prm.occ_strength, normals, depth, occ);
        HIPCALL(c, hipGetLastError());
        return SPR_OK;
    }
    // an unknown method is a silent no-op in the reference (volumerender.py:543-547)
    return SPR_OK;
}


// Bounds of the 8x8x8 bricks of a padded volume (brick_bounds.h): thread bx of block
// (by, bz) scans brick (bx, by, bz) over the padded voxels [8 b, 8 b + 8) clipped to the
// padded extent (PX = nx + 2, PY, PZ); total gets the minimum of every brick's keys (the
// volume's bounds), one atomic per wave.  A NaN voxel gives its brick the widest keys.
template <typename T>
__global__ __launch_bounds__(64) void k_brick_bounds(const T* __restrict__ pad, uint32_t pitch,
                                                     uint32_t PX, uint32_t PY, uint32_t PZ,
                                                     uint32_t nbx, uint32_t nby,
                                                     BrickKey* __restrict__ keys,
                                                     BrickKey* __restrict__ total) {
    const uint32_t bx = blockIdx.x * 64u + threadIdx.x, by = blockIdx.y, bz = blockIdx.z;
    uint32_t kmin = 0xFFFFFFFFu, nkmax = 0xFFFFFFFFu;
    if (bx < nbx) {
        const uint32_t x0 = bx * kBrick, x1 = min(x0 + kBrick, PX);
        const uint32_t y1 = min(by * kBrick + kBrick, PY), z1 = min(bz * kBrick + kBrick, PZ);
        for (uint32_t z = bz * kBrick; z < z1; ++z)
            for (uint32_t y = by * kBrick; y < y1; ++y) {
                const T* row = pad + ((size_t)z * PY + y) * pitch;
                for (uint32_t x = x0; x < x1; ++x) {
                    const T v = row[x];
                    if (v != v) {  // NaN (float voxels only)
                        kmin = nkmax = 0u;
                    } else {
                        const uint32_t k = order_key(v);
                        kmin = min(kmin, k);
                        nkmax = min(nkmax, ~k);
                    }
                }
            }
        keys[((size_t)bz * nby + by) * nbx + bx] = BrickKey{kmin, nkmax};
    }
    for (int off = 32; off > 0; off >>= 1) {
        kmin = min(kmin, (uint32_t)__shfl_xor((int)kmin, off));
        nkmax = min(nkmax, (uint32_t)__shfl_xor((int)nkmax, off));
    }
    if (threadIdx.x == 0) {
        atomicMin(&total->kmin, kmin);
        atomicMin(&total->nkmax, nkmax);
    }
}

template <typename T>
void launch_bricks(const VolSlot& vs, BrickKey* total, hipStream_t s) {
    const uint32_t PX = (uint32_t)vs.nx + 2u, PY = (uint32_t)vs.ny + 2u, PZ = (uint32_t)vs.nz + 2u;
    const dim3 grid((vs.nbx + 63u) / 64u, vs.nby, vs.nbz);
    hipLaunchKernelGGL(k_brick_bounds<T>, grid, dim3(64), 0, s, (const T*)vs.d_pad,
                       padded_pitch<T>(vs.nx), PX, PY, PZ, vs.nbx, vs.nby, vs.d_bricks, total);
}

template <typename T>
void launch_pad(const void* src, void* dst, int nx, int ny, int nz, hipStream_t s) {
    dim3 grid((padded_pitch<T>(nx) + 255) / 256, ny + 2, nz + 2);
    hipLaunchKernelGGL(k_pad_volume<T>, grid, dim3(256), 0, s, (const T*)src, (T*)dst, nx, ny, nz);
}

// out(x, y, z) = max of in over [c - r, c + r] along AXIS (0 x, 1 y, 2 z), coordinates
// clamped to the volume; in is addressed in[((z + o) py + y + o) px + x + o] (o = 1, the
// padded layout's apron; o = 0, a plain (nz, ny, nx) array); out is a plain array
template <typename T, int AXIS>
__global__ __launch_bounds__(256) void k_max_filter(const T* __restrict__ in, T* __restrict__ out,
                                                    int nx, int ny, int nz, int px, int py, int o,
                                                    int r) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= nx) return;
    T m = in[((size_t)(z + o) * py + y + o) * px + x + o];
    for (int k = -r; k <= r; ++k) {
        int xx = x, yy = y, zz = z;
        if (AXIS == 0) xx = min(max(x + k, 0), nx - 1);
        if (AXIS == 1) yy = min(max(y + k, 0), ny - 1);
        if (AXIS == 2) zz = min(max(z + k, 0), nz - 1);
        const T v = in[((size_t)(zz + o) * py + yy + o) * px + xx + o];
        m = v > m ? v : m;
    }
    out[((size_t)z * ny + y) * nx + x] = m;
}

// out (ceil(nz/2), ceil(ny/2), ceil(nx/2)) = mean (MODE 0, rounded to T for integer
// types) or max (MODE 1) of each 2x2x2 block of in (coordinates clamped to the volume);
// in addressed as in k_max_filter
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_pool2_volume(const T* __restrict__ in, T* __restrict__ out,
                                                      int nx, int ny, int nz, int px, int py, int o) {
    const int ox = (nx + 1) / 2, oy = (ny + 1) / 2;
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= ox) return;
    float acc = MODE == 0 ? 0.0f : -INFINITY;
    for (int k = 0; k < 8; ++k) {
        const int xx = min(2 * x + (k & 1), nx - 1), yy = min(2 * y + ((k >> 1) & 1), ny - 1),
                  zz = min(2 * z + (k >> 2), nz - 1);
        const float v = (float)in[((size_t)(zz + o) * py + yy + o) * px + xx + o];
        acc = MODE == 0 ? acc + v : fmaxf(acc, v);
    }
    if (MODE == 0) acc *= 0.125f;
    T r;
    if constexpr (std::is_same<T, float>::value)
        r = acc;
    else
        r = (T)__float2uint_rn(acc);
    out[((size_t)z * oy + y) * ox + x] = r;
}

template <typename T>
void launch_pool2(const VolSlot& vs, void* dst, int mode, hipStream_t s) {
    const dim3 grid(((vs.nx + 1) / 2 + 255) / 256, (vs.ny + 1) / 2, (vs.nz + 1) / 2), block(256);
    const bool pad = vs.layout == SPR_LAYOUT_PADDED;
    const T* src = (const T*)(pad ? vs.d_pad : vs.d_vol);
    const int px = pad ? (int)padded_pitch<T>(vs.nx) : vs.nx, py = pad ? vs.ny + 2 : vs.ny;
    if (mode == 0)
        hipLaunchKernelGGL((k_pool2_volume<T, 0>), grid, block, 0, s, src, (T*)dst, vs.nx, vs.ny, vs.nz,
                           px, py, pad ? 1 : 0);
    else
        hipLaunchKernelGGL((k_pool2_volume<T, 1>), grid, block, 0, s, src, (T*)dst, vs.nx, vs.ny, vs.nz,
                           px, py, pad ? 1 : 0);
}

// out (ceil(nz/2), ceil(ny/2), ceil(nx/2)) pairs: channel 0 = k_pool2_volume's mean,
// channel 1 = its max, of a uint16 volume (the same arithmetic, so each channel equals the
// one-channel pooled volume)
__global__ __launch_bounds__(256) void k_pool2_pair(const uint16_t* __restrict__ in,
                                                    U16Pair* __restrict__ out, int nx, int ny,
                                                    int nz, int px, int py, int o) {
    const int ox = (nx + 1) / 2, oy = (ny + 1) / 2;
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= ox) return;
    float sum = 0.0f, mx = -INFINITY;
    for (int k = 0; k < 8; ++k) {
        const int xx = min(2 * x + (k & 1), nx - 1), yy = min(2 * y + ((k >> 1) & 1), ny - 1),
                  zz = min(2 * z + (k >> 2), nz - 1);
        const float v = (float)in[((size_t)(zz + o) * py + yy + o) * px + xx + o];
        sum = sum + v;
        mx = fmaxf(mx, v);
    }
    sum *= 0.125f;
    U16Pair r;
    r.bits = (uint32_t)__float2uint_rn(sum) | ((uint32_t)__float2uint_rn(mx) << 16);
    out[((size_t)z * oy + y) * ox + x] = r;
}

template <typename T>
void launch_max_filter3(const VolSlot& vs, void* t1, void* t2, int r, hipStream_t s) {
    const dim3 grid((vs.nx + 255) / 256, vs.ny, vs.nz), block(256);
    const bool pad = vs.layout == SPR_LAYOUT_PADDED;
    const T* src = (const T*)(pad ? vs.d_pad : vs.d_vol);
    const int px = pad ? (int)padded_pitch<T>(vs.nx) : vs.nx, py = pad ? vs.ny + 2 : vs.ny;
    hipLaunchKernelGGL((k_max_filter<T, 0>), grid, block, 0, s, src, (T*)t1, vs.nx, vs.ny, vs.nz,
                       px, py, pad ? 1 : 0, r);
    hipLaunchKernelGGL((k_max_filter<T, 1>), grid, block, 0, s, (const T*)t1, (T*)t2, vs.nx, vs.ny,
                       vs.nz, vs.nx, vs.ny, 0, r);
    hipLaunchKernelGGL((k_max_filter<T, 2>), grid, block, 0, s, (const T*)t2, (T*)t1, vs.nx, vs.ny,
                       vs.nz, vs.nx, vs.ny, 0, r);
}

// fill one slot from host or device memory in the requested layout
int fill_slot(spr_ctx* c, VolSlot& vs, const void* src, bool src_on_device, int dtype, int nx,
              int ny, int nz) {
    const size_t es = dtype_size(dtype);
    if (!es) return fail(c, SPR_ERR_UNSUPPORTED_DTYPE, "unsupported voxel type code %d", dtype);
    if (!src || nx <= 0 || ny <= 0 || nz <= 0)
        return fail(c, SPR_ERR_INVALID_ARG, "invalid volume (%d, %d, %d)", nz, ny, nx);
    if (nx > 65000 || ny > 65000 || nz > 65000)
        return fail(c, SPR_ERR_INVALID_ARG, "volume dimension above 65000");
    const size_t nvox = (size_t)nx * ny * nz;
    if (nvox > 0xFFFFFFFFull)
        return fail(c, SPR_ERR_INVALID_ARG, "volume of %zu voxels exceeds 2^32-1", nvox);
    DeviceGuard g(c->device);
    const size_t bytes = nvox * es;
    const size_t pitch = (size_t)(nx + 2 + 16 / es - 1) / (16 / es) * (16 / es);  // padded_pitch
    const size_t pbytes = pitch * (ny + 2) * (nz + 2) * es;
    // the padded layout is addressed with 32-bit buffer offsets built from 24-bit products
    const bool pad_fits = pbytes <= 0xFFFFFFFFull && (size_t)(nz + 2) * (ny + 2) < (1u << 24) &&
                          pitch * es < (1u << 24);
    // above 4 GiB (AUTO layout) the volume stays linear for every kernel and also gets a
    // padded copy that the staged max projection addresses per z-plane ("wide")
    const bool wide = c->layout_req == SPR_LAYOUT_AUTO && !pad_fits &&
                      (size_t)(nz + 2) * (ny + 2) < (1u << 24) && pitch * es < (1u << 24);
    int layout = c->layout_req;
    if (layout == SPR_LAYOUT_AUTO) layout = pad_fits ? SPR_LAYOUT_PADDED : SPR_LAYOUT_LINEAR;
    if (layout == SPR_LAYOUT_PADDED && !pad_fits)
        return fail(c, SPR_ERR_INVALID_ARG, "padded volume of %zu bytes exceeds 4 GiB", pbytes);
    const bool want_pad = layout == SPR_LAYOUT_PADDED || wide;
    HIPCALL(c, hipStreamSynchronize(c->stream));
    // a device source may still be in production on another stream (the context's stream
    // does not synchronise with the null stream): wait for all prior work on the device
    if (src_on_device) HIPCALL(c, hipDeviceSynchronize());

    // what a batch graph captured of this slot (device buffers, shape, type, layout): a
    // refill that keeps all of it keeps the slot's generation
    const VolSlot before = vs;
    bool realloc = false;
    // linear copy: kept as the resident volume (LINEAR), or staged for padding (host source)
    const bool need_linear = layout == SPR_LAYOUT_LINEAR || !src_on_device;
    if (vs.d_pad && (!want_pad || vs.pad_bytes != pbytes)) {
        realloc = true;
        HIPCALL(c, hipFree(vs.d_pad));
        vs.d_pad = nullptr;
        vs.pad_bytes = 0;
    }
    if (vs.d_vol && (!need_linear || vs.vol_bytes != bytes)) {
        realloc = true;
        HIPCALL(c, hipFree(vs.d_vol));
        vs.d_vol = nullptr;
        vs.vol_bytes = 0;
    }
    vs.used = false;
    if (need_linear && !vs.d_vol) {
        realloc = true;
        HIPCALL(c, hipMalloc(&vs.d_vol, bytes));
        vs.vol_bytes = bytes;
    }
    if (need_linear)
        HIPCALL(c, hipMemcpyAsync(vs.d_vol, src, bytes,
                                  src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                  c->stream));
    if (want_pad) {
        if (!vs.d_pad) {
            realloc = true;
            HIPCALL(c, hipMalloc(&vs.d_pad, pbytes));
            vs.pad_bytes = pbytes;
        }
        const void* from = need_linear ? vs.d_vol : src;
        switch (dtype) {
            case SPR_DTYPE_F32: launch_pad<float>(from, vs.d_pad, nx, ny, nz, c->stream); break;
            case SPR_DTYPE_U16: launch_pad<uint16_t>(from, vs.d_pad, nx, ny, nz, c->stream); break;
            case SPR_DTYPE_U16X2: launch_pad<U16Pair>(from, vs.d_pad, nx, ny, nz, c->stream); break;
            default: launch_pad<uint8_t>(from, vs.d_pad, nx, ny, nz, c->stream); break;
        }
        HIPCALL(c, hipGetLastError());
    }
    // per-brick bounds of the padded copy (one-channel volumes) and the volume's bounds
    const uint32_t nbx = want_pad ? ((uint32_t)nx + 2u + kBrick - 1u) / kBrick : 0u;
    const uint32_t nby = want_pad ? ((uint32_t)ny + 2u + kBrick - 1u) / kBrick : 0u;
    const uint32_t nbz = want_pad ? ((uint32_t)nz + 2u + kBrick - 1u) / kBrick : 0u;
    const bool want_bricks = want_pad && dtype != SPR_DTYPE_U16X2;
    if (vs.d_bricks && (!want_bricks || vs.nbx != nbx || vs.nby != nby || vs.nbz != nbz)) {
        realloc = true;
        HIPCALL(c, hipFree(vs.d_bricks));
        vs.d_bricks = nullptr;
    }
    vs.nbx = nbx;
    vs.nby = nby;
    vs.nbz = nbz;
    if (want_bricks) {
        const size_t nbricks = (size_t)nbx * nby * nbz;
        if (!vs.d_bricks) {
            realloc = true;
            // one more entry: the volume's bounds
            HIPCALL(c, hipMalloc(&vs.d_bricks, (nbricks + 1) * sizeof(BrickKey)));
        }
        BrickKey* total = vs.d_bricks + nbricks;
        vs.dtype = dtype;
        vs.nx = nx;
        vs.ny = ny;
        vs.nz = nz;
        HIPCALL(c, hipMemsetAsync(total, 0xFF, sizeof(BrickKey), c->stream));
        switch (dtype) {
            case SPR_DTYPE_F32: launch_bricks<float>(vs, total, c->stream); break;
            case SPR_DTYPE_U16: launch_bricks<uint16_t>(vs, total, c->stream); break;
            default: launch_bricks<uint8_t>(vs, total, c->stream); break;
        }
        HIPCALL(c, hipGetLastError());
        BrickKey t;
        HIPCALL(c, hipMemcpyAsync(&t, total, sizeof(t), hipMemcpyDeviceToHost, c->stream));
        HIPCALL(c, hipStreamSynchronize(c->stream));
        switch (dtype) {
            case SPR_DTYPE_F32:
                vs.vmin = key_value<float>(t.kmin);
                vs.vmax = key_value<float>(~t.nkmax);
                break;
            default:
                vs.vmin = (float)t.kmin;
                vs.vmax = (float)~t.nkmax;
                break;
        }
    }
    HIPCALL(c, hipStreamSynchronize(c->stream));
    if (layout == SPR_LAYOUT_PADDED && vs.d_vol) {  // staging copy no longer needed
        HIPCALL(c, hipFree(vs.d_vol));
        vs.d_vol = nullptr;
        vs.vol_bytes = 0;
    }
    vs.used = true;
    vs.layout = layout;
    vs.pad_wide = wide;
    vs.dtype = dtype;
    vs.nx = nx;
    vs.ny = ny;
    vs.nz = nz;
    if (realloc || !before.used || before.d_pad != vs.d_pad || before.d_vol != vs.d_vol ||
        before.layout != layout || before.dtype != dtype || before.nx != nx || before.ny != ny ||
        before.nz != nz)
        vs.generation = c->next_generation++;
    return SPR_OK;
}

template <typename T>
void launch_minmax(const void* data, size_t n, float* partial, int blocks, hipStream_t s) {
    hipLaunchKernelGGL((k_minmax<T>), dim3(blocks), dim3(256), 0, s, (const T*)data, n, partial);
}

// ---------------------------------------------------------------------------------- batch

// Converts the caller's jobs into device job records; checks that every slot is padded and
// of one voxel type.  *pow2 = every volume's dimensions are powers of two.
// Converts the caller's jobs into device job records in out[0 .. n-1]; checks that every
// slot is padded and of one voxel type.  *pow2 = every volume's dimensions are powers of
// two; *nmax = the largest volume dimension.
int build_jobs_into(spr_ctx* c, const spr_job* jobs, int n, BatchJob* out, int* dtype, bool* pow2,
                    int* nmax) {
    if (!jobs || n <= 0) return fail(c, SPR_ERR_INVALID_ARG, "empty batch");
    *dtype = -1;
    *pow2 = true;
    *nmax = 1;
    for (int i = 0; i < n; ++i) {
        const VolSlot* vs = slot_of(c, jobs[i].slot);
        if (!vs) {
            if (any_slot_of(c, jobs[i].slot))
                return fail(c, SPR_ERR_INVALID_ARG, "job %d: slot %d holds a channel pair", i,
                            jobs[i].slot);
            return fail(c, SPR_ERR_NO_VOLUME, "job %d: no volume in slot %d", i, jobs[i].slot);
        }
        if (vs->layout != SPR_LAYOUT_PADDED)
            return fail(c, SPR_ERR_INVALID_ARG, "job %d: batched rendering needs the padded layout", i);
        if (*dtype >= 0 && vs->dtype != *dtype)
            return fail(c, SPR_ERR_UNSUPPORTED_DTYPE, "job %d: mixed voxel types in one batch", i);
        if (!jobs[i].d_out || !jobs[i].d_alpha)
            return fail(c, SPR_ERR_INVALID_ARG, "job %d: NULL output", i);
        *dtype = vs->dtype;
        *pow2 = *pow2 && slot_pow2(*vs);
        *nmax = std::max(*nmax, std::max(vs->nx, std::max(vs->ny, vs->nz)));
        BatchJob& b = out[i];
        b = BatchJob{};
        b.base = vs->d_pad;
        b.nbytes = (uint32_t)vs->pad_bytes;
        b.px_bytes = (uint32_t)(((size_t)vs->nx + 2 + 16 / dtype_size(vs->dtype) - 1) /
                                (16 / dtype_size(vs->dtype)) * 16);
        b.py = (uint32_t)vs->ny + 2;
        b.nx = vs->nx;
        b.ny = vs->ny;
        b.nz = vs->nz;
        memcpy(b.invP, jobs[i].inv_p, sizeof(b.invP));
        memcpy(b.invM, jobs[i].inv_m, sizeof(b.invM));
        b.out = jobs[i].d_out;
        b.alpha = jobs[i].d_alpha;
    }
    return SPR_OK;
}

int build_jobs(spr_ctx* c, const spr_job* jobs, int n, std::vector<BatchJob>& out, int* dtype,
               bool* pow2, int* nmax) {
    if (!jobs || n <= 0) return fail(c, SPR_ERR_INVALID_ARG, "empty batch");
    out.resize(n);
    return build_jobs_into(c, jobs, n, out.data(), dtype, pow2, nmax);
}

// tile rows of a frame of nty tile rows that fall in queue k (rows k, k + 8, ...)
uint32_t queue_rows(uint32_t nty, uint32_t k) { return k < nty ? (nty - k + 7u) / 8u : 0u; }

BatchQueues make_queues(const RayParams& p, int n) {
    BatchQueues q;
    for (uint32_t k = 0; k < 8; ++k) {
        q.tiles[k] = queue_rows(p.nty, k) * p.ntx;
        q.count[k] = q.tiles[k] * (uint32_t)n;
    }
    q.mixed = 0;
    q.njobs = (uint32_t)n;
    return q;
}

// ch: samples per staged box (chunk_for of the batch's longest ray step)
hipError_t dispatch_batch(const spr_ctx* c, const spr_params& prm, int dtype, bool pow2, int ch,
                          const BatchJob* d_jobs, const RayParams& p, const BatchQueues& q,
                          uint32_t* counters, hipStream_t s) {
    const LaunchTuning& t = c->tune;
    const bool nearest = prm.filter == SPR_FILTER_NEAREST;
    // 0: staged max projection, 3: staged opacity mode (per-wave 19 KiB slabs), 1 / 2: every
    // sample from global memory (linear / nearest)
    const int mode = nearest ? 2
                     : t.slab_kb <= 0 ? 1
                     : prm.alpha_pow == 0.0f ? 0
                     : t.staging == 0 ? 3 : 1;
    switch (dtype) {
        case SPR_DTYPE_F32:
            return pow2 ? TypedLaunch<float, true>::batch(t, mode, ch, d_jobs, p, q, counters, s)
                        : TypedLaunch<float, false>::batch(t, mode, ch, d_jobs, p, q, counters, s);
        case SPR_DTYPE_U16:
            return pow2 ? TypedLaunch<uint16_t, true>::batch(t, mode, ch, d_jobs, p, q, counters, s)
                        : TypedLaunch<uint16_t, false>::batch(t, mode, ch, d_jobs, p, q, counters, s);
        default:
            return pow2 ? TypedLaunch<uint8_t, true>::batch(t, mode, ch, d_jobs, p, q, counters, s)
                        : TypedLaunch<uint8_t, false>::batch(t, mode, ch, d_jobs, p, q, counters, s);
    }
}

// Enqueues the copy of a batch's job table to d_jobs on stream s from pinned host memory
// (the ring of the context): the call returns without waiting for the copy; a ring buffer
// is refilled only after the copy that last read it has completed.
hipError_t stage_jobs(spr_ctx* c, const BatchJob* jobs, size_t bytes, BatchJob* d_jobs,
                      hipStream_t s) {
    std::lock_guard<std::mutex> lock(c->stage_mu);
    spr_ctx::Stage& st = c->stage[c->stage_next];
    c->stage_next = (c->stage_next + 1) % (int)(sizeof(c->stage) / sizeof(c->stage[0]));
    hipError_t e = hipSuccess;
    if (st.done) e = hipEventSynchronize(st.done);
    if (e == hipSuccess && st.bytes < bytes) {
        if (st.host) (void)hipHostFree(st.host);
        st.host = nullptr;
        st.bytes = 0;
        e = hipHostMalloc(&st.host, bytes, hipHostMallocDefault);
        if (e == hipSuccess) st.bytes = bytes;
    }
    if (e == hipSuccess && !st.done) e = hipEventCreateWithFlags(&st.done, hipEventDisableTiming);
    if (e != hipSuccess) return e;
    memcpy(st.host, jobs, bytes);
    e = hipMemcpyAsync(d_jobs, st.host, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventRecord(st.done, s);
    return e;
}

// the (slot, generation) pairs a graph's jobs read (spr_batch_graph_launch checks them)
void record_refs(const spr_ctx* c, const spr_job* jobs, int n,
                 std::vector<std::pair<int, uint32_t>>& refs) {
    refs.clear();
    for (int i = 0; i < n; ++i) {
        const std::pair<int, uint32_t> r(jobs[i].slot, c->vols[jobs[i].slot].generation);
        if (std::find(refs.begin(), refs.end(), r) == refs.end()) refs.push_back(r);
    }
}


// scratch's address on `device` (the current device), looked up once per device
hipError_t noalloc_scratch_address(int device, void** ptr) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" {

int spr_abi_version(void) { return SPR_ABI_VERSION; }

int spr_device_count(int* count) {
    if (!count) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL count");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return SPR_OK;
}

int spr_device_memory(int device, size_t* free_bytes, size_t* total_bytes) {
    if (!free_bytes || !total_bytes) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL output");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(nullptr, SPR_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n)
        return fail(nullptr, SPR_ERR_INVALID_ARG, "device %d out of range [0, %d)", device, n);
    DeviceGuard g(device);
    HIPCALL(nullptr, hipMemGetInfo(free_bytes, total_bytes));
    return SPR_OK;
}

int spr_create(int device, spr_ctx** out) {
    if (!out) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL output pointer");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(nullptr, SPR_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= n)
        return fail(nullptr, SPR_ERR_INVALID_ARG, "device %d out of range [0, %d)", device, n);
    if (device >= kMaxDevices)
        return fail(nullptr, SPR_ERR_INVALID_ARG, "device %d beyond the library's %d", device, kMaxDevices);
    spr_ctx* c = new (std::nothrow) spr_ctx();
    if (!c) return fail(nullptr, SPR_ERR_OUT_OF_MEMORY, "out of host memory");
    c->device = device;
    c->vols.reserve(kMaxSlots);
    c->vols.emplace_back();
    DeviceGuard g(device);
    hipError_t e = hipDeviceGetAttribute(&c->tune.num_cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->d_count, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&c->d_partial, 2 * 1024 * sizeof(float));
    if (e == hipSuccess) e = noalloc_scratch_address(device, &c->d_scratch);
    if (e != hipSuccess) {
        int rc = fail(nullptr, SPR_ERR_HIP, "context creation failed: %s", hipGetErrorString(e));
        spr_destroy(c);
        return rc;
    }
    *out = c;
    return SPR_OK;
}

int spr_destroy(spr_ctx* c) {
    if (!c) return SPR_OK;
    {
        DeviceGuard g(c->device);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        free_outputs(c);
        free_scratch(c);
        for (VolSlot& v : c->vols) free_slot(v);
        if (c->d_count) (void)hipFree(c->d_count);
        if (c->d_partial) (void)hipFree(c->d_partial);
        for (spr_ctx::Stage& st : c->stage) {
            if (st.done) (void)hipEventSynchronize(st.done);
            if (st.done) (void)hipEventDestroy(st.done);
            if (st.host) (void)hipHostFree(st.host);
        }
        if (c->stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return SPR_OK;
}

const char* spr_last_error(const spr_ctx* c) { return g_err_ctx == c ? g_err.c_str() : ""; }

int spr_set_layout(spr_ctx* c, int layout) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (layout != SPR_LAYOUT_AUTO && layout != SPR_LAYOUT_LINEAR && layout != SPR_LAYOUT_PADDED)
        return fail(c, SPR_ERR_INVALID_ARG, "unknown layout %d", layout);
    c->layout_req = layout;
    return SPR_OK;
}

int spr_get_layout(spr_ctx* c, int* layout) {
    if (!c || !layout) return fail(c, SPR_ERR_INVALID_ARG, "NULL argument");
    if (!c->vols[0].used) return fail(c, SPR_ERR_NO_VOLUME, "no volume");
    *layout = c->vols[0].layout;
    return SPR_OK;
}

int spr_set_tuning(spr_ctx* c, int key, int value) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    switch (key) {
        case SPR_TUNE_TILE_MAP:
            if (value < 0 || value > (int)kTileMapCostXcd)
                return fail(c, SPR_ERR_INVALID_ARG, "tile map %d (0 .. 3)", value);
            c->tune.tile_map = value;
            return SPR_OK;
        case SPR_TUNE_SLAB_KB:
            if (value != 0 && value != 8 && value != 12 && value != 16 && value != 19)
                return fail(c, SPR_ERR_INVALID_ARG, "slab size %d KiB (0, 8, 12, 16 or 19)", value);
            c->tune.slab_kb = value;
            return SPR_OK;
        case SPR_TUNE_STAGING:
            if (value != 0 && value != 1) return fail(c, SPR_ERR_INVALID_ARG, "staging %d (0 or 1)", value);
            c->tune.staging = value;
            return SPR_OK;
        case SPR_TUNE_CHUNK:
            if (value != 0 && value != 4 && value != 8 && value != 16 && value != 32 && value != 64)
                return fail(c, SPR_ERR_INVALID_ARG, "chunk %d (0, 4, 8, 16, 32 or 64)", value);
            c->tune.chunk = value;
            return SPR_OK;
        case SPR_TUNE_ROW_PAD:
            if (value != 0 && value != 1) return fail(c, SPR_ERR_INVALID_ARG, "row pad %d (0 or 1)", value);
            c->tune.row_pad = value;
            return SPR_OK;
        case SPR_TUNE_WG_WAVES:
            if (value != 0 && value != 1 && value != 4)
                return fail(c, SPR_ERR_INVALID_ARG, "waves per workgroup %d (0, 1 or 4)", value);
            c->tune.wg_waves = value;
            return SPR_OK;
        case SPR_TUNE_SKIP:
            if (value != 0 && value != 1) return fail(c, SPR_ERR_INVALID_ARG, "skip %d (0 or 1)", value);
            c->tune.skip = value;
            return SPR_OK;
        default: return fail(c, SPR_ERR_INVALID_ARG, "unknown tuning key %d", key);
    }
}

int spr_get_tuning(spr_ctx* c, int key, int* value) {
    if (!c || !value) return fail(c, SPR_ERR_INVALID_ARG, "NULL argument");
    switch (key) {
        case SPR_TUNE_TILE_MAP: *value = c->tune.tile_map; return SPR_OK;
        case SPR_TUNE_SLAB_KB: *value = c->tune.slab_kb; return SPR_OK;
        case SPR_TUNE_STAGING: *value = c->tune.staging; return SPR_OK;
        case SPR_TUNE_CHUNK: *value = c->tune.chunk; return SPR_OK;
        case SPR_TUNE_ROW_PAD: *value = c->tune.row_pad; return SPR_OK;
        case SPR_TUNE_WG_WAVES: *value = c->tune.wg_waves; return SPR_OK;
        case SPR_TUNE_SKIP: *value = c->tune.skip; return SPR_OK;
        default: return fail(c, SPR_ERR_INVALID_ARG, "unknown tuning key %d", key);
    }
}

int spr_set_volume(spr_ctx* c, const void* host, int dtype, int nx, int ny, int nz) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (dtype == SPR_DTYPE_U16X2)
        return fail(c, SPR_ERR_UNSUPPORTED_DTYPE, "the renderer's volume has one channel");
    return fill_slot(c, c->vols[0], host, false, dtype, nx, ny, nz);
}

int spr_set_volume_device(spr_ctx* c, const void* dev, int dtype, int nx, int ny, int nz) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (dtype == SPR_DTYPE_U16X2)
        return fail(c, SPR_ERR_UNSUPPORTED_DTYPE, "the renderer's volume has one channel");
    return fill_slot(c, c->vols[0], dev, true, dtype, nx, ny, nz);
}

int spr_volume_add(spr_ctx* c, const void* src, int src_on_device, int dtype, int nx, int ny,
                   int nz, int* slot) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!slot) return fail(c, SPR_ERR_INVALID_ARG, "NULL slot output");
    std::lock_guard<std::mutex> lock(c->table_mu);
    size_t k = 1;
    while (k < c->vols.size() && c->vols[k].used) ++k;
    if (k == c->vols.size()) {
        if (k == kMaxSlots) return fail(c, SPR_ERR_INVALID_ARG, "all %zu volume slots in use", k);
        c->vols.emplace_back();
    }
    int rc = fill_slot(c, c->vols[k], src, src_on_device != 0, dtype, nx, ny, nz);
    if (rc) {
        free_slot(c->vols[k]);
        return rc;
    }
    *slot = (int)k;
    return SPR_OK;
}

int spr_volume_add_dilated(spr_ctx* c, int src_slot, int radius, int* slot) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!slot) return fail(c, SPR_ERR_INVALID_ARG, "NULL slot output");
    const VolSlot* vs = slot_of(c, src_slot);
    if (!vs) return no_volume(c, src_slot);
    if (radius < 1 || radius > 16) return fail(c, SPR_ERR_INVALID_ARG, "radius %d not in [1, 16]", radius);
    DeviceGuard g(c->device);
    const size_t bytes = (size_t)vs->nx * vs->ny * vs->nz * dtype_size(vs->dtype);
    void *t1 = nullptr, *t2 = nullptr;
    hipError_t e = hipMalloc(&t1, bytes);
    if (e == hipSuccess) e = hipMalloc(&t2, bytes);
    if (e == hipSuccess) {
        switch (vs->dtype) {
            case SPR_DTYPE_F32: launch_max_filter3<float>(*vs, t1, t2, radius, c->stream); break;
            case SPR_DTYPE_U16: launch_max_filter3<uint16_t>(*vs, t1, t2, radius, c->stream); break;
            default: launch_max_filter3<uint8_t>(*vs, t1, t2, radius, c->stream); break;
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    int rc = SPR_OK;
    if (e != hipSuccess)
        rc = fail(c, e == hipErrorOutOfMemory ? SPR_ERR_OUT_OF_MEMORY : SPR_ERR_HIP,
                  "max filter failed: %s", hipGetErrorString(e));
    else
        rc = spr_volume_add(c, t1, 1, vs->dtype, vs->nx, vs->ny, vs->nz, slot);
    if (t1) (void)hipFree(t1);
    if (t2) (void)hipFree(t2);
    return rc;
}

int spr_volume_add_pooled(spr_ctx* c, int src_slot, int mode, int* slot) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!slot) return fail(c, SPR_ERR_INVALID_ARG, "NULL slot output");
    const VolSlot* vs = slot_of(c, src_slot);
    if (!vs) return no_volume(c, src_slot);
    if (mode != 0 && mode != 1) return fail(c, SPR_ERR_INVALID_ARG, "pooling mode %d (0 mean, 1 max)", mode);
    DeviceGuard g(c->device);
    const int ox = (vs->nx + 1) / 2, oy = (vs->ny + 1) / 2, oz = (vs->nz + 1) / 2;
    const size_t bytes = (size_t)ox * oy * oz * dtype_size(vs->dtype);
    void* t = nullptr;
    hipError_t e = hipMalloc(&t, bytes);
    if (e == hipSuccess) {
        switch (vs->dtype) {
            case SPR_DTYPE_F32: launch_pool2<float>(*vs, t, mode, c->stream); break;
            case SPR_DTYPE_U16: launch_pool2<uint16_t>(*vs, t, mode, c->stream); break;
            default: launch_pool2<uint8_t>(*vs, t, mode, c->stream); break;
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    int rc = SPR_OK;
    if (e != hipSuccess)
        rc = fail(c, e == hipErrorOutOfMemory ? SPR_ERR_OUT_OF_MEMORY : SPR_ERR_HIP, "pooling failed: %s",
                  hipGetErrorString(e));
    else
        rc = spr_volume_add(c, t, 1, vs->dtype, ox, oy, oz, slot);
    if (t) (void)hipFree(t);
    return rc;
}

int spr_volume_add_pooled_pair(spr_ctx* c, int src_slot, int* slot) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!slot) return fail(c, SPR_ERR_INVALID_ARG, "NULL slot output");
    const VolSlot* vs = slot_of(c, src_slot);
    if (!vs) return no_volume(c, src_slot);
    if (vs->dtype != SPR_DTYPE_U16)
        return fail(c, SPR_ERR_UNSUPPORTED_DTYPE, "pooled pairs need a uint16 volume");
    DeviceGuard g(c->device);
    const int ox = (vs->nx + 1) / 2, oy = (vs->ny + 1) / 2, oz = (vs->nz + 1) / 2;
    void* t = nullptr;
    hipError_t e = hipMalloc(&t, (size_t)ox * oy * oz * sizeof(U16Pair));
    if (e == hipSuccess) {
        const dim3 grid((ox + 255) / 256, oy, oz), block(256);
        const bool pad = vs->layout == SPR_LAYOUT_PADDED;
        const int px = pad ? (int)padded_pitch<uint16_t>(vs->nx) : vs->nx, py = pad ? vs->ny + 2 : vs->ny;
        hipLaunchKernelGGL(k_pool2_pair, grid, block, 0, c->stream,
                           (const uint16_t*)(pad ? vs->d_pad : vs->d_vol), (U16Pair*)t, vs->nx, vs->ny,
                           vs->nz, px, py, pad ? 1 : 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    int rc = SPR_OK;
    if (e != hipSuccess) {
        rc = fail(c, e == hipErrorOutOfMemory ? SPR_ERR_OUT_OF_MEMORY : SPR_ERR_HIP, "pooling failed: %s",
                  hipGetErrorString(e));
    } else {
        // padded whatever the layout request: spr_render_pair samples the padded layout
        const int req = c->layout_req;
        c->layout_req = SPR_LAYOUT_PADDED;
        rc = spr_volume_add(c, t, 1, SPR_DTYPE_U16X2, ox, oy, oz, slot);
        c->layout_req = req;
    }
    if (t) (void)hipFree(t);
    return rc;
}

int spr_volume_read(spr_ctx* c, int slot, void* host, size_t bytes) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    const VolSlot* vs = any_slot_of(c, slot);
    if (!vs) return fail(c, SPR_ERR_NO_VOLUME, "no volume in slot %d", slot);
    const size_t es = dtype_size(vs->dtype), row = (size_t)vs->nx * es;
    const size_t need = row * vs->ny * vs->nz;
    if (!host || bytes != need)
        return fail(c, SPR_ERR_INVALID_ARG, "host buffer of %zu bytes, the volume has %zu", bytes, need);
    DeviceGuard g(c->device);
    if (vs->layout == SPR_LAYOUT_PADDED) {
        const size_t pitch = (size_t)(vs->dtype == SPR_DTYPE_F32     ? padded_pitch<float>(vs->nx)
                                      : vs->dtype == SPR_DTYPE_U16   ? padded_pitch<uint16_t>(vs->nx)
                                      : vs->dtype == SPR_DTYPE_U16X2 ? padded_pitch<U16Pair>(vs->nx)
                                                                     : padded_pitch<uint8_t>(vs->nx)) *
                             es;
        for (int z = 0; z < vs->nz; ++z) {  // interior rows of each padded plane
            const char* src = (const char*)vs->d_pad + (((size_t)z + 1) * (vs->ny + 2) + 1) * pitch + es;
            HIPCALL(c, hipMemcpy2DAsync((char*)host + (size_t)z * vs->ny * row, row, src, pitch, row,
                                        vs->ny, hipMemcpyDeviceToHost, c->stream));
        }
    } else {
        HIPCALL(c, hipMemcpyAsync(host, vs->d_vol, need, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return SPR_OK;
}

int spr_volume_release(spr_ctx* c, int slot) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (slot <= 0 || !any_slot_of(c, slot))
        return fail(c, SPR_ERR_INVALID_ARG, "slot %d is not an added volume", slot);
    std::lock_guard<std::mutex> lock(c->table_mu);
    DeviceGuard g(c->device);
    HIPCALL(c, hipDeviceSynchronize());  // no in-flight batch may still read it
    free_slot(c->vols[slot]);
    return SPR_OK;
}

int spr_volume_info(spr_ctx* c, int slot, int* dtype, int* nx, int* ny, int* nz,
                    size_t* device_bytes) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    const VolSlot* vs = any_slot_of(c, slot);
    if (!vs) return fail(c, SPR_ERR_NO_VOLUME, "no volume in slot %d", slot);
    if (dtype) *dtype = vs->dtype;
    if (nx) *nx = vs->nx;
    if (ny) *ny = vs->ny;
    if (nz) *nz = vs->nz;
    if (device_bytes) *device_bytes = vs->vol_bytes + vs->pad_bytes;
    return SPR_OK;
}

int spr_volume_minmax(spr_ctx* c, float* mn, float* mx) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    const VolSlot& vs = c->vols[0];
    if (!vs.used) return fail(c, SPR_ERR_NO_VOLUME, "no volume");
    if (!mn || !mx) return fail(c, SPR_ERR_INVALID_ARG, "NULL output");
    DeviceGuard g(c->device);
    // the apron replicates edge voxels, so min/max over the padded copy are the volume's
    const void* data = vs.d_pad ? vs.d_pad : vs.d_vol;
    const size_t n = (vs.d_pad ? vs.pad_bytes : vs.vol_bytes) / dtype_size(vs.dtype);
    const int blocks = 1024;
    switch (vs.dtype) {
        case SPR_DTYPE_F32: launch_minmax<float>(data, n, c->d_partial, blocks, c->stream); break;
        case SPR_DTYPE_U16: launch_minmax<uint16_t>(data, n, c->d_partial, blocks, c->stream); break;
        default: launch_minmax<uint8_t>(data, n, c->d_partial, blocks, c->stream); break;
    }
    HIPCALL(c, hipGetLastError());
    std::vector<float> part(2 * blocks);
    HIPCALL(c, hipMemcpyAsync(part.data(), c->d_partial, part.size() * sizeof(float),
                              hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    float a = INFINITY, b = -INFINITY;
    for (int i = 0; i < blocks; ++i) {
        a = fminf(a, part[2 * i]);
        b = fmaxf(b, part[2 * i + 1]);
    }
    *mn = a;
    *mx = b;
    return SPR_OK;
}

int spr_set_params(spr_ctx* c, const spr_params* p) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!p) return fail(c, SPR_ERR_INVALID_ARG, "NULL params");
    int rc = check_params(c, *p);
    if (rc) return rc;
    c->params = *p;
    c->has_params = true;
    return SPR_OK;
}

int spr_render(spr_ctx* c, int method, int num_parts, int current_part, int w, int h, float* out,
               float* alpha, float* depth, float* normals, float* occlusion) {
    int rc = check_ready(c, w, h);
    if (rc) return rc;
    DeviceGuard g(c->device);
    rc = ensure_outputs(c, w, h);
    if (rc) return rc;
    if (method == SPR_METHOD_ISO_SURFACE) {
        rc = ensure_scratch(c, (size_t)w * h);
        if (rc) return rc;
    }
    rc = render_impl(c, c->vols[0], c->params, method, num_parts, current_part, w, h, c->d_out,
                     c->d_alpha, c->d_depth, c->d_normals, c->d_occ, c->d_tmp, c->d_tmp_vec,
                     c->stream);
    if (rc) return rc;
    const size_t px = (size_t)w * h;
    if (out) HIPCALL(c, hipMemcpyAsync(out, c->d_out, px * 4, hipMemcpyDeviceToHost, c->stream));
    if (alpha)
        HIPCALL(c, hipMemcpyAsync(alpha, c->d_alpha, px * 4, hipMemcpyDeviceToHost, c->stream));
    if (depth)
        HIPCALL(c, hipMemcpyAsync(depth, c->d_depth, px * 4, hipMemcpyDeviceToHost, c->stream));
    if (normals)
        HIPCALL(c, hipMemcpyAsync(normals, c->d_normals, px * 12, hipMemcpyDeviceToHost, c->stream));
    if (occlusion)
        HIPCALL(c, hipMemcpyAsync(occlusion, c->d_occ, px * 4, hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return SPR_OK;
}

int spr_render_device(spr_ctx* c, int method, int num_parts, int current_part, int w, int h,
                      float* d_out, float* d_alpha, float* d_depth, float* d_normals,
                      float* d_occlusion, void* stream) {
    int rc = check_ready(c, w, h);
    if (rc) return rc;
    DeviceGuard g(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    if (method == SPR_METHOD_ISO_SURFACE) {
        rc = ensure_scratch(c, (size_t)w * h);
        if (rc) return rc;
    }
    return render_impl(c, c->vols[0], c->params, method, num_parts, current_part, w, h, d_out,
                       d_alpha, d_depth, d_normals, d_occlusion, c->d_tmp, c->d_tmp_vec, s);
}

int spr_render_slot(spr_ctx* c, int slot, const spr_params* params, int method, int num_parts,
                    int current_part, int w, int h, float* d_out, float* d_alpha, float* d_depth,
                    float* d_normals, float* d_occlusion, void* stream) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    const VolSlot* vs = slot_of(c, slot);
    if (!vs) return no_volume(c, slot);
    if (!params) return fail(c, SPR_ERR_INVALID_ARG, "NULL params");
    int rc = check_params(c, *params);
    if (rc) return rc;
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768)
        return fail(c, SPR_ERR_INVALID_ARG, "invalid output size %d x %d", w, h);
    DeviceGuard g(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    float* scratch = nullptr;
    if (method == SPR_METHOD_ISO_SURFACE)  // per-call scratch: concurrent calls do not share it
        HIPCALL(c, hipMallocAsync((void**)&scratch, (size_t)w * h * 16, s));
    rc = render_impl(c, *vs, *params, method, num_parts, current_part, w, h, d_out, d_alpha, d_depth,
                     d_normals, d_occlusion, scratch, scratch ? scratch + (size_t)w * h : nullptr, s);
    if (scratch) (void)hipFreeAsync(scratch, s);
    return rc;
}

int spr_render_slot_if(spr_ctx* c, int slot, const spr_params* params, int w, int h,
                       float* d_out, float* d_alpha, const unsigned int* d_flag, void* stream) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    const VolSlot* vs = slot_of(c, slot);
    if (!vs) return no_volume(c, slot);
    if (!params || !d_flag) return fail(c, SPR_ERR_INVALID_ARG, "NULL params or flag");
    int rc = check_params(c, *params);
    if (rc) return rc;
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768)
        return fail(c, SPR_ERR_INVALID_ARG, "invalid output size %d x %d", w, h);
    DeviceGuard g(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    return render_impl(c, *vs, *params, SPR_METHOD_MAX_PROJECT, 1, 0, w, h, d_out, d_alpha, nullptr,
                       nullptr, nullptr, nullptr, nullptr, s, (const uint32_t*)d_flag);
}

int spr_render_batch(spr_ctx* c, const spr_params* params, const spr_job* jobs, int n, int w, int h,
                     int num_parts, int current_part, void* stream) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    const spr_params& prm = params ? *params : c->params;
    if (!params && !c->has_params) return fail(c, SPR_ERR_INVALID_ARG, "no parameters");
    int rc = check_params(c, prm);
    if (rc) return rc;
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768)
        return fail(c, SPR_ERR_INVALID_ARG, "invalid output size %d x %d", w, h);
    rc = check_parts(c, prm, num_parts);
    if (rc) return rc;
    std::vector<BatchJob> host;
    int dtype, nmax;
    bool pow2;
    rc = build_jobs(c, jobs, n, host, &dtype, &pow2, &nmax);
    if (rc) return rc;
    DeviceGuard g(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    // per-call job table and queue counters from the stream-ordered allocator, so batches
    // submitted concurrently on different streams do not share them
    const size_t jbytes = (size_t)n * sizeof(BatchJob);
    void* buf = nullptr;
    HIPCALL(c, hipMallocAsync(&buf, jbytes + 64, s));
    BatchJob* d_jobs = (BatchJob*)buf;
    uint32_t* d_counters = (uint32_t*)((char*)buf + jbytes);
    hipError_t e = stage_jobs(c, host.data(), jbytes, d_jobs, s);
    RayParams p = make_params(prm, c->tune.tile_map, w, h);
    staged_tuning(c->tune, p);
    p.nouter = (prm.max_steps / num_parts) / kLoopUnroll;
    p.current_part = current_part;
    const BatchQueues q = make_queues(p, n);
    if (e == hipSuccess)
        e = dispatch_batch(c, prm, dtype, pow2, chunk_for(c->tune, p.nouter, nmax), d_jobs, p, q,
                           d_counters, s);
    (void)hipFreeAsync(buf, s);
    HIPCALL(c, e);
    return SPR_OK;
}

}  // extern "C"

namespace {

// The job table of a batch of frames of their own sizes and settings (first[] left 0).
int frames_to_jobs(spr_ctx* c, const spr_frame* frames, int n, std::vector<BatchJob>& host,
                   int* dtype, bool* pow2, int* nmax) {
    if (!frames || n <= 0) return fail(c, SPR_ERR_INVALID_ARG, "empty batch");
    std::vector<spr_job> jobs((size_t)n);
    for (int i = 0; i < n; ++i) {
        const spr_frame& f = frames[i];
        if (f.w <= 0 || f.h <= 0 || f.w > 32768 || f.h > 32768)
            return fail(c, SPR_ERR_INVALID_ARG, "frame %d: invalid output size %d x %d", i, f.w, f.h);
        if (f.max_steps < kLoopUnroll)
            return fail(c, SPR_ERR_INVALID_ARG, "frame %d: max_steps %d leaves no 16-sample block",
                        i, f.max_steps);
        jobs[i].slot = f.slot;
        memcpy(jobs[i].inv_p, f.inv_p, sizeof(f.inv_p));
        memcpy(jobs[i].inv_m, f.inv_m, sizeof(f.inv_m));
        jobs[i].d_out = f.d_out;
        jobs[i].d_alpha = f.d_alpha;
    }
    const int rc = build_jobs(c, jobs.data(), n, host, dtype, pow2, nmax);
    if (rc) return rc;
    for (int i = 0; i < n; ++i) {
        const spr_frame& f = frames[i];
        BatchJob& b = host[i];
        memcpy(b.box, f.box, sizeof(b.box));
        b.min_val = f.min_val;
        b.max_val = f.max_val;
        b.gamma = f.gamma;
        b.nouter = f.max_steps / kLoopUnroll;
        b.W = (uint32_t)f.w;
        b.H = (uint32_t)f.h;
    }
    return SPR_OK;
}

}  // namespace

extern "C" {

int spr_render_frames(spr_ctx* c, const spr_frame* frames, int n, float alpha_pow, int filter,
                      void* stream) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (filter != SPR_FILTER_LINEAR && filter != SPR_FILTER_NEAREST)
        return fail(c, SPR_ERR_INVALID_ARG, "unknown filter %d", filter);
    std::vector<BatchJob> host;
    int dtype, nmax;
    bool pow2;
    int rc = frames_to_jobs(c, frames, n, host, &dtype, &pow2, &nmax);
    if (rc) return rc;
    // each frame's first item in each queue; the box length follows the frame with the
    // longest ray step
    BatchQueues q{};
    int ch = kLoopUnroll;
    for (int i = 0; i < n; ++i) {
        BatchJob& b = host[i];
        const uint32_t ntx = (b.W + 15u) / 16u, nty = (b.H + 15u) / 16u;
        for (uint32_t k = 0; k < 8; ++k) {
            b.first[k] = q.count[k];
            q.count[k] += queue_rows(nty, k) * ntx;
        }
        ch = std::min(ch, chunk_for(c->tune, b.nouter, std::max(b.nx, std::max(b.ny, b.nz))));
    }
    q.mixed = 1;
    q.njobs = (uint32_t)n;
    spr_params prm{};
    prm.alpha_pow = alpha_pow;
    prm.filter = filter;
    prm.max_steps = frames[0].max_steps;
    DeviceGuard g(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    RayParams p = make_params(prm, c->tune.tile_map, frames[0].w, frames[0].h);
    staged_tuning(c->tune, p);
    p.nouter = host[0].nouter;
    p.current_part = 0;
    const size_t jbytes = (size_t)n * sizeof(BatchJob);
    void* buf = nullptr;
    HIPCALL(c, hipMallocAsync(&buf, jbytes + 64, s));
    BatchJob* d_jobs = (BatchJob*)buf;
    uint32_t* d_counters = (uint32_t*)((char*)buf + jbytes);
    hipError_t e = stage_jobs(c, host.data(), jbytes, d_jobs, s);
    if (e == hipSuccess) e = dispatch_batch(c, prm, dtype, pow2, ch, d_jobs, p, q, d_counters, s);
    (void)hipFreeAsync(buf, s);
    HIPCALL(c, e);
    return SPR_OK;
}

int spr_render_frames_split(spr_ctx* c, const spr_frame* frames, int n, int seg,
                            const unsigned int* d_gates, void* stream) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (seg == 0) seg = kSplitSegDefault;
    if (seg != 32 && seg != 64 && seg != 128)
        return fail(c, SPR_ERR_INVALID_ARG, "segment length %d (32, 64 or 128)", seg);
    std::vector<BatchJob> host;
    int dtype, nmax;
    bool pow2;
    int rc = frames_to_jobs(c, frames, n, host, &dtype, &pow2, &nmax);
    if (rc) return rc;
    // job k's first work item (8 x 8 rays x one segment) and first pixel; staged in LDS
    // slabs (tuning slab_kb > 0) when every frame's ray step allows 16-sample boxes
    uint64_t items = 0, pix = 0;
    int ch = kLoopUnroll;
    for (int i = 0; i < n; ++i) {
        BatchJob& b = host[i];
        ch = std::min(ch, chunk_for(c->tune, b.nouter, std::max(b.nx, std::max(b.ny, b.nz))));
        const uint64_t total = ((uint64_t)b.nouter + 1u) * kLoopUnroll;
        const uint64_t groups = (uint64_t)((b.W + 7u) / 8u) * ((b.H + 7u) / 8u);
        b.first[0] = (uint32_t)items;
        b.first[1] = (uint32_t)pix;
        items += groups * ((total + (uint64_t)seg - 1u) / (uint64_t)seg);
        pix += (uint64_t)b.W * b.H;
        if (items >= (1ull << 31) || pix >= (1ull << 31))
            return fail(c, SPR_ERR_INVALID_ARG, "batch too large for one split launch");
    }
    const bool staged = c->tune.slab_kb > 0 && ch == kLoopUnroll;
    DeviceGuard g(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    spr_params prm{};
    prm.max_steps = frames[0].max_steps;
    RayParams base = make_params(prm, c->tune.tile_map, frames[0].w, frames[0].h);
    staged_tuning(c->tune, base);
    const size_t jbytes = (size_t)n * sizeof(BatchJob);
    BatchJob* d_jobs = nullptr;
    HIPCALL(c, hipMallocAsync((void**)&d_jobs, jbytes, s));
    hipError_t e = stage_jobs(c, host.data(), jbytes, d_jobs, s);
    if (e == hipSuccess) {
        const uint32_t nj = (uint32_t)n, ni = (uint32_t)items, np = (uint32_t)pix;
        const uint32_t* gates = reinterpret_cast<const uint32_t*>(d_gates);
        switch (dtype) {
            case SPR_DTYPE_F32:
                e = pow2 ? TypedLaunch<float, true>::split(seg, staged, d_jobs, nj, ni, np, base, gates, s)
                         : TypedLaunch<float, false>::split(seg, staged, d_jobs, nj, ni, np, base, gates, s);
                break;
            case SPR_DTYPE_U16:
                e = pow2 ? TypedLaunch<uint16_t, true>::split(seg, staged, d_jobs, nj, ni, np, base, gates, s)
                         : TypedLaunch<uint16_t, false>::split(seg, staged, d_jobs, nj, ni, np, base, gates, s);
                break;
            default:
                e = pow2 ? TypedLaunch<uint8_t, true>::split(seg, staged, d_jobs, nj, ni, np, base, gates, s)
                         : TypedLaunch<uint8_t, false>::split(seg, staged, d_jobs, nj, ni, np, base, gates, s);
        }
    }
    (void)hipFreeAsync(d_jobs, s);
    HIPCALL(c, e);
    return SPR_OK;
}

int spr_render_batch_noalloc(spr_ctx* c, const spr_params* params, const spr_job* jobs, int n,
                             int w, int h, int num_parts, int current_part, void* stream) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (n > SPR_NOALLOC_MAX_JOBS)
        return fail(c, SPR_ERR_INVALID_ARG,
                    "%d jobs exceed the no-allocation batch's maximum of %d (SPR_NOALLOC_MAX_JOBS)",
                    n, SPR_NOALLOC_MAX_JOBS);
    const spr_params& prm = params ? *params : c->params;
    if (!params && !c->has_params) return fail(c, SPR_ERR_INVALID_ARG, "no parameters");
    int rc = check_params(c, prm);
    if (rc) return rc;
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768)
        return fail(c, SPR_ERR_INVALID_ARG, "invalid output size %d x %d", w, h);
    rc = check_parts(c, prm, num_parts);
    if (rc) return rc;
    if (!c->d_scratch) return fail(c, SPR_ERR_HIP, "no device scratch for this context");
    // one call at a time: the host staging table and each device's scratch are shared
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    int dtype, nmax;
    bool pow2;
    rc = build_jobs_into(c, jobs, n, g_noalloc_jobs, &dtype, &pow2, &nmax);
    if (rc) return rc;
    DeviceGuard g(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    uint32_t* d_counters = (uint32_t*)c->d_scratch;  // cleared by the launch's reset kernel
    BatchJob* d_jobs = (BatchJob*)((uint8_t*)c->d_scratch + kScratchCounterBytes);
    HIPCALL(c, hipMemcpyAsync(d_jobs, g_noalloc_jobs, (size_t)n * sizeof(BatchJob),
                              hipMemcpyHostToDevice, s));
    RayParams p = make_params(prm, c->tune.tile_map, w, h);
    staged_tuning(c->tune, p);
    p.nouter = (prm.max_steps / num_parts) / kLoopUnroll;
    p.current_part = current_part;
    const BatchQueues q = make_queues(p, n);
    HIPCALL(c, dispatch_batch(c, prm, dtype, pow2, chunk_for(c->tune, p.nouter, nmax), d_jobs, p, q, d_counters, s));
    // blocking: the next call may overwrite the staging table and the scratch
    HIPCALL(c, hipStreamSynchronize(s));
    return SPR_OK;
}

int spr_render_coarse_pair(spr_ctx* c, int slot_a, int slot_b, const spr_params* params, int w,
                           int h, float* d_out, float* d_alpha, void* stream) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    const VolSlot* a = slot_of(c, slot_a);
    const VolSlot* b = slot_of(c, slot_b);
    if (!a || !b) return fail(c, SPR_ERR_NO_VOLUME, "no volume in slot %d or %d", slot_a, slot_b);
    if (!params) return fail(c, SPR_ERR_INVALID_ARG, "NULL params");
    int rc = check_params(c, *params);
    if (rc) return rc;
    rc = check_parts(c, *params, 1);
    if (rc) return rc;
    if (a->layout != SPR_LAYOUT_PADDED || b->layout != SPR_LAYOUT_PADDED || a->dtype != b->dtype ||
        a->nx != b->nx || a->ny != b->ny || a->nz != b->nz)
        return fail(c, SPR_ERR_INVALID_ARG, "coarse pairs need two padded volumes of one type and shape");
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768)
        return fail(c, SPR_ERR_INVALID_ARG, "invalid output size %d x %d", w, h);
    if (!d_out || !d_alpha) return fail(c, SPR_ERR_INVALID_ARG, "NULL output");
    DeviceGuard g(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    RayParams p = make_params(*params, c->tune.tile_map, w, h);
    staged_tuning(c->tune, p);
    p.nouter = params->max_steps / kLoopUnroll;
    p.current_part = 0;
    const bool p2 = slot_pow2(*a);
    hipError_t e;
    switch (a->dtype) {
        case SPR_DTYPE_F32:
            e = p2 ? TypedLaunch<float, true>::coarse_pair(*a, *b, p, d_out, d_alpha, s)
                   : TypedLaunch<float, false>::coarse_pair(*a, *b, p, d_out, d_alpha, s);
            break;
        case SPR_DTYPE_U16:
            e = p2 ? TypedLaunch<uint16_t, true>::coarse_pair(*a, *b, p, d_out, d_alpha, s)
                   : TypedLaunch<uint16_t, false>::coarse_pair(*a, *b, p, d_out, d_alpha, s);
            break;
        default:
            e = p2 ? TypedLaunch<uint8_t, true>::coarse_pair(*a, *b, p, d_out, d_alpha, s)
                   : TypedLaunch<uint8_t, false>::coarse_pair(*a, *b, p, d_out, d_alpha, s);
            break;
    }
    HIPCALL(c, e);
    return SPR_OK;
}

int spr_render_pair(spr_ctx* c, int slot, const spr_params* params, int num_parts,
                    int current_part, int w, int h, float* d_out, float* d_alpha, void* stream) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    const VolSlot* vs = any_slot_of(c, slot);
    if (!vs) return fail(c, SPR_ERR_NO_VOLUME, "no volume in slot %d", slot);
    if (vs->dtype != SPR_DTYPE_U16X2 || vs->layout != SPR_LAYOUT_PADDED)
        return fail(c, SPR_ERR_INVALID_ARG, "slot %d is not a padded channel-pair volume", slot);
    if (!params) return fail(c, SPR_ERR_INVALID_ARG, "NULL params");
    int rc = check_params(c, *params);
    if (rc) return rc;
    if (params->filter != SPR_FILTER_LINEAR || params->alpha_pow != 0.0f)
        return fail(c, SPR_ERR_INVALID_ARG, "pair renders are plain max projections (linear filter, alpha_pow 0)");
    rc = check_parts(c, *params, num_parts);
    if (rc) return rc;
    if (current_part < 0 || current_part >= num_parts)
        return fail(c, SPR_ERR_INVALID_ARG, "current part %d of %d", current_part, num_parts);
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768)
        return fail(c, SPR_ERR_INVALID_ARG, "invalid output size %d x %d", w, h);
    if (!d_out || !d_alpha) return fail(c, SPR_ERR_INVALID_ARG, "NULL output");
    DeviceGuard g(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    RayParams p = make_params(*params, c->tune.tile_map, w, h);
    staged_tuning(c->tune, p);
    p.nouter = (params->max_steps / num_parts) / kLoopUnroll;
    p.current_part = current_part;
    HIPCALL(c, slot_pow2(*vs) ? launch_pair<true>(c->tune, *vs, p, d_out, d_alpha, s)
                              : launch_pair<false>(c->tune, *vs, p, d_out, d_alpha, s));
    return SPR_OK;
}

int spr_batch_graph_create(spr_ctx* c, const spr_job* jobs, int n, int w, int h,
                           spr_batch_graph** out) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!out) return fail(c, SPR_ERR_INVALID_ARG, "NULL output");
    *out = nullptr;
    int rc = check_frame(c, w, h);
    if (rc) return rc;
    std::vector<BatchJob> host;
    int dtype, nmax;
    bool pow2;
    rc = build_jobs(c, jobs, n, host, &dtype, &pow2, &nmax);
    if (rc) return rc;
    DeviceGuard g(c->device);
    spr_batch_graph* bg = new (std::nothrow) spr_batch_graph();
    if (!bg) return fail(c, SPR_ERR_OUT_OF_MEMORY, "out of host memory");
    bg->n = n;
    bg->w = w;
    bg->h = h;
    bg->dtype = dtype;
    bg->pow2 = pow2;
    record_refs(c, jobs, n, bg->refs);
    auto cleanup = [&](int code) {
        spr_batch_graph_destroy(c, bg);
        return code;
    };
    if (hipMalloc(&bg->d_jobs, (size_t)n * sizeof(BatchJob)) != hipSuccess ||
        hipMalloc(&bg->d_counters, 8 * sizeof(uint32_t)) != hipSuccess)
        return cleanup(fail(c, SPR_ERR_OUT_OF_MEMORY, "graph buffers"));
    if (hipMemcpy(bg->d_jobs, host.data(), (size_t)n * sizeof(BatchJob), hipMemcpyHostToDevice) !=
        hipSuccess)
        return cleanup(fail(c, SPR_ERR_HIP, "job upload failed"));
    rc = check_parts(c, c->params, 1);
    if (rc) return cleanup(rc);
    RayParams p = make_params(c, w, h);
    staged_tuning(c->tune, p);
    p.nouter = c->params.max_steps / kLoopUnroll;
    p.current_part = 0;
    const BatchQueues q = make_queues(p, n);
    hipStream_t cs;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(c, SPR_ERR_HIP, "capture stream"));
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    hipError_t e2 = hipSuccess;
    if (e == hipSuccess) {
        // the box length follows the captured jobs' volumes (speed only, never results)
        e2 = dispatch_batch(c, c->params, dtype, pow2, chunk_for(c->tune, p.nouter, nmax), bg->d_jobs, p, q,
                            bg->d_counters, cs);
        e = hipStreamEndCapture(cs, &bg->graph);
    }
    (void)hipStreamDestroy(cs);
    if (e != hipSuccess || e2 != hipSuccess)
        return cleanup(fail(c, SPR_ERR_HIP, "graph capture failed: %s",
                            hipGetErrorString(e != hipSuccess ? e : e2)));
    if (hipGraphInstantiate(&bg->exec, bg->graph, nullptr, nullptr, 0) != hipSuccess)
        return cleanup(fail(c, SPR_ERR_HIP, "graph instantiation failed"));
    *out = bg;
    return SPR_OK;
}

int spr_batch_graph_update(spr_ctx* c, spr_batch_graph* bg, const spr_job* jobs, int n) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!bg) return fail(c, SPR_ERR_INVALID_ARG, "NULL graph");
    if (n != bg->n) return fail(c, SPR_ERR_INVALID_ARG, "graph holds %d jobs, got %d", bg->n, n);
    std::vector<BatchJob> host;
    int dtype, nmax;
    bool pow2;
    int rc = build_jobs(c, jobs, n, host, &dtype, &pow2, &nmax);
    if (rc) return rc;
    if (dtype != bg->dtype)
        return fail(c, SPR_ERR_UNSUPPORTED_DTYPE, "voxel type differs from the captured batch");
    // the captured kernel's texel rule (one fma for power-of-two dimensions, two roundings
    // otherwise) must match every new volume, or its samples would differ from single renders
    if (pow2 != bg->pow2)
        return fail(c, SPR_ERR_INVALID_ARG,
                    "the captured batch's volumes %s power-of-two dimensions and the new ones %s "
                    "(capture a new graph)", bg->pow2 ? "all have" : "do not all have",
                    pow2 ? "do" : "do not");
    DeviceGuard g(c->device);
    if (bg->last_stream) HIPCALL(c, hipStreamSynchronize(bg->last_stream));
    HIPCALL(c, hipMemcpy(bg->d_jobs, host.data(), (size_t)n * sizeof(BatchJob),
                         hipMemcpyHostToDevice));
    record_refs(c, jobs, n, bg->refs);
    return SPR_OK;
}

int spr_batch_graph_launch(spr_ctx* c, spr_batch_graph* bg, void* stream) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!bg || !bg->exec) return fail(c, SPR_ERR_INVALID_ARG, "NULL graph");
    for (const auto& r : bg->refs) {
        const VolSlot* vs = slot_of(c, r.first);
        if (!vs || vs->generation != r.second)
            return fail(c, SPR_ERR_INVALID_ARG,
                        "slot %d was released or refilled since the graph's jobs were set "
                        "(spr_batch_graph_update or a new graph)", r.first);
    }
    DeviceGuard g(c->device);
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    HIPCALL(c, hipGraphLaunch(bg->exec, s));
    bg->last_stream = s;
    return SPR_OK;
}

int spr_batch_graph_destroy(spr_ctx* c, spr_batch_graph* bg) {
    if (!bg) return SPR_OK;
    if (c) {
        DeviceGuard g(c->device);
        if (bg->last_stream) (void)hipStreamSynchronize(bg->last_stream);
    }
    if (bg->exec) (void)hipGraphExecDestroy(bg->exec);
    if (bg->graph) (void)hipGraphDestroy(bg->graph);
    if (bg->d_jobs) (void)hipFree(bg->d_jobs);
    if (bg->d_counters) (void)hipFree(bg->d_counters);
    delete bg;
    return SPR_OK;
}

int spr_synchronize(spr_ctx* c) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    DeviceGuard g(c->device);
    HIPCALL(c, hipStreamSynchronize(c->stream));
    return SPR_OK;
}

int spr_get_stream(spr_ctx* c, void** stream) {
    if (!c || !stream) return fail(c, SPR_ERR_INVALID_ARG, "NULL argument");
    *stream = (void*)c->stream;
    return SPR_OK;
}

int spr_count_samples(spr_ctx* c, int num_parts, int w, int h, double* samples) {
    int rc = check_ready(c, w, h);
    if (rc) return rc;
    if (!samples) return fail(c, SPR_ERR_INVALID_ARG, "NULL output");
    rc = check_parts(c, c->params, num_parts);
    if (rc) return rc;
    DeviceGuard g(c->device);
    RayParams p = make_params(c, w, h);
    staged_tuning(c->tune, p);
    HIPCALL(c, hipMemsetAsync(c->d_count, 0, sizeof(unsigned long long), c->stream));
    hipLaunchKernelGGL(k_count_hits, dim3(tile_grid(p.tile_map, p.ntx, p.nty)), dim3(256), 0,
                       c->stream, p, c->d_count);
    HIPCALL(c, hipGetLastError());
    unsigned long long n = 0;
    HIPCALL(c, hipMemcpyAsync(&n, c->d_count, sizeof(n), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    const int per_ray = ((c->params.max_steps / num_parts) / kLoopUnroll + 1) * kLoopUnroll;
    *samples = (double)n * (double)per_ray;
    return SPR_OK;
}

}  // extern "C"
