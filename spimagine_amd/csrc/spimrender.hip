// spimrender.hip -- libspimrender.so: C ABI (include/spimrender.h) over the gfx950
// ray-casting kernels.  Host side of spimagine's VolumeRenderer
// (spimagine/volumerender/volumerender.py): volume upload and layout, parameter state,
// kernel dispatch in the reference's launch order, and the device-to-host output copies.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "mip_staged.h"
#include "render_kernels.h"
#include "spimrender.h"

using namespace spr;

struct spr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // resident volume: (nz, ny, nx) C order, x fastest, in one of two layouts
    int layout_req = SPR_LAYOUT_AUTO;  // requested for the next upload
    int layout = SPR_LAYOUT_AUTO;      // layout of the resident volume
    void* d_vol = nullptr;             // SPR_LAYOUT_LINEAR: the volume as uploaded
    size_t vol_bytes = 0;
    void* d_pad = nullptr;             // SPR_LAYOUT_PADDED: (nz+2, ny+2, nx+2) apron copy
    size_t pad_bytes = 0;
    int dtype = -1, nx = 0, ny = 0, nz = 0;
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
    // launch tuning (spr_set_tuning)
    int tile_map = (int)kTileMapInterleaved;
    int slab_kb = 12;  // per-wave LDS slab of the staged max projection, 0 = off
};

namespace {

thread_local std::string g_err;

int fail(spr_ctx* ctx, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    if (ctx)
        ctx->err = buf;
    else
        g_err = buf;
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

void free_volume(spr_ctx* c) {
    if (c->d_vol) (void)hipFree(c->d_vol);
    if (c->d_pad) (void)hipFree(c->d_pad);
    c->d_vol = c->d_pad = nullptr;
    c->vol_bytes = c->pad_bytes = 0;
}

bool has_volume(const spr_ctx* c) { return c->d_vol || c->d_pad; }

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

RayParams make_params(const spr_ctx* c, int w, int h) {
    RayParams p;
    memcpy(p.invP, c->params.inv_p, sizeof(p.invP));
    memcpy(p.invM, c->params.inv_m, sizeof(p.invM));
    memcpy(p.box, c->params.box, sizeof(p.box));
    p.min_val = c->params.min_val;
    p.max_val = c->params.max_val;
    p.gamma = c->params.gamma;
    p.alpha_pow = c->params.alpha_pow;
    p.iso_val = c->params.max_val / 2.0f;
    p.max_steps = c->params.max_steps;
    p.nouter = 0;
    p.current_part = 0;
    p.W = (uint32_t)w;
    p.H = (uint32_t)h;
    p.ntx = (uint32_t)((w + 15) / 16);
    p.nty = (uint32_t)((h + 15) / 16);
    p.ntiles = p.ntx * p.nty;
    p.tile_map = (uint32_t)c->tile_map;
    return p;
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

bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

template <typename T>
LinearVolume<T> linear_volume(const spr_ctx* c) {
    LinearVolume<T> v;
    v.data = (const T*)c->d_vol;
    v.nx = c->nx;
    v.ny = c->ny;
    v.nz = c->nz;
    v.fnx = (float)c->nx;
    v.fny = (float)c->ny;
    v.fnz = (float)c->nz;
    return v;
}

template <typename T, bool POW2>
PaddedVolume<T, POW2> padded_volume(const spr_ctx* c) {
    PaddedVolume<T, POW2> v;
    v.base = c->d_pad;
    v.nbytes = (uint32_t)c->pad_bytes;
    const uint32_t PY = (uint32_t)c->ny + 2;
    v.px_bytes = padded_pitch<T>(c->nx) * (uint32_t)sizeof(T);
    v.py = PY;
    v.soff_y = v.px_bytes;
    v.soff_z = v.px_bytes * PY;
    v.soff_yz = v.soff_y + v.soff_z;
    v.nx = c->nx;
    v.ny = c->ny;
    v.nz = c->nz;
    v.fnx = (float)c->nx;
    v.fny = (float)c->ny;
    v.fnz = (float)c->nz;
    v.mx = (float)(c->nx - 1);
    v.my = (float)(c->ny - 1);
    v.mz = (float)(c->nz - 1);
    return v;
}

// Calls f(volume accessor) with the accessor matching the resident volume's voxel type,
// layout and (padded layout) whether every dimension is a power of two.
template <typename T, typename F>
hipError_t visit_typed(const spr_ctx* c, F&& f) {
    if (c->layout == SPR_LAYOUT_LINEAR) return f(linear_volume<T>(c));
    if (is_pow2(c->nx) && is_pow2(c->ny) && is_pow2(c->nz)) return f(padded_volume<T, true>(c));
    return f(padded_volume<T, false>(c));
}

template <typename F>
hipError_t visit_volume(const spr_ctx* c, F&& f) {
    switch (c->dtype) {
        case SPR_DTYPE_F32: return visit_typed<float>(c, f);
        case SPR_DTYPE_U16: return visit_typed<uint16_t>(c, f);
        default: return visit_typed<uint8_t>(c, f);
    }
}

int check_ready(spr_ctx* c, int w, int h) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!has_volume(c)) return fail(c, SPR_ERR_NO_VOLUME, "no data provided, set_data(data) before");
    if (!c->has_params) return fail(c, SPR_ERR_INVALID_ARG, "spr_set_params was not called");
    if (w <= 0 || h <= 0 || w > 32768 || h > 32768)
        return fail(c, SPR_ERR_INVALID_ARG, "invalid output size %d x %d", w, h);
    return SPR_OK;
}

// render(method, numParts, currentPart) into device buffers; the reference's launch
// order (volumerender.py:366-386 for max_project, :451-497 for iso_surface)
int render_impl(spr_ctx* c, int method, int num_parts, int current_part, int w, int h, float* out,
                float* alpha, float* depth, float* normals, float* occ, hipStream_t s) {
    int rc = check_ready(c, w, h);
    if (rc) return rc;
    RayParams p = make_params(c, w, h);
    const dim3 grid(tile_grid(p.tile_map, p.ntx, p.nty)), block(256);
    const bool nearest = c->params.filter == SPR_FILTER_NEAREST;
    if (method == SPR_METHOD_MAX_PROJECT) {
        if (num_parts < 1 || (c->params.max_steps / num_parts) / kLoopUnroll < 1)
            return fail(c, SPR_ERR_INVALID_ARG,
                        "numParts=%d leaves no 16-sample block of max_steps=%d "
                        "(the reference divides by zero here)",
                        num_parts, c->params.max_steps);
        if (!out || !alpha) return fail(c, SPR_ERR_INVALID_ARG, "out/alpha buffers are required");
        p.nouter = (c->params.max_steps / num_parts) / kLoopUnroll;
        p.current_part = current_part;
        const bool staged = !nearest && c->params.alpha_pow == 0.0f && c->slab_kb > 0;
        hipError_t e = visit_volume(c, [&](auto v) {
            using V = decltype(v);
            if constexpr (is_padded<V>::value) {
                if (staged) {
                    using T = typename V::value_type;
                    constexpr bool P2 = V::kPow2;
                    if (c->slab_kb == 8)
                        hipLaunchKernelGGL((k_max_project_staged<T, P2, 8192>), grid, block, 0, s, v,
                                           p, out, alpha);
                    else if (c->slab_kb == 16)
                        hipLaunchKernelGGL((k_max_project_staged<T, P2, 16384>), grid, block, 0, s,
                                           v, p, out, alpha);
                    else
                        hipLaunchKernelGGL((k_max_project_staged<T, P2, 12288>), grid, block, 0, s,
                                           v, p, out, alpha);
                    return hipGetLastError();
                }
            }
            if (nearest)
                hipLaunchKernelGGL((k_max_project<V, 1>), grid, block, 0, s, v, p, out, alpha);
            else
                hipLaunchKernelGGL((k_max_project<V, 0>), grid, block, 0, s, v, p, out, alpha);
            return hipGetLastError();
        });
        HIPCALL(c, e);
        return SPR_OK;
    }
    if (method == SPR_METHOD_ISO_SURFACE) {
        if (c->params.max_steps < 2)
            return fail(c, SPR_ERR_INVALID_ARG, "iso_surface needs max_steps >= 2");
        if (!out || !alpha || !depth || !normals || !occ)
            return fail(c, SPR_ERR_INVALID_ARG, "iso_surface needs all five output buffers");
        if (c->params.occ_n_points < 1 || c->params.occ_radius < 0)
            return fail(c, SPR_ERR_INVALID_ARG, "invalid occlusion parameters");
        rc = ensure_scratch(c, (size_t)w * h);
        if (rc) return rc;
        hipError_t e = visit_volume(c, [&](auto v) {
            using V = decltype(v);
            if (nearest)
                hipLaunchKernelGGL((k_iso_surface<V, 1>), grid, block, 0, s, v, p, out, alpha,
                                   depth, normals);
            else
                hipLaunchKernelGGL((k_iso_surface<V, 0>), grid, block, 0, s, v, p, out, alpha,
                                   depth, normals);
            return hipGetLastError();
        });
        HIPCALL(c, e);
        dim3 g2((w + 15) / 16, (h + 15) / 16), b2(256);
        const ConvWeights wv = conv_weights(7, 5.0f), ws = conv_weights(5, 10.0f);
        // _convolve_vec(buf_normals, 7) (volumerender.py:405-416, 470)
        hipLaunchKernelGGL((k_conv<0, 3>), g2, b2, 0, s, normals, c->d_tmp_vec, w, h, wv);
        hipLaunchKernelGGL((k_conv<1, 3>), g2, b2, 0, s, c->d_tmp_vec, normals, w, h, wv);
        // occlusion (:472-481)
        hipLaunchKernelGGL(k_occlusion, g2, b2, 0, s, occ, w, h, c->params.occ_radius,
                           c->params.occ_n_points, depth);
        // _convolve_scalar(buf_occlusion, 5) (:392-403, 483)
        hipLaunchKernelGGL((k_conv<0, 1>), g2, b2, 0, s, occ, c->d_tmp, w, h, ws);
        hipLaunchKernelGGL((k_conv<1, 1>), g2, b2, 0, s, c->d_tmp, occ, w, h, ws);
        // shading overwrites the output (:485-497)
        hipLaunchKernelGGL(k_shading, g2, b2, 0, s, out, p, c->params.occ_strength, normals, depth,
                           occ);
        HIPCALL(c, hipGetLastError());
        return SPR_OK;
    }
    // an unknown method is a silent no-op in the reference (volumerender.py:543-547)
    return SPR_OK;
}

template <typename T>
void launch_pad(const void* src, void* dst, int nx, int ny, int nz, hipStream_t s) {
    dim3 grid((padded_pitch<T>(nx) + 255) / 256, ny + 2, nz + 2);
    hipLaunchKernelGGL(k_pad_volume<T>, grid, dim3(256), 0, s, (const T*)src, (T*)dst, nx, ny, nz);
}

int set_volume_impl(spr_ctx* c, const void* src, bool src_on_device, int dtype, int nx, int ny,
                    int nz) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
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
    int layout = c->layout_req;
    if (layout == SPR_LAYOUT_AUTO) layout = pad_fits ? SPR_LAYOUT_PADDED : SPR_LAYOUT_LINEAR;
    if (layout == SPR_LAYOUT_PADDED && !pad_fits)
        return fail(c, SPR_ERR_INVALID_ARG, "padded volume of %zu bytes exceeds 4 GiB", pbytes);
    HIPCALL(c, hipStreamSynchronize(c->stream));
    // a device source may still be in production on another stream (the context's stream
    // does not synchronise with the null stream): wait for all prior work on the device
    if (src_on_device) HIPCALL(c, hipDeviceSynchronize());

    // linear copy: kept as the resident volume (LINEAR), or staged for padding (host source)
    const bool need_linear = layout == SPR_LAYOUT_LINEAR || !src_on_device;
    if (c->d_pad && (layout != SPR_LAYOUT_PADDED || c->pad_bytes != pbytes)) {
        HIPCALL(c, hipFree(c->d_pad));
        c->d_pad = nullptr;
        c->pad_bytes = 0;
    }
    if (c->d_vol && (!need_linear || c->vol_bytes != bytes)) {
        HIPCALL(c, hipFree(c->d_vol));
        c->d_vol = nullptr;
        c->vol_bytes = 0;
    }
    if (need_linear && !c->d_vol) {
        HIPCALL(c, hipMalloc(&c->d_vol, bytes));
        c->vol_bytes = bytes;
    }
    if (need_linear)
        HIPCALL(c, hipMemcpyAsync(c->d_vol, src, bytes,
                                  src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                                  c->stream));
    if (layout == SPR_LAYOUT_PADDED) {
        if (!c->d_pad) {
            HIPCALL(c, hipMalloc(&c->d_pad, pbytes));
            c->pad_bytes = pbytes;
        }
        const void* from = need_linear ? c->d_vol : src;
        switch (dtype) {
            case SPR_DTYPE_F32: launch_pad<float>(from, c->d_pad, nx, ny, nz, c->stream); break;
            case SPR_DTYPE_U16: launch_pad<uint16_t>(from, c->d_pad, nx, ny, nz, c->stream); break;
            default: launch_pad<uint8_t>(from, c->d_pad, nx, ny, nz, c->stream); break;
        }
        HIPCALL(c, hipGetLastError());
    }
    HIPCALL(c, hipStreamSynchronize(c->stream));
    if (layout == SPR_LAYOUT_PADDED && c->d_vol) {  // staging copy no longer needed
        HIPCALL(c, hipFree(c->d_vol));
        c->d_vol = nullptr;
        c->vol_bytes = 0;
    }
    c->layout = layout;
    c->dtype = dtype;
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    return SPR_OK;
}

template <typename T>
void launch_minmax(const void* data, size_t n, float* partial, int blocks, hipStream_t s) {
    hipLaunchKernelGGL((k_minmax<T>), dim3(blocks), dim3(256), 0, s, (const T*)data, n, partial);
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
    spr_ctx* c = new (std::nothrow) spr_ctx();
    if (!c) return fail(nullptr, SPR_ERR_OUT_OF_MEMORY, "out of host memory");
    c->device = device;
    DeviceGuard g(device);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->d_count, sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc(&c->d_partial, 2 * 1024 * sizeof(float));
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
        free_volume(c);
        if (c->d_count) (void)hipFree(c->d_count);
        if (c->d_partial) (void)hipFree(c->d_partial);
        if (c->stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return SPR_OK;
}

const char* spr_last_error(const spr_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int spr_set_layout(spr_ctx* c, int layout) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (layout != SPR_LAYOUT_AUTO && layout != SPR_LAYOUT_LINEAR && layout != SPR_LAYOUT_PADDED)
        return fail(c, SPR_ERR_INVALID_ARG, "unknown layout %d", layout);
    c->layout_req = layout;
    return SPR_OK;
}

int spr_get_layout(spr_ctx* c, int* layout) {
    if (!c || !layout) return fail(c, SPR_ERR_INVALID_ARG, "NULL argument");
    if (!has_volume(c)) return fail(c, SPR_ERR_NO_VOLUME, "no volume");
    *layout = c->layout;
    return SPR_OK;
}

int spr_set_tuning(spr_ctx* c, int key, int value) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    switch (key) {
        case SPR_TUNE_TILE_MAP:
            if (value < 0 || value > 2) return fail(c, SPR_ERR_INVALID_ARG, "tile map %d", value);
            c->tile_map = value;
            return SPR_OK;
        case SPR_TUNE_SLAB_KB:
            if (value != 0 && value != 8 && value != 12 && value != 16)
                return fail(c, SPR_ERR_INVALID_ARG, "slab size %d KiB (0, 8, 12 or 16)", value);
            c->slab_kb = value;
            return SPR_OK;
        default: return fail(c, SPR_ERR_INVALID_ARG, "unknown tuning key %d", key);
    }
}

int spr_get_tuning(spr_ctx* c, int key, int* value) {
    if (!c || !value) return fail(c, SPR_ERR_INVALID_ARG, "NULL argument");
    switch (key) {
        case SPR_TUNE_TILE_MAP: *value = c->tile_map; return SPR_OK;
        case SPR_TUNE_SLAB_KB: *value = c->slab_kb; return SPR_OK;
        default: return fail(c, SPR_ERR_INVALID_ARG, "unknown tuning key %d", key);
    }
}

int spr_set_volume(spr_ctx* c, const void* host, int dtype, int nx, int ny, int nz) {
    return set_volume_impl(c, host, false, dtype, nx, ny, nz);
}

int spr_set_volume_device(spr_ctx* c, const void* dev, int dtype, int nx, int ny, int nz) {
    return set_volume_impl(c, dev, true, dtype, nx, ny, nz);
}

int spr_volume_minmax(spr_ctx* c, float* mn, float* mx) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!has_volume(c)) return fail(c, SPR_ERR_NO_VOLUME, "no volume");
    if (!mn || !mx) return fail(c, SPR_ERR_INVALID_ARG, "NULL output");
    DeviceGuard g(c->device);
    // the apron replicates edge voxels, so min/max over the padded copy are the volume's
    const void* data = c->d_pad ? c->d_pad : c->d_vol;
    const size_t n = (c->d_pad ? c->pad_bytes : c->vol_bytes) / dtype_size(c->dtype);
    const int blocks = 1024;
    switch (c->dtype) {
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
    if (p->filter != SPR_FILTER_LINEAR && p->filter != SPR_FILTER_NEAREST)
        return fail(c, SPR_ERR_INVALID_ARG, "unknown filter %d", p->filter);
    if (p->max_steps < 1)
        return fail(c, SPR_ERR_INVALID_ARG, "max_steps must be positive (got %d)", p->max_steps);
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
    rc = render_impl(c, method, num_parts, current_part, w, h, c->d_out, c->d_alpha, c->d_depth,
                     c->d_normals, c->d_occ, c->stream);
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
    return render_impl(c, method, num_parts, current_part, w, h, d_out, d_alpha, d_depth, d_normals,
                       d_occlusion, s);
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
    if (num_parts < 1 || (c->params.max_steps / num_parts) / kLoopUnroll < 1)
        return fail(c, SPR_ERR_INVALID_ARG, "invalid numParts %d", num_parts);
    DeviceGuard g(c->device);
    RayParams p = make_params(c, w, h);
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
