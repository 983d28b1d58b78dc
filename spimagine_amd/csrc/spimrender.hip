// spimrender.hip -- libspimrender.so: C ABI (include/spimrender.h) over the gfx950
// ray-casting kernels.  Host side of spimagine's VolumeRenderer
// (spimagine/volumerender/volumerender.py): volume upload, parameter state, kernel
// dispatch in the reference's launch order, and the device-to-host output copies.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "render_kernels.h"
#include "spimrender.h"

using namespace spr;

struct spr_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // resident volume, (nz, ny, nx) C order, x fastest
    void* d_vol = nullptr;
    size_t vol_bytes = 0;
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
    p.ntiles = p.ntx * (uint32_t)((h + 15) / 16);
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

template <typename T>
hipError_t launch_mip(const spr_ctx* c, const RayParams& p, float* out, float* alpha,
                      hipStream_t s) {
    LinearVolume<T> v = linear_volume<T>(c);
    dim3 grid(p.ntiles), block(256);
    if (c->params.filter == SPR_FILTER_NEAREST)
        hipLaunchKernelGGL((k_max_project<T, 1>), grid, block, 0, s, v, p, out, alpha);
    else
        hipLaunchKernelGGL((k_max_project<T, 0>), grid, block, 0, s, v, p, out, alpha);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_iso(const spr_ctx* c, const RayParams& p, float* out, float* alpha, float* depth,
                      float* normals, hipStream_t s) {
    LinearVolume<T> v = linear_volume<T>(c);
    dim3 grid(p.ntiles), block(256);
    if (c->params.filter == SPR_FILTER_NEAREST)
        hipLaunchKernelGGL((k_iso_surface<T, 1>), grid, block, 0, s, v, p, out, alpha, depth,
                           normals);
    else
        hipLaunchKernelGGL((k_iso_surface<T, 0>), grid, block, 0, s, v, p, out, alpha, depth,
                           normals);
    return hipGetLastError();
}

int check_ready(spr_ctx* c, int w, int h) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!c->d_vol) return fail(c, SPR_ERR_NO_VOLUME, "no data provided, set_data(data) before");
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
    hipError_t e = hipSuccess;
    if (method == SPR_METHOD_MAX_PROJECT) {
        if (num_parts < 1 || (c->params.max_steps / num_parts) / kLoopUnroll < 1)
            return fail(c, SPR_ERR_INVALID_ARG,
                        "numParts=%d leaves no 16-sample block of max_steps=%d "
                        "(the reference divides by zero here)",
                        num_parts, c->params.max_steps);
        if (!out || !alpha) return fail(c, SPR_ERR_INVALID_ARG, "out/alpha buffers are required");
        p.nouter = (c->params.max_steps / num_parts) / kLoopUnroll;
        p.current_part = current_part;
        switch (c->dtype) {
            case SPR_DTYPE_F32: e = launch_mip<float>(c, p, out, alpha, s); break;
            case SPR_DTYPE_U16: e = launch_mip<uint16_t>(c, p, out, alpha, s); break;
            default: e = launch_mip<uint8_t>(c, p, out, alpha, s); break;
        }
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
        switch (c->dtype) {
            case SPR_DTYPE_F32: e = launch_iso<float>(c, p, out, alpha, depth, normals, s); break;
            case SPR_DTYPE_U16: e = launch_iso<uint16_t>(c, p, out, alpha, depth, normals, s); break;
            default: e = launch_iso<uint8_t>(c, p, out, alpha, depth, normals, s); break;
        }
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

int set_volume_impl(spr_ctx* c, const void* src, bool src_on_device, int dtype, int nx, int ny,
                    int nz) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    const size_t es = dtype_size(dtype);
    if (!es) return fail(c, SPR_ERR_UNSUPPORTED_DTYPE, "unsupported voxel type code %d", dtype);
    if (!src || nx <= 0 || ny <= 0 || nz <= 0)
        return fail(c, SPR_ERR_INVALID_ARG, "invalid volume (%d, %d, %d)", nz, ny, nx);
    const size_t nvox = (size_t)nx * ny * nz;
    if (nvox > 0xFFFFFFFFull)
        return fail(c, SPR_ERR_INVALID_ARG, "volume of %zu voxels exceeds 2^32-1", nvox);
    DeviceGuard g(c->device);
    const size_t bytes = nvox * es;
    if (c->d_vol && c->vol_bytes != bytes) {
        HIPCALL(c, hipStreamSynchronize(c->stream));
        HIPCALL(c, hipFree(c->d_vol));
        c->d_vol = nullptr;
        c->vol_bytes = 0;
    }
    if (!c->d_vol) {
        HIPCALL(c, hipMalloc(&c->d_vol, bytes));
        c->vol_bytes = bytes;
    }
    HIPCALL(c, hipMemcpyAsync(c->d_vol, src, bytes,
                              src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                              c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    c->dtype = dtype;
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    return SPR_OK;
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
        if (c->d_vol) (void)hipFree(c->d_vol);
        if (c->d_count) (void)hipFree(c->d_count);
        if (c->d_partial) (void)hipFree(c->d_partial);
        if (c->stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return SPR_OK;
}

const char* spr_last_error(const spr_ctx* c) { return c ? c->err.c_str() : g_err.c_str(); }

int spr_set_volume(spr_ctx* c, const void* host, int dtype, int nx, int ny, int nz) {
    return set_volume_impl(c, host, false, dtype, nx, ny, nz);
}

int spr_set_volume_device(spr_ctx* c, const void* dev, int dtype, int nx, int ny, int nz) {
    return set_volume_impl(c, dev, true, dtype, nx, ny, nz);
}

int spr_volume_minmax(spr_ctx* c, float* mn, float* mx) {
    if (!c) return fail(nullptr, SPR_ERR_INVALID_ARG, "NULL context");
    if (!c->d_vol) return fail(c, SPR_ERR_NO_VOLUME, "no volume");
    if (!mn || !mx) return fail(c, SPR_ERR_INVALID_ARG, "NULL output");
    DeviceGuard g(c->device);
    const size_t n = (size_t)c->nx * c->ny * c->nz;
    const int blocks = 1024;
    switch (c->dtype) {
        case SPR_DTYPE_F32:
            hipLaunchKernelGGL((k_minmax<float>), dim3(blocks), dim3(256), 0, c->stream,
                               (const float*)c->d_vol, n, c->d_partial);
            break;
        case SPR_DTYPE_U16:
            hipLaunchKernelGGL((k_minmax<uint16_t>), dim3(blocks), dim3(256), 0, c->stream,
                               (const uint16_t*)c->d_vol, n, c->d_partial);
            break;
        default:
            hipLaunchKernelGGL((k_minmax<uint8_t>), dim3(blocks), dim3(256), 0, c->stream,
                               (const uint8_t*)c->d_vol, n, c->d_partial);
            break;
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
    hipLaunchKernelGGL(k_count_hits, dim3(p.ntiles), dim3(256), 0, c->stream, p, c->d_count);
    HIPCALL(c, hipGetLastError());
    unsigned long long n = 0;
    HIPCALL(c, hipMemcpyAsync(&n, c->d_count, sizeof(n), hipMemcpyDeviceToHost, c->stream));
    HIPCALL(c, hipStreamSynchronize(c->stream));
    const int per_ray = ((c->params.max_steps / num_parts) / kLoopUnroll + 1) * kLoopUnroll;
    *samples = (double)n * (double)per_ray;
    return SPR_OK;
}

}  // extern "C"
