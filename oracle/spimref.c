/*
 * spimref.c -- CPU restatement of spimagine's OpenCL ray-casting kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle that the HIP product path
 * is checked against, and the timed CPU baseline.  Only tests/, the smoke check
 * in __graft_entry__.py and bench.py's cpu_baseline leg load it.  The product
 * library (spimagine_amd/csrc) never links or calls it.
 *
 * Reference files restated here (paths relative to the reference checkout):
 *   spimagine/volumerender/kernels/utils.cl          intersectBox, mult, random, rand_int
 *   spimagine/volumerender/kernels/volume_kernel.cl  max_project_float / max_project_short
 *   spimagine/volumerender/kernels/iso_kernel.cl     iso_surface (:17-225), shading (:505-588)
 *   spimagine/volumerender/kernels/convolve_2d.cl    conv_x/conv_y/conv_vec_x/conv_vec_y
 *   spimagine/volumerender/kernels/occlusion.cl      occlusion (:41-82)
 *   OpenCL 1.2 spec section 8.2 (sampler: normalized coords, CLAMP_TO_EDGE, LINEAR/NEAREST)
 *
 * Floating-point semantics.  The reference leaves several operations to the OpenCL
 * implementation (dot order, normalize, min/max on inf, sampler rounding, u16 with a
 * LINEAR sampler).  This restatement fixes one choice for each, and the HIP kernels
 * implement the *same* float32 operation sequence, so the two agree bit for bit on
 * the ray set-up and the sample positions:
 *   - compiled with -ffp-contract=off; every fused multiply-add is an explicit fmaf();
 *   - dot(a,b) = ((a.x*b.x + a.y*b.y) + a.z*b.z) + a.w*b.w;
 *   - normalize(v) = v / sqrt(dot(v,v)) per component, v itself when dot(v,v)==0;
 *   - min(x,y) = (y<x)?y:x and max(x,y) = (x<y)?y:x  (OpenCL 1.2 section 6.12.4);
 *   - LINEAR sampler: u = s*N, u' = u-0.5 (two roundings), i0 = floor(u'), a = u'-i0,
 *     corners clamped to [0,N-1], then lerp(x) -> lerp(y) -> lerp(z) with
 *     lerp(p,q,t) = fmaf(t, q-p, p);  integer volumes are interpolated in float32
 *     (no rounding back to an integer);
 *   - NEAREST sampler: i = clamp(floor(s*N), 0, N-1);
 *   - IEEE division, sqrt and infinities (no fast-math);
 *   - sin/cos in the occlusion pass use spr_sinf/spr_cosf below (Cephes polynomials),
 *     the conv weights use glibc expf on a float32 argument.
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SPR_F32 0
#define SPR_U16 1
#define SPR_U8 2
#define SPR_LINEAR 0
#define SPR_NEAREST 1

typedef struct { float x, y, z, w; } f4;

static inline f4 mk4(float x, float y, float z, float w) { f4 r; r.x = x; r.y = y; r.z = z; r.w = w; return r; }
static inline f4 add4(f4 a, f4 b) { return mk4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
static inline f4 sub4(f4 a, f4 b) { return mk4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
static inline f4 scl4(float s, f4 a) { return mk4(s * a.x, s * a.y, s * a.z, s * a.w); }
static inline f4 mul4s(f4 a, float s) { return mk4(a.x * s, a.y * s, a.z * s, a.w * s); }

static inline float dot4(f4 a, f4 b) {
    float s = a.x * b.x;
    s = s + a.y * b.y;
    s = s + a.z * b.z;
    s = s + a.w * b.w;
    return s;
}

static inline f4 normalize4(f4 v) {
    float l2 = dot4(v, v);
    if (l2 == 0.0f) return v;
    float l = sqrtf(l2);
    return mk4(v.x / l, v.y / l, v.z / l, v.w / l);
}

/* utils.cl:63-70  res.i = dot(v, row i of M) */
static inline f4 mult(const float *M, f4 v) {
    f4 r;
    r.x = dot4(v, mk4(M[0], M[1], M[2], M[3]));
    r.y = dot4(v, mk4(M[4], M[5], M[6], M[7]));
    r.z = dot4(v, mk4(M[8], M[9], M[10], M[11]));
    r.w = dot4(v, mk4(M[12], M[13], M[14], M[15]));
    return r;
}

static inline float cl_min(float x, float y) { return (y < x) ? y : x; }
static inline float cl_max(float x, float y) { return (x < y) ? y : x; }

/* utils.cl:41-60 slab test; only x,y,z enter tnear/tfar */
static inline int intersect_box(f4 ro, f4 rd, f4 bmin, f4 bmax, float *tnear, float *tfar) {
    float ix = 1.0f / rd.x, iy = 1.0f / rd.y, iz = 1.0f / rd.z;
    float tbx = ix * (bmin.x - ro.x), tby = iy * (bmin.y - ro.y), tbz = iz * (bmin.z - ro.z);
    float ttx = ix * (bmax.x - ro.x), tty = iy * (bmax.y - ro.y), ttz = iz * (bmax.z - ro.z);
    float mnx = cl_min(ttx, tbx), mny = cl_min(tty, tby), mnz = cl_min(ttz, tbz);
    float mxx = cl_max(ttx, tbx), mxy = cl_max(tty, tby), mxz = cl_max(ttz, tbz);
    float lt = cl_max(cl_max(mnx, mny), cl_max(mnx, mnz));
    float st = cl_min(cl_min(mxx, mxy), cl_min(mxx, mxz));
    *tnear = lt;
    *tfar = st;
    return st > lt;
}

/* volume_kernel.cl:48-77 (identical in iso_kernel.cl:42-73 and shading :21-43) */
static inline void ray_setup(const float *invP, const float *invM, uint32_t x, uint32_t y,
                             uint32_t W, uint32_t H, f4 *orig_o, f4 *direc_o) {
    float u = ((float)x / (float)W) * 2.0f - 1.0f;
    float v = ((float)y / (float)H) * 2.0f - 1.0f;
    f4 front = mk4(u, v, -1.0f, 1.0f);
    f4 back = mk4(u, v, 1.0f, 1.0f);
    f4 orig0 = mult(invP, front);
    orig0 = mul4s(orig0, 1.0f / orig0.w);
    f4 orig = mult(invM, orig0);
    orig = mul4s(orig, 1.0f / orig.w);
    f4 temp = mult(invP, back);
    temp = mul4s(temp, 1.0f / temp.w);
    f4 direc = mult(invM, normalize4(sub4(temp, orig0)));
    direc.w = 0.0f;
    *orig_o = orig;
    *direc_o = direc;
}

/* ---------------------------------------------------------------- sampler */

typedef struct {
    const void *data;
    int dtype, nx, ny, nz, filter;
} vol_t;

static inline float voxel(const vol_t *v, int x, int y, int z) {
    size_t i = ((size_t)z * (size_t)v->ny + (size_t)y) * (size_t)v->nx + (size_t)x;
    switch (v->dtype) {
    case SPR_U16: return (float)((const uint16_t *)v->data)[i];
    case SPR_U8: return (float)((const uint8_t *)v->data)[i];
    default: return ((const float *)v->data)[i];
    }
}

static inline int clampi(int a, int lo, int hi) { return a < lo ? lo : (a > hi ? hi : a); }
static inline float lerpf(float p, float q, float t) { return fmaf(t, q - p, p); }

/* floor of a texel coordinate, kept in [-1, N] before the int conversion; any value
   outside that range clamps to the same corner pair, so the result is unchanged */
static inline int floor_idx(float f, int n) {
    f = fminf(fmaxf(f, -1.0f), (float)n);
    return (int)f;
}

/* OpenCL 1.2 section 8.2, CLK_NORMALIZED_COORDS_TRUE | CLK_ADDRESS_CLAMP_TO_EDGE */
static inline float sample(const vol_t *v, f4 p) {
    if (v->filter == SPR_NEAREST) {
        float ux = p.x * (float)v->nx, uy = p.y * (float)v->ny, uz = p.z * (float)v->nz;
        int ix = clampi(floor_idx(floorf(ux), v->nx), 0, v->nx - 1);
        int iy = clampi(floor_idx(floorf(uy), v->ny), 0, v->ny - 1);
        int iz = clampi(floor_idx(floorf(uz), v->nz), 0, v->nz - 1);
        return voxel(v, ix, iy, iz);
    }
    float ux = p.x * (float)v->nx;
    float uy = p.y * (float)v->ny;
    float uz = p.z * (float)v->nz;
    ux = ux - 0.5f;
    uy = uy - 0.5f;
    uz = uz - 0.5f;
    float fx = floorf(ux), fy = floorf(uy), fz = floorf(uz);
    float ax = ux - fx, ay = uy - fy, az = uz - fz;
    int ix = floor_idx(fx, v->nx), iy = floor_idx(fy, v->ny), iz = floor_idx(fz, v->nz);
    int x0 = clampi(ix, 0, v->nx - 1), x1 = clampi(ix + 1, 0, v->nx - 1);
    int y0 = clampi(iy, 0, v->ny - 1), y1 = clampi(iy + 1, 0, v->ny - 1);
    int z0 = clampi(iz, 0, v->nz - 1), z1 = clampi(iz + 1, 0, v->nz - 1);
    float c00 = lerpf(voxel(v, x0, y0, z0), voxel(v, x1, y0, z0), ax);
    float c10 = lerpf(voxel(v, x0, y1, z0), voxel(v, x1, y1, z0), ax);
    float c01 = lerpf(voxel(v, x0, y0, z1), voxel(v, x1, y0, z1), ax);
    float c11 = lerpf(voxel(v, x0, y1, z1), voxel(v, x1, y1, z1), ax);
    float c0 = lerpf(c00, c10, ay);
    float c1 = lerpf(c01, c11, ay);
    return lerpf(c0, c1, az);
}

/* single sample, exported so tests can compare the sampler with a library routine */
float spimref_sample(const void *data, int dtype, int nx, int ny, int nz, int filter,
                     float sx, float sy, float sz) {
    vol_t v = {data, dtype, nx, ny, nz, filter};
    return sample(&v, mk4(sx, sy, sz, 0.0f));
}

/* ------------------------------------------------------------- max project */

#define LOOPUNROLL 16

/* volume_kernel.cl:20-185 (float) and :190-345 (short; also used for uint8).
   Writes rows [y_begin, y_end) of out/alpha (row-major, index x + W*y). */
static void max_project_pixel(const vol_t *v, const float *invP, const float *invM, const float *box,
                              float min_val, float max_val, float gamma, float alpha_pow,
                              int max_steps, int num_parts, int current_part,
                              uint32_t x, uint32_t y, uint32_t W, uint32_t H,
                              float *out, float *alpha) {
    const int is_short = v->dtype != SPR_F32;
    size_t o = (size_t)x + (size_t)W * (size_t)y;
    f4 orig, direc;
    ray_setup(invP, invM, x, y, W, H, &orig, &direc);
    f4 bmin = mk4(box[0], box[2], box[4], 1.0f);
    f4 bmax = mk4(box[1], box[3], box[5], 1.0f);
    float tnear, tfar;
    int hit = intersect_box(orig, direc, bmin, bmax, &tnear, &tfar);
    if (!hit) {
        out[o] = 0.0f;
        alpha[o] = is_short ? 0.0f : -1.0f;
        return;
    }
    if (tnear < 0.0f) tnear = 0.0f;

    float colVal = 0.0f;
    const int reduced = max_steps / num_parts;
    const float dt = fabsf(tfar - tnear) / (float)((reduced / LOOPUNROLL) * LOOPUNROLL);
    orig = add4(orig, scl4((float)current_part * dt, direc));
    f4 delta = scl4(0.5f * dt, direc);
    f4 pos = scl4(0.5f, add4(mk4(1.0f + orig.x, 1.0f + orig.y, 1.0f + orig.z, 1.0f + orig.w),
                             scl4(tnear, direc)));
    const int nouter = reduced / LOOPUNROLL;

    if (alpha_pow == 0.0f) {
        for (int i = 0; i <= nouter; ++i)
            for (int j = 0; j < LOOPUNROLL; ++j) {
                float nv = sample(v, pos);
                colVal = fmaxf(colVal, nv);
                pos = add4(pos, delta);
            }
        colVal = (max_val == 0.0f) ? colVal : (colVal - min_val) / (max_val - min_val);
    } else {
        float cumsum = 1.0f;
        const float a2 = alpha_pow * alpha_pow;
        const float a2s = 0.1f * alpha_pow * alpha_pow;
        for (int i = 0; i <= nouter; ++i)
            for (int j = 0; j < LOOPUNROLL; ++j) {
                float nv = sample(v, pos);
                nv = (max_val == 0.0f) ? nv : (nv - min_val) / (max_val - min_val);
                colVal = fmaxf(colVal, cumsum * nv);
                if (is_short)
                    cumsum *= (1.0f - a2s * nv);
                else
                    cumsum *= (1.0f - a2 * fminf(fmaxf(nv, 0.0f), 1.0f));
                pos = add4(pos, delta);
                if (cumsum <= 0.01f) break; /* leaves the inner loop only (volume_kernel.cl:148) */
            }
    }
    /* clamp(pow(colVal, gamma), 0, 1) (volume_kernel.cl:162); pow(x, 1) == x */
    colVal = fminf(fmaxf((gamma == 1.0f) ? colVal : powf(colVal, gamma), 0.0f), 1.0f);
    float alphaVal = is_short ? tnear : 1.0f;
    if (current_part == 0) {
        out[o] = colVal;
        alpha[o] = alphaVal;
    } else {
        out[o] = fmaxf(colVal, out[o]);
        alpha[o] = fmaxf(alphaVal, alpha[o]);
    }
}

int spimref_max_project(const void *data, int dtype, int nx, int ny, int nz, int filter,
                        const float *invP, const float *invM, const float *box,
                        float min_val, float max_val, float gamma, float alpha_pow,
                        int max_steps, int num_parts, int current_part,
                        int W, int H, int y_begin, int y_end,
                        float *out, float *alpha, int nthreads) {
    if (num_parts < 1 || (max_steps / num_parts) / LOOPUNROLL < 1) return 1;
    vol_t v = {data, dtype, nx, ny, nz, filter};
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int y = y_begin; y < y_end; ++y)
        for (int x = 0; x < W; ++x)
            max_project_pixel(&v, invP, invM, box, min_val, max_val, gamma, alpha_pow, max_steps,
                              num_parts, current_part, (uint32_t)x, (uint32_t)y, (uint32_t)W,
                              (uint32_t)H, out, alpha);
    return 0;
}

/* ------------------------------------------------------------- iso surface */

static inline float rd(const vol_t *v, f4 p) { return sample(v, p); }

/* iso_kernel.cl:17-225 */
static void iso_pixel(const vol_t *v, const float *invP, const float *invM, const float *box,
                      float iso_val, float gamma, int max_steps, uint32_t x, uint32_t y,
                      uint32_t W, uint32_t H, float *out, float *alpha, float *depth,
                      float *normals) {
    size_t o = (size_t)x + (size_t)W * (size_t)y;
    f4 orig, direc;
    ray_setup(invP, invM, x, y, W, H, &orig, &direc);
    f4 bmin = mk4(box[0], box[2], box[4], 1.0f);
    f4 bmax = mk4(box[1], box[3], box[5], 1.0f);
    float tnear, tfar;
    int hit = intersect_box(orig, direc, bmin, bmax, &tnear, &tfar);
    if (!hit) {
        out[o] = 0.0f;
        alpha[o] = 0.0f;
        depth[o] = INFINITY;
        normals[3 * o] = normals[3 * o + 1] = normals[3 * o + 2] = 0.0f;
        return;
    }
    if (tnear < 0.0f) tnear = 0.0f;
    float dt = 1.0f * (tfar - tnear) / ((float)max_steps - 1.0f);
    f4 delta = scl4(0.5f * dt, direc);
    f4 pos0 = scl4(0.5f, add4(mk4(1.0f + orig.x, 1.0f + orig.y, 1.0f + orig.z, 1.0f + orig.w),
                              scl4(tnear, direc)));
    f4 pos = pos0;
    float nv = rd(v, pos);
    int isGreater = nv > iso_val;
    int hitIso = 0;
    float t_hit = INFINITY;
    int i;
    for (i = 1; i < max_steps; i++) {
        pos = add4(pos, delta);
        t_hit = tnear + (float)i * dt;
        nv = rd(v, pos);
        if ((nv > iso_val) != isGreater) {
            hitIso = 1;
            break;
        }
    }
    if (!hitIso) {
        out[o] = 0.0f;
        alpha[o] = 0.0f;
        depth[o] = INFINITY;
        normals[3 * o] = normals[3 * o + 1] = normals[3 * o + 2] = 0.0f;
        return;
    }
    const int maxBisect = 10;
    f4 delta2 = mk4(delta.x / (float)maxBisect, delta.y / (float)maxBisect,
                    delta.z / (float)maxBisect, delta.w / (float)maxBisect);
    float dt2 = dt / (float)maxBisect;
    pos = add4(pos0, scl4((float)(i - 1), delta));
    for (int j = 1; j <= maxBisect; j++) {
        nv = rd(v, pos);
        pos = add4(pos, delta2);
        t_hit += dt2;
        if ((nv > iso_val) != isGreater) break;
    }

    f4 light = normalize4(mult(invM, mk4(2.0f, -1.0f, -2.0f, 0.0f)));
    const float c_ambient = 0.3f, c_diffuse = 0.4f, c_specular = 0.3f;
    float h = dt;
    h *= gamma * gamma; /* pow(gamma, 2), iso_kernel.cl:180 */
    float h2 = 2.0f * h;
    f4 n;
    n.x = 2.0f * rd(v, add4(pos, mk4(h, 0, 0, 0))) - 2.0f * rd(v, add4(pos, mk4(-h, 0, 0, 0)));
    n.x = n.x + rd(v, add4(pos, mk4(h2, 0, 0, 0)));
    n.x = n.x - rd(v, add4(pos, mk4(-h2, 0, 0, 0)));
    n.y = 2.0f * rd(v, add4(pos, mk4(0, h, 0, 0))) - 2.0f * rd(v, add4(pos, mk4(0, -h, 0, 0)));
    n.y = n.y + rd(v, add4(pos, mk4(0, h2, 0, 0)));
    n.y = n.y - rd(v, add4(pos, mk4(0, -h2, 0, 0)));
    n.z = rd(v, add4(pos, mk4(0, 0, h, 0))) - rd(v, add4(pos, mk4(0, 0, -h, 0)));
    n.z = n.z + rd(v, add4(pos, mk4(0, 0, h2, 0)));
    n.z = n.z - rd(v, add4(pos, mk4(0, 0, -h2, 0)));
    n.w = 0.0f;
    n = scl4(1.0f - (float)(2 * isGreater), normalize4(n));
    f4 refl = sub4(scl4(2.0f * dot4(light, n), n), light);
    float diffuse = fmaxf(0.0f, dot4(light, n));
    float specular = powf(fmaxf(0.0f, dot4(normalize4(refl), normalize4(direc))), 10.0f);
    float col = c_ambient + c_diffuse * diffuse;
    col = col + ((float)(diffuse > 0.0f) * c_specular) * specular;
    out[o] = col;
    alpha[o] = tnear;
    depth[o] = t_hit;
    normals[3 * o] = n.x;
    normals[3 * o + 1] = n.y;
    normals[3 * o + 2] = n.z;
}

int spimref_iso_surface(const void *data, int dtype, int nx, int ny, int nz, int filter,
                        const float *invP, const float *invM, const float *box, float iso_val,
                        float gamma, int max_steps, int W, int H, int y_begin, int y_end,
                        float *out, float *alpha, float *depth, float *normals, int nthreads) {
    if (max_steps < 2) return 1;
    vol_t v = {data, dtype, nx, ny, nz, filter};
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int y = y_begin; y < y_end; ++y)
        for (int x = 0; x < W; ++x)
            iso_pixel(&v, invP, invM, box, iso_val, gamma, max_steps, (uint32_t)x, (uint32_t)y,
                      (uint32_t)W, (uint32_t)H, out, alpha, depth, normals);
    return 0;
}

/* ------------------------------------------------------ 2D post passes */

/* convolve_2d.cl weight: exp((float)(-k*(ht-Nh/2.f)*(ht-Nh/2.f)/Nh/Nh)) */
void spimref_conv_weights(int Nh, float k, float *w) {
    for (int ht = 0; ht < Nh; ++ht) {
        float d = (float)ht - (float)Nh / 2.0f;
        float a = -k * d;
        a = a * d;
        a = a / (float)Nh;
        a = a / (float)Nh;
        w[ht] = expf(a);
    }
}

static inline void tap_range(int i, int N, int Nh, int *hs, int *he) {
    *hs = ((i - Nh / 2) < 0) ? Nh / 2 - i : 0;
    *he = ((i + Nh / 2) >= N) ? Nh - (i + Nh / 2 - N + 1) : Nh;
}

/* convolve_2d.cl:6-61 (scalar, k=10) */
static void conv_x_rows(const float *in, float *out, int W, int H, int Nh, int j0, int j1) {
    float w[256];
    if (Nh > 256) return;
    spimref_conv_weights(Nh, 10.0f, w);
    for (int j = j0; j < j1; ++j)
        for (int i = 0; i < W; ++i) {
            int hs, he;
            tap_range(i, W, Nh, &hs, &he);
            float res = 0.0f, s = 0.0f;
            int start = i - Nh / 2;
            for (int ht = hs; ht < he; ++ht) {
                s += w[ht];
                res += w[ht] * in[start + ht + j * W];
            }
            out[i + j * W] = res / s;
        }
}

static void conv_y_rows(const float *in, float *out, int W, int H, int Nh, int j0, int j1) {
    float w[256];
    if (Nh > 256) return;
    spimref_conv_weights(Nh, 10.0f, w);
    for (int j = j0; j < j1; ++j)
        for (int i = 0; i < W; ++i) {
            int hs, he;
            tap_range(j, H, Nh, &hs, &he);
            float res = 0.0f, s = 0.0f;
            int start = j - Nh / 2;
            for (int ht = hs; ht < he; ++ht) {
                s += w[ht];
                res += w[ht] * in[i + (start + ht) * W];
            }
            out[i + j * W] = res / s;
        }
}

/* convolve_2d.cl:65-137 (3-vector, k=5) */
static void conv_vec_x_rows(const float *in, float *out, int W, int H, int Nh, int j0, int j1) {
    float w[256];
    if (Nh > 256) return;
    spimref_conv_weights(Nh, 5.0f, w);
    for (int j = j0; j < j1; ++j)
        for (int i = 0; i < W; ++i) {
            int hs, he;
            tap_range(i, W, Nh, &hs, &he);
            float rx = 0.0f, ry = 0.0f, rz = 0.0f, s = 0.0f;
            int start = i - Nh / 2;
            for (int ht = hs; ht < he; ++ht) {
                size_t q = 3 * ((size_t)(start + ht) + (size_t)j * W);
                s += w[ht];
                rx += w[ht] * in[q];
                ry += w[ht] * in[q + 1];
                rz += w[ht] * in[q + 2];
            }
            size_t o = 3 * ((size_t)i + (size_t)j * W);
            out[o] = rx / s;
            out[o + 1] = ry / s;
            out[o + 2] = rz / s;
        }
}

static void conv_vec_y_rows(const float *in, float *out, int W, int H, int Nh, int j0, int j1) {
    float w[256];
    if (Nh > 256) return;
    spimref_conv_weights(Nh, 5.0f, w);
    for (int j = j0; j < j1; ++j)
        for (int i = 0; i < W; ++i) {
            int hs, he;
            tap_range(j, H, Nh, &hs, &he);
            float rx = 0.0f, ry = 0.0f, rz = 0.0f, s = 0.0f;
            int start = j - Nh / 2;
            for (int ht = hs; ht < he; ++ht) {
                size_t q = 3 * ((size_t)i + (size_t)(start + ht) * W);
                s += w[ht];
                rx += w[ht] * in[q];
                ry += w[ht] * in[q + 1];
                rz += w[ht] * in[q + 2];
            }
            size_t o = 3 * ((size_t)i + (size_t)j * W);
            out[o] = rx / s;
            out[o + 1] = ry / s;
            out[o + 2] = rz / s;
        }
}

void spimref_conv_x(const float *in, float *out, int W, int H, int Nh) { conv_x_rows(in, out, W, H, Nh, 0, H); }
void spimref_conv_y(const float *in, float *out, int W, int H, int Nh) { conv_y_rows(in, out, W, H, Nh, 0, H); }
void spimref_conv_vec_x(const float *in, float *out, int W, int H, int Nh) { conv_vec_x_rows(in, out, W, H, Nh, 0, H); }
void spimref_conv_vec_y(const float *in, float *out, int W, int H, int Nh) { conv_vec_y_rows(in, out, W, H, Nh, 0, H); }

/* utils.cl:10-23: uint32 LCG, 10 rounds, mod 79197919 */
float spimref_random(uint32_t x, uint32_t y) {
    uint32_t a = 4421u + (1u + x) * (1u + y) + x + y;
    for (int i = 0; i < 10; i++) a = (1664525u * a + 1013904223u) % 79197919u;
    return ((float)a * 1.0f) / (float)79197919;
}

/* utils.cl:25-38 (declared float, returns an int value) */
float spimref_rand_int(uint32_t x, uint32_t y, int start, int end) {
    float rnd = spimref_random(x, y);
    return (float)(int)((float)start + rnd * (float)(end - start));
}

/* Cephes single-precision sin/cos (S. L. Moshier, Cephes Math Library, sinf.c):
   Cody-Waite reduction by pi/2 and degree-7/8 minimax polynomials.  The HIP kernel
   evaluates the identical float32 sequence. */
static inline void spr_sincosf(float x, float *s, float *c) {
    const float TWO_OVER_PI = 0.636619772367581343f;
    const float DP1 = 1.5703125f, DP2 = 4.837512969970703125e-4f, DP3 = 7.54978995489188216e-8f;
    float k = rintf(x * TWO_OVER_PI);
    int q = (int)k;
    float r = fmaf(-k, DP1, x);
    r = fmaf(-k, DP2, r);
    r = fmaf(-k, DP3, r);
    float z = r * r;
    float sp = fmaf(fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f), z * r, r);
    float cp = fmaf(fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f), z * z, fmaf(-0.5f, z, 1.0f));
    switch (q & 3) {
    case 0: *s = sp; *c = cp; break;
    case 1: *s = cp; *c = -sp; break;
    case 2: *s = -sp; *c = -cp; break;
    default: *s = -cp; *c = sp; break;
    }
}

void spimref_sincos(float x, float *s, float *c) { spr_sincosf(x, s, c); }

/* occlusion.cl:41-82 */
static void occlusion_rows(float *out, int W, int H, int radius, int n_points, const float *depth,
                           int y0, int y1) {
    const float MPI_2 = 6.2831853071795f;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < W; ++x) {
            float depth0 = depth[x + y * W];
            float occ = 0.0f;
            for (uint32_t i = 0; i < (uint32_t)n_points; ++i) {
                uint32_t ii = i * i;
                float ra = (float)x + spimref_rand_int(i, ii, 0, 1000);
                float rb = (float)y + spimref_rand_int(ii, i, 294, 97701);
                float r = (float)(uint32_t)radius * spimref_random((uint32_t)ra, (uint32_t)rb);
                float pa = (float)x + spimref_rand_int(ii, i, 0, 1997);
                float pb = (float)y + spimref_rand_int(i, ii, 569, 17633);
                float phi = MPI_2 * spimref_random((uint32_t)pa, (uint32_t)pb);
                float sn, cs;
                spr_sincosf(phi, &sn, &cs);
                int x2 = clampi((int)((float)x + r * cs), 0, W - 1);
                int y2 = clampi((int)((float)y + r * sn), 0, H - 1);
                float d = depth[x2 + y2 * W];
                occ += (d < depth0) ? 1.0f : 0.0f;
            }
            out[x + W * y] = occ / (float)(uint32_t)n_points;
        }
}

void spimref_occlusion(float *out, int W, int H, int radius, int n_points, const float *depth) {
    occlusion_rows(out, W, H, radius, n_points, depth, 0, H);
}

/* iso_kernel.cl:505-588 */
static void shading_rows(float *out, int W, int H, const float *invP, const float *invM,
                         float occ_strength, const float *normals, const float *depth,
                         const float *occl, int y0, int y1) {
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < W; ++x) {
            f4 orig, direc;
            ray_setup(invP, invM, (uint32_t)x, (uint32_t)y, (uint32_t)W, (uint32_t)H, &orig, &direc);
            f4 light = normalize4(mult(invM, mk4(2.0f, -1.0f, -2.0f, 0.0f)));
            const float c_ambient = 0.5f, c_diffuse = 0.3f, c_specular = 0.2f;
            size_t o = (size_t)x + (size_t)W * (size_t)y;
            f4 n = mk4(normals[3 * o], normals[3 * o + 1], normals[3 * o + 2], 0.0f);
            float occ = occl[o];
            float d = depth[o];
            n = normalize4(n);
            f4 refl = sub4(scl4(2.0f * dot4(light, n), n), light);
            float diffuse = fmaxf(0.0f, dot4(light, n));
            float specular = powf(fmaxf(0.0f, dot4(normalize4(refl), normalize4(direc))), 10.0f);
            float col = c_ambient + c_diffuse * diffuse;
            col = col + ((float)(diffuse > 0.0f) * c_specular) * specular;
            col = (1.0f - occ_strength) * col + (occ_strength * col) * (1.0f - occ);
            /* iso_kernel.cl:579 multiplies by the double literal 1.0 */
            col = (float)((double)((1.0f - occ_strength) * col) + (1.0 * (double)occ_strength) * (double)col);
            col *= (d < INFINITY) ? 1.0f : 0.0f;
            out[o] = col;
        }
}

void spimref_shading(float *out, int W, int H, const float *invP, const float *invM,
                     float occ_strength, const float *normals, const float *depth,
                     const float *occl) {
    shading_rows(out, W, H, invP, invM, occ_strength, normals, depth, occl, 0, H);
}

/* volumerender.py:446-506 pipeline: iso_surface -> conv_vec(7) -> occlusion -> conv(5) -> shading */
int spimref_render_iso(const void *data, int dtype, int nx, int ny, int nz, int filter,
                       const float *invP, const float *invM, const float *box, float iso_val,
                       float gamma, int max_steps, float occ_strength, int occ_radius,
                       int occ_n_points, int W, int H, float *out, float *alpha, float *depth,
                       float *normals, float *occl, float *tmp_vec, float *tmp, int nthreads) {
    int rc = spimref_iso_surface(data, dtype, nx, ny, nz, filter, invP, invM, box, iso_val, gamma,
                                 max_steps, W, H, 0, H, out, alpha, depth, normals, nthreads);
    if (rc) return rc;
    spimref_conv_vec_x(normals, tmp_vec, W, H, 7);
    spimref_conv_vec_y(tmp_vec, normals, W, H, 7);
    spimref_occlusion(occl, W, H, occ_radius, occ_n_points, depth);
    spimref_conv_x(occl, tmp, W, H, 5);
    spimref_conv_y(tmp, occl, W, H, 5);
    spimref_shading(out, W, H, invP, invM, occ_strength, normals, depth, occl);
    return 0;
}

/* The same pipeline's output, alpha, depth, normals and occlusion on rows [y_begin, y_end)
   only.  A row's result depends on the iso march of the rows within max(3, occ_radius + 3)
   of it (normal blur +-3; occlusion samples within occ_radius, blurred +-2), so only
   those rows are marched; the arrays' other rows must hold a miss (depth +inf, normals 0)
   or anything else: they are read but cannot reach rows [y_begin, y_end). */
int spimref_render_iso_rows(const void *data, int dtype, int nx, int ny, int nz, int filter,
                            const float *invP, const float *invM, const float *box,
                            float iso_val, float gamma, int max_steps, float occ_strength,
                            int occ_radius, int occ_n_points, int W, int H, int y_begin,
                            int y_end, float *out, float *alpha, float *depth, float *normals,
                            float *occl, float *tmp_vec, float *tmp, int nthreads) {
    if (y_begin < 0 || y_end > H || y_begin >= y_end || occ_radius < 0) return 1;
    const int m = occ_radius + 3 > 3 ? occ_radius + 3 : 3;
    const int ma = y_begin - m < 0 ? 0 : y_begin - m, mb = y_end + m > H ? H : y_end + m;
    int rc = spimref_iso_surface(data, dtype, nx, ny, nz, filter, invP, invM, box, iso_val, gamma,
                                 max_steps, W, H, ma, mb, out, alpha, depth, normals, nthreads);
    if (rc) return rc;
    const int na = y_begin - 3 < 0 ? 0 : y_begin - 3, nb = y_end + 3 > H ? H : y_end + 3;
    const int oa = y_begin - 2 < 0 ? 0 : y_begin - 2, ob = y_end + 2 > H ? H : y_end + 2;
    conv_vec_x_rows(normals, tmp_vec, W, H, 7, na, nb);
    conv_vec_y_rows(tmp_vec, normals, W, H, 7, y_begin, y_end);
    occlusion_rows(occl, W, H, occ_radius, occ_n_points, depth, oa, ob);
    conv_x_rows(occl, tmp, W, H, 5, oa, ob);
    conv_y_rows(tmp, occl, W, H, 5, y_begin, y_end);
    shading_rows(out, W, H, invP, invM, occ_strength, normals, depth, occl, y_begin, y_end);
    return 0;
}
