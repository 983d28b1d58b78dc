// ray_math.h -- device-side float32 ray set-up shared by every render kernel.
//
// The reference leaves dot order, normalize, min/max on infinities and division
// accuracy to the OpenCL implementation (spimagine/volumerender/kernels/utils.cl:41-70,
// volume_kernel.cl:48-93).  This file fixes one float32 operation sequence; the library
// is compiled with -ffp-contract=off and correctly rounded division/sqrt, and every fused
// multiply-add below is an explicit fmaf, so the sample positions are reproducible bit for
// bit on the host.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spr {

struct F4 {
    float x, y, z, w;
};

__device__ __forceinline__ F4 mk4(float x, float y, float z, float w) { return F4{x, y, z, w}; }
__device__ __forceinline__ F4 add4(F4 a, F4 b) { return mk4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ F4 sub4(F4 a, F4 b) { return mk4(a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w); }
__device__ __forceinline__ F4 scl4(float s, F4 a) { return mk4(s * a.x, s * a.y, s * a.z, s * a.w); }
__device__ __forceinline__ F4 mul4s(F4 a, float s) { return mk4(a.x * s, a.y * s, a.z * s, a.w * s); }

__device__ __forceinline__ float dot4(F4 a, F4 b) {
    float s = a.x * b.x;
    s = s + a.y * b.y;
    s = s + a.z * b.z;
    s = s + a.w * b.w;
    return s;
}

// v / |v| per component; the zero vector is returned unchanged
__device__ __forceinline__ F4 normalize4(F4 v) {
    float l2 = dot4(v, v);
    if (l2 == 0.0f) return v;
    float l = sqrtf(l2);
    return mk4(v.x / l, v.y / l, v.z / l, v.w / l);
}

// utils.cl:63-70: component i = dot(v, row i of the row-major 4x4 M)
__device__ __forceinline__ F4 mult(const float* M, F4 v) {
    F4 r;
    r.x = dot4(v, mk4(M[0], M[1], M[2], M[3]));
    r.y = dot4(v, mk4(M[4], M[5], M[6], M[7]));
    r.z = dot4(v, mk4(M[8], M[9], M[10], M[11]));
    r.w = dot4(v, mk4(M[12], M[13], M[14], M[15]));
    return r;
}

// OpenCL 1.2 section 6.12.4 min/max: min(x,y) = y<x ? y : x, max(x,y) = x<y ? y : x
__device__ __forceinline__ float cl_min(float x, float y) { return (y < x) ? y : x; }
__device__ __forceinline__ float cl_max(float x, float y) { return (x < y) ? y : x; }

// utils.cl:41-60 slab test (IEEE infinities for zero direction components)
__device__ __forceinline__ bool intersect_box(F4 ro, F4 rd, const float* box, float* tnear,
                                              float* tfar) {
    float ix = 1.0f / rd.x, iy = 1.0f / rd.y, iz = 1.0f / rd.z;
    float tbx = ix * (box[0] - ro.x), tby = iy * (box[2] - ro.y), tbz = iz * (box[4] - ro.z);
    float ttx = ix * (box[1] - ro.x), tty = iy * (box[3] - ro.y), ttz = iz * (box[5] - ro.z);
    float mnx = cl_min(ttx, tbx), mny = cl_min(tty, tby), mnz = cl_min(ttz, tbz);
    float mxx = cl_max(ttx, tbx), mxy = cl_max(tty, tby), mxz = cl_max(ttz, tbz);
    float lt = cl_max(cl_max(mnx, mny), cl_max(mnx, mnz));
    float st = cl_min(cl_min(mxx, mxy), cl_min(mxx, mxz));
    *tnear = lt;
    *tfar = st;
    return st > lt;
}

// volume_kernel.cl:48-77: pixel-corner mapping u = x/W*2-1 (not pixel centres)
__device__ __forceinline__ void ray_setup(const float* invP, const float* invM, uint32_t x,
                                          uint32_t y, uint32_t W, uint32_t H, F4* orig_o,
                                          F4* direc_o) {
    float u = ((float)x / (float)W) * 2.0f - 1.0f;
    float v = ((float)y / (float)H) * 2.0f - 1.0f;
    F4 front = mk4(u, v, -1.0f, 1.0f);
    F4 back = mk4(u, v, 1.0f, 1.0f);
    F4 orig0 = mult(invP, front);
    orig0 = mul4s(orig0, 1.0f / orig0.w);
    F4 orig = mult(invM, orig0);
    orig = mul4s(orig, 1.0f / orig.w);
    F4 temp = mult(invP, back);
    temp = mul4s(temp, 1.0f / temp.w);
    F4 direc = mult(invM, normalize4(sub4(temp, orig0)));
    direc.w = 0.0f;
    *orig_o = orig;
    *direc_o = direc;
}

// fused multiply-add lerp used by the trilinear sampler
__device__ __forceinline__ float lerpf(float p, float q, float t) { return fmaf(t, q - p, p); }

// utils.cl:10-23 uint32 LCG, 10 rounds, modulo 79197919
__device__ __forceinline__ float lcg_random(uint32_t x, uint32_t y) {
    uint32_t a = 4421u + (1u + x) * (1u + y) + x + y;
#pragma unroll
    for (int i = 0; i < 10; i++) a = (1664525u * a + 1013904223u) % 79197919u;
    return ((float)a * 1.0f) / (float)79197919;
}

// utils.cl:25-38 (declared float, returns an int value)
__device__ __forceinline__ float lcg_rand_int(uint32_t x, uint32_t y, int start, int end) {
    float rnd = lcg_random(x, y);
    return (float)(int)((float)start + rnd * (float)(end - start));
}

// Cephes single-precision sin/cos (S. L. Moshier, Cephes Math Library, sinf.c):
// Cody-Waite reduction by pi/2 and minimax polynomials on [-pi/4, pi/4].
__device__ __forceinline__ void spr_sincosf(float x, float* s, float* c) {
    const float TWO_OVER_PI = 0.636619772367581343f;
    const float DP1 = 1.5703125f, DP2 = 4.837512969970703125e-4f, DP3 = 7.54978995489188216e-8f;
    float k = rintf(x * TWO_OVER_PI);
    int q = (int)k;
    float r = fmaf(-k, DP1, x);
    r = fmaf(-k, DP2, r);
    r = fmaf(-k, DP3, r);
    float z = r * r;
    float sp = fmaf(fmaf(fmaf(-1.9515295891e-4f, z, 8.3321608736e-3f), z, -1.6666654611e-1f), z * r, r);
    float cp = fmaf(fmaf(fmaf(2.443315711809948e-5f, z, -1.388731625493765e-3f), z, 4.166664568298827e-2f),
                    z * z, fmaf(-0.5f, z, 1.0f));
    switch (q & 3) {
        case 0: *s = sp; *c = cp; break;
        case 1: *s = cp; *c = -sp; break;
        case 2: *s = -sp; *c = -cp; break;
        default: *s = -cp; *c = sp; break;
    }
}

}  // namespace spr
