"""Pins of the CPU oracle (oracle/spimref.c) against what the mathematics and the
reference's own fixtures fix -- not against a retyped copy of the oracle.

Each test states the quirk or formula of the reference it pins (file:line).
"""
import math
import os

import numpy as np
import pytest
from scipy.ndimage import correlate1d, map_coordinates

import oracle
from spimagine_amd import synth
from spimagine_amd.utils import (mat4_identity, mat4_ortho, mat4_perspective, mat4_rotation,
                                 mat4_translate)
from spimagine_amd.volumerender.volumerender import render_matrices

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ORTHO = mat4_ortho(-1, 1, -1, 1, -1, 1)


def ortho_mats(shape_zyx):
    nz, ny, nx = shape_zyx
    return render_matrices(mat4_identity(), ORTHO, (nx, ny, nz))


# ---------------------------------------------------------------- sampler (a6)

@pytest.mark.parametrize("dtype", [np.float32, np.uint16, np.uint8])
def test_linear_sampler_is_clamped_trilinear(oracle_lib, dtype):
    """OpenCL 1.2 s8.2 LINEAR + CLAMP_TO_EDGE == map_coordinates(order=1, mode='nearest')
    at texel coordinate s*N - 0.5."""
    rng = np.random.default_rng(1)
    vol = rng.uniform(0, 200, (5, 6, 7)).astype(dtype)
    nz, ny, nx = vol.shape
    pts = rng.uniform(-0.3, 1.3, (400, 3)).astype(np.float32)
    got = np.array([oracle.sample(vol, *p) for p in pts])
    n = np.array([nx, ny, nz], np.float32)
    u = (pts * n).astype(np.float32) - np.float32(0.5)
    exp = map_coordinates(vol.astype(np.float64), [u[:, 2], u[:, 1], u[:, 0]], order=1,
                          mode="nearest")
    np.testing.assert_allclose(got, exp, rtol=2e-6, atol=2e-4)


def test_nearest_sampler_is_clamped_floor(oracle_lib):
    rng = np.random.default_rng(2)
    vol = rng.uniform(-5, 5, (4, 9, 6)).astype(np.float32)
    nz, ny, nx = vol.shape
    pts = rng.uniform(-0.5, 1.5, (400, 3)).astype(np.float32)
    got = np.array([oracle.sample(vol, *p, interpolation="nearest") for p in pts])
    n = np.array([nx, ny, nz], np.float32)
    idx = np.floor((pts * n).astype(np.float32)).astype(int)
    idx = np.clip(idx, 0, np.array([nx, ny, nz]) - 1)
    np.testing.assert_array_equal(got, vol[idx[:, 2], idx[:, 1], idx[:, 0]])


def test_integer_voxels_interpolate_in_float(oracle_lib):
    """u16 + LINEAR is undefined in OpenCL 1.2; the documented choice is float32
    trilinear without rounding back to an integer: halfway between 1 and 2 gives 1.5."""
    vol = np.array([[[1, 2]]], np.uint16)
    assert oracle.sample(vol, 0.5, 0.5, 0.5) == 1.5


# ------------------------------------------------------- max projection (a4/a5)

@pytest.mark.parametrize("dtype", [np.float32, np.uint16])
def test_constant_volume_rend_const(oracle_lib, dtype):
    """tests/test_rendering/test_simple_rendering.py:85-95: 100^3 of 123, maxVal 200,
    4x4, identity modelView (eye inside the box), default perspective: every pixel is
    (123-0)/(200-0); alpha 1 for float (volume_kernel.cl:168), tnear (clamped to 0) for
    short (:329)."""
    data = (123 * np.ones((100, 100, 100))).astype(dtype)
    invP, invM = render_matrices(mat4_identity(), mat4_perspective(), (100, 100, 100))
    out, alpha = oracle.max_project(data, invP, invM, max_val=200., width=4, height=4)
    np.testing.assert_array_equal(out, np.float32(123.) / np.float32(200.))
    np.testing.assert_array_equal(alpha, 1.0 if dtype == np.float32 else 0.0)


def _ramp():
    vol = np.fromfile(os.path.join(GOLDEN, "raw_64_65_66.raw"), dtype=np.uint16)
    return vol.reshape(66, 65, 64)


def test_ramp_fixture_is_closed_form():
    """tests/data/raw_64_65_66.raw == floor(i * 10000 / (N-1)) over the flat index."""
    vol = _ramp().ravel().astype(np.int64)
    i = np.arange(vol.size)
    np.testing.assert_array_equal(vol, np.floor(i * 10000 / (vol.size - 1)).astype(np.int64))


@pytest.mark.parametrize("as_float", [False, True])
def test_ramp_fixture_kat(oracle_lib, as_float):
    """KAT on the reference's ramp fixture (66,65,64), orthographic view along -z.
    The ramp grows with z fastest, so the MIP of the clamped trilinear field is the
    first sample, on the top layer z=65, bilinear in (x, y):
    out = bilerp(layer 65, px*64-0.5, py*65-0.5) / 10000 with px = (1 + u*66/64)/2."""
    vol = _ramp()
    if as_float:
        vol = vol.astype(np.float32)
    W = H = 64
    invP, invM = ortho_mats(vol.shape)
    out, alpha = oracle.max_project(vol, invP, invM, min_val=0., max_val=10000., width=W,
                                    height=H, max_steps=128)
    u = np.arange(W) / W * 2 - 1
    v = np.arange(H) / H * 2 - 1
    ox, oy = u * 66 / 64, v * 66 / 65
    hitx, hity = np.abs(ox) < 1, np.abs(oy) < 1
    px, py = .5 * (1 + ox), .5 * (1 + oy)
    PX, PY = np.meshgrid(px * 64 - .5, py * 65 - .5)
    exp = map_coordinates(vol[65].astype(np.float64), [PY, PX], order=1, mode="nearest") / 1e4
    hit = hity[:, None] & hitx[None, :]
    assert hit.sum() > 3000
    np.testing.assert_allclose(out[hit], exp[hit], atol=2e-6)
    np.testing.assert_array_equal(out[~hit], 0.)
    miss_alpha = -1. if as_float else 0.
    np.testing.assert_array_equal(alpha[~hit], miss_alpha)
    np.testing.assert_array_equal(alpha[hit], 1. if as_float else 0.)


@pytest.mark.parametrize("dtype", [np.float32, np.uint16])
def test_nearest_ortho_is_numpy_max(oracle_lib, dtype):
    """Nearest + orthographic + axis-aligned: each odd pixel's ray walks one voxel column
    and every z layer (4 samples per layer), so the MIP is max(0, np.max(vol, axis=0))
    (Q3: the running max starts at 0), normalised by (min, max)."""
    N = 16
    rng = np.random.default_rng(3)
    if dtype == np.float32:
        vol = rng.uniform(-1, 1, (N, N, N)).astype(np.float32)
        lo, hi = 0., 0.
    else:
        vol = rng.integers(0, 4000, (N, N, N)).astype(np.uint16)
        lo, hi = 100., 4000.
    invP, invM = ortho_mats(vol.shape)
    out, _ = oracle.max_project(vol, invP, invM, min_val=lo, max_val=hi, width=2 * N,
                                height=2 * N, max_steps=64, interpolation="nearest")
    m = np.maximum(vol.astype(np.float64).max(axis=0), 0)
    if hi != 0:
        m = (m - lo) / (hi - lo)
    exp = np.clip(m, 0, 1)
    np.testing.assert_allclose(out[1::2, 1::2], exp, rtol=1e-6, atol=1e-7)


def _layered(values_by_z, N=16):
    vol = np.zeros((N, N, N), np.float32)
    for z in range(N):
        vol[z] = values_by_z(z)
    return vol


def test_opacity_mode_inner_break_quirk(oracle_lib):
    """volume_kernel.cl:136-157 (float): cumsum *= 1 - a^2 clamp(v,0,1) and
    `if (cumsum <= 0.01) break` leaves only the inner 16-loop (Q8).

    Nearest/ortho, N=16, S=64: 5 outer x 16 samples, sample k on layer floor(16-k/4).
    Layers z>=10 hold 0.5, z=9 holds 80, z<=8 hold 500; a^2 = 0.34 gives a per-sample
    factor f = 0.83 in the front, so cumsum = f^(k+1) after sample k and the break fires
    after sample 24 (outer 1).  Outer iterations 2,3,4 then take one sample each: k=25,26,27,
    all on layer 9.  Closed forms:
      inner-only break (reference)  max(0.5, 80 f^25)                          = 0.754
      break out of both loops       0.5
      no break                      max(80 f^25, 500 f^25 (1-0.34)^4)          = 0.894
    """
    vol = _layered(lambda z: 0.5 if z >= 10 else (80. if z == 9 else 500.))
    a = math.sqrt(0.34)
    invP, invM = ortho_mats(vol.shape)
    out, alpha = oracle.max_project(vol, invP, invM, min_val=0., max_val=1., alpha_pow=a,
                                    width=32, height=32, max_steps=64, interpolation="nearest")
    a2 = float(np.float32(np.float32(a) * np.float32(a)))
    f = 1 - a2 * 0.5
    reference = max(0.5, 80 * f ** 25)
    full_break = 0.5
    no_break = max(80 * f ** 25, 500 * f ** 25 * (1 - a2) ** 4)
    assert abs(reference - full_break) > .1 and abs(no_break - reference) > .1
    np.testing.assert_allclose(out[1::2, 1::2], reference, rtol=2e-5)
    np.testing.assert_array_equal(alpha[1::2, 1::2], 1.)


def test_opacity_mode_short_formula(oracle_lib):
    """volume_kernel.cl:312 (short): cumsum *= 1 - 0.1 a^2 v with v normalised and NOT
    clamped (Q7).  Layers z>=12 hold raw 50, z<12 raw 200, min 0, max 100, a = 1:
    front factor 1 - 0.1*0.5 = 0.95, so out = max(0.5, 2 * 0.95^17) = 0.836; the float
    formula (factor 1 - clamp(0.5) = 0.5) would give 0.5."""
    vol = _layered(lambda z: 50 if z >= 12 else 200).astype(np.uint16)
    invP, invM = ortho_mats(vol.shape)
    out, alpha = oracle.max_project(vol, invP, invM, min_val=0., max_val=100., alpha_pow=1.,
                                    width=32, height=32, max_steps=64, interpolation="nearest")
    f = 1 - np.float32(0.1) * 0.5
    np.testing.assert_allclose(out[1::2, 1::2], 2 * f ** 17, rtol=2e-5)
    np.testing.assert_array_equal(alpha[1::2, 1::2], 0.)


def test_multipass_shift_and_merge(oracle_lib):
    """volume_kernel.cl:100-106,172-183: part p shifts the ray origin by p*dt and parts>0
    merge with fmax.  Nearest/ortho, 64^3, box z in [0,1], S=64, numParts=2: dt = 1/32,
    48 samples, sample k of part p lies on layer 64-p-k, so part p's MIP is
    max(0, vol[17-p:64].max(axis=0))."""
    N = 64
    rng = np.random.default_rng(4)
    vol = rng.uniform(0, 1, (N, N, N)).astype(np.float32)
    vol[16] += 2.0  # only part 1 reaches layer 16
    invP, invM = ortho_mats(vol.shape)
    box = (-1, 1, -1, 1, 0, 1)
    kw = dict(box=box, width=2 * N, height=2 * N, max_steps=64, num_parts=2,
              interpolation="nearest", min_val=0., max_val=4.)
    parts = []
    for p in range(2):
        o, a = oracle.max_project(vol, invP, invM, current_part=p,
                                  out=np.full((2 * N, 2 * N), -np.inf, np.float32),
                                  alpha=np.full((2 * N, 2 * N), -np.inf, np.float32), **kw)
        parts.append(o)
        np.testing.assert_allclose(o[1::2, 1::2], vol[17 - p:].max(axis=0) / 4., rtol=1e-6)
    assert (parts[1] > parts[0] + .2)[1::2, 1::2].all()
    merged, _ = oracle.max_project(vol, invP, invM, current_part=0, **kw)
    merged, _ = oracle.max_project(vol, invP, invM, current_part=1, out=merged,
                                   alpha=np.zeros_like(merged), **kw)
    np.testing.assert_array_equal(merged, np.maximum(parts[0], parts[1]))
    with pytest.raises(ValueError):
        oracle.max_project(vol, invP, invM, max_steps=64, num_parts=5)


def test_perspective_float64_brute_force(oracle_lib):
    """Independent float64 march (vectorised numpy, scipy trilinear sampler) on a tiny
    volume under perspective + rotation + anisotropic units: same pixel-corner rays,
    ((S/16)+1)*16 samples from max(tnear,0), fixed dt = |tfar-tnear|/((S/16)*16)."""
    rng = np.random.default_rng(5)
    vol = rng.uniform(0, 1, (12, 10, 14)).astype(np.float32)
    nz, ny, nx = vol.shape
    units = (1., 1.2, 2.)
    P = mat4_perspective(45, 1., .1, 10)
    M = np.dot(mat4_translate(.1, -.2, -3.5), mat4_rotation(.7, .3, 1, .2))
    invP, invM = render_matrices(M, P, (nx, ny, nz), units)
    W, H, S = 24, 20, 40
    out, _ = oracle.max_project(vol, invP, invM, width=W, height=H, max_steps=S)
    iP, iM = invP.astype(np.float64).reshape(4, 4), invM.astype(np.float64).reshape(4, 4)
    exp = np.zeros((H, W))
    n = np.array([nx, ny, nz], np.float64)
    nsamp = (S // 16 + 1) * 16
    for y in range(H):
        for x in range(W):
            u, v = x / W * 2 - 1, y / H * 2 - 1
            o0 = iP @ [u, v, -1, 1]
            o0 /= o0[3]
            o = iM @ o0
            o /= o[3]
            t = iP @ [u, v, 1, 1]
            t /= t[3]
            d = t - o0
            d = iM @ (d / np.linalg.norm(d))
            with np.errstate(divide="ignore", invalid="ignore"):
                t0 = (-1 - o[:3]) / d[:3]
                t1 = (1 - o[:3]) / d[:3]
            tn = np.max(np.minimum(t0, t1))
            tf = np.min(np.maximum(t0, t1))
            if not tf > tn:
                continue
            tn = max(tn, 0.)
            dt = abs(tf - tn) / ((S // 16) * 16)
            k = np.arange(nsamp)
            pos = .5 * (1 + o[:3, None] + (tn + k * dt) * d[:3, None])
            uu = pos * n[:, None] - .5
            vals = map_coordinates(vol.astype(np.float64), [uu[2], uu[1], uu[0]], order=1,
                                   mode="nearest")
            exp[y, x] = np.clip(max(0., vals.max()), 0, 1)
    assert (exp > 0).sum() > 100
    np.testing.assert_allclose(out, exp, atol=2e-4)


# ------------------------------------------------------------ iso surface (a8)

def _linear_field(N=64, a=1., b=2., c=3.):
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    return (a * x + b * y + c * z).astype(np.float32)


def test_iso_linear_field_depth_bias_and_normal(oracle_lib):
    """iso_kernel.cl on a linear field (reproduced exactly by trilinear filtering),
    orthographic along -z.  Closed forms:
      crossing  t* = (63.5 - (iso - a ux - b uy)/c) / 32  (uz' = 63.5 - 32 t)
      depth     t_hit - t* in [1.0, 1.2) dt: t_hit accumulates from tnear + i dt although
                the refinement restarts at step i-1 (Q10)
      normal    -(8a, 8b, 6c)/|.|: weights (2,1) in x,y but (1,1) in z (Q11), flipped
                because the first sample is above the iso value
      colour    Phong 0.3 + 0.4 diffuse + 0.3 spec^10, light = normalize(2,-1,-2)."""
    a, b, c = 1., 2., 3.
    N = 64
    vol = _linear_field(N, a, b, c)
    iso = 200.
    S = 300
    invP, invM = ortho_mats(vol.shape)
    W = H = 64
    out, alpha, depth, normals = oracle.iso_surface(vol, invP, invM, iso_val=iso, max_steps=S,
                                                    width=W, height=H)
    dt = 2. / (S - 1)
    xs = np.arange(W)
    ux = (.5 * (1 + (xs / W * 2 - 1))) * N - .5
    UY, UX = np.meshgrid(ux, ux, indexing="ij")
    tstar = (63.5 - (iso - a * UX - b * UY) / c) / 32.
    inner = (UX > 3) & (UY > 3) & (UX < 60) & (UY < 60) & (tstar > .05) & (tstar < 1.9)
    assert inner.sum() > 1500
    bias = (depth - tstar)[inner] / dt
    assert bias.min() > 0.99 and bias.max() < 1.21, (bias.min(), bias.max())
    n_exp = -np.array([8 * a, 8 * b, 6 * c]) / np.linalg.norm([8 * a, 8 * b, 6 * c])
    np.testing.assert_allclose(normals[inner], np.broadcast_to(n_exp, normals[inner].shape),
                               atol=2e-5)
    light = np.array([2., -1., -2.]) / 3.
    ln = light @ n_exp
    r = 2 * ln * n_exp - light
    d = np.array([0., 0., -1.])
    spec = max(0., (r / np.linalg.norm(r)) @ d) ** 10
    col = .3 + .4 * max(0., ln) + (ln > 0) * .3 * spec
    np.testing.assert_allclose(out[inner], col, atol=2e-6)
    np.testing.assert_array_equal(alpha[inner], 0.)
    # no crossing anywhere: out 0, depth +inf, normal 0
    out2, _, depth2, n2 = oracle.iso_surface(vol, invP, invM, iso_val=1e6, max_steps=S, width=8,
                                             height=8)
    assert np.all(out2 == 0) and np.all(np.isinf(depth2)) and np.all(n2 == 0)


# ------------------------------------------------------------ post passes (a9-a12)

def _weights(Nh, k):
    ht = np.arange(Nh)
    return np.exp(-k * (ht - Nh / 2.) ** 2 / Nh / Nh)


@pytest.mark.parametrize("axis", ["x", "y"])
@pytest.mark.parametrize("vec,Nh,k", [(False, 5, 10.), (True, 7, 5.)])
def test_conv_is_normalised_correlation(oracle_lib, axis, vec, Nh, k):
    """convolve_2d.cl: taps at offsets ht - Nh/2 (int), weights centred at Nh/2.f (Q13),
    normalised by the in-range weight sum == correlate(img, w) / correlate(1, w) with zero
    padding."""
    rng = np.random.default_rng(6)
    H, W = 23, 31
    img = rng.uniform(-1, 1, (H, W, 3) if vec else (H, W)).astype(np.float32)
    got = oracle.conv(img, Nh, axis, vec=vec)
    w = _weights(Nh, k)
    assert abs(w[0] - w[-1]) > 1e-3  # the half-pixel shift makes the kernel asymmetric
    ax = 1 if axis == "x" else 0
    num = correlate1d(img.astype(np.float64), w, axis=ax, mode="constant", cval=0.)
    den = correlate1d(np.ones(img.shape), w, axis=ax, mode="constant", cval=0.)
    np.testing.assert_allclose(got, num / den, atol=2e-6)


def _py_random(x, y):
    """utils.cl:39-52 with Python integers and explicit 32-bit wrap-around."""
    m = 0xFFFFFFFF
    a = (4421 + (1 + x) * (1 + y) + x + y) & m
    for _ in range(10):
        a = ((1664525 * a + 1013904223) & m) % 79197919
    return np.float32(np.float32(a) / np.float32(79197919))


def test_lcg_matches_integer_arithmetic(oracle_lib):
    rng = np.random.default_rng(7)
    pts = list(rng.integers(0, 2 ** 20, (300, 2))) + [(0, 0), (2 ** 16 - 1, 2 ** 16 + 7),
                                                      (4000000000, 3)]
    for x, y in pts:
        assert oracle.random(int(x), int(y)) == _py_random(int(x), int(y))
    for i in range(40):
        r = _py_random(i, i * i)
        assert oracle.rand_int(i, i * i, 294, 97701) == float(int(np.float32(294) + r * np.float32(97701 - 294)))


def test_sincos_accuracy(oracle_lib):
    xs = np.linspace(0, 2 * np.pi, 2001).astype(np.float32)
    s, c = np.array([oracle.sincos(float(x)) for x in xs]).T
    np.testing.assert_allclose(s, np.sin(xs.astype(np.float64)), atol=3e-7)
    np.testing.assert_allclose(c, np.cos(xs.astype(np.float64)), atol=3e-7)


def test_occlusion_geometry(oracle_lib):
    """occlusion.cl:41-82 on depth = x (resp. y): a sample counts iff its offset
    r cos(phi) (resp. r sin(phi)) is negative, so occ = #{i: cos(phi_i) < 0} / n with
    phi_i = 2 pi random(x + rand_int(i^2, i, 0, 1997), y + rand_int(i, i^2, 569, 17633))
    computed from the integer LCG; a constant depth map has no occlusion."""
    W, H, n, radius = 20, 12, 30, 21
    X, Y = np.meshgrid(np.arange(W, dtype=np.float32), np.arange(H, dtype=np.float32))

    def py_rand_int(x, y, s, e):
        return float(int(np.float32(s) + _py_random(x, y) * np.float32(e - s)))

    occ_x = oracle.occlusion(X, radius, n)
    occ_y = oracle.occlusion(Y, radius, n)
    for yy in range(H):
        for xx in range(1, W):
            cnt_x = cnt_y = 0
            for i in range(n):
                ra = int(np.float32(xx) + np.float32(py_rand_int(i, i * i, 0, 1000)))
                rb = int(np.float32(yy) + np.float32(py_rand_int(i * i, i, 294, 97701)))
                r = radius * float(_py_random(ra, rb))
                pa = int(np.float32(xx) + np.float32(py_rand_int(i * i, i, 0, 1997)))
                pb = int(np.float32(yy) + np.float32(py_rand_int(i, i * i, 569, 17633)))
                phi = 6.2831853071795 * float(_py_random(pa, pb))
                cnt_x += (r > 0) and math.cos(phi) < 0
                if yy >= 1:
                    cnt_y += (r > 0) and math.sin(phi) < 0
            assert occ_x[yy, xx] == np.float32(cnt_x) / np.float32(n)
            if yy >= 1:
                assert occ_y[yy, xx] == np.float32(cnt_y) / np.float32(n)
    assert np.all(occ_x[:, 0] == 0) and np.all(occ_y[0] == 0)
    assert np.all(oracle.occlusion(np.full((H, W), 3., np.float32), radius, n) == 0)


def test_shading_phong_closed_form(oracle_lib):
    """iso_kernel.cl:505-588: Phong (0.5, 0.3, 0.2, shininess 10) from the normalised
    normal, AO blend (1-s)c + s c (1-occ), the identity line (1-s)c + s c, and masking by
    depth < inf; per-pixel view direction from the un-projected near/far points."""
    rng = np.random.default_rng(8)
    H, W = 9, 11
    normals = rng.normal(size=(H, W, 3)).astype(np.float32)
    normals[0, 0] = 0
    depth = rng.uniform(1, 3, (H, W)).astype(np.float32)
    depth[2, 3] = np.inf
    occ = rng.uniform(0, 1, (H, W)).astype(np.float32)
    P = mat4_perspective(50, 1., .5, 8)
    M = np.dot(mat4_translate(0, 0, -4), mat4_rotation(.4, 1, 1, 0))
    invP, invM = render_matrices(M, P, (10, 10, 10))
    s = .35
    got = oracle.shading(normals, depth, occ, invP, invM, occ_strength=s)
    iP, iM = invP.astype(np.float64).reshape(4, 4), invM.astype(np.float64).reshape(4, 4)
    light = iM @ [2, -1, -2, 0]
    light = light[:3] / np.linalg.norm(light[:3])
    for y in range(H):
        for x in range(W):
            u, v = x / W * 2 - 1, y / H * 2 - 1
            o0 = iP @ [u, v, -1, 1]
            o0 /= o0[3]
            t = iP @ [u, v, 1, 1]
            t /= t[3]
            d = iM @ ((t - o0) / np.linalg.norm(t - o0))
            d = d[:3] / np.linalg.norm(d[:3])
            nrm = normals[y, x].astype(np.float64)
            nn = np.linalg.norm(nrm)
            nrm = nrm / nn if nn > 0 else nrm
            ln = light @ nrm
            r = 2 * ln * nrm - light
            spec = max(0., (r / np.linalg.norm(r)) @ d) ** 10
            col = .5 + .3 * max(0., ln) + (ln > 0) * .2 * spec
            col = (1 - s) * col + s * col * (1 - occ[y, x])
            col = (1 - s) * col + s * col
            col *= np.isfinite(depth[y, x])
            assert abs(got[y, x] - col) < 3e-6, (y, x, got[y, x], col)


def test_iso_pipeline_order(oracle_lib):
    """volumerender.py:446-506: output = shading(conv_vec_y(conv_vec_x(normals, 7)),
    depth, conv_y(conv_x(occlusion(depth), 5))) -- render_iso equals the composition."""
    vol = _linear_field(24) / 10
    P = mat4_perspective(45, 1., .1, 10)
    M = np.dot(mat4_translate(0, 0, -3.5), mat4_rotation(.5, 0, 1, 0))
    invP, invM = render_matrices(M, P, (24, 24, 24))
    kw = dict(iso_val=6., max_steps=80, width=40, height=36)
    out, alpha, depth, normals, occ = oracle.render_iso(vol, invP, invM, occ_strength=.3,
                                                        occ_radius=7, occ_n_points=12, **kw)
    o1, a1, d1, n1 = oracle.iso_surface(vol, invP, invM, **kw)
    n2 = oracle.conv(oracle.conv(n1, 7, "x", vec=True), 7, "y", vec=True)
    oc = oracle.conv(oracle.conv(oracle.occlusion(d1, 7, 12), 5, "x"), 5, "y")
    np.testing.assert_array_equal(alpha, a1)
    np.testing.assert_array_equal(depth, d1)
    np.testing.assert_array_equal(normals, n2)
    np.testing.assert_array_equal(occ, oc)
    np.testing.assert_array_equal(out, oracle.shading(n2, d1, oc, invP, invM, occ_strength=.3))
    assert np.isfinite(depth).sum() > 100


@pytest.mark.parametrize("rows", [(0, 9), (41, 58), (110, 120), (5, 6)])
def test_iso_rows_pipeline_is_the_full_pipeline(oracle_lib, rows):
    """render_iso_rows (the band form the full-size GPU tests compare against) equals
    render_iso on its rows: it marches only the rows within max(3, occ_radius + 3) of the
    band, so a margin too small (or a wrong row offset) would change the blurred normals,
    the occlusion or the shading of the band's edge rows."""
    vol = synth.blob_volume((40, 48, 44), n_blobs=12, sigma=(3., 6.), amp=(500., 3000.),
                            background=100., poisson=True, read_noise=5., dtype="uint16", seed=3)
    W, H = 90, 120
    M = np.dot(mat4_translate(0, 0, -3.2), mat4_rotation(.7, .3, 1, -.2))
    invP, invM = render_matrices(M, mat4_perspective(45, W / H, .1, 10), vol.shape[::-1])
    lo, hi = float(vol.min()), float(vol.max())
    kw = dict(iso_val=np.float32(lo + .3 * (hi - lo)), max_steps=120, occ_strength=.3,
              occ_radius=11, occ_n_points=20, width=W, height=H)
    full = oracle.render_iso(vol, invP, invM, **kw)
    band = oracle.render_iso_rows(vol, invP, invM, rows, **kw)
    r0, r1 = rows
    for a, b in zip(full, band):
        np.testing.assert_array_equal(a[r0:r1], b[r0:r1])
    assert np.isfinite(full[2]).sum() > 1000


# ------------------------------------------------------------------ gamma != 1

@pytest.mark.parametrize("gamma", [.5, 1.7, 2.])
def test_mip_gamma_closed_form(oracle_lib, gamma):
    """volume_kernel.cl:162: out = clamp(pow((v - min)/(max - min), gamma), 0, 1).  On the
    rend_const volume (123, maxVal 200) every pixel is (123/200)^gamma (float64 closed
    form); gamma 1 would give 0.615 and an exponent of gamma^2 or 1/gamma other values."""
    data = (123 * np.ones((40, 40, 40))).astype(np.uint16)
    invP, invM = render_matrices(mat4_identity(), mat4_perspective(), (40, 40, 40))
    out, _ = oracle.max_project(data, invP, invM, max_val=200., gamma=gamma, width=6, height=5)
    want = (123. / 200.) ** gamma
    np.testing.assert_allclose(out, want, rtol=3e-7)
    for wrong in ((123. / 200.) ** (gamma * gamma), (123. / 200.) ** (1. / gamma), .615):
        assert abs(want - wrong) > 1e-3


def test_iso_gradient_step_is_dt_gamma_squared(oracle_lib):
    """iso_kernel.cl:178-196: the normal is a 4-tap difference at h = dt * gamma^2 (in
    normalised texture units) with weights (2, 1) in x, y and (1, 1) in z.  Field
    f = q(x) + b y + c z with q(x) = x^3/64: trilinear filtering reproduces the y and z
    terms exactly and q piecewise linearly (np.interp of the voxel values), so along x the
    4-tap difference depends on h -- the float64 closed form with h = dt gamma^2 matches
    the oracle's normals, h = dt gamma or h = dt does not.  Orthographic view along -z: x
    and y are constant along each ray."""
    N, b, c = 32, 4., 30.
    z, y, x = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    vol = (x ** 3 / 64. + b * y + c * z).astype(np.float32)
    gamma, S = 1.6, 40
    invP, invM = ortho_mats(vol.shape)
    W = H = 32
    iso = 600.
    out, alpha, depth, normals = oracle.iso_surface(vol, invP, invM, iso_val=iso, gamma=gamma,
                                                    max_steps=S, width=W, height=H)
    dt = 2. / (S - 1)
    q = np.arange(N, dtype=np.float64) ** 3 / 64.

    def F(s):  # q interpolated at normalised coordinate s (texel s N - 0.5, clamped)
        return np.interp(s * N - .5, np.arange(N), q)

    xs = np.arange(W)
    sx = .5 * (1 + (xs / W * 2 - 1))       # normalised x of pixel column xs
    inner = (sx * N - .5 > 10) & (sx * N - .5 < 21)
    rows = slice(12, 21)  # y taps (up to 2 h = 8.4 voxels) stay inside the volume
    hit = np.isfinite(depth[rows][:, inner])
    assert hit.all()

    def normal(h, flip):
        dx = 2 * (F(sx + h) - F(sx - h)) + F(sx + 2 * h) - F(sx - 2 * h)
        dy = np.full_like(dx, 8 * h * N * b)
        dz = np.full_like(dx, 6 * h * N * c)
        n = np.stack([dx, dy, dz], -1)
        return flip * n / np.linalg.norm(n, axis=-1, keepdims=True)

    # the first sample (front face, z = N - 1) lies above the iso value: the normal flips
    want = normal(dt * gamma * gamma, -1.)[inner]
    got = normals[rows][:, inner]
    np.testing.assert_allclose(got, np.broadcast_to(want, got.shape), atol=3e-5)
    for h in (dt * gamma, dt):
        assert np.abs(normal(h, -1.)[inner] - want).max() > 1e-3
