"""GPU: empty-space skipping and early ray termination by brick bounds (SPR_TUNE_SKIP)
leave every output bit-identical to the full march, and to the oracle within the bars.

Volumes with empty space (blobs on a constant background) make the skip paths run: boxes
of background behind a blob's maximum (max projection), boxes wholly below the iso value
(iso march), rays whose transparency is used up (opacity mode with a strong alpha_pow).
"""
import sys

import numpy as np
import pytest

import oracle
from spimagine_amd import _lib, synth
from spimagine_amd.utils import mat4_perspective, mat4_rotation, mat4_translate

pytestmark = pytest.mark.gpu

TOL = 1e-4


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def VR():
    assert _lib.device_count() > 0, "no HIP device"
    from spimagine_amd.volumerender import VolumeRenderer
    return VolumeRenderer


def volume(dtype, shape, seed, nan=False):
    if dtype == np.uint8:
        amp, bg = (60., 200.), 6.
    elif dtype == np.uint16:
        amp, bg = (500., 3000.), 100.
    else:
        amp, bg = (.3, 1.), .02
    v = synth.blob_volume(shape, n_blobs=10, sigma=(2., 4.), amp=amp, background=bg,
                          poisson=False, read_noise=0., dtype=dtype, seed=seed)
    if nan:
        v[shape[0] // 2, shape[1] // 3, shape[2] // 4] = np.nan
    return v


POSES = [np.dot(mat4_translate(0, 0, -3.4), mat4_rotation(.7, .3, 1, -.2)),
         np.dot(mat4_translate(.2, -.1, -2.8), mat4_rotation(2.3, 1, .2, .5)),
         np.dot(mat4_translate(0, 0, -4.), mat4_rotation(0., 0, 1, 0))]


def render_pair(VR, vol, pose, W, H, steps, method, kw, tuning=None):
    """The same render with skipping off and on: {skip: [outputs]}."""
    res = {}
    for skip in (0, 1):
        rend = VR((W, H))
        rend._ctx.set_tuning("skip", skip)
        for k, v in (tuning or {}).items():
            rend._ctx.set_tuning(k, v)
        rend.set_max_steps(steps)
        rend.set_data(vol)
        rend.set_projection(mat4_perspective(45, W / float(H), .1, 10))
        rend.set_modelView(pose)
        rend.render(method=method, **kw)
        out = [rend.output.copy(), rend.output_alpha.copy()]
        if method == "iso_surface":
            out += [rend.output_depth.copy(), rend.output_normals.copy(),
                    rend.output_occlusion.copy()]
        res[skip] = (out, rend)
    return res


def assert_same(res):
    for a, b in zip(res[0][0], res[1][0]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8, np.float32])
@pytest.mark.parametrize("shape", [(45, 70, 61), (64, 64, 64)])
@pytest.mark.parametrize("pose", [0, 1, 2])
def test_skip_max_project(VR, dtype, shape, pose):
    vol = volume(dtype, shape, seed=pose)
    lo, hi = float(np.nanmin(vol)), float(np.nanmax(vol))
    W, H = 120, 96
    res = render_pair(VR, vol, POSES[pose], W, H, 160, "max_project", dict(minVal=lo, maxVal=hi))
    assert_same(res)
    rend = res[1][1]
    o, a = oracle.max_project(vol, rend._invP, rend._invM, min_val=lo, max_val=hi, max_steps=160,
                              width=W, height=H)
    assert np.abs(res[1][0][0] - o).max() <= TOL
    np.testing.assert_array_equal(res[1][0][1], a)
    assert (res[1][0][0] > .05).sum() > 200


@pytest.mark.parametrize("dtype,alpha_pow", [(np.uint16, .6), (np.uint16, 3.), (np.uint8, 2.),
                                             (np.float32, .6), (np.float32, .95)])
@pytest.mark.parametrize("pose", [0, 1])
def test_skip_opacity(VR, dtype, alpha_pow, pose):
    """Strong alpha_pow: rays use up their transparency inside a blob, and the wave stops
    once no later sample can raise any ray's colour."""
    vol = volume(dtype, (48, 64, 56), seed=10 + pose)
    lo, hi = float(vol.min()), float(vol.max())
    W, H = 112, 90
    kw = dict(minVal=lo, maxVal=hi)
    res = {}
    for skip in (0, 1):
        rend = VR((W, H))
        rend._ctx.set_tuning("skip", skip)
        rend.set_max_steps(128)
        rend.set_data(vol)
        rend.set_alpha_pow(alpha_pow)
        rend.set_projection(mat4_perspective(45, W / float(H), .1, 10))
        rend.set_modelView(POSES[pose])
        rend.render(**kw)
        res[skip] = ([rend.output.copy(), rend.output_alpha.copy()], rend)
    assert_same(res)
    rend = res[1][1]
    o, a = oracle.max_project(vol, rend._invP, rend._invM, min_val=lo, max_val=hi,
                              alpha_pow=alpha_pow, max_steps=128, width=W, height=H)
    assert np.abs(res[1][0][0] - o).max() <= TOL
    np.testing.assert_array_equal(res[1][0][1], a)


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8, np.float32])
@pytest.mark.parametrize("pose", [0, 1, 2])
def test_skip_iso(VR, dtype, pose):
    vol = volume(dtype, (50, 66, 58), seed=20 + pose)
    lo, hi = float(vol.min()), float(vol.max())
    W, H = 120, 100
    iso_max = 2 * (lo + .3 * (hi - lo))
    res = render_pair(VR, vol, POSES[pose], W, H, 180, "iso_surface", dict(maxVal=iso_max))
    assert_same(res)
    rend = res[1][1]
    want = oracle.render_iso(vol, rend._invP, rend._invM, iso_val=np.float32(iso_max) / 2,
                             max_steps=180, occ_strength=rend.occ_strength,
                             occ_radius=rend.occ_radius, occ_n_points=rend.occ_n_points,
                             width=W, height=H)
    got = res[1][0]
    hit = np.isfinite(want[2])
    assert hit.sum() > 300
    np.testing.assert_array_equal(np.isfinite(got[2]), hit)
    np.testing.assert_array_equal(got[1], want[1])
    assert np.abs(got[3] - want[3]).max() <= TOL
    assert np.abs(got[0] - want[0]).max() <= TOL


@pytest.mark.parametrize("dtype", [np.uint16, np.float32])
@pytest.mark.parametrize("interp", ["linear", "nearest"])
def test_iso_one_and_four_waves(VR, dtype, interp):
    """The iso march with one wave per workgroup (the default) and with four: all five
    outputs bit for bit, skipping on and off, three poses; 16-, 8- and 4-sample boxes (180,
    48 and 16 steps through the volume)."""
    vol = volume(dtype, (50, 66, 58), seed=41)
    lo, hi = float(vol.min()), float(vol.max())
    iso_max = 2 * (lo + .3 * (hi - lo))
    W, H = 136, 104  # ragged: 9 x 7 tiles, partial quarters
    for k, pose in enumerate(POSES):
        steps = (180, 48, 16)[k]
        outs = {}
        for waves in (1, 4):
            for skip in (0, 1):
                rend = VR((W, H), interpolation=interp)
                rend._ctx.set_tuning("wg_waves", waves)
                rend._ctx.set_tuning("skip", skip)
                rend.set_max_steps(steps)
                rend.set_data(vol)
                rend.set_projection(mat4_perspective(45, W / float(H), .1, 10))
                rend.set_modelView(pose)
                rend.render(method="iso_surface", maxVal=iso_max)
                outs[waves, skip] = [rend.output.copy(), rend.output_alpha.copy(),
                                     rend.output_depth.copy(), rend.output_normals.copy(),
                                     rend.output_occlusion.copy()]
        assert np.isfinite(outs[4, 0][2]).sum() > 300, (k, steps)
        for key in ((1, 0), (1, 1), (4, 1)):
            for a, b in zip(outs[key], outs[4, 0]):
                np.testing.assert_array_equal(a, b, err_msg="pose %d waves/skip %s" % (k, key))


def test_skip_four_waves_multipass_and_nan(VR):
    """Four waves per workgroup (SPR_TUNE_WG_WAVES 4), multipass parts, and a
    float volume holding a NaN voxel (its brick never skips): skipping changes nothing."""
    vol = volume(np.uint16, (40, 56, 50), seed=31)
    lo, hi = float(vol.min()), float(vol.max())
    for parts in (1, 3):
        outs = {}
        for skip in (0, 1):
            rend = VR((100, 84))
            rend._ctx.set_tuning("skip", skip)
            rend._ctx.set_tuning("wg_waves", 4)
            rend.set_max_steps(144)
            rend.set_data(vol)
            rend.set_modelView(POSES[0])
            for part in range(parts):
                rend.render(minVal=lo, maxVal=hi, numParts=parts, currentPart=part)
            outs[skip] = (rend.output.copy(), rend.output_alpha.copy())
        np.testing.assert_array_equal(outs[0][0], outs[1][0])
        np.testing.assert_array_equal(outs[0][1], outs[1][1])
    fvol = volume(np.float32, (40, 56, 50), seed=32, nan=True)
    for method, kw in (("max_project", dict(minVal=0., maxVal=1.)),
                       ("iso_surface", dict(maxVal=.6))):
        res = render_pair(VR, fvol, POSES[1], 96, 80, 150, method, kw)
        assert_same(res)
