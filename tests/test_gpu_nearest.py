"""GPU: the nearest filter staged through the per-wave LDS slabs (max projection and
opacity mode, brick skipping on and off) gives every output bit for bit as the unstaged
nearest kernel (slab_kb 0: every sample gathered from global memory), for every voxel
type, power-of-two and other volume shapes, several poses and step counts; and matches
the oracle (interpolation="nearest").
"""
import sys

import numpy as np
import pytest

import oracle
from spimagine_amd import _lib, synth
from spimagine_amd.utils import mat4_perspective, mat4_rotation, mat4_translate

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def VR():
    assert _lib.device_count() > 0, "no HIP device"
    from spimagine_amd.volumerender import VolumeRenderer
    return VolumeRenderer


POSES = [np.dot(mat4_translate(0, 0, -3.4), mat4_rotation(.7, .3, 1, -.2)),
         np.dot(mat4_translate(.2, -.1, -2.8), mat4_rotation(2.3, 1, .2, .5)),
         np.dot(mat4_translate(0, 0, -4.), mat4_rotation(0., 0, 1, 0))]


def _volume(dtype, shape, seed):
    amp = (.3, 1.) if dtype == np.float32 else (500., 3000.) if dtype == np.uint16 else (60., 200.)
    return synth.blob_volume(shape, n_blobs=10, sigma=(2., 5.), amp=amp, background=amp[0] / 10,
                             poisson=False, read_noise=0., dtype=dtype, seed=seed)


def _render(VR, vol, pose, W, H, steps, alpha_pow, tuning):
    rend = VR((W, H), interpolation="nearest")
    for k, v in tuning.items():
        rend._ctx.set_tuning(k, v)
    rend.set_max_steps(steps)
    rend.set_data(vol)
    rend.set_alpha_pow(alpha_pow)
    rend.set_projection(mat4_perspective(45, W / float(H), .1, 10))
    rend.set_modelView(pose)
    lo, hi = float(vol.min()), float(vol.max())
    rend.render(minVal=lo, maxVal=hi)
    return rend.output.copy(), rend.output_alpha.copy(), rend


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8, np.float32])
@pytest.mark.parametrize("shape", [(45, 70, 61), (64, 64, 64)])
@pytest.mark.parametrize("alpha_pow", [0., .8])
def test_nearest_staged_equals_global(VR, dtype, shape, alpha_pow):
    say("nearest: %s %s alpha_pow %s" % (np.dtype(dtype).name, shape, alpha_pow))
    vol = _volume(dtype, shape, seed=len(shape) + int(10 * alpha_pow))
    for k, pose in enumerate(POSES):
        steps = (160, 96, 224)[k]
        ref = _render(VR, vol, pose, 120, 96, steps, alpha_pow, {"slab_kb": 0})
        for skip in (0, 1):
            got = _render(VR, vol, pose, 120, 96, steps, alpha_pow, {"skip": skip})
            np.testing.assert_array_equal(got[0], ref[0], err_msg="pose %d skip %d" % (k, skip))
            np.testing.assert_array_equal(got[1], ref[1])
        assert (ref[0] > .05).sum() > 200
    rend = got[2]
    lo, hi = float(vol.min()), float(vol.max())
    o, a = oracle.max_project(vol, rend._invP, rend._invM, min_val=lo, max_val=hi,
                              alpha_pow=alpha_pow, max_steps=224, width=120, height=96,
                              interpolation="nearest")
    assert np.abs(got[0] - o).max() <= 1e-4
    np.testing.assert_array_equal(got[1], a)


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8, np.float32])
def test_nearest_iso_staged_equals_global(VR, dtype):
    """The iso-surface march with the nearest filter staged (brick skipping on and off)
    against the unstaged kernel: all five outputs bit for bit; and the oracle."""
    vol = _volume(dtype, (50, 66, 58), seed=7)
    lo, hi = float(vol.min()), float(vol.max())
    iso_max = 2 * (lo + .3 * (hi - lo))
    W, H = 120, 100

    def iso(pose, tuning):
        rend = VR((W, H), interpolation="nearest")
        for k, v in tuning.items():
            rend._ctx.set_tuning(k, v)
        rend.set_max_steps(180)
        rend.set_data(vol)
        rend.set_projection(mat4_perspective(45, W / float(H), .1, 10))
        rend.set_modelView(pose)
        rend.render(method="iso_surface", maxVal=iso_max)
        return [rend.output.copy(), rend.output_alpha.copy(), rend.output_depth.copy(),
                rend.output_normals.copy(), rend.output_occlusion.copy()], rend

    for k, pose in enumerate(POSES):
        ref, _ = iso(pose, {"slab_kb": 0})
        for skip in (0, 1):
            got, rend = iso(pose, {"skip": skip})
            for a, b in zip(got, ref):
                np.testing.assert_array_equal(a, b, err_msg="pose %d skip %d" % (k, skip))
    want = oracle.render_iso(vol, rend._invP, rend._invM, iso_val=np.float32(iso_max) / 2,
                             max_steps=180, occ_strength=rend.occ_strength,
                             occ_radius=rend.occ_radius, occ_n_points=rend.occ_n_points,
                             width=W, height=H, interpolation="nearest")
    hit = np.isfinite(want[2])
    assert hit.sum() > 150
    np.testing.assert_array_equal(np.isfinite(got[2]), hit)
    np.testing.assert_array_equal(got[1], want[1])
    assert np.abs(got[3] - want[3]).max() <= 1e-4
    assert np.abs(got[0] - want[0]).max() <= 1e-4
