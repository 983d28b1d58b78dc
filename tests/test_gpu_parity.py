"""GPU parity: libspimrender.so (through the C ABI, driven by VolumeRenderer) against
the CPU oracle on the same seeded volumes and poses.

Tolerances: `output` max |diff| <= 1e-4 in normalised output units;
`output_alpha` exact; iso `output_depth` on the same first-hit voxel, normals <= 1e-4.
The kernels evaluate the oracle's float32 operation sequence, so most of these are in
fact bit-identical; the tolerance covers powf/libm differences (gamma != 1, specular).
"""
import os
import sys

import numpy as np
import pytest

import oracle
from spimagine_amd import _lib, synth
from spimagine_amd.utils import (mat4_identity, mat4_ortho, mat4_perspective, mat4_rotation,
                                  mat4_translate)

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


def oracle_mip(rend, vol, num_parts=1, current_part=0, out=None, alpha=None):
    return oracle.max_project(vol, rend._invP, rend._invM, box=rend.boxBounds,
                              min_val=rend.minVal, max_val=rend.maxVal, gamma=rend.gamma,
                              alpha_pow=rend.alphaPow, max_steps=rend.max_steps,
                              num_parts=num_parts, current_part=current_part, width=rend.width,
                              height=rend.height, interpolation=rend.interpolation, out=out,
                              alpha=alpha)


def test_native_library_is_the_one_loaded(VR):
    L = _lib.load()
    assert os.path.dirname(os.path.abspath(L._name)) == os.path.dirname(os.path.abspath(_lib.__file__))
    assert _lib.device_count() >= 1


def test_parity_config_mip(VR):
    """128^3 float32 blobs -> 256^2, perspective 45 deg, translate(0,0,-4).rotY(30 deg),
    max_steps 224 (the CPU-runnable parity workload)."""
    vol = synth.config_volume("parity")
    M, P = synth.headline_camera()
    rend = VR((256, 256))
    rend.set_max_steps(224)
    rend.set_data(vol)
    rend.set_projection(P)
    rend.set_modelView(M)
    rend.render(minVal=float(vol.min()), maxVal=float(vol.max()))
    out, alpha = oracle_mip(rend, vol)
    d = np.abs(rend.output - out)
    say("parity config: max|diff| %.3g, bit-identical %.4f" % (d.max(), np.mean(d == 0)))
    assert d.max() <= TOL
    np.testing.assert_array_equal(rend.output_alpha, alpha)
    assert (alpha == 1).sum() > 10000


CASES = [
    # dtype, interpolation, alpha_pow, gamma
    (np.float32, "linear", 0., 1.),
    (np.uint16, "linear", 0., 1.),
    (np.uint8, "linear", 0., 1.),
    (np.float32, "nearest", 0., 1.),
    (np.uint16, "nearest", 0., 1.),
    (np.float32, "linear", .6, 1.),
    (np.uint16, "linear", .6, 1.),
    (np.uint8, "linear", .3, .8),
    (np.float32, "linear", 0., .7),
]


@pytest.mark.parametrize("dtype,interp,alpha_pow,gamma", CASES)
def test_mip_modes(VR, dtype, interp, alpha_pow, gamma):
    """Ragged shapes (volume 37x50x29, frame 97x83), anisotropic units, a shrunken box
    (so the 16-sample overshoot reads real data outside it), random pose."""
    shape = (29, 50, 37)
    if dtype == np.uint8:
        amp, bg = (60., 150.), 20.
    elif dtype == np.uint16:
        amp, bg = (500., 3000.), 100.
    else:
        amp, bg = (.3, 1.), .05
    vol = synth.blob_volume(shape, n_blobs=12, sigma=(2., 5.), amp=amp, background=bg,
                            poisson=dtype != np.float32, read_noise=0., dtype=dtype, seed=11)
    rend = VR((97, 83), interpolation=interp)
    rend.set_max_steps(96)
    rend.set_data(vol)
    rend.set_units([1., .8, 1.7])
    rend.set_box_boundaries([-.9, .8, -1, .7, -.6, 1])
    rend.set_alpha_pow(alpha_pow)
    rend.set_projection(mat4_perspective(50, 97 / 83., .1, 10))
    rend.set_modelView(np.dot(mat4_translate(.1, 0, -3.2), mat4_rotation(1.1, .3, 1, -.4)))
    rend.render(minVal=float(vol.min()), maxVal=float(vol.max()), gamma=gamma)
    out, alpha = oracle_mip(rend, vol)
    d = np.abs(rend.output - out)
    say("%s %s a=%s g=%s: max|diff| %.3g, bit-identical %.4f" % (
        np.dtype(dtype).name, interp, alpha_pow, gamma, d.max(), np.mean(d == 0)))
    assert d.max() <= TOL
    np.testing.assert_array_equal(rend.output_alpha, alpha)


@pytest.mark.parametrize("dtype", [np.float32, np.uint16, np.uint8])
@pytest.mark.parametrize("shape", [(29, 50, 37), (32, 64, 16), (1, 1, 1), (2, 3, 1)])
def test_layouts_are_identical(VR, dtype, shape):
    """The padded (apron) layout and the linear layout give bit-identical frames for every
    voxel type, both filters and both render modes, including degenerate volumes."""
    rng = np.random.default_rng(13)
    vol = (rng.uniform(0, 200, shape)).astype(dtype)
    outs = {}
    for layout in (_lib.LAYOUT_LINEAR, _lib.LAYOUT_PADDED):
        res = []
        for interp in ("linear", "nearest"):
            rend = VR((61, 45), interpolation=interp)
            rend._ctx.set_layout(layout)
            rend.set_max_steps(64)
            rend.set_data(vol)
            assert rend._ctx.layout() == layout
            rend.set_modelView(np.dot(mat4_translate(0, .1, -3.), mat4_rotation(.9, 1, .5, .2)))
            rend.render(maxVal=200.)
            res.append(rend.output.copy())
            rend.render(maxVal=200., method="iso_surface")
            res += [rend.output.copy(), rend.output_depth.copy(), rend.output_normals.copy()]
        outs[layout] = res
    for a, b in zip(outs[_lib.LAYOUT_LINEAR], outs[_lib.LAYOUT_PADDED]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8, np.float32])
@pytest.mark.parametrize("shape,pose", [((64, 128, 96), 0), ((33, 70, 51), 1), ((128, 64, 64), 2),
                                        ((16, 16, 16), 3)])
def test_staged_mip_matches_global(VR, dtype, shape, pose):
    """The LDS-staged max projection (per-wave slabs, with the global fallback for boxes
    larger than the slab) is bit-identical to the unstaged kernel for every slab size,
    including multipass parts and poses whose rays graze the box."""
    vol = synth.blob_volume(shape, n_blobs=15, sigma=(2., 6.), amp=(100., 250.), background=5.,
                            poisson=False, read_noise=0., dtype=dtype, seed=pose)
    M = [np.dot(mat4_translate(0, 0, -4), mat4_rotation(np.pi / 6, 0, 1, 0)),
         np.dot(mat4_translate(.3, -.2, -2.6), mat4_rotation(2.1, 1, .4, -.3)),
         np.dot(mat4_translate(0, 0, -1.2), mat4_rotation(.8, .2, 1, .1)),  # eye close to the box
         mat4_identity()][pose]                                          # eye inside the box
    frames = {}
    # (slab KiB, workgroup staging, row pad, chunk, waves per workgroup, tile map)
    for key in ((0, 0, 0, 0, 4, 0), (8, 0, 0, 0, 4, 0), (12, 0, 0, 0, 4, 0), (16, 0, 0, 0, 4, 0),
                (19, 0, 0, 0, 4, 0), (8, 1, 0, 0, 4, 0), (12, 1, 0, 0, 4, 0), (16, 0, 1, 0, 4, 0),
                (12, 1, 1, 0, 4, 0), (16, 0, 0, 32, 4, 0), (16, 0, 0, 64, 4, 0),
                (12, 1, 0, 32, 4, 0), (19, 1, 0, 0, 4, 0), (19, 0, 0, 0, 1, 0), (16, 0, 0, 0, 1, 0),
                (19, 0, 0, 32, 1, 0), (19, 0, 0, 0, 4, 3), (19, 0, 0, 0, 1, 3), (16, 0, 0, 64, 1, 3),
                (12, 0, 0, 0, 4, 3), (12, 1, 0, 0, 1, 3)):
        slab, staging, pad, chunk, waves, tile_map = key
        rend = VR((130, 97))
        rend._ctx.set_tuning("slab_kb", slab)
        rend._ctx.set_tuning("staging", staging)
        rend._ctx.set_tuning("row_pad", pad)
        rend._ctx.set_tuning("chunk", chunk)
        rend._ctx.set_tuning("wg_waves", waves)
        rend._ctx.set_tuning("tile_map", tile_map)
        rend.set_max_steps(160)
        rend.set_data(vol)
        rend.set_modelView(M)
        rend.render(minVal=float(vol.min()), maxVal=float(vol.max()))
        res = [rend.output.copy(), rend.output_alpha.copy()]
        for part in range(3):
            rend.render(numParts=3, currentPart=part)
        res.append(rend.output.copy())
        frames[key] = res
    ref = frames[0, 0, 0, 0, 4, 0]
    for key in frames:
        for a, b in zip(ref, frames[key]):
            np.testing.assert_array_equal(a, b, err_msg=str(key))
    assert (ref[0] > 0).sum() > 200


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8, np.float32])
@pytest.mark.parametrize("max_steps", [32, 64, 96])
def test_staged_mip_long_steps(VR, dtype, max_steps):
    """Ray steps of 1-3 voxels (coarse renders): the staged kernel stages 16, 8 or 4
    samples per box and stays bit-identical to the unstaged kernel."""
    vol = synth.blob_volume((80, 96, 88), n_blobs=20, sigma=(2., 6.), amp=(100., 250.),
                            background=5., poisson=False, read_noise=0., dtype=dtype, seed=7)
    M = np.dot(mat4_translate(0, .1, -3.5), mat4_rotation(.7, .3, 1, .2))
    frames = {}
    for slab, staging, chunk in ((0, 0, 0), (12, 0, 0), (12, 1, 0), (12, 0, 16), (12, 1, 4),
                                 (8, 0, 8), (16, 0, 32), (16, 0, 64), (19, 0, 0), (19, 0, 4)):
        rend = VR((120, 104))
        rend._ctx.set_tuning("slab_kb", slab)
        rend._ctx.set_tuning("staging", staging)
        rend._ctx.set_tuning("chunk", chunk)
        rend.set_max_steps(max_steps)
        rend.set_data(vol)
        rend.set_modelView(M)
        rend.render(minVal=float(vol.min()), maxVal=float(vol.max()))
        frames[slab, staging, chunk] = (rend.output.copy(), rend.output_alpha.copy())
    for key in frames:
        np.testing.assert_array_equal(frames[0, 0, 0][0], frames[key][0])
        np.testing.assert_array_equal(frames[0, 0, 0][1], frames[key][1])
    assert (frames[0, 0, 0][0] > 0).sum() > 500


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8, np.float32])
@pytest.mark.parametrize("max_steps,pose", [(160, 0), (64, 1), (224, 2)])
def test_staged_opacity_matches_global(VR, dtype, max_steps, pose):
    """The opacity mode (alpha_pow > 0) staged in LDS (per-wave slabs; boxes split when
    large; rays whose cumulative transparency falls to 0.01 break their 16-sample group
    and take one sample per group after that) is bit-identical to the unstaged kernel,
    multipass included, with bright volumes that make most rays break."""
    vol = synth.blob_volume((64, 96, 80), n_blobs=25, sigma=(3., 8.), amp=(150., 250.),
                            background=20., poisson=False, read_noise=0., dtype=dtype, seed=pose)
    M = [np.dot(mat4_translate(0, 0, -4), mat4_rotation(np.pi / 6, 0, 1, 0)),
         np.dot(mat4_translate(.3, -.2, -2.6), mat4_rotation(2.1, 1, .4, -.3)),
         np.dot(mat4_translate(0, 0, -1.2), mat4_rotation(.8, .2, 1, .1))][pose]
    frames = {}
    # (slab KiB, alpha_pow, waves per workgroup, tile map)
    for slab, ap, waves, tile_map in ((0, .6, 4, 0), (19, .6, 4, 0), (16, .6, 4, 0), (0, 3., 4, 0),
                                      (19, 3., 4, 0), (19, .6, 1, 0), (19, 3., 1, 3),
                                      (19, .6, 4, 3)):
        rend = VR((130, 97))
        rend._ctx.set_tuning("slab_kb", slab)
        rend._ctx.set_tuning("wg_waves", waves)
        rend._ctx.set_tuning("tile_map", tile_map)
        rend.set_max_steps(max_steps)
        rend.set_data(vol)
        rend.set_modelView(M)
        rend.set_alpha_pow(ap)
        rend.render(minVal=float(vol.min()), maxVal=float(vol.max()))
        res = [rend.output.copy(), rend.output_alpha.copy()]
        for part in range(2):
            rend.render(numParts=2, currentPart=part)
        res.append(rend.output.copy())
        frames[slab, ap, waves, tile_map] = res
    for key in frames:
        for a, b in zip(frames[0, key[1], 4, 0], frames[key]):
            np.testing.assert_array_equal(a, b, err_msg=str(key))
    assert (frames[0, .6, 4, 0][0] > 0).sum() > 200


def test_tuning_does_not_change_results(VR):
    vol = synth.blob_volume((40, 48, 56), n_blobs=10, sigma=(2., 5.), amp=(500., 3000.),
                            background=100., poisson=True, read_noise=5., dtype="uint16", seed=4)
    frames = {}
    # 203 x 150: 13 x 10 tiles, tile rows not a multiple of the 8 XCDs
    for tile_map in (0, 1, 2, 3):
        for waves in (4, 1):
            rend = VR((203, 150))
            rend._ctx.set_tuning("tile_map", tile_map)
            rend._ctx.set_tuning("wg_waves", waves)
            assert rend._ctx.get_tuning("tile_map") == tile_map
            assert rend._ctx.get_tuning("wg_waves") == waves
            rend.set_data(vol)
            rend.set_modelView(np.dot(mat4_translate(0, 0, -3.), mat4_rotation(.6, 0, 1, 0)))
            res = []
            for ap in (0., .7):
                rend.set_alpha_pow(ap)
                rend.render(minVal=float(vol.min()), maxVal=float(vol.max()))
                res += [rend.output.copy(), rend.output_alpha.copy()]
            rend.render(maxVal=1500., method="iso_surface")
            res.append(rend.output.copy())
            frames[tile_map, waves] = res
    for key in frames:
        for a, b in zip(frames[0, 4], frames[key]):
            np.testing.assert_array_equal(a, b, err_msg=str(key))
    with pytest.raises(_lib.SprError):
        rend._ctx.set_tuning("tile_map", 9)
    with pytest.raises(_lib.SprError):
        rend._ctx.set_tuning("wg_waves", 2)


def test_multipass_parts_merge(VR):
    vol = synth.blob_volume((40, 40, 40), n_blobs=10, sigma=(2., 4.), amp=(.3, 1.),
                            background=.05, poisson=False, read_noise=0., dtype="float32", seed=2)
    rend = VR((64, 64))
    rend.set_max_steps(96)
    rend.set_data(vol)
    rend.set_modelView(np.dot(mat4_translate(0, 0, -3.5), mat4_rotation(.4, 1, 1, 0)))
    out = np.zeros((64, 64), np.float32)
    alpha = np.zeros_like(out)
    for part in range(3):
        rend.render(maxVal=1., numParts=3, currentPart=part)
        oracle_mip(rend, vol, num_parts=3, current_part=part, out=out, alpha=alpha)
        assert np.abs(rend.output - out).max() <= TOL
        np.testing.assert_array_equal(rend.output_alpha, alpha)
    with pytest.raises(ValueError):
        rend.render(numParts=7)


def test_constant_volume_rend_const(VR):
    """tests/test_rendering/test_simple_rendering.py:85-95 (rend_const)."""
    data = (123 * np.ones((100, 100, 100))).astype(np.float32)
    rend = VR((4, 4))
    rend.render(data, maxVal=200.)
    np.testing.assert_array_equal(rend.output, np.float32(123.) / np.float32(200.))
    np.testing.assert_array_equal(rend.output_alpha, 1.)


def test_nearest_ortho_is_numpy_max(VR):
    N = 16
    vol = np.random.default_rng(3).uniform(-1, 1, (N, N, N)).astype(np.float32)
    rend = VR((2 * N, 2 * N), interpolation="nearest")
    rend.set_max_steps(64)
    rend.set_data(vol)
    rend.set_projection(mat4_ortho(-1, 1, -1, 1, -1, 1))
    rend.set_modelView(mat4_identity())
    rend.render()
    np.testing.assert_allclose(rend.output[1::2, 1::2], np.clip(vol.max(axis=0), 0, 1), rtol=1e-6)


def first_hit_voxel(depth, rend, shape_zyx):
    """floor(p(t_hit) * N) per hit pixel, p(t) = (1 + orig + t direc)/2 (float64)."""
    iP = rend._invP.astype(np.float64).reshape(4, 4)
    iM = rend._invM.astype(np.float64).reshape(4, 4)
    H, W = depth.shape
    nz, ny, nx = shape_zyx
    vox = np.full((H, W, 3), -1, np.int64)
    for y in range(H):
        for x in range(W):
            t = depth[y, x]
            if not np.isfinite(t):
                continue
            u, v = x / W * 2 - 1, y / H * 2 - 1
            o0 = iP @ [u, v, -1, 1]
            o0 /= o0[3]
            o = iM @ o0
            o /= o[3]
            b = iP @ [u, v, 1, 1]
            b /= b[3]
            d = iM @ ((b - o0) / np.linalg.norm(b - o0))
            p = .5 * (1 + o[:3] + t * d[:3])
            vox[y, x] = np.floor(p * [nx, ny, nz])
    return vox


@pytest.mark.parametrize("dtype", [np.uint16, np.float32])
def test_iso_surface_pipeline(VR, dtype):
    """render(method="iso_surface"): iso march + normal blur + AO + AO blur + shading."""
    shape = (48, 56, 52)
    amp, bg = ((500., 3000.), 100.) if dtype == np.uint16 else ((.3, 1.), .05)
    vol = synth.blob_volume(shape, n_blobs=14, sigma=(3., 6.), amp=amp, background=bg,
                            poisson=dtype == np.uint16, read_noise=0., dtype=dtype, seed=5)
    W, H = 120, 100
    rend = VR((W, H))
    rend.set_max_steps(160)
    rend.set_data(vol)
    rend.set_occ_strength(.3)
    rend.set_occ_radius(11)
    rend.set_occ_n_points(20)
    rend.set_projection(mat4_perspective(45, W / float(H), .1, 10))
    rend.set_modelView(np.dot(mat4_translate(0, 0, -3.3), mat4_rotation(.8, -.2, 1, .3)))
    lo, hi = float(vol.min()), float(vol.max())
    iso_max = 2 * (lo + .3 * (hi - lo))
    rend.render(maxVal=iso_max, method="iso_surface")
    out, alpha, depth, normals, occ = oracle.render_iso(
        vol, rend._invP, rend._invM, box=rend.boxBounds, iso_val=np.float32(iso_max) / 2,
        gamma=rend.gamma, max_steps=rend.max_steps, occ_strength=rend.occ_strength,
        occ_radius=rend.occ_radius, occ_n_points=rend.occ_n_points, width=W, height=H)
    hit = np.isfinite(depth)
    say("iso %s: hits %d, out max|diff| %.3g, normals %.3g, occ %.3g, depth-exact %.4f" % (
        np.dtype(dtype).name, hit.sum(), np.abs(rend.output - out).max(),
        np.abs(rend.output_normals - normals).max(), np.abs(rend.output_occlusion - occ).max(),
        np.mean(rend.output_depth[hit] == depth[hit])))
    assert hit.sum() > 500
    np.testing.assert_array_equal(np.isfinite(rend.output_depth), hit)
    np.testing.assert_array_equal(rend.output_alpha, alpha)
    same_voxel = np.all(first_hit_voxel(rend.output_depth, rend, shape)
                        == first_hit_voxel(depth, rend, shape), axis=-1)
    assert same_voxel[hit].all()
    assert np.abs(rend.output_normals - normals).max() <= TOL
    assert np.abs(rend.output_occlusion - occ).max() <= TOL
    assert np.abs(rend.output - out).max() <= TOL


def test_volume_minmax_and_errors(VR, capsys):
    vol = synth.blob_volume((20, 30, 40), n_blobs=5, sigma=(2., 3.), amp=(500., 3000.),
                            background=100., poisson=True, read_noise=5., dtype="uint16", seed=1)
    rend = VR((16, 16))
    assert rend.render() is None
    assert "no data provided" in capsys.readouterr().out
    rend.set_data(vol)
    assert rend.dataImg.shape == (40, 30, 20)
    assert rend.dataImg.min_max() == (float(vol.min()), float(vol.max()))
    with pytest.raises(KeyError):
        rend.rebuild_program("cubic")
    with pytest.raises(NotImplementedError):
        rend.set_data(vol.astype(np.float64), autoConvert=False)
    rend.set_data(vol.astype(np.int32))  # auto-converted to the current dtype (uint16)
    assert rend.dtype == np.uint16
    before = rend.output.copy()
    rend.render(method="unknown")  # silent no-op (volumerender.py:543-547)
    assert np.array_equal(rend.output, before)


def test_render_device_matches_host_render(VR):
    import torch
    vol = synth.blob_volume((32, 32, 32), n_blobs=6, sigma=(2., 4.), amp=(500., 3000.),
                            background=100., poisson=True, read_noise=5., dtype="uint16", seed=9)
    rend = VR((48, 40))
    rend.set_data(vol)
    rend.set_modelView(np.dot(mat4_translate(0, 0, -3), mat4_rotation(.3, 0, 1, 0)))
    rend.render(minVal=float(vol.min()), maxVal=float(vol.max()))
    o = torch.empty((40, 48), dtype=torch.float32, device="cuda")
    a = torch.empty_like(o)
    rend._push_params()
    s = torch.cuda.current_stream()
    rend._ctx.render_device(_lib.METHOD_MAX_PROJECT, 1, 0, 48, 40, o.data_ptr(), a.data_ptr(),
                            stream=s.cuda_stream)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(o.cpu().numpy(), rend.output)
    np.testing.assert_array_equal(a.cpu().numpy(), rend.output_alpha)


def test_each_render_returns_fresh_arrays(VR):
    """render() leaves fresh arrays in output / output_alpha / output_depth (the reference
    assigns buffer.get() copies, volumerender.py:388-390, 502-506): a later render of
    another pose changes none of the arrays an earlier one returned, and the arrays are
    writable and independent (they live in pinned host memory where torch can pin)."""
    vol = synth.blob_volume((36, 40, 44), n_blobs=8, sigma=(2., 4.), amp=(500., 3000.),
                            background=100., poisson=False, read_noise=0., dtype="uint16",
                            seed=12)
    rend = VR((64, 48))
    rend.set_data(vol)
    lo, hi = float(vol.min()), float(vol.max())
    poses = [np.dot(mat4_translate(0, 0, -3.2), mat4_rotation(a, .2, 1, .1)) for a in (.3, 1.4)]
    got = []
    for M in poses:
        rend.set_modelView(M)
        rend.render(minVal=lo, maxVal=hi)
        got.append((rend.output, rend.output_alpha, rend.output_depth))
    first = [a.copy() for a in got[0]]
    assert not np.array_equal(got[0][0], got[1][0])
    for a, b in zip(got[0], first):
        np.testing.assert_array_equal(a, b)
    for a in got[0] + got[1]:
        assert a.dtype == np.float32 and a.flags.writeable and a.flags.c_contiguous
    assert not np.shares_memory(got[0][0], got[1][0])
    assert not np.shares_memory(got[1][0], got[1][1])
    got[1][0][:] = -1.  # the caller owns the array
    rend.set_modelView(poses[0])
    rend.render(minVal=lo, maxVal=hi)
    np.testing.assert_array_equal(rend.output, first[0])
    np.testing.assert_array_equal(rend.output_alpha, first[1])
    want = oracle_mip(rend, vol)
    assert np.abs(rend.output - want[0]).max() <= TOL
    # iso: all five planes fresh as well
    rend.render(method="iso_surface", maxVal=2 * (lo + .3 * (hi - lo)))
    iso1 = [rend.output, rend.output_alpha, rend.output_depth, rend.output_normals,
            rend.output_occlusion]
    keep = [a.copy() for a in iso1]
    rend.set_modelView(poses[1])
    rend.render(method="iso_surface", maxVal=2 * (lo + .3 * (hi - lo)))
    for a, b in zip(iso1, keep):
        np.testing.assert_array_equal(a, b)
    assert rend.output_normals.shape == (48, 64, 3)


def test_headline_size_rows_parity():
    """1024x1024x512 uint16 (8000 blobs) -> 1024^2, max_steps 1536: GPU frame against the
    oracle on 12 full rows spread over the frame (12 x 1024 rays x 1552 samples)."""
    import torch
    cfg = synth.CONFIGS["headline"]
    say("generating headline volume on the GPU")
    vt = synth.config_volume("headline", device="cuda")
    ctx = _lib.Context(0)
    ctx.set_volume_device(vt.data_ptr(), np.uint16, cfg["shape"])
    lo, hi = ctx.volume_minmax()
    vol = synth.device_volume_to_numpy(vt, np.uint16)
    del vt
    assert (lo, hi) == (float(vol.min()), float(vol.max()))
    from spimagine_amd.volumerender.volumerender import render_matrices
    M, P = synth.headline_camera()
    nz, ny, nx = cfg["shape"]
    invP, invM = render_matrices(M, P, (nx, ny, nz))
    W, H = cfg["out"]
    ctx.set_params(invP, invM, (-1, 1, -1, 1, -1, 1), lo, hi, 1., 0., .1, 21, 30, cfg["max_steps"], 0)
    out = np.empty((H, W), np.float32)
    alpha = np.empty_like(out)
    ctx.render(_lib.METHOD_MAX_PROJECT, 1, 0, W, H, out, alpha)
    assert ctx.count_samples(1, W, H) == int((alpha > 0).sum()) * 1552
    rows = np.linspace(40, H - 40, 12).astype(int)
    for r in rows:
        o, a = oracle.max_project(vol, invP, invM, min_val=lo, max_val=hi, max_steps=1536,
                                  width=W, height=H, rows=(r, r + 1))
        d = np.abs(out[r] - o[r]).max()
        assert d <= TOL, (r, d)
        np.testing.assert_array_equal(alpha[r], a[r])
    say("headline rows parity ok")


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8, np.float32])
@pytest.mark.parametrize("pose,steps", [(0, 160), (1, 96), (2, 224)])
def test_staged_iso_march_matches_global(VR, dtype, pose, steps):
    """The iso march staged in LDS (crossed rays leave the box; the wave stops when no ray
    marches) gives every output of the unstaged march bit for bit: colour, alpha, depth,
    normals, occlusion."""
    shape = (48, 64, 56)
    amp, bg = ((500., 3000.), 100.) if dtype != np.float32 else ((.3, 1.), .05)
    if dtype == np.uint8:
        amp, bg = (60., 200.), 10.
    vol = synth.blob_volume(shape, n_blobs=16, sigma=(3., 7.), amp=amp, background=bg,
                            poisson=False, read_noise=0., dtype=dtype, seed=pose)
    M = [np.dot(mat4_translate(0, 0, -3.3), mat4_rotation(.8, -.2, 1, .3)),
         np.dot(mat4_translate(.3, -.2, -2.6), mat4_rotation(2.1, 1, .4, -.3)),
         np.dot(mat4_translate(0, 0, -1.2), mat4_rotation(.8, .2, 1, .1))][pose]
    lo, hi = float(vol.min()), float(vol.max())
    res = {}
    for slab in (0, 19):
        rend = VR((120, 100))
        rend._ctx.set_tuning("slab_kb", slab)
        rend.set_max_steps(steps)
        rend.set_data(vol)
        rend.set_modelView(M)
        rend.render(maxVal=2 * (lo + .3 * (hi - lo)), method="iso_surface")
        res[slab] = [rend.output.copy(), rend.output_alpha.copy(), rend.output_depth.copy(),
                     rend.output_normals.copy(), rend.output_occlusion.copy()]
    for a, b in zip(res[0], res[19]):
        np.testing.assert_array_equal(a, b)
    assert np.isfinite(res[0][2]).sum() > 300
