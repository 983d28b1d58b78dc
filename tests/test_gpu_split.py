"""GPU: frames of scattered rays with each ray's samples split into segments
(spr_render_frames_split) equal the same frames rendered one by one (spr_render_slot) and
in one mixed persistent launch (spr_render_frames), for every voxel type, segment length,
frame size and transfer setting, including rays that miss the box; and the oracle.

The learned renderer's judge renders sparse grids of pixels (one in 8 x 8 or 16 x 16 of
a frame: sparse_params) this way; full frames go through the same kernels.
"""
import sys

import numpy as np
import pytest

import oracle
from spimagine_amd import _lib, learned, synth
from spimagine_amd.utils import mat4_perspective, mat4_rotation, mat4_translate
from spimagine_amd.volumerender.volumerender import render_matrices

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert _lib.device_count() > 0, "no HIP device"
    return t


def _volume(dtype, shape, seed):
    amp = (.3, 1.) if dtype == np.float32 else (500., 3000.) if dtype == np.uint16 else (60., 200.)
    return synth.blob_volume(shape, n_blobs=12, sigma=(2., 5.), amp=amp, background=amp[0] / 10,
                             poisson=False, read_noise=0., dtype=dtype, seed=seed)


POSES = [np.dot(mat4_translate(0, 0, -3.2), mat4_rotation(.7, .3, 1, -.2)),
         np.dot(mat4_translate(.5, -.3, -2.6), mat4_rotation(2.3, 1, .2, .5)),  # rays miss
         np.dot(mat4_translate(0, 0, -4.), mat4_rotation(0., 0, 1, 0))]


def _frames(torch, slot, vol, specs):
    """specs: (w, h, pose, max_steps, (min, max, gamma), box, sparse stride or 0)."""
    nz, ny, nx = vol.shape
    out = []
    for w, h, k, steps, (lo, hi, g), box, stride in specs:
        P = mat4_perspective(45, w / float(h), .1, 10)
        iP, iM = render_matrices(POSES[k], P, (nx, ny, nz))
        p = _lib.make_params(iP, iM, box, lo, hi, g, 0., .1, 21, 30, steps, _lib.FILTER_LINEAR)
        fw, fh = w, h
        if stride:
            p = learned.sparse_params(p, w, h, stride=stride)
            fw, fh = w // stride, h // stride
        planes = [torch.full((fh, fw), -7., device="cuda") for _ in range(6)]
        out.append((p, fw, fh, planes))
    return out


@pytest.mark.parametrize("dtype", [np.uint16, np.uint8, np.float32])
@pytest.mark.parametrize("seg", [32, 64, 128])
@pytest.mark.parametrize("slab_kb", [19, 0])
def test_split_equals_single_and_mixed(torch, dtype, seg, slab_kb):
    """slab_kb 19: segments staged in LDS slabs (k_split_staged); 0: global gathers."""
    say("split: %s seg %d slab_kb %d" % (np.dtype(dtype).name, seg, slab_kb))
    ctx = _lib.Context(0)
    ctx.set_tuning("slab_kb", slab_kb)
    vol = _volume(dtype, (44, 61, 53), seed=seg)
    slot = ctx.volume_add(vol)
    lo, hi = float(vol.min()), float(vol.max())
    specs = [(128, 96, 0, 160, (lo, hi, 1.), (-1, 1, -1, 1, -1, 1), 8),
             (256, 128, 1, 96, (0., 0., 1.), (-1, 1, -1, 1, -1, 1), 16),
             (77, 45, 2, 208, (lo, hi, .7), (-1, 1, -.6, .8, -1, .5), 0),
             (64, 64, 1, 16, (lo, .5 * hi, 1.6), (-1, 1, -1, 1, -1, 1), 0)]
    frames = _frames(torch, slot, vol, specs)
    s = torch.cuda.Stream()
    for p, fw, fh, pl in frames:
        ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, fw, fh, pl[0].data_ptr(),
                        pl[1].data_ptr(), stream=s.cuda_stream)
    ctx.render_frames([_lib.make_frame(slot, p, fw, fh, pl[2].data_ptr(), pl[3].data_ptr())
                       for p, fw, fh, pl in frames], stream=s.cuda_stream)
    split = [_lib.make_frame(slot, p, fw, fh, pl[4].data_ptr(), pl[5].data_ptr())
             for p, fw, fh, pl in frames]
    ctx.render_frames_split(split, seg=seg, stream=s.cuda_stream)
    s.synchronize()
    hits = 0
    for k, (p, fw, fh, pl) in enumerate(frames):
        want, wa = pl[0].cpu().numpy(), pl[1].cpu().numpy()
        np.testing.assert_array_equal(pl[2].cpu().numpy(), want, err_msg="mixed frame %d" % k)
        np.testing.assert_array_equal(pl[3].cpu().numpy(), wa)
        np.testing.assert_array_equal(pl[4].cpu().numpy(), want, err_msg="split frame %d" % k)
        np.testing.assert_array_equal(pl[5].cpu().numpy(), wa)
        hits += int((want > 0).sum())
    assert hits > 2000
    # a missed ray's outputs are written too
    assert (frames[1][3][5].cpu().numpy() == (-1. if dtype == np.float32 else 0.)).any()
    # and again into the same planes (the accumulators are cleared on every call)
    ctx.render_frames_split(split, seg=seg, stream=s.cuda_stream)
    s.synchronize()
    for p, fw, fh, pl in frames:
        np.testing.assert_array_equal(pl[4].cpu().numpy(), pl[0].cpu().numpy())
    # device gates: frames 1 and 3 closed (their planes untouched), 0 and 2 rendered
    gates = torch.tensor([1, 0, 7, 0], dtype=torch.int32, device="cuda")
    for p, fw, fh, pl in frames:
        pl[4].fill_(-3.)
        pl[5].fill_(-3.)
    torch.cuda.synchronize()
    ctx.render_frames_split(split, seg=seg, stream=s.cuda_stream, d_gates=gates.data_ptr())
    s.synchronize()
    for k, (p, fw, fh, pl) in enumerate(frames):
        if k in (0, 2):
            np.testing.assert_array_equal(pl[4].cpu().numpy(), pl[0].cpu().numpy())
            np.testing.assert_array_equal(pl[5].cpu().numpy(), pl[1].cpu().numpy())
        else:
            assert (pl[4].cpu().numpy() == -3.).all() and (pl[5].cpu().numpy() == -3.).all()


def test_split_matches_oracle(torch):
    ctx = _lib.Context(0)
    vol = _volume(np.uint16, (40, 56, 48), seed=3)
    slot = ctx.volume_add(vol)
    lo, hi = float(vol.min()), float(vol.max())
    W, H = 96, 80
    (p, fw, fh, pl), = _frames(torch, slot, vol, [(W, H, 0, 144, (lo, hi, 1.),
                                                   (-1, 1, -1, 1, -1, 1), 0)])
    s = torch.cuda.Stream()
    ctx.render_frames_split([_lib.make_frame(slot, p, fw, fh, pl[0].data_ptr(), pl[1].data_ptr())],
                            stream=s.cuda_stream)
    s.synchronize()
    iP = np.asarray(list(p.inv_p), np.float32).reshape(4, 4)
    iM = np.asarray(list(p.inv_m), np.float32).reshape(4, 4)
    o, a = oracle.max_project(vol, iP, iM, min_val=lo, max_val=hi, max_steps=144, width=W,
                              height=H)
    assert np.abs(pl[0].cpu().numpy() - o).max() <= 1e-4
    np.testing.assert_array_equal(pl[1].cpu().numpy(), a)


def test_split_rejects(torch):
    ctx = _lib.Context(0)
    vol = _volume(np.uint8, (20, 24, 28), seed=1)
    slot = ctx.volume_add(vol)
    (p, fw, fh, pl), = _frames(torch, slot, vol, [(32, 32, 0, 64, (0., 0., 1.),
                                                   (-1, 1, -1, 1, -1, 1), 0)])
    f = _lib.make_frame(slot, p, fw, fh, pl[0].data_ptr(), pl[1].data_ptr())
    with pytest.raises(_lib.SprError, match="segment"):
        ctx.render_frames_split([f], seg=48)
    ctx.set_layout(_lib.LAYOUT_LINEAR)
    lin = ctx.volume_add(vol)
    g = _lib.make_frame(lin, p, fw, fh, pl[0].data_ptr(), pl[1].data_ptr())
    with pytest.raises(_lib.SprError, match="padded"):
        ctx.render_frames_split([f, g])
