"""GPU: batched frames (many resident volumes x poses in one persistent launch, and the
same batch captured in a HIP graph) give exactly the frames of separate renders."""
import numpy as np
import pytest

from spimagine_amd import _lib, synth
from spimagine_amd.utils import mat4_perspective, mat4_rotation, mat4_translate
from spimagine_amd.volumerender.volumerender import render_matrices

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert _lib.device_count() > 0
    return t


def poses(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        axis = rng.normal(size=3)
        out.append(np.dot(mat4_translate(rng.uniform(-.2, .2), rng.uniform(-.2, .2),
                                         -rng.uniform(2.5, 4.)),
                          mat4_rotation(rng.uniform(0, 6.28), *axis)))
    return out


def setup(torch, dtype, shapes, n_frames, W, H, alpha_pow=0., interp="linear", seed=0):
    ctx = _lib.Context(0)
    vols = []
    for k, sh in enumerate(shapes):
        amp = (.3, 1.) if dtype == np.float32 else (300., 3000.) if dtype == np.uint16 else (60., 200.)
        vols.append(synth.blob_volume(sh, n_blobs=10, sigma=(2., 5.), amp=amp, background=amp[0] / 10,
                                      poisson=False, read_noise=0., dtype=dtype, seed=seed + k))
    slots = [ctx.volume_add(v) for v in vols]
    P = mat4_perspective(50, W / float(H), .1, 10)
    frames = []
    for f, M in enumerate(poses(n_frames, seed)):
        k = f % len(vols)
        nz, ny, nx = vols[k].shape
        iP, iM = render_matrices(M, P, (nx, ny, nz))
        o = torch.full((H, W), -7., device="cuda")
        a = torch.full((H, W), -7., device="cuda")
        frames.append(dict(vol=k, slot=slots[k], iP=iP, iM=iM, out=o, alpha=a))
    filt = _lib.FILTER_NEAREST if interp == "nearest" else _lib.FILTER_LINEAR
    ctx.set_params(frames[0]["iP"], frames[0]["iM"], (-1, 1, -1, 1, -1, 1), 0., 0., 1., alpha_pow,
                   .1, 21, 30, 128, filt)
    return ctx, vols, frames, filt


def single(torch, vols, f, W, H, alpha_pow, filt, tuning=None):
    """The frame rendered alone through spr_render_device in its own context."""
    c = _lib.Context(0)
    for k, v in (tuning or {}).items():
        c.set_tuning(k, v)
    c.set_volume(vols[f["vol"]])
    c.set_params(f["iP"], f["iM"], (-1, 1, -1, 1, -1, 1), 0., 0., 1., alpha_pow, .1, 21, 30, 128, filt)
    o = torch.empty((H, W), device="cuda")
    a = torch.empty((H, W), device="cuda")
    c.render_device(_lib.METHOD_MAX_PROJECT, 1, 0, W, H, o.data_ptr(), a.data_ptr())
    c.synchronize()
    torch.cuda.synchronize()
    return o.cpu().numpy(), a.cpu().numpy()


@pytest.mark.parametrize("dtype,alpha_pow,interp,staging,chunk", [
    (np.uint16, 0., "linear", 0, 0), (np.uint8, 0., "linear", 0, 0),
    (np.float32, 0., "linear", 0, 0), (np.uint16, .5, "linear", 0, 0),
    (np.float32, 0., "nearest", 0, 0), (np.uint16, 0., "linear", 1, 0),
    (np.float32, 0., "linear", 1, 0), (np.uint16, 0., "linear", 0, 32),
    (np.uint8, 0., "linear", 0, 64), (np.float32, .5, "linear", 0, 0),
    (np.uint8, 2., "linear", 0, 0), (np.uint16, .5, "linear", 1, 0)])
def test_batch_equals_separate_renders(torch, dtype, alpha_pow, interp, staging, chunk):
    """max_steps 128: 9 groups of 16 samples, so boxes of 2 or 4 groups (chunk 32, 64) end
    with a shorter box."""
    W, H = 150, 117
    shapes = [(40, 64, 48), (33, 50, 61), (64, 64, 64)]
    ctx, vols, frames, filt = setup(torch, dtype, shapes, 11, W, H, alpha_pow, interp)
    ctx.set_tuning("staging", staging)
    ctx.set_tuning("chunk", chunk)
    jobs = [(f["slot"], f["iP"], f["iM"], f["out"].data_ptr(), f["alpha"].data_ptr()) for f in frames]
    ctx.render_batch(jobs, W, H)
    torch.cuda.synchronize()
    for f in frames:
        o, a = single(torch, vols, f, W, H, alpha_pow, filt)
        np.testing.assert_array_equal(f["out"].cpu().numpy(), o)
        np.testing.assert_array_equal(f["alpha"].cpu().numpy(), a)
    assert sum((f["out"].cpu().numpy() > 0).sum() for f in frames) > 5000


def test_batch_graph_replays_and_updates(torch):
    W, H = 96, 80
    ctx, vols, frames, filt = setup(torch, np.uint16, [(32, 48, 40), (40, 40, 40)], 6, W, H, seed=3)
    jobs = [(f["slot"], f["iP"], f["iM"], f["out"].data_ptr(), f["alpha"].data_ptr()) for f in frames]
    g = ctx.batch_graph(jobs, W, H)
    want = [single(torch, vols, f, W, H, 0., filt) for f in frames]
    s = torch.cuda.Stream()
    for rep in range(4):
        for f in frames:  # eager work between replays: scribble over the outputs
            f["out"].fill_(-3.)
            f["alpha"].fill_(-3.)
        torch.cuda.synchronize()
        g.launch(s.cuda_stream)
        s.synchronize()
        for f, (o, a) in zip(frames, want):
            np.testing.assert_array_equal(f["out"].cpu().numpy(), o)
            np.testing.assert_array_equal(f["alpha"].cpu().numpy(), a)
    # retarget the captured batch to new poses
    P = mat4_perspective(50, W / float(H), .1, 10)
    for f, M in zip(frames, poses(len(frames), 99)):
        nz, ny, nx = vols[f["vol"]].shape
        f["iP"], f["iM"] = render_matrices(M, P, (nx, ny, nz))
    g.update([(f["slot"], f["iP"], f["iM"], f["out"].data_ptr(), f["alpha"].data_ptr()) for f in frames])
    g.launch(s.cuda_stream)
    s.synchronize()
    for f in frames:
        o, a = single(torch, vols, f, W, H, 0., filt)
        np.testing.assert_array_equal(f["out"].cpu().numpy(), o)
    g.close()


def test_batch_errors(torch):
    ctx = _lib.Context(0)
    s16 = ctx.volume_add(np.ones((8, 8, 8), np.uint16))
    s8 = ctx.volume_add(np.ones((8, 8, 8), np.uint8))
    o = torch.zeros((16, 16), device="cuda")
    eye = np.eye(4, dtype=np.float32)
    ctx.set_params(eye, eye, (-1, 1, -1, 1, -1, 1), 0., 0., 1., 0., .1, 21, 30, 64, 0)
    with pytest.raises(_lib.SprError):  # mixed voxel types
        ctx.render_batch([(s16, eye, eye, o.data_ptr(), o.data_ptr()),
                          (s8, eye, eye, o.data_ptr(), o.data_ptr())], 16, 16)
    with pytest.raises(_lib.SprError):  # unknown slot
        ctx.render_batch([(17, eye, eye, o.data_ptr(), o.data_ptr())], 16, 16)
    info = ctx.volume_info(s16)
    assert info["shape"] == (8, 8, 8) and info["bytes"] > 0
    ctx.volume_release(s8)
    with pytest.raises(_lib.SprError):
        ctx.volume_info(s8)


def test_graph_refuses_released_or_refilled_slots(torch):
    """A captured batch records (slot, generation) per job: after its slot is released or
    reused, spr_batch_graph_launch returns an error before launching anything; an update
    to live slots makes it launchable again."""
    W, H = 48, 40
    ctx, vols, frames, filt = setup(torch, np.uint16, [(24, 32, 28), (20, 20, 20)], 4, W, H, seed=5)
    jobs = [(f["slot"], f["iP"], f["iM"], f["out"].data_ptr(), f["alpha"].data_ptr()) for f in frames]
    g = ctx.batch_graph(jobs, W, H)
    g.launch()
    ctx.synchronize()
    gone = frames[1]["slot"]
    ctx.volume_release(gone)
    with pytest.raises(_lib.SprError, match="released or refilled"):
        g.launch()
    # the slot is reused by a new volume of another shape: still refused
    again = ctx.volume_add(vols[0])
    assert again == gone
    with pytest.raises(_lib.SprError, match="released or refilled"):
        g.launch()
    # rewrite the jobs against live slots (the reused slot now holds vols[0])
    for f in frames:
        if f["slot"] == gone:
            f["vol"] = 0
            nz, ny, nx = vols[0].shape
            P = mat4_perspective(50, W / float(H), .1, 10)
            f["iP"], f["iM"] = render_matrices(poses(1, 7)[0], P, (nx, ny, nz))
    g.update([(f["slot"], f["iP"], f["iM"], f["out"].data_ptr(), f["alpha"].data_ptr()) for f in frames])
    s = torch.cuda.Stream()
    g.launch(s.cuda_stream)
    s.synchronize()
    for f in frames:
        o, a = single(torch, vols, f, W, H, 0., filt)
        np.testing.assert_array_equal(f["out"].cpu().numpy(), o)
        np.testing.assert_array_equal(f["alpha"].cpu().numpy(), a)
    g.close()


def test_graph_update_keeps_the_texel_rule(torch):
    """A graph captured on power-of-two volumes runs the one-fma texel rule; an update to a
    volume with other dimensions is refused (its samples would differ from single renders)."""
    W, H = 40, 32
    ctx, vols, frames, filt = setup(torch, np.uint8, [(32, 32, 32)], 2, W, H, seed=8)
    jobs = [(f["slot"], f["iP"], f["iM"], f["out"].data_ptr(), f["alpha"].data_ptr()) for f in frames]
    g = ctx.batch_graph(jobs, W, H)
    odd = ctx.volume_add(synth.blob_volume((24, 40, 48), n_blobs=4, sigma=(2., 4.), amp=(60., 200.),
                                           background=6., poisson=False, read_noise=0.,
                                           dtype=np.uint8, seed=2))
    with pytest.raises(_lib.SprError, match="power-of-two"):
        g.update([(odd, f["iP"], f["iM"], f["out"].data_ptr(), f["alpha"].data_ptr())
                  for f in frames])
    g.launch()
    ctx.synchronize()
    for f in frames:
        o, a = single(torch, vols, f, W, H, 0., filt)
        np.testing.assert_array_equal(f["out"].cpu().numpy(), o)
    g.close()
