"""GPU: volumes whose padded copy exceeds 4 GiB (2048x2048x512 uint16: 4.33 GB padded).
The staged max projection addresses such a copy through buffer resources rebased at each
box's first z-plane and 64-bit gathers; its frames must equal the linear layout's bit for
bit, and the slot keeps the linear volume for the other kernels."""
import numpy as np
import pytest

from spimagine_amd import _lib, synth
from spimagine_amd.utils import mat4_perspective
from spimagine_amd.volumerender.volumerender import render_matrices

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert _lib.device_count() > 0
    return t


def test_wide_padded_volume_matches_linear(torch):
    shape = (512, 2048, 2048)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    # smooth-ish structure: a coarse random field upsampled, plus noise
    coarse = torch.rand((1, 1, 32, 64, 64), generator=gen, device="cuda") * 3000.
    vol = torch.nn.functional.interpolate(coarse, size=shape, mode="trilinear",
                                          align_corners=False)[0, 0]
    vol += torch.rand(shape, generator=gen, device="cuda") * 200.
    v16 = torch.clamp(vol, 0, 65535).to(torch.int32)
    del vol, coarse
    v16 = torch.where(v16 >= 32768, v16 - 65536, v16).to(torch.int16)
    torch.cuda.synchronize()
    wide = _lib.Context(0)
    sw = wide.volume_add(dev_ptr=v16.data_ptr(), dtype=np.uint16, shape=shape)
    nbytes = 2048 * 2048 * 512 * 2
    info = wide.volume_info(sw)
    assert info["bytes"] > nbytes + (1 << 32) - 1  # linear volume + the wide padded copy
    lin = _lib.Context(0)
    lin.set_layout(_lib.LAYOUT_LINEAR)
    sl = lin.volume_add(dev_ptr=v16.data_ptr(), dtype=np.uint16, shape=shape)
    assert lin.volume_info(sl)["bytes"] == nbytes
    del v16
    P = mat4_perspective(45, 1., .1, 10)
    s = torch.cuda.Stream()
    W, H = 192, 160
    for k, M in enumerate(synth.random_poses(3, seed=11)):
        iP, iM = render_matrices(M, P, shape[::-1])
        for steps in (96, 256):
            p = _lib.make_params(iP, iM, (-1, 1, -1, 1, -1, 1), 0., 3300., 1., 0., .1, 21, 30, steps,
                                 _lib.FILTER_LINEAR)
            outs = []
            for ctx, slot in ((wide, sw), (lin, sl)):
                o = torch.empty((H, W), device="cuda")
                a = torch.empty_like(o)
                ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, W, H, o.data_ptr(), a.data_ptr(),
                                stream=s.cuda_stream)
                outs.append((o, a))
            s.synchronize()
            assert torch.equal(outs[0][0], outs[1][0]), (k, steps)
            assert torch.equal(outs[0][1], outs[1][1])
            assert (outs[0][0] > 0).sum().item() > 5000
    # the other kernels run on the slot's linear volume
    p = _lib.make_params(iP, iM, (-1, 1, -1, 1, -1, 1), 0., 3300., 1., .5, .1, 21, 30, 96,
                         _lib.FILTER_LINEAR)
    outs = []
    for ctx, slot in ((wide, sw), (lin, sl)):
        o = torch.empty((H, W), device="cuda")
        a = torch.empty_like(o)
        ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, W, H, o.data_ptr(), a.data_ptr(),
                        stream=s.cuda_stream)
        outs.append(o)
    s.synchronize()
    assert torch.equal(outs[0], outs[1])
    # a wide slot is not batchable (the batch kernels address 32-bit offsets)
    o = torch.empty((H, W), device="cuda")
    with pytest.raises(_lib.SprError):
        wide.render_batch([(sw, iP, iM, o.data_ptr(), o.data_ptr())], W, H,
                          params=_lib.make_params(iP, iM, max_steps=96))
    wide.volume_release(sw)
    lin.volume_release(sl)
