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
    # 3000 seeded blobs over a background of 100 plus read noise (uint16 bit pattern)
    v16 = next(synth.blob_series_device(shape, 3000, (3., 10.), (500., 3000.), 100., False, 5.,
                                        1, seed=3))
    torch.cuda.synchronize()
    assert int(v16.max()) > 2000
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
    # the wide slot with one and with four waves per workgroup, then the linear layout
    runs = ((wide, sw, 1), (wide, sw, 4), (lin, sl, 1))
    W, H = 192, 160
    for k, M in enumerate(synth.random_poses(3, seed=11)):
        iP, iM = render_matrices(M, P, shape[::-1])
        for steps in (96, 256):
            p = _lib.make_params(iP, iM, (-1, 1, -1, 1, -1, 1), 0., 3300., 1., 0., .1, 21, 30, steps,
                                 _lib.FILTER_LINEAR)
            outs = []
            for ctx, slot, waves in runs:
                ctx.set_tuning("wg_waves", waves)
                o = torch.empty((H, W), device="cuda")
                a = torch.empty_like(o)
                ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, W, H, o.data_ptr(), a.data_ptr(),
                                stream=s.cuda_stream)
                outs.append((o, a))
            s.synchronize()
            for o, a in outs[:2]:
                assert torch.equal(o, outs[2][0]), (k, steps)
                assert torch.equal(a, outs[2][1])
            assert (outs[0][0] > 0).sum().item() > 5000
    # the opacity mode (staged on the wide copy) and the nearest filter (the slot's linear
    # volume) equal the linear layout's unstaged kernels
    for ap, filt in ((.5, _lib.FILTER_LINEAR), (3., _lib.FILTER_LINEAR), (0., _lib.FILTER_NEAREST)):
        p = _lib.make_params(iP, iM, (-1, 1, -1, 1, -1, 1), 0., 3300., 1., ap, .1, 21, 30, 96, filt)
        outs = []
        for ctx, slot, waves in runs:
            ctx.set_tuning("wg_waves", waves)
            o = torch.empty((H, W), device="cuda")
            a = torch.empty_like(o)
            ctx.render_slot(slot, p, _lib.METHOD_MAX_PROJECT, W, H, o.data_ptr(), a.data_ptr(),
                            stream=s.cuda_stream)
            outs.append((o, a))
        s.synchronize()
        for o, a in outs[:2]:
            assert torch.equal(o, outs[2][0]), (ap, filt)
            assert torch.equal(a, outs[2][1])
    # the iso-surface march staged on the wide copy: all five outputs as the linear layout's
    p = _lib.make_params(iP, iM, (-1, 1, -1, 1, -1, 1), 0., 2 * 1500., 1., 0., .1, 21, 30, 160,
                         _lib.FILTER_LINEAR)
    res = []
    for ctx, slot, waves in runs:
        ctx.set_tuning("wg_waves", waves)
        bufs = [torch.empty((H, W), device="cuda") for _ in range(3)]
        nrm = torch.empty((H, W, 3), device="cuda")
        occ = torch.empty((H, W), device="cuda")
        ctx.render_slot(slot, p, _lib.METHOD_ISO_SURFACE, W, H, bufs[0].data_ptr(),
                        bufs[1].data_ptr(), bufs[2].data_ptr(), nrm.data_ptr(), occ.data_ptr(),
                        stream=s.cuda_stream)
        res.append(bufs + [nrm, occ])
    s.synchronize()
    for r in res[:2]:
        for a, b in zip(r, res[2]):
            assert torch.equal(a, b)
    assert torch.isfinite(res[0][2]).sum().item() > 200
    # enough steps for 16-sample boxes (a step of at most ~0.8 voxel): the one-wave staged
    # max projection and iso march on the wide copy, against four waves and the linear layout
    # (the iso march at 3360 steps, whose skipping tests four chunks at a time first)
    for method in (_lib.METHOD_MAX_PROJECT, _lib.METHOD_ISO_SURFACE):
        mip = method == _lib.METHOD_MAX_PROJECT
        p = _lib.make_params(iP, iM, (-1, 1, -1, 1, -1, 1), 0., 3300. if mip else 3000., 1., 0.,
                             .1, 21, 30, 2688 if mip else 3360, _lib.FILTER_LINEAR)
        res = []
        for ctx, slot, waves in runs:
            ctx.set_tuning("wg_waves", waves)
            bufs = [torch.empty((H, W), device="cuda") for _ in range(4)]
            nrm = torch.empty((H, W, 3), device="cuda")
            ctx.render_slot(slot, p, method, W, H, bufs[0].data_ptr(), bufs[1].data_ptr(),
                            bufs[2].data_ptr(), nrm.data_ptr(), bufs[3].data_ptr(),
                            stream=s.cuda_stream)
            res.append(bufs[:2] if method == _lib.METHOD_MAX_PROJECT else bufs + [nrm])
        s.synchronize()
        for r in res[:2]:
            for a, b in zip(r, res[2]):
                assert torch.equal(a, b), method
        if method == _lib.METHOD_MAX_PROJECT:
            assert (res[0][0] > 0).sum().item() > 5000
        else:
            assert torch.isfinite(res[0][2]).sum().item() > 200
    # a wide slot is not batchable (the batch kernels address 32-bit offsets)
    o = torch.empty((H, W), device="cuda")
    with pytest.raises(_lib.SprError):
        wide.render_batch([(sw, iP, iM, o.data_ptr(), o.data_ptr())], W, H,
                          params=_lib.make_params(iP, iM, max_steps=96))
    wide.volume_release(sw)
    lin.volume_release(sl)
