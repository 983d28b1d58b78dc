"""GPU: spr_render_batch_noalloc (job table and queue counters in the library's fixed
device scratch, no allocation) gives the frames of spr_render_batch bit for bit, on two
calls in a row (the second smaller), and refuses batches above its maximum."""
import ctypes
import sys

import numpy as np
import pytest

from spimagine_amd import _lib, synth
from spimagine_amd.utils import mat4_perspective, mat4_rotation, mat4_translate
from spimagine_amd.volumerender.volumerender import render_matrices

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _poses(n, seed):
    rng = np.random.default_rng(seed)
    return [np.dot(mat4_translate(rng.uniform(-.2, .2), rng.uniform(-.2, .2), -rng.uniform(2.5, 4.)),
                   mat4_rotation(rng.uniform(0, 6.28), *rng.normal(size=3))) for _ in range(n)]


def _case(torch, ctx, slots, vols, n, W, H, seed, params):
    P = mat4_perspective(50, W / float(H), .1, 10)
    jobs, planes = [], []
    for f, M in enumerate(_poses(n, seed)):
        k = f % len(slots)
        nz, ny, nx = vols[k].shape
        iP, iM = render_matrices(M, P, (nx, ny, nz))
        a = torch.full((2, H, W), -5., device="cuda")
        b = torch.full((2, H, W), -6., device="cuda")
        planes.append((a, b))
        jobs.append((slots[k], iP, iM, a[0].data_ptr(), a[1].data_ptr(),
                     b[0].data_ptr(), b[1].data_ptr()))
    ref = [(j[0], j[1], j[2], j[3], j[4]) for j in jobs]
    new = [(j[0], j[1], j[2], j[5], j[6]) for j in jobs]
    s = torch.cuda.Stream()
    ctx.render_batch(ref, W, H, stream=s.cuda_stream, params=params)
    ctx.render_batch_noalloc(new, W, H, stream=s.cuda_stream, params=params)
    s.synchronize()
    hits = 0
    for a, b in planes:
        np.testing.assert_array_equal(b.cpu().numpy(), a.cpu().numpy())
        hits += int((a[1] > 0).sum())
    return hits


@pytest.mark.parametrize("dtype,alpha_pow", [(np.uint16, 0.), (np.float32, .5)])
def test_noalloc_batch_equals_batch(dtype, alpha_pow):
    import torch
    assert _lib.device_count() > 0
    say("noalloc: %s alpha_pow %s" % (np.dtype(dtype).name, alpha_pow))
    ctx = _lib.Context(0)
    amp = (.3, 1.) if dtype == np.float32 else (300., 3000.)
    vols = [synth.blob_volume(sh, n_blobs=10, sigma=(2., 5.), amp=amp, background=amp[0] / 10,
                              poisson=False, read_noise=0., dtype=dtype, seed=k)
            for k, sh in enumerate([(40, 64, 48), (33, 50, 61), (64, 64, 64)])]
    slots = [ctx.volume_add(v) for v in vols]
    eye = np.eye(4, dtype=np.float32)
    params = _lib.make_params(eye, eye, min_val=0., max_val=0., alpha_pow=alpha_pow, max_steps=128)
    # first call: 9 frames of 150 x 117; second, smaller: 2 frames of 64 x 48
    hits = _case(torch, ctx, slots, vols, 9, 150, 117, 1, params)
    assert hits > 2000
    hits = _case(torch, ctx, slots[:1], vols[:1], 2, 64, 48, 2, params)
    assert hits > 200


def test_noalloc_batch_limit():
    assert _lib.device_count() > 0
    ctx = _lib.Context(0)
    n = _lib.NOALLOC_MAX_JOBS + 1
    jobs = (_lib.SprJob * n)()
    eye = np.eye(4, dtype=np.float32)
    params = _lib.make_params(eye, eye, max_steps=64)
    with pytest.raises(_lib.SprError, match="maximum"):
        ctx.render_batch_noalloc(jobs, 16, 16, params=params)
