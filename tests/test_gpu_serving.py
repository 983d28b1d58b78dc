"""GPU: the render server.  Every request's frame is identical whether it is rendered alone
or inside a continuously batched mix of volumes, voxel types, sizes and methods, also when
the HBM budget forces volumes to be evicted and uploaded again."""
import numpy as np
import pytest

from spimagine_amd import _lib, synth
from spimagine_amd.serving import RenderServer, Request
from spimagine_amd.utils import mat4_perspective, mat4_rotation, mat4_translate

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch():
    import torch as t
    assert _lib.device_count() > 0
    return t


VOLUMES = {
    "a16": ((48, 64, 56), np.uint16),
    "b16": ((40, 50, 61), np.uint16),
    "c08": ((32, 48, 48), np.uint8),
    "d32": ((40, 40, 40), np.float32),
}


def make_volume(shape, dtype, seed):
    amp = (.3, 1.) if dtype == np.float32 else (300., 3000.) if dtype == np.uint16 else (60., 200.)
    return synth.blob_volume(shape, n_blobs=8, sigma=(2., 5.), amp=amp, background=amp[0] / 10,
                             poisson=False, read_noise=0., dtype=dtype, seed=seed)


def requests(n, seed):
    rng = np.random.default_rng(seed)
    names = sorted(VOLUMES)
    out = []
    for i in range(n):
        name = names[rng.integers(len(names))]
        size = [(96, 80), (64, 64)][rng.integers(2)]
        M = np.dot(mat4_translate(0, 0, -rng.uniform(2.5, 4.)),
                   mat4_rotation(rng.uniform(0, 6.28), *rng.normal(size=3)))
        P = mat4_perspective(50, size[0] / float(size[1]), .1, 10)
        iso = name == "d32" and i % 3 == 0
        dt = VOLUMES[name][1]
        top = 1. if dt == np.float32 else 3000. if dt == np.uint16 else 200.
        out.append(Request(name, M, P, size=size, method="iso_surface" if iso else "max_project",
                           min_val=0., max_val=top if iso else 0., gamma=1.,
                           alpha_pow=.5 if i % 5 == 0 else 0., max_steps=96))
    return out


def check_same(res, ref, method):
    np.testing.assert_array_equal(res.output, ref.output)
    np.testing.assert_array_equal(res.alpha, ref.alpha)
    if method == "iso_surface":
        np.testing.assert_array_equal(res.depth, ref.depth)
        np.testing.assert_array_equal(res.normals, ref.normals)
        np.testing.assert_array_equal(res.occlusion, ref.occlusion)


@pytest.mark.parametrize("budget_kib", [None, 600])
def test_same_alone_and_in_the_mix(torch, budget_kib):
    # 600 KiB holds at most two of the four padded volumes: requests force evictions
    server = RenderServer(0, hbm_budget=None if budget_kib is None else budget_kib << 10,
                          max_batch=8, workers=2)
    try:
        for k, (name, (shape, dt)) in enumerate(sorted(VOLUMES.items())):
            server.register(name, make_volume(shape, dt, k))
        reqs = requests(40, 7)
        futs = [server.submit(r) for r in reqs]
        got = [f.result(timeout=120) for f in futs]
        assert server.resident_bytes <= server.hbm_budget
        batched = sum(1 for g in got if g.batch_size > 1)
        if budget_kib is None:
            assert batched > 0  # requests were really batched
        for r, g in zip(reqs, got):
            check_same(g, server.render_alone(r), r.method)
            assert g.latency_s >= g.queued_s >= 0
        assert sum((g.output > 0).sum() for g in got) > 2000
    finally:
        server.close()


def test_server_errors(torch):
    server = RenderServer(0, hbm_budget=1 << 20, workers=1)
    try:
        with pytest.raises(NotImplementedError):
            server.register("x", np.zeros((4, 4, 4), np.int32))
        with pytest.raises(KeyError):
            server.submit(Request("missing", np.eye(4), np.eye(4)))
        server.register("big", np.zeros((64, 64, 256), np.uint16))  # 2 MiB > 1 MiB budget
        fut = server.submit(Request("big", np.eye(4), np.eye(4), size=(16, 16)))
        with pytest.raises(MemoryError):
            fut.result(timeout=60)
    finally:
        server.close()
