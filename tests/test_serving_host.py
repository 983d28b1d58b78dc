"""Render-server scheduling on the host (no GPU): HBM-budget admission, LRU eviction of idle
volumes only, batch keys, and the device-size estimate, with a context stand-in that
records slot traffic."""
import numpy as np
import pytest

from spimagine_amd.serving import RenderServer, Request, device_bytes, padded_bytes


class FakeContext(object):
    def __init__(self):
        self.slots = {}
        self.next = 1
        self.log = []

    def volume_add(self, data):
        s = self.next
        self.next += 1
        self.slots[s] = data
        self.log.append(("add", s))
        return s

    def volume_info(self, slot):
        d = self.slots[slot]
        return {"dtype": 1, "shape": d.shape, "bytes": device_bytes(d.shape, d.dtype)}

    def volume_release(self, slot):
        del self.slots[slot]
        self.log.append(("release", slot))

    def close(self):
        pass


def server(budget):
    return RenderServer(hbm_budget=budget, workers=0, ctx=FakeContext(), max_batch=4)


def req(name, size=(64, 64), **kw):
    return Request(name, np.eye(4), np.eye(4), size=size, **kw)


def test_device_size_estimates():
    assert padded_bytes((10, 20, 30), np.uint16) == 12 * 22 * 32 * 2   # rows padded to 16 B
    assert padded_bytes((10, 20, 30), np.uint8) == 12 * 22 * 32
    assert padded_bytes((10, 20, 30), np.float32) == 12 * 22 * 32 * 4
    big = (512, 2048, 2048)
    assert device_bytes(big, np.uint16) == padded_bytes(big, np.uint16) + 2 * 512 * 2048 * 2048


def test_admission_evicts_least_recently_used_idle_volumes():
    vol = np.zeros((30, 30, 30), np.uint16)
    one = padded_bytes(vol.shape, vol.dtype)
    srv = server(2 * one + 10)          # room for two volumes
    for n in "abc":
        srv.register(n, vol)
    srv.submit(req("a", size=(64, 64)))  # different frame sizes: separate batches
    srv.submit(req("b", size=(32, 32)))
    batch = srv._take_batch()           # admits a
    assert [r.volume for r, _ in batch] == ["a"]
    for r, _ in batch:
        srv.volumes[r.volume].users -= 1
    batch = srv._take_batch()           # admits b
    assert [r.volume for r, _ in batch] == ["b"]
    assert srv.resident_bytes == 2 * one
    # c needs room: a (idle, used first) is evicted, b has a running request and stays
    srv.submit(req("c", size=(48, 48)))
    batch = srv._take_batch()
    assert [r.volume for r, _ in batch] == ["c"]
    assert srv.volumes["a"].slot is None and srv.volumes["b"].slot is not None
    assert srv.resident_bytes <= srv.hbm_budget
    assert ("release", 1) in srv.ctx.log


def test_busy_volumes_are_never_evicted_and_requests_wait():
    vol = np.zeros((30, 30, 30), np.uint16)
    one = padded_bytes(vol.shape, vol.dtype)
    srv = server(one + 10)              # room for one volume
    srv.register("a", vol)
    srv.register("b", vol)
    srv.submit(req("a"))
    srv.submit(req("b", size=(32, 32)))
    first = srv._take_batch()
    assert [r.volume for r, _ in first] == ["a"]
    assert srv._take_batch() is None    # b cannot be admitted while a's request runs
    srv.volumes["a"].users -= 1         # a's request finishes
    second = srv._take_batch()
    assert [r.volume for r, _ in second] == ["b"] and srv.volumes["a"].slot is None


def test_batches_group_same_key_requests_only():
    """A batch mixes volumes, frame sizes and transfer settings (spr_render_frames takes
    them per frame) but not alpha_pow, method or filter."""
    vol = np.zeros((30, 30, 30), np.uint16)
    srv = server(1 << 30)
    srv.register("a", vol)
    srv.register("b", vol)
    for n, size, ap, mx in [("a", (64, 64), 0., 1.), ("b", (64, 64), 0., 2.), ("a", (32, 32), 0., 1.),
                            ("b", (64, 64), .5, 1.), ("a", (64, 48), 0., 3.), ("b", (64, 64), 0., 1.),
                            ("a", (64, 64), 0., 1.)]:
        srv.submit(req(n, size=size, alpha_pow=ap, max_val=mx))
    srv.submit(req("a", method="iso_surface"))
    # only volumes already resident join a batch (b is admitted when its own request is
    # the oldest): a's four plain projections, of three sizes and two transfer ranges
    batch = srv._take_batch()
    assert len(batch) == 4  # max_batch
    assert [(r.volume, r.size) for r, _ in batch] == [("a", (64, 64)), ("a", (32, 32)),
                                                      ("a", (64, 48)), ("a", (64, 64))]
    assert all(r.settings["alpha_pow"] == 0. for r, _ in batch)
    rest = [(r.volume, r.settings["alpha_pow"], r.method) for r, _ in srv.waiting]
    assert rest == [("b", 0., "max_project"), ("b", .5, "max_project"), ("b", 0., "max_project"),
                    ("a", 0., "iso_surface")]
    assert [len(srv._take_batch()) for _ in range(3)] == [2, 1, 1]
    assert srv._take_batch() is None


def test_oversized_request_behind_an_admissible_one():
    """An oversized request queued behind an admissible one fails on its own; the other
    is served and the scheduler keeps going."""
    srv = server(1 << 20)
    srv.register("ok", np.zeros((16, 16, 16), np.uint16))
    srv.register("big", np.zeros((64, 64, 256), np.uint16))  # 2 MiB > 1 MiB budget
    f_ok = srv.submit(req("ok"))
    f_big = srv.submit(req("big"))
    f_ok2 = srv.submit(req("ok", size=(32, 32)))
    batch = srv._take_batch()
    assert [r.volume for r, _ in batch] == ["ok", "ok"]
    assert srv._take_batch() is None
    with pytest.raises(MemoryError):
        f_big.result(timeout=1)
    assert not f_ok.done() and not f_ok2.done()  # the batch is the caller's to render
    assert srv.volumes["big"].users == 0


def test_oversized_and_unknown_requests_fail():
    srv = server(1000)
    srv.register("big", np.zeros((64, 64, 64), np.uint16))
    fut = srv.submit(req("big"))
    assert srv._take_batch() is None
    with pytest.raises(MemoryError):
        fut.result(timeout=1)
    with pytest.raises(KeyError):
        srv.submit(req("nope"))
    with pytest.raises(ValueError):
        srv.submit(req("big", method="volume_render"))
    with pytest.raises(NotImplementedError):
        srv.register("x", np.zeros((2, 2, 2), np.int64))
