"""Render server: many resident volumes, a stream of render requests, continuous batching.

The reference renders one frame at a time from the GUI thread, and its only residency
logic is the DataModel prefetch cache of timepoints on the host
(spimagine/models/data_model.py:600-757).  Here volumes live in HBM slots of one
libspimrender context:

- admission: a volume is uploaded into a slot on first use if it fits the HBM budget;
  otherwise least-recently-used volumes that no queued or running request needs are
  released first; a request whose volume cannot be admitted waits in the queue.  The
  scheduler reserves the budget under its lock and the worker uploads (and releases the
  evicted volumes) outside it, so other workers keep taking batches meanwhile;
- continuous batching: each worker thread (own HIP stream) repeatedly takes the oldest
  request whose volume is resident plus every other waiting request with the same batch
  key -- voxel type, method, alpha_pow and filter; frame size, pose, box, transfer range,
  gamma and step count may all differ -- whose volume is resident, up to ``max_batch``,
  and renders max projections of padded-layout volumes in one ``spr_render_frames``
  launch (iso-surface requests and volumes above 4 GiB padded are rendered one by one with
  ``spr_render_slot``).  Requests that arrive while batches run are picked up by the next
  free worker, so work from several batches overlaps on the device;
- every request records its queueing and end-to-end latency.

Every result is bit-identical to rendering the request alone (tested): batching changes
which frames share a launch, never a frame's samples.
"""
import collections
import threading
import time
from concurrent.futures import Future

import numpy as np

from . import _lib
from .volumerender.volumerender import render_matrices


class Request(object):
    """One frame: a registered volume, a pose and the render settings."""

    def __init__(self, volume, model_view, projection, size=(512, 512), method="max_project",
                 min_val=0., max_val=0., gamma=1., alpha_pow=0., max_steps=200,
                 interpolation="linear", units=(1., 1., 1.), box=(-1, 1, -1, 1, -1, 1),
                 occ_strength=.1, occ_radius=21, occ_n_points=30):
        self.volume = volume
        self.model_view = np.asarray(model_view)
        self.projection = np.asarray(projection)
        self.size = (int(size[0]), int(size[1]))
        self.method = method
        self.settings = dict(min_val=float(min_val), max_val=float(max_val), gamma=float(gamma),
                             alpha_pow=float(alpha_pow), max_steps=int(max_steps),
                             interpolation=interpolation, units=tuple(float(u) for u in units),
                             box=tuple(float(b) for b in box), occ_strength=float(occ_strength),
                             occ_radius=int(occ_radius), occ_n_points=int(occ_n_points))
        self.t_submit = None


class Result(object):
    def __init__(self, output, alpha, depth=None, normals=None, occlusion=None):
        self.output = output
        self.alpha = alpha
        self.depth = depth
        self.normals = normals
        self.occlusion = occlusion
        self.queued_s = 0.
        self.latency_s = 0.
        self.batch_size = 1


class _Volume(object):
    def __init__(self, name, data):
        self.name = name
        self.data = data
        self.slot = None
        self.bytes = 0
        self.users = 0          # queued + running requests
        self.last_use = 0.
        self.loading = False    # an upload (admitted, budget reserved) is under way
        # the padded layout alone (batchable); above 4 GiB a slot is linear + a padded copy
        self.padded = padded_bytes(data.shape, data.dtype) <= 0xFFFFFFFF


class RenderServer(object):
    METHODS = {"max_project": _lib.METHOD_MAX_PROJECT, "iso_surface": _lib.METHOD_ISO_SURFACE}

    def __init__(self, device=0, hbm_budget=None, max_batch=16, workers=2, ctx=None):
        """ctx: a context to serve from (default: a new one on `device`); workers = 0 starts
        no worker threads (the scheduler is then driven by the caller)."""
        if workers:
            import torch
            self.torch = torch
        self.ctx = ctx if ctx is not None else _lib.Context(device)
        if hbm_budget is None:
            free, _ = _lib.device_memory(device)
            hbm_budget = int(.6 * free)
        self.hbm_budget = int(hbm_budget)
        self.max_batch = int(max_batch)
        self.volumes = {}
        self.resident_bytes = 0
        self.waiting = collections.deque()
        self.cv = threading.Condition()
        self.table_lock = threading.Lock()  # slot table changes (add / release)
        self.stop = False
        self.batches = 0
        self._alone_stream = None
        self.threads = [threading.Thread(target=self._worker, daemon=True) for _ in range(workers)]
        for t in self.threads:
            t.start()

    # ------------------------------------------------------------------ client API
    def register(self, name, data):
        """A volume (numpy (nz, ny, nx) float32/uint16/uint8) requests may name."""
        data = np.ascontiguousarray(data)
        if data.dtype not in _lib.DTYPE_CODES:
            raise NotImplementedError("unsupported voxel type %s" % data.dtype)
        with self.cv:
            self.volumes[name] = _Volume(name, data)

    def submit(self, req):
        if req.volume not in self.volumes:
            raise KeyError("unknown volume %r" % req.volume)
        if req.method not in self.METHODS:
            raise ValueError("unknown method %r" % req.method)
        fut = Future()
        req.t_submit = time.perf_counter()
        with self.cv:
            self.volumes[req.volume].users += 1
            self.waiting.append((req, fut))
            self.cv.notify()
        return fut

    def render_alone(self, req):
        """The request rendered by itself, synchronously (reference for the batched result)."""
        vol = self.volumes[req.volume]
        with self.cv:
            vol.users += 1
            if self._alone_stream is None:
                self._alone_stream = self.torch.cuda.Stream()
        try:
            self._admit_blocking(vol)
            # a stream of its own: torch's default stream has handle 0, which the C ABI
            # reads as the context's stream, unordered with torch's copies
            res = self._render_single(req, vol, self._alone_stream)
        finally:
            with self.cv:
                vol.users -= 1
        return res

    def close(self):
        with self.cv:
            self.stop = True
            self.cv.notify_all()
        for t in self.threads:
            t.join()
        self.ctx.close()

    # ------------------------------------------------------------------ residency
    def _reserve(self, vol):
        """Admission of `vol` (caller holds cv): "ready" when resident, "wait" when it is
        being uploaded or cannot be admitted now, else the list of idle volumes to evict
        for it -- its budget is then reserved and the caller must run _load(vol, victims)
        outside the lock.  MemoryError when it can never fit."""
        if vol.slot is not None:
            return "ready"
        if vol.loading:
            return "wait"
        need = device_bytes(vol.data.shape, vol.data.dtype)
        if need > self.hbm_budget:
            raise MemoryError("volume %r larger than the HBM budget" % vol.name)
        idle = sorted((v for v in self.volumes.values() if v.slot is not None and v.users == 0
                       and not v.loading and v is not vol), key=lambda v: v.last_use)
        victims = []
        free = self.hbm_budget - self.resident_bytes
        while free < need and idle:
            v = idle.pop(0)
            victims.append((v.slot, v.bytes))
            free += v.bytes
        if free < need:
            return "wait"
        for v in self.volumes.values():
            if any(v.slot == slot for slot, _ in victims):
                self.resident_bytes -= v.bytes
                v.slot = None
                v.bytes = 0
        vol.loading = True
        self.resident_bytes += need
        return victims

    def _load(self, vol, victims):
        """Releases the evicted slots and uploads `vol` (no scheduler lock held)."""
        need = device_bytes(vol.data.shape, vol.data.dtype)
        try:
            with self.table_lock:
                for slot, _ in victims:
                    self.ctx.volume_release(slot)
                slot = self.ctx.volume_add(vol.data)
                info = self.ctx.volume_info(slot)
        except Exception as e:
            with self.cv:
                self.resident_bytes -= need
                vol.loading = False
                self._fail_volume(vol, e)
                self.cv.notify_all()
            return False
        with self.cv:
            vol.slot = slot
            vol.bytes = info["bytes"]
            self.resident_bytes += vol.bytes - need
            vol.loading = False
            self.cv.notify_all()
        return True

    def _fail_volume(self, vol, err):
        """Every waiting request of `vol` fails with err (caller holds cv)."""
        for item in [it for it in self.waiting if self.volumes[it[0].volume] is vol]:
            self.waiting.remove(item)
            vol.users -= 1
            item[1].set_exception(err)

    def _admit_blocking(self, vol):
        while True:
            with self.cv:
                state = self._reserve(vol)
                if state == "ready":
                    return
                if state == "wait":
                    self.cv.wait(.01)
                    continue
            if not self._load(vol, state):
                raise RuntimeError("upload of volume %r failed" % vol.name)

    # ------------------------------------------------------------------ scheduling
    def _key(self, req, vol):
        """Requests with equal keys share a launch: voxel type, method, alpha_pow and filter
        (frame size, pose, box, transfer range, gamma and steps are per frame)."""
        s = req.settings
        batchable = req.method == "max_project" and vol.padded
        return (batchable, vol.data.dtype.str, req.method, s["alpha_pow"], s["interpolation"])

    def _next_task(self):
        """The next piece of work (caller holds cv): ("batch", [(req, future), ...]) -- the
        oldest request whose volume is resident plus same-key followers whose volumes are
        resident -- or ("load", vol, victims) for the oldest request whose volume can be
        admitted, or None."""
        for (req, fut) in list(self.waiting):
            vol = self.volumes[req.volume]
            try:
                state = self._reserve(vol)
            except Exception as e:  # can never be admitted
                self.waiting.remove((req, fut))
                vol.users -= 1
                fut.set_exception(e)
                continue
            if state == "wait":
                continue
            if state != "ready":
                return ("load", vol, state)
            key = self._key(req, vol)
            batch = [(req, fut)]
            if key[0]:
                for (r2, f2) in list(self.waiting):
                    if len(batch) >= self.max_batch:
                        break
                    if r2 is req:
                        continue
                    v2 = self.volumes[r2.volume]
                    if v2.slot is not None and self._key(r2, v2) == key:
                        batch.append((r2, f2))
            for item in batch:
                self.waiting.remove(item)
            now = time.perf_counter()
            for r, _ in batch:
                self.volumes[r.volume].last_use = now
            return ("batch", batch)
        return None

    def _take_batch(self):
        """The next batch, uploading volumes on the way (drives the scheduler without worker
        threads); None when nothing can run now."""
        while True:
            with self.cv:
                task = self._next_task()
            if task is None:
                return None
            if task[0] == "batch":
                return task[1]
            self._load(task[1], task[2])

    def _worker(self):
        torch = self.torch
        stream = torch.cuda.Stream()
        while True:
            with self.cv:
                task = None
                while not self.stop:
                    task = self._next_task()
                    if task:
                        break
                    self.cv.wait(.005)
                if self.stop and not task:
                    return
            if task[0] == "load":
                self._load(task[1], task[2])
                continue
            batch = task[1]
            try:
                self._run(batch, stream)
            except Exception as e:  # deliver the error to every request of the batch
                for _, fut in batch:
                    if not fut.done():
                        fut.set_exception(e)
            finally:
                with self.cv:
                    for r, _ in batch:
                        self.volumes[r.volume].users -= 1
                    self.batches += 1
                    self.cv.notify_all()

    # ------------------------------------------------------------------ rendering
    def _params(self, req, vol):
        s = req.settings
        nz, ny, nx = vol.data.shape
        iP, iM = render_matrices(req.model_view, req.projection, (nx, ny, nz), s["units"])
        filt = _lib.FILTER_NEAREST if s["interpolation"] == "nearest" else _lib.FILTER_LINEAR
        return _lib.make_params(iP, iM, s["box"], s["min_val"], s["max_val"], s["gamma"],
                                s["alpha_pow"], s["occ_strength"], s["occ_radius"],
                                s["occ_n_points"], s["max_steps"], filt)

    def _render_single(self, req, vol, stream):
        torch = self.torch
        assert stream.cuda_stream, "render on a non-default stream"
        w, h = req.size
        p = self._params(req, vol)
        extra = {}
        with torch.cuda.stream(stream):  # buffers allocated on the stream that writes them
            out = torch.empty((h, w), device="cuda")
            alpha = torch.empty_like(out)
            if req.method == "iso_surface":
                depth = torch.empty_like(out)
                occ = torch.empty_like(out)
                normals = torch.empty((h, w, 3), device="cuda")
                self.ctx.render_slot(vol.slot, p, _lib.METHOD_ISO_SURFACE, w, h, out.data_ptr(),
                                     alpha.data_ptr(), depth.data_ptr(), normals.data_ptr(),
                                     occ.data_ptr(), stream=stream.cuda_stream)
                extra = dict(depth=depth, normals=normals, occlusion=occ)
            else:
                self.ctx.render_slot(vol.slot, p, _lib.METHOD_MAX_PROJECT, w, h, out.data_ptr(),
                                     alpha.data_ptr(), stream=stream.cuda_stream)
            host = self._download([out, alpha] + list(extra.values()), stream)
        return Result(host[0], host[1], **dict(zip(extra, host[2:])))

    def _download(self, tensors, stream):
        """Device frames -> host arrays: pinned host tensors (torch's caching host
        allocator), copied asynchronously on the render stream, one wait for all of them.
        Pinned memory is what lets the copy run at DMA rate: a copy into pageable memory
        is staged through a bounce buffer and, for a 2048^2 iso-surface result (112 MB),
        took longer than the render."""
        torch = self.torch
        with torch.cuda.stream(stream):
            host = [torch.empty(t.shape, dtype=t.dtype, pin_memory=True) for t in tensors]
            for h, t in zip(host, tensors):
                h.copy_(t, non_blocking=True)
        stream.synchronize()
        return [h.numpy() for h in host]

    def _run(self, batch, stream):
        torch = self.torch
        t0 = time.perf_counter()
        req0 = batch[0][0]
        vol0 = self.volumes[req0.volume]
        if len(batch) == 1 or not self._key(req0, vol0)[0]:
            results = [self._render_single(r, self.volumes[r.volume], stream) for r, _ in batch]
        else:
            with torch.cuda.stream(stream):
                planes, frames = [], []
                for r, _ in batch:
                    w, h = r.size
                    pl = torch.empty((2, h, w), device="cuda")
                    planes.append(pl)
                    vol = self.volumes[r.volume]
                    frames.append(_lib.make_frame(vol.slot, self._params(r, vol), w, h,
                                                  pl[0].data_ptr(), pl[1].data_ptr()))
                filt = (_lib.FILTER_NEAREST if req0.settings["interpolation"] == "nearest"
                        else _lib.FILTER_LINEAR)
                self.ctx.render_frames(frames, req0.settings["alpha_pow"], filt,
                                       stream=stream.cuda_stream)
            host = self._download(planes, stream)
            results = [Result(hp[0], hp[1]) for hp in host]
        t1 = time.perf_counter()
        for (r, fut), res in zip(batch, results):
            res.queued_s = t0 - r.t_submit
            res.latency_s = t1 - r.t_submit
            res.batch_size = len(batch)
            fut.set_result(res)


def padded_bytes(shape, dtype):
    """Device bytes of a (nz, ny, nx) volume in the padded layout: one voxel of apron on
    every side, rows padded to 16 bytes (the layout spr_volume_add picks when it fits)."""
    nz, ny, nx = shape
    es = np.dtype(dtype).itemsize
    per16 = 16 // es
    return (nz + 2) * (ny + 2) * ((nx + 2 + per16 - 1) // per16) * 16


def device_bytes(shape, dtype):
    """Device bytes spr_volume_add takes for a volume: the padded copy, plus the linear
    volume when the padded copy exceeds 4 GiB."""
    pb = padded_bytes(shape, dtype)
    if pb <= 0xFFFFFFFF:
        return pb
    return pb + int(np.prod(shape)) * np.dtype(dtype).itemsize


def poisson_load(server, requests, qps, seed=0):
    """Submit `requests` with exponential inter-arrival times at `qps`; returns the futures."""
    rng = np.random.default_rng(seed)
    futs = []
    t_next = time.perf_counter()
    for req in requests:
        t_next += rng.exponential(1. / qps)
        delay = t_next - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        futs.append(server.submit(req))
    return futs
