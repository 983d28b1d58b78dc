"""Render server: many resident volumes, a stream of render requests, continuous batching.

The reference renders one frame at a time from the GUI thread, and its only residency
logic is the DataModel prefetch cache of timepoints on the host
(spimagine/models/data_model.py:600-757).  Here volumes live in HBM slots of one
libspimrender context:

- admission: a volume is uploaded into a slot on first use if it fits the HBM budget;
  otherwise least-recently-used volumes that no queued or running request needs are
  released first; a request whose volume cannot be admitted waits in the queue;
- continuous batching: each worker thread (own HIP stream) repeatedly takes the oldest
  waiting request plus every other waiting request with the same batch key (voxel type,
  frame size, method and all parameters except the pose) up to ``max_batch``, and renders
  max projections of padded-layout volumes in one ``spr_render_batch`` launch
  (iso-surface requests and volumes too large for the padded layout are rendered one by
  one with ``spr_render_slot``).  Requests that arrive while batches run are picked up by
  the next free worker, so work from several batches overlaps on the device;
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
        self.padded = None


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
    def _try_admit(self, vol):
        """Make `vol` resident if the budget allows (evicting idle volumes). Caller holds cv."""
        if vol.slot is not None:
            return True
        need = device_bytes(vol.data.shape, vol.data.dtype)
        if need > self.hbm_budget:
            raise MemoryError("volume %r larger than the HBM budget" % vol.name)
        idle = sorted((v for v in self.volumes.values() if v.slot is not None and v.users == 0
                       and v is not vol), key=lambda v: v.last_use)
        while self.resident_bytes + need > self.hbm_budget and idle:
            self._evict(idle.pop(0))
        if self.resident_bytes + need > self.hbm_budget:
            return False
        with self.table_lock:
            vol.slot = self.ctx.volume_add(vol.data)
            info = self.ctx.volume_info(vol.slot)
        vol.bytes = info["bytes"]
        # the padded layout alone (batchable); above 4 GiB a slot is linear + a padded copy
        vol.padded = vol.bytes == padded_bytes(vol.data.shape, vol.data.dtype)
        self.resident_bytes += vol.bytes
        return True

    def _evict(self, vol):
        with self.table_lock:
            self.ctx.volume_release(vol.slot)
        self.resident_bytes -= vol.bytes
        vol.slot = None
        vol.bytes = 0

    def _admit_blocking(self, vol):
        with self.cv:
            while not self._try_admit(vol):
                self.cv.wait(.01)

    # ------------------------------------------------------------------ scheduling
    def _key(self, req, vol):
        s = req.settings
        batchable = req.method == "max_project" and vol.padded
        return (batchable, vol.data.dtype.str, req.size, req.method, s["min_val"], s["max_val"],
                s["gamma"], s["alpha_pow"], s["max_steps"], s["interpolation"], s["box"])

    def _take_batch(self):
        """Oldest admissible request + same-key followers. Caller holds cv."""
        for (req, fut) in list(self.waiting):
            vol = self.volumes[req.volume]
            try:
                if not self._try_admit(vol):
                    continue
            except Exception as e:  # can never be admitted (or the upload failed)
                self.waiting.remove((req, fut))
                vol.users -= 1
                fut.set_exception(e)
                continue
            key = self._key(req, vol)
            batch = [(req, fut)]
            if key[0]:
                for (r2, f2) in list(self.waiting):
                    if len(batch) >= self.max_batch:
                        break
                    if r2 is req:
                        continue
                    v2 = self.volumes[r2.volume]
                    if v2.slot is None and not self._try_admit(v2):
                        continue
                    if self._key(r2, v2) == key:
                        batch.append((r2, f2))
            for item in batch:
                self.waiting.remove(item)
            now = time.perf_counter()
            for r, _ in batch:
                self.volumes[r.volume].last_use = now
            return batch
        return None

    def _worker(self):
        torch = self.torch
        stream = torch.cuda.Stream()
        while True:
            with self.cv:
                batch = None
                while not self.stop:
                    batch = self._take_batch()
                    if batch:
                        break
                    self.cv.wait(.005)
                if self.stop and not batch:
                    return
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
            return Result(out.cpu().numpy(), alpha.cpu().numpy(),
                          **{k: v.cpu().numpy() for k, v in extra.items()})

    def _run(self, batch, stream):
        torch = self.torch
        t0 = time.perf_counter()
        req0 = batch[0][0]
        vol0 = self.volumes[req0.volume]
        if len(batch) == 1 or not self._key(req0, vol0)[0]:
            results = [self._render_single(r, self.volumes[r.volume], stream) for r, _ in batch]
        else:
            w, h = req0.size
            with torch.cuda.stream(stream):
                planes = torch.empty((len(batch), 2, h, w), device="cuda")
                jobs = []
                for k, (r, _) in enumerate(batch):
                    p = self._params(r, self.volumes[r.volume])
                    jobs.append((self.volumes[r.volume].slot, list(p.inv_p), list(p.inv_m),
                                 planes[k, 0].data_ptr(), planes[k, 1].data_ptr()))
                self.ctx.render_batch(jobs, w, h, stream=stream.cuda_stream,
                                      params=self._params(req0, vol0))
                host = planes.cpu().numpy()
            results = [Result(host[k, 0], host[k, 1]) for k in range(len(batch))]
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
